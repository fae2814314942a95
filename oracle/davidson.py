"""Davidson eigensolver (oracle restatement of source/tools/pdvdson.py:22-390).

Single process, real arithmetic, no excited-state projection, no start noise
(the reference enables its unseeded start noise only when noise > 1e-10,
mpo_dmrg_dot.py:78).  Test infrastructure only.
"""
import numpy as np
import scipy.linalg


class DavidsonError(RuntimeError):
    pass


def dvdson_ortho(vbas, rbas, crit_indp):
    """Orthonormalise rbas against vbas, 5 projection passes (pdvdson.py:350-390)."""
    rbas = rbas.copy()
    nres = rbas.shape[0]
    vbas2 = np.zeros_like(rbas)
    nindp = 0
    for _ in range(5):
        rbas = rbas - (rbas @ vbas.T) @ vbas
    for i in range(nres):
        rvec = rbas[i].copy()
        rii = np.linalg.norm(rvec)
        if rii <= crit_indp:
            continue
        rvec = rvec / rii
        rvec = rvec / np.linalg.norm(rvec)
        vbas2[nindp] = rvec
        nindp += 1
        for _ in range(5):
            rbas[i:] -= (rbas[i:] @ vbas.T) @ vbas
            rbas[i:] -= (rbas[i:] @ vbas2[:nindp].T) @ vbas2[:nindp]
    return nindp, vbas2[:nindp].copy()


def gen_ortho_bas(vbas, crit_indp):
    """pdvdson.py:335-345."""
    vbas[0] = vbas[0] / np.linalg.norm(vbas[0])
    if vbas.shape[0] != 1:
        nindp, v2 = dvdson_ortho(vbas[0:1], vbas[1:], crit_indp)
        if nindp + 1 != vbas.shape[0]:
            raise DavidsonError('insufficient orthonormal basis')
        vbas[1:] = v2
    return vbas


def gen_v0(diag, neig):
    """Unit vectors at the lowest diagonal entries with a 1e-12 ramp tie-break (pdvdson.py:81-87)."""
    n = len(diag)
    d = diag + 1e-12 * np.arange(1, 1 + n) / float(n)
    index = np.argsort(d, kind='stable')[:neig]
    v0 = np.zeros((neig, n))
    v0[np.arange(neig), index] = 1.0
    return v0


def solve_iter(hvec, diag, neig=1, v0=None, crit_e=1e-8, crit_vec=1e-8,
               crit_indp=1e-12, maxcycle=20, lshift=1e-4, nfac=10, iop=4):
    """LOBPCG-like Davidson with |diag-e|+lshift preconditioner (pdvdson.py:90-330).

    ``hvec(v)`` maps one vector.  Returns (eigs, vecs, nmvp, niter).
    """
    ndim = len(diag)
    if neig > ndim:
        raise DavidsonError('neig > ndim')
    nmvp = 0

    def hvecs(vb):
        nonlocal nmvp
        nmvp += vb.shape[0]
        return np.array([hvec(v) for v in vb])

    vbas = gen_v0(diag, neig) if v0 is None else np.array(v0, dtype=np.float64).copy()
    vbas = gen_ortho_bas(vbas, crit_indp)
    wbas = hvecs(vbas)
    eigs = np.zeros(neig) + 1e3
    nbas = neig
    ifconv = False
    niter = 0
    vr = None
    for niter in range(1, maxcycle):
        iden = vbas @ vbas.T
        if np.linalg.norm(iden - np.identity(nbas)) > 1e-10:
            raise DavidsonError('basis not orthonormal')
        tmpH = vbas @ wbas.T
        if np.linalg.norm(tmpH - tmpH.T) > 1e-10:
            raise DavidsonError('projected H not Hermitian')
        tmpH = 0.5 * (tmpH + tmpH.T)
        eig, vr = scipy.linalg.eigh(tmpH)
        vr = vr[:, :neig].T.copy()
        teig = eig[:neig]
        econv = [abs(teig[i] - eigs[i]) <= crit_e for i in range(neig)]
        eigs = teig.copy()
        rbas = vr @ wbas - np.diag(eigs) @ (vr @ vbas)
        rindx = []
        rconv = []
        for i in range(neig):
            rn = np.linalg.norm(rbas[i])
            rconv.append(rn <= crit_vec)
            if not rconv[i] and not econv[i]:
                rindx.append(i)
        ifconv = len(rindx) == 0
        if ifconv or niter == maxcycle - 1:
            break
        if nbas > nfac * neig:
            pr = np.identity(nbas)[:neig, :] - vr
            nindp, vr2 = dvdson_ortho(vr, pr[rindx, :], crit_indp)
            if nindp != 0:
                vr = np.vstack((vr, vr2))
            vbas = vr @ vbas
            wbas = vr @ wbas
        for i in range(neig):
            if rconv[i]:
                continue
            rbas[i] = rbas[i] / (np.abs(diag - eigs[i]) + lshift)
        nindp, vbas2 = dvdson_ortho(vbas, rbas[rindx, :], crit_indp)
        if nindp == 0:
            raise DavidsonError('unable to generate new direction')
        wbas2 = hvecs(vbas2)
        vbas = np.vstack((vbas, vbas2))
        wbas = np.vstack((wbas, wbas2))
        nbas = vbas.shape[0]
    return eigs, vr @ vbas, nmvp, niter
