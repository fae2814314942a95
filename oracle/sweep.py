"""Two-site sweep driver, local solver and decimation (oracle).

Restates source/mpo_dmrg_class.py (build :279-297, default :329-434),
source/mpo_dmrg_opt.py (sweep :31-95, checkOpers :153-193, updateSite
:313-344, updateBoundary :347-377), source/mpo_dmrg_dot.py (ci_solver :24-112),
source/mpo_dmrg_dotutil.py (setupQlst :81-95, symmetrySpaceInfo :148-189,
setupQtmpCI :192-203, decimation :288-495, genGuess :498-559),
source/mpo_dmrg_qparser.py (rdm_blkdiag :101-180) and source/mpo_dmrg_util.py
(mps0occun :118-155, initOps :193-214), for ncsite=2, maxslc=1, one process,
noise=0 and the 'Fixed2' schedule (mpo_dmrg_schedule.py:100-122).
Test infrastructure only.

Deterministic truncation rule (shared with the device path; the reference's
rule depends on Python-2 dict order and an unstable argsort and is not
reproducible, mpo_dmrg_qparser.py:220-233,259): weights sorted descending;
weights <= 1e-12 count as zero; runs of weights equal to within 1e-10
relative form a tie cluster, ordered inside by (N, Sz) ascending then by position; keep the
first min(Dcut, n) states; re-sort the kept states by (N, Sz), stable.  The
reference keeps zero-weight states (thresh = -1e-14), whose eigenvectors are
an arbitrary basis of the RDM null space; here that basis is fixed
(canonical_null) so that two implementations agree beyond rounding.
"""
import numpy as np

from . import davidson
from .kernels import HDiagQt, HVecQt, boundary_env, genHops, renorm_site
from .mpo import build_mpo, qphys_spatial
from .qtensor import QTensor, fromQnums


# RDM weights (trace-normalised) at or below this are numerically zero: they
# are ordered by QN only, and their eigenvectors get the canonical basis.
WZERO = 1e-12


def dpt(q1, q2):
    """Direct-product QNs, second index fastest (mpo_dmrg_qphys.py:63-68)."""
    return np.array([np.asarray(a) + np.asarray(b) for a in q1 for b in q2])


def truncation_order(sigs, qsyms):
    """Descending weights with the deterministic tie-cluster rule."""
    n = len(sigs)
    w = np.where(sigs > WZERO, sigs, 0.0)
    order = list(np.argsort(-w, kind='stable'))
    out = []
    i = 0
    while i < n:
        j = i + 1
        w0 = w[order[i]]
        while j < n and abs(w[order[j]] - w0) <= 1e-10 * abs(w0):
            j += 1
        cl = order[i:j]
        cl.sort(key=lambda k: (qsyms[k][0], qsyms[k][1], k))
        out += cl
        i = j
    return np.array(out, dtype=np.int64)


def canonical_null(u, s2, tol):
    """Replace eigenvectors with weight <= tol by a fixed basis of their span:
    unit vectors e_0, e_1, ... projected out of the kept vectors and of each
    other, accepted when the projected norm exceeds 0.5 (then 1e-6)."""
    keep = s2 > tol
    nnull = int(np.sum(~keep))
    if nnull == 0:
        return u
    basis = [u[:, keep]]
    out = []
    n = u.shape[0]
    for cut in (0.5, 1e-6):
        for j in range(n):
            if len(out) == nnull:
                break
            v = np.zeros(n)
            v[j] = 1.0
            for _ in range(2):
                for b in basis + ([np.array(out).T] if out else []):
                    v = v - b @ (b.T @ v)
            nv = np.linalg.norm(v)
            if nv > cut:
                out.append(v / nv)
    u2 = u.copy()
    u2[:, ~keep] = np.array(out).T
    return u2


def rdm_blkdiag(rho, classes, thresh, Dcut):
    """Per-sector eigh of the RDM and truncation (mpo_dmrg_qparser.py:101-180)."""
    groups = {}
    keys = []
    for idx, val in enumerate(classes):
        k = (float(val[0]), float(val[1]))
        if k not in groups:
            groups[k] = []
            keys.append(k)
        groups[k].append(idx)
    nrows = rho.shape[0]
    rot, sigs, qsym = [], [], []
    for k in keys:
        rows = groups[k]
        s2, u = np.linalg.eigh(rho[np.ix_(rows, rows)])
        s2 = np.abs(s2)
        u = canonical_null(u, s2, WZERO)
        uc = np.zeros((nrows, len(rows)))
        uc[rows, :] = u
        rot.append(uc)
        sigs.append(s2)
        qsym += [k] * len(rows)
    rot = np.hstack(rot)
    sigs = np.hstack(sigs)
    sigs = sigs / np.sum(sigs)
    indx = truncation_order(sigs, qsym)
    nres = len(sigs)
    for i in range(nres):
        if sigs[indx[i]] < thresh * 1.01:
            nres = i
            break
    if nres == 0:
        nres = 1
    if Dcut > 0:
        nres = min(nres, Dcut)
    indx = indx[:nres]
    sigs = sigs[indx]
    dwts = 1.0 - np.sum(sigs)
    qsym = [qsym[i] for i in indx]
    rot = rot[:, indx]
    re = sorted(range(nres), key=lambda x: qsym[x])
    return dwts, np.array([qsym[i] for i in re]), rot[:, re].copy(), sigs[re].copy()


def mps0occun(K, occun):
    """Product-state MPS from a spin-orbital occupation (mpo_dmrg_util.py:118-155)."""
    qnuml = [np.array([[0., 0.]])]
    sites = []
    ne, ms = 0.0, 0.0
    for i in range(K):
        na, nb = occun[2 * i], occun[2 * i + 1]
        t = np.zeros(4)
        if na < 0.5 and nb < 0.5:
            t[0] = 1.0
        elif na < 0.5:
            ne += 1.0
            ms -= 0.5
            t[1] = 1.0
        elif nb < 0.5:
            ne += 1.0
            ms += 0.5
            t[2] = 1.0
        else:
            ne += 2.0
            t[3] = 1.0
        sites.append(t.reshape(1, 4, 1))
        qnuml.append(np.array([[ne, ms]]))
    qnumr = [np.array([[ne - q[0][0], ms - q[0][1]]]) for q in qnuml]
    return qnuml, qnumr, sites


class State:
    """The subset of the reference's mpo_dmrg object used on this path (class.py:47-192)."""

    def __init__(self, int1e, int2e, nelec, sz=0.0, Dmax=100, tol=1e-8, occun=None,
                 crit_vec=1e-8):
        self.sbas = int1e.shape[0]
        self.nsite = self.sbas // 2
        self.nops = self.sbas
        self.opers = list(range(self.sbas))
        self.hpwts = np.ones(self.nops)
        self.mpo = build_mpo(int1e, int2e)
        K = self.nsite
        self.qphys = [qphys_spatial() for _ in range(K)]
        self.qsectors = {(float(nelec), float(sz)): 1}
        self.Dmax = Dmax
        self.crit_e = tol
        self.crit_vec = crit_vec
        self.thresh = -1e-14
        self.psi0 = None
        self.lmps, self.rmps = {}, {}
        if occun is None:
            occun = default_occupation(K, nelec, sz)
        self.qnuml, self.qnumr, sites0 = mps0occun(K, occun)
        for i in range(K):
            qt = QTensor([False, False, True])
            qt.fromDenseTensor(sites0[i], [self.qnuml[i], self.qphys[i], self.qnuml[i + 1]])
            self.lmps[i] = qt
        # initOps: right boundary + all left environments (util.py:193-214)
        self.envR = {K: [boundary_env() for _ in range(self.nops)]}
        self.envL, _ = genHops(self, self.lmps, 'L')
        self.nmvp = 0


def default_occupation(K, nelec, sz=0.0):
    """Aufbau occupation, alpha/beta alternating in the open shell (the reference's
    samples pass an explicit occun, e.g. samps/example_conversion/0.h6_nqt/main_spmps.py:37)."""
    na = int(round(nelec / 2 + sz))
    nb = int(round(nelec / 2 - sz))
    occ = np.zeros(2 * K)
    occ[0:2 * na:2] = 1.0
    occ[1:2 * nb:2] = 1.0
    return occ


def ci_solver(st, isite, status, actlst, hvec_fn=None, stats=None):
    """One two-site step: template, HDiag, Davidson, decimation, guess (dot.py:24-112)."""
    K = st.nsite
    lenv, renv = st.envL[isite - 1], st.envR[isite + 2]
    ldim = int(lenv[0].shape[2])
    rdim = int(renv[0].shape[2])
    st.dims = (ldim, 16, rdim, ldim * 16 * rdim)
    st.qlst = [st.qnuml[isite], st.qphys[isite], st.qphys[isite + 1], st.qnumr[isite + 2]]
    civecs, eigs, nmvps = [], [], 0
    for key, neig in st.qsectors.items():
        st.qkey = np.array(key)
        qnums = [st.qlst[0], st.qlst[1], st.qlst[2], st.qkey - st.qlst[3]]
        st.qtmp = QTensor([False, False, False, True])
        rank, qsyms, ndims, idlst = fromQnums(qnums)
        st.qtmp.fromQsyms(rank, qsyms, ndims)
        st.idlstCI = idlst
        prjmap = st.qtmp.prjmap(idlst)
        ndim0 = st.qtmp.size_allowed
        if ndim0 == 0:
            raise RuntimeError('no basis in the symmetry sector %s' % (key,))
        diag = HDiagQt(st, lenv, renv, isite, prjmap)
        fn = hvec_fn or (lambda v: HVecQt(v, st, lenv, renv, isite, stats))
        v0 = None if st.psi0 is None else st.psi0[:neig, prjmap].copy()
        eig, vec, nmvp, _ = davidson.solve_iter(fn, diag, neig, v0, crit_e=st.crit_e,
                                                crit_vec=st.crit_vec)
        civec = np.zeros((neig, st.dims[3]))
        civec[:, prjmap] = vec
        civecs.append(civec)
        eigs += list(eig)
        nmvps += nmvp
    civecs = np.vstack(civecs)
    sigs, dwts, qred, site, srotR = decimation(st, civecs, isite, status)
    gen_guess(st, isite, actlst, civecs, srotR, status)
    return nmvps, np.array(eigs), sigs, dwts, qred, site, srotR


def decimation(st, civecs, isite, status):
    """Dense RDM of the half superblock, per-sector eigh, truncation (dotutil.py:288-495)."""
    ldim, cdim, rdim, _ = st.dims
    neig = civecs.shape[0]
    mats = civecs.reshape(neig, ldim * 4, 4 * rdim)
    if status == 'L':
        rdm = sum(m @ m.T for m in mats)
        classes = dpt(st.qnuml[isite], st.qphys[isite])
    else:
        rdm = sum(m.T @ m for m in mats)
        classes = dpt(st.qphys[isite + 1], st.qnumr[isite + 2])
    tr = np.trace(rdm)
    if tr < -1e-12:
        raise RuntimeError('RDM trace negative')
    elif tr < 1e-12:
        rdm = rdm + np.identity(rdm.shape[0]) / float(rdm.shape[0])
    else:
        rdm = rdm / tr
    dwts, qred, rot, sigs = rdm_blkdiag(rdm, classes, st.thresh, st.Dmax)
    nres = len(qred)
    if status == 'L':
        site = rot.reshape(ldim, 4, nres).copy()
        srotR = np.array([(rot.T @ m).reshape(nres, 4, rdim) for m in mats])
    else:
        site = rot.reshape(4, rdim, nres).transpose(2, 0, 1).copy()
        srotR = np.array([(m @ rot).reshape(ldim, 4, nres) for m in mats])
    return sigs, dwts, qred, site, srotR


def gen_guess(st, isite, actlst, civecs, srotR, status):
    """Next-step guess (dotutil.py:498-559, ncsite=2)."""
    if isite == actlst[-1]:
        st.psi0 = civecs.copy()
        return
    neig = civecs.shape[0]
    out = []
    for ie in range(neig):
        if status == 'R':
            t = st.lmps[isite - 1].toDenseTensor()
            t = np.tensordot(t, srotR[ie], axes=([2], [0]))
        else:
            t = st.rmps[isite + 2].toDenseTensor()
            t = np.tensordot(srotR[ie], t, axes=([2], [0]))
        out.append(t.reshape(-1))
    st.psi0 = np.array(out)


def update_site(st, isite, qred, site, status):
    """opt.py:313-344."""
    if status == 'L':
        st.qnuml[isite + 1] = qred.copy()
        qt = QTensor([False, False, True])
        qt.fromDenseTensor(site, [st.qlst[0], st.qlst[1], qred])
        st.lmps[isite] = qt
    else:
        st.qnumr[isite + 1] = qred.copy()
        qt = QTensor([True, False, False])
        qt.fromDenseTensor(site, [qred, st.qlst[-2], st.qlst[-1]])
        st.rmps[isite + 1] = qt
    return qt


def update_boundary(st, qred, srotR, status):
    """opt.py:347-377."""
    K = st.nsite
    smat = srotR[0] / np.linalg.norm(srotR[0])
    key = list(st.qsectors.keys())[0]
    qsym = np.array([key])
    if status == 'L':
        qt = QTensor([False, False, True])
        qt.fromDenseTensor(smat, [qred, st.qphys[K - 1], qsym])
        st.lmps[K - 1] = qt
        st.qnuml[K] = qsym
    else:
        qt = QTensor([True, False, False])
        qt.fromDenseTensor(smat, [qsym, st.qphys[0], qred])
        st.rmps[0] = qt
        st.qnumr[0] = qsym


def sweep(st, status, stats=None, max_steps=None):
    """One half sweep (opt.py:31-95).  status 'R' runs right-to-left."""
    K = st.nsite
    actlst = list(range(K - 2, -1, -1)) if status == 'R' else list(range(0, K - 1))
    elst, dlst, nmvp = [], [], 0
    qred = srotR = None
    for n, isite in enumerate(actlst):
        if max_steps is not None and n >= max_steps:
            break
        nm, eigs, sigs, dwts, qred, site, srotR = ci_solver(st, isite, status, actlst,
                                                           stats=stats)
        nmvp += nm
        elst.append(eigs)
        dlst.append(dwts)
        qsite = update_site(st, isite, qred, site, status)
        if status == 'L':
            st.envL[isite] = renorm_site(st, st.envL[isite - 1], qsite, isite, 'L')
        else:
            st.envR[isite + 1] = renorm_site(st, st.envR[isite + 2], qsite, isite + 1, 'R')
    if max_steps is None:
        update_boundary(st, qred, srotR, status)
    eavs = np.array([e[0] for e in elst])
    indx = int(np.argmin(eavs))
    return nmvp, indx, float(eavs[indx]), elst, dlst


def run_dmrg(st, nsweeps, tol=None):
    """Fixed2 sweeps 'R' then 'L'; energy = min of the two half-sweep minima (class.py:385-430)."""
    tol = st.crit_e if tol is None else tol
    eold = 1e3
    energies = []
    for isweep in range(nsweeps):
        r1 = sweep(st, 'R')
        r2 = sweep(st, 'L')
        eav = min(r1[2], r2[2])
        energies.append(eav)
        st.nmvp += r1[0] + r2[0]
        if abs(eav - eold) < tol:
            break
        eold = eav
    return energies
