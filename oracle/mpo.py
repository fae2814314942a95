"""MPO site tensors H_p[i] = a_p^dagger V_p (oracle restatement).

Dense factors follow source/mpo_dmrg_opers.py:33-260 with the constant
matrices of source/mpo_dmrg_const.py:12-20; quantum numbers and the Qt form
follow source/qtensor/qtensor_opers.py:56-193 and :358-408.  Test
infrastructure only.
"""
import numpy as np

from .qtensor import QTensor

# mpo_dmrg_const.py:12-20 (2x2 single spin-orbital matrices, basis |0>,|1>)
SGN = np.array([[1., 0.], [0., -1.]])
IDN = np.array([[1., 0.], [0., 1.]])
IDNT = np.array([[1., 0.], [0., -1.]])
CRE = np.array([[0., 0.], [1., 0.]])
CRET = np.array([[0., 0.], [1., 0.]])
ANN = np.array([[0., 1.], [0., 0.]])
ANNT = np.array([[0., -1.], [0., 0.]])
NII = np.array([[0., 0.], [0., 1.]])
NIIT = np.array([[0., 0.], [0., -1.]])


def link_two_opers(op1, op2):
    """W1[a,b,i,j] W2[b,c,k,l] -> W[a,c,(ik),(jl)]  (mpo_dmrg_opers.py:43-49)."""
    op12 = np.tensordot(op1, op2, axes=([1], [0])).transpose(0, 3, 1, 4, 2, 5)
    s = op12.shape
    return op12.reshape(s[0], s[1], s[2] * s[3], s[4] * s[5])


def elem_spin_mat(p, isite):
    """a_p^dagger on spin-orbital isite with JW string (mpo_dmrg_opers.py:56-66)."""
    if isite < p:
        return SGN
    if isite == p:
        return CRE
    return IDN


def wfac_spin(nsite, isite, hq, vqrs):
    """Three-index V_p factor on spin-orbital site isite (mpo_dmrg_opers.py:91-194).

    Bond layout (both bonds, length 2(nsite+1)): [0] | q1 block (right sites) |
    q2 block (right sites) | q3 block (left sites) | q4 block (left sites) | last.
    The first / last site keep only row 0 / the last column.
    """
    nb = 2 * (nsite + 1)
    diml, dimr = isite, nsite - isite - 1
    w = np.zeros((nb, nb, 2, 2))
    last = nb - 1
    rs = range(isite + 1, nsite)
    # row 0
    w[0, 0] = IDNT
    for r, rsite in enumerate(rs):
        w[0, 1 + dimr + r] = vqrs[isite, isite, rsite] * NIIT
    w[0, 1 + 2 * dimr + diml] = -ANN
    w[0, 2 + 2 * dimr + 2 * diml] = CRE
    w[0, last] = hq[isite] * ANN
    # q1 (left) rows
    w[1, last] = CRE
    for r in range(dimr):
        w[2 + r, 1 + r] = IDNT
    # q2 (left) rows
    w[2 + dimr, last] = ANN
    for r in range(dimr):
        w[3 + dimr + r, 1 + dimr + r] = IDNT
    # q3 (left) rows
    off = 3 + 2 * dimr
    for l in range(diml):
        for r, rsite in enumerate(rs):
            w[off + l, 1 + r] = -vqrs[rsite, l, isite] * ANNT
            w[off + l, 1 + dimr + r] = vqrs[isite, l, rsite] * CRET
        w[off + l, 1 + 2 * dimr + l] = IDN
        w[off + l, last] = vqrs[isite, l, isite] * NII
    # q4 (left) rows
    off4 = 3 + 2 * dimr + diml
    for l in range(diml):
        for r, rsite in enumerate(rs):
            w[off4 + l, 1 + dimr + r] = vqrs[l, isite, rsite] * ANNT
        w[off4 + l, 2 + 2 * dimr + diml + l] = IDN
    w[last, last] = IDN
    if isite == 0:
        w = w[:1]
    if isite == nsite - 1:
        w = w[:, last:]
    return w


def hfac_spin(p, nsite, isite, hq, vqrs):
    """a_p^dagger applied on top of V_p (mpo_dmrg_opers.py:197-203)."""
    return np.einsum('ij,abjk->abik', elem_spin_mat(p, isite), wfac_spin(nsite, isite, hq, vqrs))


def hfac_spatial(p, nsite, isite, hq, vqrs):
    """Spatial site = link of its alpha and beta spin orbitals (mpo_dmrg_opers.py:216-219)."""
    return link_two_opers(hfac_spin(p, nsite, 2 * isite, hq, vqrs),
                          hfac_spin(p, nsite, 2 * isite + 1, hq, vqrs))


# ---------------------------------------------------------------- QNs
SPINS = (0.5, -0.5)


def qphys_spatial():
    """Local states |0>,|b>,|a>,|ab> (qtensor_opers.py:95-109, isym=2)."""
    return np.array([[0., 0.], [1., -0.5], [1., 0.5], [2., 0.]])


def elem_qnums(p, isite):
    """QN shift of a_p^dagger across spin site isite (qtensor_opers.py:113-142)."""
    spin = 0.5 if p % 2 == 0 else -0.5
    if isite < p:
        return np.array([[0., 0.]]), np.array([[0., 0.]])
    if isite == p:
        return np.array([[0., 0.]]), np.array([[1., spin]])
    return np.array([[1., spin]]), np.array([[1., spin]])


def wfac_qnums_right(nsite, isite, sz):
    """QNs of the right bond of spin site isite (qtensor_opers.py:161-172)."""
    q = [[0., 0.]]
    q += [[-2., -sz - SPINS[i % 2]] for i in range(isite + 1, nsite)]
    q += [[0., -sz + SPINS[i % 2]] for i in range(isite + 1, nsite)]
    q += [[-1., -SPINS[i % 2]] for i in range(isite + 1)]
    q += [[1., SPINS[i % 2]] for i in range(isite + 1)]
    q += [[-1., -sz]]
    return np.array(q)


def wfac_qnums(nsite, isite, isz):
    """(qtensor_opers.py:175-193), status 'L'."""
    sz = SPINS[isz]
    if isite == 0:
        return np.array([[0., 0.]]), wfac_qnums_right(nsite, isite, sz)
    if isite == nsite - 1:
        return wfac_qnums_right(nsite, isite - 1, sz), np.array([[-1., -sz]])
    return wfac_qnums_right(nsite, isite - 1, sz), wfac_qnums_right(nsite, isite, sz)


def sort_qnums(qnums):
    """Stable sort of bond indices by the printed QN (qtensor_opers.py:56-58).

    The key is numpy's str() of the QN row under this numpy version; the
    reference used its own numpy's format (a block-layout-only hazard).
    """
    memo = {}
    keys = []
    for row in qnums:
        t = np.asarray(row, dtype=np.float64).tobytes()
        if t not in memo:
            memo[t] = str(np.asarray(row))
        keys.append(memo[t])
    return np.array(sorted(range(len(qnums)), key=lambda k: keys[k]))


def mpo_site(p, nsite, isite, int1e, int2e):
    """Qt MPO tensor of H_p on spatial site isite (qtensor_opers.py:358-408).

    nsite = number of spin orbitals (2K); int1e[p] / int2e[p] are the rows of
    the spin-orbital integrals.  Legs [w_left, w_right, u, d], status
    [F, T, F, T], physical legs uncollected (ifcollect=[1,1,0,0]).
    """
    hq, vqrs = int1e[p], int2e[p]
    site = hfac_spatial(p, nsite, isite, hq, vqrs)
    isz = p % 2
    qu = qphys_spatial()
    ql1e, _ = elem_qnums(p, 2 * isite)
    _, qr2e = elem_qnums(p, 2 * isite + 1)
    ql1w, _ = wfac_qnums(nsite, 2 * isite, isz)
    _, qr2w = wfac_qnums(nsite, 2 * isite + 1, isz)
    ql = np.array([a + b for a in ql1e for b in ql1w])
    qr = np.array([a + b for a in qr2e for b in qr2w])
    il, ir = sort_qnums(ql), sort_qnums(qr)
    ql, qr = ql[il], qr[ir]
    site = site[np.ix_(il, ir)].copy()
    qt = QTensor([False, True, False, True])
    qt.fromDenseTensor(site, [ql, qr, qu, qu], ifcollect=[1, 1, 0, 0])
    return qt


def build_mpo(int1e, int2e, nops=None):
    """All H_p site tensors: mpo[isite][p] (io.py:168-206, maxslc=1)."""
    sbas = int1e.shape[0]
    K = sbas // 2
    nops = sbas if nops is None else nops
    return [[mpo_site(p, sbas, i, int1e, int2e) for p in range(nops)] for i in range(K)]
