"""Synthetic block-sparse MPS with a DMRG-like quantum-number structure.

Used to build realistic environments for benchmarking a middle-bond sigma
at sizes where a converged sweep would take hours: bond sectors (N, Sz) get
a dimension proportional to exp(-(N-N0)^2/(2 s_N^2) - Sz^2/(2 s_S^2)),
capped by what the previous bond and the four local states can reach; site
tensors are random with orthonormal columns per sector (left-canonical from
the left end, right-canonical from the right end).
"""
import numpy as np

from .mpo import qphys
from .qtensor import QTensor


def _bond_dims(prev, nsite_block, ntot_block_max, nelec_left_mean, D, sN, sS, nelec_total,
               sites_total):
    """prev: dict (N, 2Sz) -> dim.  Returns the next bond's dict."""
    avail = {}
    for (n, s2), d in prev.items():
        for (dn, ds2) in ((0, 0), (1, -1), (1, 1), (2, 0)):
            key = (n + dn, s2 + ds2)
            avail[key] = avail.get(key, 0) + d
    # electrons still placeable on the remaining sites
    rem = sites_total - nsite_block
    keep = {k: v for k, v in avail.items()
            if 0 <= nelec_total - k[0] <= 2 * rem and abs(k[1]) <= 2 * rem + 2 * nsite_block}
    keys = sorted(keep)
    w = np.array([np.exp(-(k[0] - nelec_left_mean) ** 2 / (2 * sN ** 2)
                         - (0.5 * k[1]) ** 2 / (2 * sS ** 2)) for k in keys])
    cap = np.array([keep[k] for k in keys])
    total = min(D, int(cap.sum()))
    dims = np.zeros(len(keys), dtype=np.int64)
    # water-filling allocation proportional to w, respecting caps
    left = total
    active = np.ones(len(keys), bool)
    while left > 0 and active.any():
        share = w * active
        share = share / share.sum()
        add = np.floor(share * left).astype(np.int64)
        if add.sum() == 0:
            add[np.argmax(share)] = 1
        add = np.minimum(add, cap - dims)
        dims += add
        left -= int(add.sum())
        active &= dims < cap
    return {k: int(d) for k, d in zip(keys, dims) if d > 0}


def _qlist(dims):
    q = []
    for (n, s2) in sorted(dims):
        q += [[float(n), 0.5 * s2]] * dims[(n, s2)]
    return np.array(q)


def _orth_site(qt, rng, left):
    """Fill qt with random values orthonormal per (n-summed) sector column/row."""
    val = rng.standard_normal(qt.size_allowed)
    qt.value = val
    dense = qt.toDenseTensor()
    l, p, r = dense.shape
    if left:
        m = dense.reshape(l * p, r)
        q_cols = qt.qsyms[2]
        out = np.zeros_like(m)
        for s in range(len(qt.ndims[2])):
            cols = qt.idlst[2][s]
            sub = m[:, cols]
            rows = np.flatnonzero(np.any(sub != 0, axis=1))
            if len(rows) >= len(cols):
                qmat, _ = np.linalg.qr(sub[rows])
                out[np.ix_(rows, cols)] = qmat
        dense = out.reshape(l, p, r)
    else:
        m = dense.reshape(l, p * r)
        out = np.zeros_like(m)
        for s in range(len(qt.ndims[0])):
            rws = qt.idlst[0][s]
            sub = m[rws]
            cols = np.flatnonzero(np.any(sub != 0, axis=0))
            if len(cols) >= len(rws):
                qmat, _ = np.linalg.qr(sub[:, cols].T)
                out[np.ix_(rws, cols)] = qmat.T
        dense = out.reshape(l, p, r)
    qt.fromDense(dense, qt.idlst)
    return qt


def synthetic_mps(K, nelec, D, upto_left, from_right, seed=0, sN=None, sS=None):
    """Left-canonical sites 0..upto_left and right-canonical sites from_right..K-1.

    Returns (lsites, rsites, qnuml, qnumr) with qnuml[i] the QNs of the bond left
    of site i (left convention) and qnumr[i] those of the block i..K-1.
    """
    rng = np.random.default_rng(seed)
    sN = sN if sN is not None else 0.5 + 0.08 * K
    sS = sS if sS is not None else 0.5 + 0.03 * K
    qp = qphys()
    # left bonds
    dims = {(0, 0): 1}
    qnuml = {0: np.array([[0., 0.]])}
    lsites = {}
    for i in range(upto_left + 1):
        nxt = _bond_dims(dims, i + 1, 2 * (i + 1), nelec * (i + 1) / K, D, sN, sS, nelec, K)
        qnuml[i + 1] = _qlist(nxt)
        qt = QTensor([False, False, True])
        qt.fromDenseTensor(np.zeros((len(qnuml[i]), 4, len(qnuml[i + 1]))),
                           [qnuml[i], qp, qnuml[i + 1]])
        lsites[i] = _orth_site(qt, rng, True)
        dims = nxt
    # right bonds (right-block QNs)
    dims = {(0, 0): 1}
    qnumr = {K: np.array([[0., 0.]])}
    rsites = {}
    for i in range(K - 1, from_right - 1, -1):
        nb = K - i
        nxt = _bond_dims(dims, nb, 2 * nb, nelec * nb / K, D, sN, sS, nelec, K)
        qnumr[i] = _qlist(nxt)
        qt = QTensor([True, False, False])
        qt.fromDenseTensor(np.zeros((len(qnumr[i]), 4, len(qnumr[i + 1]))),
                           [qnumr[i], qp, qnumr[i + 1]])
        rsites[i] = _orth_site(qt, rng, False)
        dims = nxt
    return lsites, rsites, qnuml, qnumr


# (N, Sz) sector widths of the synthetic MPS.  'dmrg' matches the middle
# bond of a real K = 36, D = 1000 DMRG state (the first D = 1000 sweep of a
# Normal2 ramp from the product state: 24 / 21 sectors, ndim0 1.20e6, 5.7
# TFLOP per sigma; synthetic: 28 / 28 sectors, ndim0 1.18e6); 'wide' is the
# generator's default, with many small sectors (116 at K = 36, ndim0 2.8e5).
STRUCTURES = {'dmrg': lambda K: (0.5 + 0.02 * K, 0.35 + 0.01 * K), 'wide': lambda K: (None, None)}


def middle_bond(K, D, seed, budget, fused=True, extra=None, structure='dmrg', log=None):
    """Middle-bond two-site sigma workload of a K-orbital, N = K, Sz = 0
    problem at bond dimension D: seeded integrals, the device MPO, a
    synthetic MPS (sector widths by ``structure``) and its left / right
    environments built on the device by the renormalisation launches.
    Returns (mpo, isite, st, L, R, prog) with st.qtmp / st.qkey the psi
    template and prog its SigmaProgram.  ``extra`` (a dict) also receives
    the site tensors, bond quantum numbers and the environments one site
    before L and R with the site tensors that renormalise them."""
    import time
    from . import device as dv
    from . import kernelQt as kq
    from . import mpo as pmpo
    from .dot import setupQtmpCI
    from .integrals import gen_integrals
    log = log or (lambda msg: None)
    t0 = time.time()
    i1, i2 = gen_integrals(K, seed=seed)
    mpo = pmpo.MPOStore(i1, i2)
    log('MPO K=%d built in %.1fs' % (K, time.time() - t0))
    isite = K // 2 - 1
    sN, sS = STRUCTURES[structure](K)
    ls, rs, ql, qr = synthetic_mps(K, K, D, isite - 1, isite + 2, seed=seed, sN=sN, sS=sS)
    if extra is not None:
        extra.update(ls=ls, rs=rs, ql=ql, qr=qr)
    log('synthetic MPS: bond dims L %s R %s' % (len(ql[isite]), len(qr[isite + 2])))
    ren_flops = 0
    t = time.time()
    env = kq.boundary_store(mpo.nops)
    for i in range(isite):
        if extra is not None and i == isite - 1:
            extra.update(Lprev=env, lsite=ls[i], lidx=i)
        env, prog = kq.renorm(mpo, i, ls[i], env, 'L')
        ren_flops += prog.flops
    L = env
    env = kq.boundary_store(mpo.nops)
    for i in range(K - 1, isite + 1, -1):
        if extra is not None and i == isite + 2:
            extra.update(Rprev=env, rsite=rs[i], ridx=i)
        env, prog = kq.renorm(mpo, i, rs[i], env, 'R')
        ren_flops += prog.flops
    R = env
    dv.sync()
    log('environments built (%d renormalisations, %.1f GFLOP, %.1fs incl. plans)' % (
        K - 2, ren_flops / 1e9, time.time() - t))

    class St:
        pass
    st = St()
    st.qlst = [ql[isite], pmpo.qphys(), pmpo.qphys(), qr[isite + 2]]
    setupQtmpCI(st, (float(K), 0.0))
    t = time.time()
    prog = kq.SigmaProgram(mpo, isite, st.qtmp, st.qkey, L, R, mpo.hpwts, budget=budget,
                           fused=fused)
    log('sigma plan: ndim0=%d flops=%.3e chunks=%d (%.1fs)' % (
        st.qtmp.size_allowed, prog.flops, len(prog.chunks), time.time() - t))
    return mpo, isite, st, L, R, prog
