"""Two-site sigma, diagonal, renormalisation and environment build (oracle).

Restates source/mpo_dmrg_kernelQt.py (HDiagQt :30-127, HVecQt :153-319,
renormQt_H00 :633-767) and source/mpo_dmrg_initQt.py (genBopsQt :49-85,
genHopsQt :88-271) for maxslc=1, with environments held in memory instead of
HDF5 files.  Test infrastructure only.

Environment store: ``envL[i]`` / ``envR[i]`` are lists (one QTensor per op)
equal to the reference files ``lop_site_i`` / ``rop_site_i``; the boundaries
are i=-1 (left) and i=K (right), stored under those keys of a dict.
"""
import numpy as np

from .mpo import link_two_opers  # noqa: F401  (re-exported for tests)
from .qtensor import QTensor


def case(porb, isite, ncsite=1):
    """Operator class of H_p relative to the dot (qtensor_opers.py:72-90)."""
    if porb // 2 < isite:
        return 0 + (porb % 2)
    if porb // 2 == isite:
        return 2 + (porb % 2)
    if ncsite == 1 or porb // 2 == isite + 1:
        return 4 + (porb % 2)
    return 6 + (porb % 2)


def boundary_env(nqnum=2):
    """1x1x1 boundary operator with zero QN (initQt.py:49-85)."""
    qt = QTensor([True, False, True])
    z = [[0.0] * nqnum]
    qt.fromDenseTensor(np.ones((1, 1, 1)), [z, z, z])
    return qt


def _flip_rop(rop, qkey):
    """R environment in left convention: [-q0, qkey-q1, qkey-q2] (kernelQt.py:297-300)."""
    r = rop.copy()
    r.qsyms = [-rop.qsyms[0], qkey - rop.qsyms[1], qkey - rop.qsyms[2]]
    return r


def _flip_cop_horizontal(cop):
    """Negate and reverse both MPO bonds for the 'R' direction (kernelQt.py:733-738)."""
    c = cop.copy()
    c.qsyms = [-cop.qsyms[0], -cop.qsyms[1], cop.qsyms[2], cop.qsyms[3]]
    c.status = cop.status.copy()
    c.status[0] = not c.status[0]
    c.status[1] = not c.status[1]
    c.idlst = cop.idlst
    return c


def HVecQt(vec0, st, lenv, renv, isite, stats=None):
    """Two-site sigma = sum_p hpwts[p] L_p W_p[i] W_p[i+1] R_p psi (kernelQt.py:153-156, 238-317).

    ``st`` carries the qtmp template, qkey, mpo and hpwts; ``lenv``/``renv``
    are the per-op environments left of isite and right of isite+1.
    ``stats``, if a dict, accumulates the algorithmic FLOPs per stage.
    """
    qvec = st.qtmp
    qvec.value = np.array(vec0, dtype=np.float64).copy()
    hvec = np.zeros(len(vec0))
    jsite = isite + 1
    # The reference rebuilds the symbolic plans on every call (kernelQt.py:167-174);
    # they depend only on the block structure, so they are kept for the whole
    # Davidson solve of one site (identical results, less Python time).
    cache = getattr(st, '_hvec_plans', None)
    if cache is not None and cache[0] is qvec and cache[1] == isite and cache[2] is lenv \
            and cache[3] is renv:
        top1, top2, top3, top4 = cache[4]
    else:
        top1, top2, top3, top4 = {}, {}, {}, {}
        st._hvec_plans = (qvec, isite, lenv, renv, (top1, top2, top3, top4))
    for iop in range(st.nops):
        p = st.opers[iop]
        c0 = case(p, isite, 2)
        lop = lenv[iop]
        cop = st.mpo[isite][iop]
        wop = st.mpo[jsite][iop]
        rop = renv[iop]
        if c0 not in top1:
            top1[c0] = QTensor()
            top1[c0].tensordotSYM(lop, qvec, axes=([2], [0]))
            top2[c0] = QTensor()
            top2[c0].tensordotSYM(top1[c0], cop, axes=([0, 2], [0, 3]))
            rflip = _flip_rop(rop, st.qkey)
            top3[c0] = QTensor()
            top3[c0].tensordotSYM(top2[c0], wop, axes=([1, 3], [3, 0]))
            top4[c0] = QTensor()
            top4[c0].tensordotSYM(top3[c0], rflip, axes=([1, 3], [2, 0]))
        if top1[c0].size_allowed == 0 or top4[c0].nblks_allowed <= 0:
            continue
        top1[c0].tensordotCAL(lop, qvec)
        top2[c0].tensordotCAL(top1[c0], cop)
        top3[c0].tensordotCAL(top2[c0], wop)
        top4[c0].tensordotCAL(top3[c0], rop)
        hvec += st.hpwts[iop] * top4[c0].value
        if stats is not None:
            for k, t in enumerate((top1, top2, top3, top4)):
                stats[k] = stats.get(k, 0) + t[c0].flops()
    return hvec


def HDiagQt(st, lenv, renv, isite, prjmap):
    """Diagonal of H_eff in the dense (l,n1,n2,r) space, then [prjmap] (kernelQt.py:30-127)."""
    ldim, cdim, rdim, ndim = st.dims
    diag = np.zeros(ndim)
    jsite = isite + 1
    for iop in range(st.nops):
        ld = lenv[iop].toDenseDiag(lenv[iop].idlst)
        cd = st.mpo[isite][iop].toDenseDiag(st.mpo[isite][iop].idlst)
        wd = st.mpo[jsite][iop].toDenseDiag(st.mpo[jsite][iop].idlst)
        rd = renv[iop].toDenseDiag(renv[iop].idlst)
        tmp = np.tensordot(ld, cd, axes=([0], [0]))      # [l, q, c]
        tmp = np.tensordot(tmp, wd, axes=([1], [0]))     # [l, c, s, w]
        tmp = np.tensordot(tmp, rd, axes=([2], [0]))     # [l, c, w, r]
        diag += st.hpwts[iop] * tmp.reshape(ndim)
    return diag[prjmap].copy()


def renorm_site(st, env_in, qsite, isite, status, bra=None):
    """One renormalisation step for all ops (kernelQt.py:633-767 / initQt.py:88-271).

    status 'L': L'[wr,r,r'] = A*[l,n,r] L[w,l,k] W[w,wr,n,d] A[k,d,r'].
    status 'R': the mirror with the MPO's horizontal QNs flipped.
    ``bra`` defaults to ``qsite`` (renormQt_H00 uses bra = ket).
    """
    bsite = qsite if bra is None else bra
    out = []
    t1, t2, t3 = {}, {}, {}
    for iop in range(st.nops):
        p = st.opers[iop]
        c0 = case(p, isite)
        cop = st.mpo[isite][iop]
        op_in = env_in[iop]
        if status == 'L':
            if c0 not in t1:
                t1[c0] = QTensor()
                t1[c0].tensordotSYM(bsite, op_in, axes=([0], [1]))
                t2[c0] = QTensor()
                t2[c0].tensordotSYM(cop, t1[c0], axes=([0, 2], [2, 0]))
                t3[c0] = QTensor()
                t3[c0].tensordotSYM(t2[c0], qsite, axes=([1, 3], [1, 0]))
            idl = [cop.idlst[1], bsite.idlst[2], qsite.idlst[2]]
        else:
            if c0 not in t1:
                cflip = _flip_cop_horizontal(cop)
                t1[c0] = QTensor()
                t1[c0].tensordotSYM(bsite, op_in, axes=([2], [1]))
                t2[c0] = QTensor()
                t2[c0].tensordotSYM(cflip, t1[c0], axes=([1, 2], [2, 1]))
                t3[c0] = QTensor()
                t3[c0].tensordotSYM(t2[c0], qsite, axes=([1, 3], [1, 2]))
            idl = [cop.idlst[0], bsite.idlst[0], qsite.idlst[0]]
        t1[c0].tensordotCAL(bsite, op_in, ifc1=True)
        t2[c0].tensordotCAL(cop, t1[c0])
        t3[c0].tensordotCAL(t2[c0], qsite)
        new = t3[c0].copy()
        new.idlst = idl
        out.append(new)
    return out


def genHops(st, sites, status):
    """Build every environment from an MPS (initQt.py:88-271).

    Returns (envs, exphop) with envs[i] the per-op list after site i and
    exphop[p] = <psi|H_p|psi> read off the far boundary (initQt.py:173,255).
    """
    K = st.nsite
    envs = {}
    if status == 'L':
        envs[-1] = [boundary_env() for _ in range(st.nops)]
        for i in range(K):
            envs[i] = renorm_site(st, envs[i - 1], sites[i], i, 'L')
        exphop = np.array([e.value[0] for e in envs[K - 1]])
    else:
        envs[K] = [boundary_env() for _ in range(st.nops)]
        for i in range(K - 1, -1, -1):
            envs[i] = renorm_site(st, envs[i + 1], sites[i], i, 'R')
        exphop = np.array([e.value[0] for e in envs[0]])
    return envs, exphop
