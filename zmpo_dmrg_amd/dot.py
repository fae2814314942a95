"""Local two-site solver: template, diagonal, Davidson, decimation, guess.

Counterpart of source/mpo_dmrg_dot.py (ci_solver :24-112) and
source/mpo_dmrg_dotutil.py (setupQlst :81-95, symmetrySpaceInfo :148-189,
setupQtmpCI :192-203, genHDiag :206-231, finalizeDot :259-285, decimation
:288-495, genGuess :498-559) for ncsite=2, one process, noise=0.

Sigma, diagonal and the Davidson vector algebra run on the device.  The
reduced-density-matrix decimation runs on the host with numpy over the QN
sectors (per-sector eigh, as mpo_dmrg_qparser.rdm_blkdiag); its truncation
rule is deterministic: weights sorted descending, weights <= 1e-12 counting as
zero; runs of weights equal to 1e-10 relative form tie clusters ordered by (N, Sz) then
position; keep the first min(Dcut, n); re-sort the kept states by (N, Sz).
Zero-weight states (kept by the reference, thresh = -1e-14) get a fixed
basis of the RDM null space (canonical_null) instead of LAPACK's arbitrary
one.  The reference's own rule depends on Python-2 dict order and an
unstable argsort (mpo_dmrg_qparser.py:220-233, 259) and cannot be
reproduced.
"""
import os
import time

import numpy as np

from . import _lib
from . import device as dv
from . import kernelQt
from .pdvdson import DavidsonError, EigenSolver
from .qtensor import QTensor, fromQnums


# ZMPO_GRAPH=1: Davidson sigma applications after the first of a step replay
# a captured HIP graph of the sigma launches (one host call instead of one per
# launch).  Off by default: measured from K = 12, D = 128 to K = 24, D = 512
# the sigma is bound by its kernels, not by host launch cost (replay and
# eager launches take the same time), and each capture costs a little.
GRAPH = os.environ.get('ZMPO_GRAPH', '0') != '0'

# RDM weights (trace-normalised) at or below this are numerically zero: they
# are ordered by QN only, and their eigenvectors get the canonical basis.
WZERO = 1e-12


def dpt(q1, q2):
    """Direct-product QNs, second index fastest (mpo_dmrg_qphys.py:63-68)."""
    q1, q2 = np.asarray(q1), np.asarray(q2)
    return (q1[:, None, :] + q2[None, :, :]).reshape(-1, q1.shape[1])


def truncation_order(sigs, qsyms):
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


def setupQlst(dmrg, isite):
    """dotutil.py:81-95 (ncsite=2)."""
    dmrg.qlst = [np.asarray(dmrg.qnuml[isite]), dmrg.qphys[isite], dmrg.qphys[isite + 1],
                 np.asarray(dmrg.qnumr[isite + 2])]


def setupQtmpCI(dmrg, key):
    """dotutil.py:192-203 (ncsite=2)."""
    dmrg.qkey = np.array(key, dtype=np.float64)
    qnums = [dmrg.qlst[0], dmrg.qlst[1], dmrg.qlst[2], dmrg.qkey - dmrg.qlst[3]]
    dmrg.qtmp = QTensor([False, False, False, True])
    rank, qsyms, ndims, idlst = fromQnums(qnums)
    dmrg.qtmp.fromQsyms(rank, qsyms, ndims, ifallocate=False)
    dmrg.idlstCI = idlst


def ci_solver(dmrg, isite, ncsite, actlst, flst, status, ifsym=True):
    """One two-site step (dot.py:24-112).  Returns (nmvps, eigs, sigs, dwts, qred, site, srotR)."""
    if ncsite != 2:
        raise NotImplementedError('device solver implements ncsite=2')
    fL, fR = flst[0][0:2]
    ldim = int(fL.meta[0].shape[2])
    rdim = int(fR.meta[0].shape[2])
    dmrg.dims = (ldim, 16, rdim, ldim * 16 * rdim)
    setupQlst(dmrg, isite)
    info = [dmrg, isite, ncsite, flst, status, ifsym]
    civecs, eigs, nmvps = [], [], 0
    tm = dmrg.__dict__.setdefault('timing', {})

    def lap(key, t0):
        """Accumulate wall time since t0 under key (device work finished)."""
        dv.sync()
        t = time.time()
        tm[key] = tm.get(key, 0.0) + t - t0
        return t
    for key, neig in dmrg.qsectors.items():
        setupQtmpCI(dmrg, key)
        prjmap = dmrg.qtmp.prjmap(dmrg.idlstCI)
        ndim0 = dmrg.qtmp.size_allowed
        if ndim0 == 0:
            raise RuntimeError('no basis in the symmetry sector %s' % (key,))
        # the sigma plans are host work: built first, without waiting for
        # the device, they overlap the previous step's renormalisation
        # launches (whose remaining device time then counts under 'diag')
        t0 = time.time()
        prog = kernelQt._sigma_program(dmrg, isite, fL, fR)
        t1 = time.time()
        tm['sigma_plan'] = tm.get('sigma_plan', 0.0) + t1 - t0
        diag_h = kernelQt.HDiagQt(info, ndim0, prjmap)
        t0 = lap('diag', t1)
        dmrg.last_step = {'ndim0': ndim0, 'sigma_gflop': prog.flops / 1e9,
                          'sectors': (len(dmrg.qtmp.ndims[0]), len(dmrg.qtmp.ndims[3])),
                          'dims': (int(dmrg.qtmp.shape[0]), int(dmrg.qtmp.shape[3]))}
        dtype = getattr(dmrg, 'dtype', 0)
        solver = EigenSolver()
        solver.maxcycle = getattr(dmrg, 'maxcycle', solver.maxcycle)
        solver.crit_e = dmrg.crit_e
        solver.crit_vec = dmrg.crit_vec
        solver.ndim = ndim0
        solver.neig = neig
        solver.diag = dv.upload(diag_h, 0)
        # low-precision operand / sigma buffers per tier (BF16 runs also use
        # FP32 sigmas between the BF16 phase and the FP64 corrections)
        lo_bufs = {}

        def lo_buffers(dt):
            if dt not in lo_bufs:
                lo_bufs[dt] = (dv.empty(ndim0, dt), dv.empty(ndim0, kernelQt.sigma_dtype(dt)))
            return lo_bufs[dt]
        psi_lo, sig_lo = lo_buffers(dtype) if dtype else (None, None)

        graphs = {}

        def graph_sigma(x, out, dt):
            """From the second application in a dtype on, replay the captured
            launches (the first one runs eagerly and allocates)."""
            if dt in graphs:
                graphs[dt].run(x, out)
                return True
            if GRAPH and dt in seen:
                graphs[dt] = kernelQt.SigmaGraph(prog, dmrg.fhop, fL, fR, dt, warmup=False)
                graphs[dt].run(x, out)
                return True
            seen.add(dt)
            return False

        seen = set()

        def hvec(x, out):
            if graph_sigma(x, out, dtype):
                pass
            elif dtype == 0:
                prog.apply(dmrg.fhop, fL, fR, x, out, 0)
            else:
                _convert(psi_lo, dtype, x, 0)
                prog.apply(dmrg.fhop, fL, fR, psi_lo, sig_lo, dtype)
                _convert(out, 0, sig_lo, kernelQt.sigma_dtype(dtype))
            dmrg.nsigma += 1
            dmrg.sigma_flops += prog.flops

        def hvec64(x, out):
            if not graph_sigma(x, out, 0):
                prog.apply(dmrg.fhop, fL, fR, x, out, 0)
            dmrg.nsigma64 = getattr(dmrg, 'nsigma64', 0) + 1
            dmrg.sigma_flops += prog.flops

        def hvec_tier(dt):
            """Sigma with operands in dt (FP64 vectors in and out); None:
            a callable that releases the intermediates of a finished tier."""
            if dt is None:
                return prog.drop_buffers
            if dt == dtype:
                return hvec

            def h(x, out):
                pl, sl = lo_buffers(dt)
                _convert(pl, dt, x, 0)
                prog.apply(dmrg.fhop, fL, fR, pl, sl, dt)
                _convert(out, 0, sl, kernelQt.sigma_dtype(dt))
                dmrg.nsigma += 1
                dmrg.nsigma_tier = getattr(dmrg, 'nsigma_tier', {})
                dmrg.nsigma_tier[dt] = dmrg.nsigma_tier.get(dt, 0) + 1
                dmrg.sigma_flops += prog.flops
            return h

        solver.HVec = hvec
        v0 = None
        if dmrg.psi0 is not None:
            if isinstance(dmrg.psi0, np.ndarray):
                v0 = np.ascontiguousarray(dmrg.psi0[:neig][:, prjmap])
            else:
                # guess left on the device by the previous step's decimation
                from .qtensor import gather
                idx = dv.upload(np.asarray(prjmap, dtype=np.int64))
                v0 = dv.empty(neig * ndim0, 0).view(neig, ndim0)
                for e in range(neig):
                    gather(v0[e], dmrg.psi0[e], idx)
        if dtype and neig == 1 and PRECISION_MODE in ('correction', 'rq'):
            eig, vec, nmvp = _defect_correction(dmrg, solver, hvec_tier, hvec64, v0, ndim0,
                                                dtype)
            for env in (fL, fR):
                env.drop_low()
            graphs.pop(dtype, None)
        elif dtype:
            # low-precision sigma: converge to residual 1e-4 (what FP32 sigma
            # resolves with margin), then
            # restart Davidson from that solution with FP64 sigma (residual
            # correction): the reported eigenvalue is an FP64 Ritz value
            solver.crit_vec = max(dmrg.crit_vec, getattr(dmrg, 'crit_vec_lo', 1e-4))
            solver.crit_e = max(dmrg.crit_e, getattr(dmrg, 'crit_e_lo', 1e-7))
            solver.herm_tol = 1e-5
            eig, vec, nmvp = solver.solve_iter(v0)
        else:
            eig, vec, nmvp = solver.solve_iter(v0)
        if dtype and not (neig == 1 and PRECISION_MODE in ('correction', 'rq')):
            # the low-precision copies of this step's environments are not
            # needed again (the next step at this bond rebuilds them): drop
            # them so that only the FP64 environments stay resident
            graphs.pop(dtype, None)
            for env in (fL, fR):
                env.drop_low()
            fine = EigenSolver()
            fine.crit_e, fine.crit_vec = dmrg.crit_e, dmrg.crit_vec
            fine.ndim, fine.neig, fine.diag = ndim0, neig, solver.diag
            fine.HVec = hvec64
            eig, vec, nmvp64 = fine.solve_iter(dv.download(vec).reshape(neig, ndim0))
            nmvp += nmvp64
        t0 = lap('davidson', t0)
        # the sigma intermediates are not needed before the next step's
        # program: free them ahead of the decimation's own buffers
        prog._buffers.clear()
        eigs += list(eig)
        nmvps += nmvp
        dmrg.niters.append(solver.niter)
        if _device_decimation(dmrg):
            dvec = (vec, prjmap)
            continue
        vec_h = dv.download(vec)
        civec = np.zeros((neig, dmrg.dims[3]))
        civec[:, prjmap] = vec_h
        civecs.append(civec)
    t0 = time.time()
    if _device_decimation(dmrg):
        from .decim import DeviceDecimation
        dd = DeviceDecimation(dmrg, dvec[0], dvec[1], isite, status)
        if status == 'L':
            classes = dpt(dmrg.qnuml[isite], dmrg.qphys[isite])
        else:
            classes = dpt(dmrg.qphys[isite + 1], dmrg.qnumr[isite + 2])
        extra = None
        noise = float(getattr(dmrg, 'noise', 0.0))
        if noise >= 1e-10:
            # the operator term of pRDMQt on the dense two-site vectors
            extra = kernelQt.prdm_noise_device(dd.C, dd.neig, info, noise)
            t0 = lap('noise', t0)
        if dd.run(classes, dmrg.thresh, dmrg.Dmax, extra):
            site, srotR = dd.site_and_srotR(dmrg.qlst)
            t0 = lap('decimation', t0)
            if isite == actlst[-1]:
                dmrg.psi0 = dd.C.view(dd.neig, -1)
            else:
                nxt = dmrg.lmps[isite - 1] if status == 'R' else dmrg.rmps[isite + 2]
                dmrg.psi0 = dd.guess(nxt)
            lap('guess', t0)
            return nmvps, np.array(eigs), dd.sigs, dd.dwts, dd.qred, site, srotR
        # an all-zero density matrix: the host path handles it
        civecs = [dv.download(dd.C).reshape(dd.neig, -1)]
    civecs = np.vstack(civecs)
    sigs, dwts, qred, site, srotR = decimation(dmrg, civecs, isite, status, info)
    t0 = lap('decimation', t0)
    genGuess(dmrg, isite, actlst, civecs, srotR, status)
    lap('guess', t0)
    return nmvps, np.array(eigs), sigs, dwts, qred, site, srotR


def _convert(dst, dst_code, src, src_code):
    """dst = src between storage types (0 FP64, 1 FP32, 2 BF16) on the
    device (zmpo_convert; BF16 rounds to nearest even)."""
    n = src.numel()
    _lib.check(_lib.lib().zmpo_convert(dst_code, dv.ptr(dst), src_code, dv.ptr(src), n,
                                       dv.stream()))
    return dst


# Decimation and guess on the device (decim.py) unless ZMPO_DECIM=host;
# several symmetry sectors take the host path.
DECIM = os.environ.get('ZMPO_DECIM', 'device')


def _device_decimation(dmrg):
    return (DECIM == 'device' and len(dmrg.qsectors) == 1
            and dv.device().type == 'cuda')


# Precision form (dmrg.dtype 1 = FP32, 2 = BF16 sigma operands), all in
# _defect_correction except 'restart':
# 'rq' (default): low-precision Davidson phases (the reference's iteration
# budget), then one FP64 sigma whose Rayleigh quotient is the reported
# eigenvalue; 'correction': the same phases, then outer FP64 corrections
# around anchored low-precision inner solves until the FP64 residual meets
# the Davidson criteria; 'restart': a low-precision Davidson to residual
# 1e-4 followed by a full FP64 Davidson from its vector.
PRECISION_MODE = os.environ.get('ZMPO_PRECISION_MODE', 'rq')
# inner solves stop when their residual falls below this fraction of the
# current FP64 residual, or after this many iterations
INNER_FRAC = 0.05
INNER_MAXCYCLE = 12
OUTER_MAX = 8
# phase tiers per run dtype (BF16 runs continue in FP32 before the FP64
# corrections) and the residual each tier's rounding still resolves, per
# unit of |theta|: measured sigma errors relative to |H psi| are 1.7e-3
# (BF16 operands) and 2.2e-7 (FP32) at K = 36, D = 1000
TIERS = {1: (1,), 2: (2, 1)}
TIER_TOL = {2: 4e-3, 1: 1e-6}
# iterations of a BF16 phase (the last tier gets the solver's own budget)
TIER_MAXCYCLE = {2: 12}
# tiers whose phases also stop on the eigenvalue change (ZMPO_TIER_ECONV=0
# turns it off: residual criterion only)
TIER_ECONV = {1: os.environ.get('ZMPO_TIER_ECONV', '1') != '0', 2: False}
TIER_HERM = {2: 2e-2, 1: 1e-5}


def _defect_correction(dmrg, solver, hvec_tier, hvec64, v0, n, dtype):
    """Lowest eigenpair of H_eff with FP64 accuracy from mostly low-precision
    sigma applications (one state).

    Phases: for each low-precision tier of the run (BF16 runs: BF16, then
    FP32; FP32 runs: FP32), one Davidson with that tier's sigma -- one
    search space, from the reference's start vector or the previous tier's
    vector -- until its residual reaches what the tier's rounding resolves
    (TIER_TOL x max(1, |theta|)) or the iteration budget ends (the last
    tier has the solver's own, as the reference's Davidson).  Mode 'rq'
    then returns the FP64 Rayleigh quotient theta = x.Hx of the phases'
    vector (one FP64 sigma; its error is of the order of |r|^2 / gap for
    the FP64 residual r).  Mode 'correction' goes on with outer FP64
    iterations from that vector: y = H x_k with FP64 sigma, theta_k = x_k.y (x_k normalised),
    r_k = y - theta_k x_k.  Stop when |r_k| <= crit_vec or |theta_k -
    theta_(k-1)| <= crit_e (the Davidson criteria of pdvdson.py:161-188, on
    FP64 quantities).  Otherwise an inner Davidson with the most precise
    low-precision sigma starts from x_k, whose first basis vector keeps its
    exact sigma y (pdvdson.EigenSolver.solve_iter(v0, w0)), and returns
    x_(k+1).  The operator error of the inner solve enters only through the
    correction x_(k+1) - x_k, so each outer step shrinks the error by about
    |H_lo - H| / gap; the reported eigenvalue is the FP64 Rayleigh quotient
    of the returned vector.  When an outer step fails to halve the FP64
    residual (the low-precision operator is too far from H for this gap),
    the solve continues as a plain FP64 Davidson from the current vector.
    Returns (eigs, vec, nmvp)."""
    from .pdvdson import _VecOps
    ops = _VecOps(n)
    if v0 is None:
        d = dv.download(solver.diag) + 1e-12 * np.arange(1, 1 + n) / float(n)
        e0 = np.zeros(n)
        e0[int(np.argsort(d, kind='stable')[0])] = 1.0
        x = dv.upload(e0, 0)
    elif isinstance(v0, np.ndarray):
        x = dv.upload(np.asarray(v0).reshape(-1), 0)
    else:
        x = ops.copy(dv.empty(n, 0), v0.reshape(-1).contiguous())
    ops.scale(x, 1.0 / ops.norm(x))
    y = dv.empty(n, 0)
    r = dv.empty(n, 0)
    nmvp = 0
    niter = 0
    tiers = TIERS[dtype]
    tm = dmrg.__dict__.setdefault('timing', {})

    def lap(key, t0):
        dv.sync()
        tm[key] = tm.get(key, 0.0) + time.time() - t0
    # ---- low-precision phases
    for tier in tiers:
        t0 = time.time()
        ph = EigenSolver()
        ph.maxcycle = TIER_MAXCYCLE.get(tier, solver.maxcycle) if tier != tiers[-1] \
            else solver.maxcycle
        # FP32 phases also stop on the eigenvalue change (crit_e), as the
        # FP64 Davidson does: the FP32 rounding moves theta = x.H'x by about
        # |H'x - Hx| / sqrt(n), below crit_e for these sizes; BF16 phases
        # stop on the residual only
        ph.crit_e = dmrg.crit_e if TIER_ECONV.get(tier) else -1.0
        ph.ndim, ph.neig, ph.diag = n, 1, solver.diag
        ph.herm_tol = TIER_HERM[tier]
        ph.HVec = hvec_tier(tier)
        # the tolerance scales with |theta| of the current vector, known
        # after the first projection: a first estimate from the diagonal
        ph.crit_vec = max(dmrg.crit_vec, TIER_TOL[tier])
        ph.tol_scale = TIER_TOL[tier]
        try:
            eigs, vec, nm = ph.solve_iter(x.view(1, -1))
        except DavidsonError:
            # the search space ran out of independent directions at this
            # tier's rounding level: go on from the current vector
            continue
        nmvp += nm
        niter += ph.niter
        ops.copy(x, vec[0])
        ops.scale(x, 1.0 / ops.norm(x))
        lap('tier%d' % tier, t0)
    # only the last tier runs again (inner solves, 'correction' mode):
    # release the others' intermediates before the FP64 ones are allocated
    for tier in (tiers if PRECISION_MODE == 'rq' else tiers[:-1]):
        hvec_tier(None)(tier)
    hvec_lo = hvec_tier(tiers[-1])
    herm = TIER_HERM[tiers[-1]]
    theta_prev = rn_prev = None
    if PRECISION_MODE == 'rq':
        # the reported eigenvalue: FP64 Rayleigh quotient of the phases' vector
        t0 = time.time()
        hvec64(x, y)
        nmvp += 1
        theta = float(ops.dots(x.view(1, -1), 1, y)[0])
        ops.copy(r, y)
        ops.axpy(r, -theta, x)
        dmrg.last_residual64 = ops.norm(r)
        lap('rq64', t0)
        solver.niter = niter
        return np.array([theta]), x.view(1, -1), nmvp

    def fallback():
        fine = EigenSolver()
        fine.maxcycle = getattr(dmrg, 'maxcycle', fine.maxcycle)
        fine.crit_e, fine.crit_vec = dmrg.crit_e, dmrg.crit_vec
        fine.ndim, fine.neig, fine.diag = n, 1, solver.diag
        fine.HVec = hvec64
        eigs, vec, nm = fine.solve_iter(x.view(1, -1), w0=y.view(1, -1))
        solver.niter = niter + fine.niter
        dmrg.nfallback = getattr(dmrg, 'nfallback', 0) + 1
        return eigs, vec, nmvp + nm
    # ---- FP64 outer corrections
    for k in range(OUTER_MAX):
        hvec64(x, y)
        nmvp += 1
        theta = float(ops.dots(x.view(1, -1), 1, y)[0])
        ops.copy(r, y)
        ops.axpy(r, -theta, x)
        rn = ops.norm(r)
        if rn <= dmrg.crit_vec or (theta_prev is not None
                                   and abs(theta - theta_prev) <= dmrg.crit_e):
            break
        if rn_prev is not None and rn > 0.5 * rn_prev:
            return fallback()
        theta_prev, rn_prev = theta, rn
        inner = EigenSolver()
        inner.crit_vec = max(dmrg.crit_vec, INNER_FRAC * rn)
        inner.crit_e = -1.0                   # residual criterion only
        inner.maxcycle = INNER_MAXCYCLE
        inner.herm_tol = herm
        inner.ndim, inner.neig, inner.diag = n, 1, solver.diag
        inner.HVec = hvec_lo
        try:
            _, vec, nm = inner.solve_iter(x.view(1, -1), w0=y.view(1, -1))
        except DavidsonError:
            # the low-precision search space ran out of independent
            # directions (small H_eff, or rounding at the level of the
            # residual): continue in FP64 from x_k
            return fallback()
        nmvp += nm
        niter += inner.niter
        ops.copy(x, vec[0])
        ops.scale(x, 1.0 / ops.norm(x))
    solver.niter = niter
    dmrg.nouter = getattr(dmrg, 'nouter', 0) + k + 1
    return np.array([theta]), x.view(1, -1), nmvp


def canonical_null(u, s2, tol):
    """Fixed basis for the RDM null space of one sector: eigenvectors with
    weight <= tol are replaced by unit vectors e_0, e_1, ... orthogonalised
    against the kept vectors and each other (twice), accepted when their
    projected norm exceeds 0.5 (second pass 1e-6)."""
    keep = s2 > tol
    nnull = int(np.sum(~keep))
    if nnull == 0:
        return u
    kept = u[:, keep]
    out = []
    n = u.shape[0]
    for cut in (0.5, 1e-6):
        for j in range(n):
            if len(out) == nnull:
                break
            v = np.zeros(n)
            v[j] = 1.0
            for _ in range(2):
                v = v - kept @ (kept.T @ v)
                if out:
                    ob = np.array(out).T
                    v = v - ob @ (ob.T @ v)
            nv = np.linalg.norm(v)
            if nv > cut:
                out.append(v / nv)
    u2 = u.copy()
    u2[:, ~keep] = np.array(out).T
    return u2


def rdm_blkdiag(mats, classes, thresh, Dcut, side, extra=None):
    """Per-sector RDM eigh and truncation (mpo_dmrg_qparser.py:101-180).

    ``mats`` are the C matrices (rows = kept side when side == 'L', columns
    when side == 'R'); the RDM of each sector is formed from its own rows only.
    ``extra`` (dense, kept side x kept side) is added to every sector's RDM:
    the noise term (dotutil.py:312-333), block diagonal in the sectors.
    """
    keys, groups = [], {}
    for idx, val in enumerate(classes):
        k = (float(val[0]), float(val[1]))
        if k not in groups:
            groups[k] = []
            keys.append(k)
        groups[k].append(idx)
    nrows = len(classes)
    tr = 0.0
    blocks = []
    for k in keys:
        rows = np.array(groups[k])
        if side == 'L':
            sub = [m[rows] for m in mats]
            rho = sum(s @ s.T for s in sub)
        else:
            sub = [m[:, rows] for m in mats]
            rho = sum(s.T @ s for s in sub)
        if extra is not None:
            rho = rho + extra[np.ix_(rows, rows)]
        tr += np.trace(rho)
        blocks.append((k, rows, rho))
    if tr < -1e-12:
        raise RuntimeError('RDM trace negative')
    rot_cols, sigs, qsym, sectors = [], [], [], []
    for b, (k, rows, rho) in enumerate(blocks):
        if tr < 1e-12:
            rho = rho + np.identity(len(rows)) / float(nrows)
        else:
            rho = rho / tr
        s2, u = np.linalg.eigh(rho)
        s2 = np.abs(s2)
        sectors.append([rows, u, s2, False])
        for c in range(len(rows)):
            rot_cols.append((b, c))
        sigs.append(s2)
        qsym += [k] * len(rows)
    sigs = np.hstack(sigs)
    sigs = sigs / np.sum(sigs)
    indx = truncation_order(sigs, qsym)
    nres = len(sigs)
    for i in range(nres):
        if sigs[indx[i]] < thresh * 1.01:
            nres = i
            break
    nres = max(nres, 1)
    if Dcut > 0:
        nres = min(nres, Dcut)
    indx = indx[:nres]
    re = sorted(range(nres), key=lambda x: qsym[indx[x]])
    indx = indx[re]
    # the null-space basis of a sector (canonical_null) only matters when
    # one of its zero-weight vectors is kept: build it for those sectors only
    for i in indx:
        b, c = rot_cols[i]
        sec = sectors[b]
        if not sec[3] and sec[2][c] <= WZERO:
            sec[1] = canonical_null(sec[1], sec[2], WZERO)
            sec[3] = True
    rot = np.zeros((nrows, nres))
    for c, i in enumerate(indx):
        b, col = rot_cols[i]
        rot[sectors[b][0], c] = sectors[b][1][:, col]
    sig_kept = sigs[indx]
    return 1.0 - np.sum(sig_kept), np.array([qsym[i] for i in indx]), rot, sig_kept


def decimation(dmrg, civecs, isite, status, info=None):
    """dotutil.py:288-495 for ncsite=2.  With dmrg.noise >= 1e-10 (and the
    solver's info) the RDM gets noise times the operator term of pRDMQt."""
    ldim, cdim, rdim, _ = dmrg.dims
    neig = civecs.shape[0]
    mats = civecs.reshape(neig, ldim * 4, 4 * rdim)
    extra = None
    noise = float(getattr(dmrg, 'noise', 0.0))
    if noise >= 1e-10 and info is not None:
        shp = (neig, ldim, 4, 4 * rdim) if status == 'L' else (neig, ldim * 4, 4, rdim)
        extra = noise * kernelQt.prdm_noise(civecs.reshape(shp), info)
    if status == 'L':
        classes = dpt(dmrg.qnuml[isite], dmrg.qphys[isite])
        dwts, qred, rot, sigs = rdm_blkdiag(mats, classes, dmrg.thresh, dmrg.Dmax, 'L', extra)
        nres = len(qred)
        site = rot.reshape(ldim, 4, nres).copy()
        srotR = np.array([_sparse_matmul(rot.T, m).reshape(nres, 4, rdim) for m in mats])
    else:
        classes = dpt(dmrg.qphys[isite + 1], dmrg.qnumr[isite + 2])
        dwts, qred, rot, sigs = rdm_blkdiag(mats, classes, dmrg.thresh, dmrg.Dmax, 'R', extra)
        nres = len(qred)
        site = rot.reshape(4, rdim, nres).transpose(2, 0, 1).copy()
        srotR = np.array([_sparse_matmul(m, rot).reshape(ldim, 4, nres) for m in mats])
    return sigs, dwts, qred, site, srotR


def genGuess(dmrg, isite, actlst, civecs, srotR, status):
    """dotutil.py:498-559 (ncsite=2) with block-sparse host products."""
    if isite == actlst[-1]:
        dmrg.psi0 = civecs.copy()
        return
    out = []
    for ie in range(civecs.shape[0]):
        if status == 'R':
            t = dmrg.lmps[isite - 1].toDenseTensor()                     # (l', 4, l)
            a = t.reshape(-1, t.shape[2])
            b = srotR[ie].reshape(t.shape[2], -1)
        else:
            t = dmrg.rmps[isite + 2].toDenseTensor()                     # (r, 4, r')
            a = srotR[ie].reshape(-1, t.shape[0])
            b = t.reshape(t.shape[0], -1)
        out.append(_sparse_matmul(a, b).reshape(-1))
    dmrg.psi0 = np.array(out)


def _sparse_matmul(a, b):
    """a @ b skipping all-zero rows/columns of the block-sparse operands."""
    rows = np.flatnonzero(np.any(a != 0, axis=1))
    mid = np.flatnonzero(np.any(a != 0, axis=0) & np.any(b != 0, axis=1))
    cols = np.flatnonzero(np.any(b != 0, axis=0))
    out = np.zeros((a.shape[0], b.shape[1]))
    if len(rows) and len(mid) and len(cols):
        out[np.ix_(rows, cols)] = a[np.ix_(rows, mid)] @ b[np.ix_(mid, cols)]
    return out
