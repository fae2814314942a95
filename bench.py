"""Benchmark: two-site Davidson sigma (H psi = L W W R psi) at the middle bond.

One "step" = one full sigma-vector application over all 2K operators at the
middle two-site block of a K-orbital, bond-dimension-D block-sparse MPS
(synthetic data: seeded integrals of the stated generator, a synthetic MPS
with DMRG-like (N, Sz) sector structure; environments are built on the device
by the renormalisation kernels before timing).  Prints one JSON line.

value = algorithmic sigma GFLOP/s (reference contraction order, block-pair
FLOPs counted from the integer plans), inputs resident in HBM.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {0: 78.6e12, 1: 157.3e12}
HBM = 8.0e12


def log(msg):
    print('[bench] ' + msg, file=sys.stderr, flush=True)


def build_workload(K, D, seed, dtype, budget, fused=True, extra=None, structure='dmrg'):
    """Middle-bond sigma workload: (mpo, isite, st, L, R, prog)
    (zmpo_dmrg_amd.synthetic.middle_bond)."""
    from zmpo_dmrg_amd import synthetic
    return synthetic.middle_bond(K, D, seed, budget, fused=fused, extra=extra,
                                 structure=structure, log=log)


def time_stages(prog, mpo, L, R, psi, sig, dtype, reps=3):
    """Per-launch average time of each sigma kernel, with HIP events recorded
    on the stream the kernels are launched on, plus its algorithmic FLOPs and
    bytes.  Launches (per op chunk): L*psi, then either the chained
    W[i]*W[i+1] launch or its two separate stages, then *R into sigma.
    dtype 2: BF16 operands, FP32 sigma (the *R launch runs as ZMPO_BF16_F32)."""
    from zmpo_dmrg_amd import device as dv
    from zmpo_dmrg_amd.kernelQt import out_dtype
    from zmpo_dmrg_amd.qtensor import run_chain, run_program
    torch = dv.torch()
    fused = prog.fused
    odt = out_dtype(dtype)
    t1, t2, t3 = prog.buffers(dtype, mid=not fused)
    Lb, Rb = L.buffer(dtype), R.buffer(dtype)
    Wi, Wj = mpo.buffer(prog.isite, dtype), mpo.buffer(prog.isite + 1, dtype)
    if fused:
        Wpi, Wpj = mpo.padded_buffer(prog.isite, dtype), mpo.padded_buffer(prog.isite + 1, dtype)
    elsize = {0: 8, 1: 4, 2: 2}[dtype]
    if fused:
        names = ['L*psi', 'W[i]*W[i+1] chained', '*R']
    else:
        names = ['L*psi', '*W[i]', '*W[i+1]', '*R']
    n = len(names)
    ms, flops, nbytes = np.zeros(n), np.zeros(n), np.zeros(n)
    stream = dv.stream_obj()
    for _ in range(reps):
        for st, ch in zip(prog.stages, prog.chains):
            if fused:
                ops = [(lambda: run_program(st[0], Lb, psi, t1, dtype, beta=0), st[0]),
                       (lambda: run_chain(ch, t1, Wpi, Wpj, t3, dtype), ch),
                       (lambda: run_program(st[3], t3, Rb, sig, odt, beta=0), st[3])]
            else:
                ops = [(lambda: run_program(st[0], Lb, psi, t1, dtype, beta=0), st[0]),
                       (lambda: run_program(st[1], t1, Wi, t2, dtype, beta=0), st[1]),
                       (lambda: run_program(st[2], t2, Wj, t3, dtype, beta=0), st[2]),
                       (lambda: run_program(st[3], t3, Rb, sig, odt, beta=0), st[3])]
            for k, (fn, pg) in enumerate(ops):
                ev0 = torch.cuda.Event(enable_timing=True)
                ev1 = torch.cuda.Event(enable_timing=True)
                ev0.record(stream)
                fn()
                ev1.record(stream)
                ev1.synchronize()
                ms[k] += ev0.elapsed_time(ev1)
                flops[k] += pg.flops
                nbytes[k] += pg.alg_bytes(elsize)
    return names, ms / reps, flops / reps, nbytes / reps


def traffic_per_launch(args, launch):
    """Measured HBM bytes per launch of the dominant stage kernel, from the
    calibrated rocprofv3 FETCH_SIZE / WRITE_SIZE summary kept in profiles/
    for this workload (None when no such measurement exists)."""
    path = os.path.join(ROOT, 'profiles', 'traffic_sigma_k%d_d%d_%s.json' % (args.K, args.D, args.dtype))
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    rec = d.get('launches', {}).get(launch)
    if not rec:
        return None
    return rec.get('bytes_per_launch')


def step_window(mpo, isite, L, R, extra, D, dtype, budget, nsteps=2, tol=1e-8, maxcycle=None):
    """nsteps full two-site steps ('L' direction from the middle bond of the
    synthetic state): Davidson solve (the first from the lowest diagonal
    element, the next from the reference's guess), decimation to D, site
    update and renormalisation, as a sweep does them.  Returns per-step
    seconds and eigenvalues, sigma counts and the phase timing."""
    from zmpo_dmrg_amd import device as dv
    from zmpo_dmrg_amd import dmrg as pd
    from zmpo_dmrg_amd import kernelQt as kq
    from zmpo_dmrg_amd import mpo as pmpo
    from zmpo_dmrg_amd.dot import ci_solver
    K = mpo.nsite
    d = pd.MPODMRG.__new__(pd.MPODMRG)
    d.fhop, d.nsite, d.sbas, d.nops = mpo, K, mpo.sbas, mpo.nops
    d.opers, d.hpwts = mpo.opers, mpo.hpwts
    d.qphys = [pmpo.qphys() for _ in range(K)]
    d.qsectors = {(float(K), 0.0): 1}
    d.Dmax, d.crit_e, d.crit_vec, d.thresh, d.noise = D, tol, tol, -1e-14, 0.0
    d.dtype, d.budget, d.psi0, d.keep_envs = dtype, budget, None, True
    d.nsigma, d.sigma_flops, d.niters = 0, 0, []
    if maxcycle:
        d.maxcycle = maxcycle
    d.timing = {'solve': 0.0, 'renorm': 0.0}
    d.qnuml, d.qnumr = dict(extra['ql']), dict(extra['qr'])
    d.lmps, d.rmps = dict(extra['ls']), dict(extra['rs'])
    d.envL = {isite - 1: L}
    d.envR = {isite + 2: R, isite + 3: extra['Rprev']}
    actlst = list(range(isite, isite + nsteps + 1))
    times, eigs = [], []
    for n in range(nsteps):
        i = isite + n
        flst = [[d.envL[i - 1], d.envR[i + 2]]]
        dv.sync()
        t = time.time()
        _, e, _, _, qred, site, _ = ci_solver(d, i, 2, actlst, flst, 'L')
        qsite = pd.sweep_updateSite(d, i, qred, site, 'L')
        flstN = [[None]]
        kq.renormQt_H00(d, i, 2, flst, flstN, qsite, 'L')
        d.envL[i] = flstN[0][0]
        dv.sync()
        times.append(time.time() - t)
        eigs.append(float(e[0]))
        log('window dtype %d step %d: E=%.10f %.2fs' % (dtype, i, eigs[-1], times[-1]))
    tiers = {str(k): v for k, v in getattr(d, 'nsigma_tier', {}).items()}
    if dtype:
        tiers[str(dtype)] = d.nsigma - sum(getattr(d, 'nsigma_tier', {}).values())
    return {'step_s': times, 'eigs': eigs, 'nsigma_low': d.nsigma if dtype else 0,
            'nsigma_f64': getattr(d, 'nsigma64', 0) if dtype else d.nsigma,
            'nsigma_by_dtype': tiers, 'fallbacks': getattr(d, 'nfallback', 0),
            'timing': {k: round(v, 3) for k, v in d.timing.items()}}


def to_oracle(q):
    """An oracle QTensor with q's tables and values (host copy; values from
    the device layout into the reference's row-major blocks)."""
    from oracle.qtensor import QTensor as OQ
    o = OQ(q.status)
    o.fromQsyms(q.rank, q.qsyms, q.ndims, ifallocate=False, idlst=q.idlst)
    o.value = np.zeros(o.size_allowed)
    v = q._host_value()
    inv = np.argsort(q.perm)
    for j in range(q.nblks_allowed):
        dm = q.blk_dims[j]
        blk = v[q.blk_off[j]:q.blk_off[j] + q.blk_size[j]].reshape(dm[list(q.perm)]).transpose(inv)
        o.value[o.ioff_allowed[q.idx_allowed[j]]:o.ioff_allowed[q.idx_allowed[j]] + blk.size] = blk.reshape(-1)
    return o


def oracle_state(mpo, sites, sample):
    """The oracle's operator state restricted to the operators in sample
    (weights 1) with the MPO site tensors of ``sites``."""
    class S:
        pass
    s = S()
    s.nops = len(sample)
    s.opers = [mpo.opers[p] for p in sample]
    s.hpwts = np.ones(len(sample))
    s.mpo = {i: [to_oracle(mpo.sites[i][p]) for p in sample] for i in sites}
    return s


def oracle_sigma(mpo, isite, st, L, R, sample, psi_host):
    """Oracle HVecQt (numpy, the reference's per-block loop) summed over the
    operators in sample.  Returns (sigma, algorithmic flops, seconds)."""
    from oracle import kernels as ok
    from oracle.qtensor import QTensor as OQ
    s = oracle_state(mpo, (isite, isite + 1), sample)
    lenv = [to_oracle(L.op_tensor(p)) for p in sample]
    renv = [to_oracle(R.op_tensor(p)) for p in sample]
    qt = OQ(st.qtmp.status)
    qt.fromQsyms(st.qtmp.rank, st.qtmp.qsyms, st.qtmp.ndims)
    s.qtmp = qt
    s.qkey = st.qkey
    stats = {}
    t = time.time()
    sig = ok.HVecQt(psi_host, s, lenv, renv, isite, stats)
    dt = time.time() - t
    return sig, sum(stats.values()), dt


def cpu_baseline(mpo, isite, st, L, R, nops_sample, psi_host):
    """Oracle HVecQt (numpy, the reference's per-block loop) on a subset of ops."""
    sample = list(range(0, mpo.nops, max(1, mpo.nops // nops_sample)))[:nops_sample]
    _, fl, dt = oracle_sigma(mpo, isite, st, L, R, sample, psi_host)
    return fl, dt, sample


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--K', type=int, default=36)
    ap.add_argument('--D', type=int, default=1000)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--dtype', choices=['f64', 'f32'], default='f64')
    ap.add_argument('--budget-gb', type=float, default=96.0)
    ap.add_argument('--cpu-ops', type=int, default=2)
    ap.add_argument('--structure', choices=['dmrg', 'wide'], default='dmrg',
                    help='sector structure of the synthetic MPS')
    ap.add_argument('--window', type=int, default=2,
                    help='full two-site steps timed per precision after the sigma (0: none)')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-f32', action='store_true',
                    help='skip the FP32-operand sigma timed after the FP64 one')
    ap.add_argument('--unfused', action='store_true',
                    help='run sigma stages 2 and 3 as two launches through HBM')
    ap.add_argument('--kernel', choices=['stream', 'shaped', 'wave', 'tile'], default=None,
                    help='grouped-GEMM schedule (default: library default)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one MI355X; --gpus %d is not supported (the operator-sharded '
              'multi-GPU design is documented, not built)' % args.gpus, file=sys.stderr)
        sys.exit(2)
    from zmpo_dmrg_amd import device as dv
    from zmpo_dmrg_amd import qtensor as pq
    dv.require()
    if args.kernel:
        pq.KERNEL = args.kernel
    torch = dv.torch()
    dtype = 0 if args.dtype == 'f64' else 1
    extra = {}
    mpo, isite, st, L, R, prog = build_workload(args.K, args.D, args.seed, dtype,
                                                int(args.budget_gb * 2 ** 30),
                                                fused=not args.unfused, extra=extra,
                                                structure=args.structure)
    rng = np.random.default_rng(1)
    psi_h = rng.standard_normal(st.qtmp.size_allowed)
    psi_h /= np.linalg.norm(psi_h)
    psi = dv.upload(psi_h, dtype)
    sig = dv.empty(st.qtmp.size_allowed, dtype)
    for _ in range(args.warmup):
        prog.apply(mpo, L, R, psi, sig, dtype)
    dv.sync()
    log('warmup done')
    t0 = time.time()
    for _ in range(args.steps):
        prog.apply(mpo, L, R, psi, sig, dtype)
    dv.sync()
    dt = time.time() - t0
    ms_step = 1e3 * dt / args.steps
    gflops = prog.flops * args.steps / dt / 1e9
    log('timed %d sigma steps: %.3f ms/step, %.1f GFLOP/s' % (args.steps, ms_step, gflops))
    names, stage_ms, stage_fl, stage_by = time_stages(prog, mpo, L, R, psi, sig, dtype)
    dom = int(np.argmax(stage_ms))
    nlaunch = len(prog.stages)
    avg_ms = stage_ms[dom] / nlaunch
    ach_b = stage_by[dom] / nlaunch / (avg_ms * 1e-3)
    ach_f = stage_fl[dom] / nlaunch / (avg_ms * 1e-3)
    # the stage GEMMs have K ~ 18-60 per block pair: arithmetic intensity
    # (FLOP per algorithmic byte) is below the FP64 ridge point, so HBM bounds
    bound = 'hbm' if stage_fl[dom] / stage_by[dom] < PEAK[dtype] / HBM else 'mfma'
    kname = 'chained grouped GEMM (zmpo_chaingemm)' if (prog.fused and dom == 1) else \
        '%s grouped GEMM' % pq.KERNEL
    # kernel dispatches behind one launch of the dominant stage (the chained
    # launch issues one dispatch per LDS size class), so the per-dispatch
    # average can be set beside rocprofv3's kernel statistics
    ndisp = sum(len(ch.launches) for ch in prog.chains) if (prog.fused and dom == 1) else nlaunch
    # algorithmic bytes of one sigma by the reference-order rule: every
    # operator's L_p, R_p, W_p[i], W_p[i+1] read once, psi read and sigma
    # written once (intermediates excluded; they are reported beside it)
    env_el = sum(int(m.size_allowed) for m in L.meta) + sum(int(m.size_allowed) for m in R.meta)
    w_el = sum(int(mpo.sites[s_][p].size_allowed) for s_ in (isite, isite + 1)
               for p in range(mpo.nops))
    sigma_bytes = 8 * (env_el + w_el + 2 * int(st.qtmp.size_allowed))
    # top1 and top3 written once and read once per sigma (FP64 elements)
    inter_bytes = 8 * 2 * sum(
        int(pg.tasks_host['M'].astype(np.int64) @ pg.tasks_host['N'].astype(np.int64))
        for stg in prog.stages for pg in (stg[0], stg[2]) if pg.ntasks)
    roof = {'kernel': '%s, sigma launch %s' % (kname, names[dom]),
            'bound': bound,
            'achieved': ach_b / 1e9 if bound == 'hbm' else ach_f / 1e12,
            'peak': HBM / 1e9 if bound == 'hbm' else PEAK[dtype] / 1e12,
            'unit': 'GB/s' if bound == 'hbm' else 'TFLOP/s',
            'frac': ach_b / HBM if bound == 'hbm' else ach_f / PEAK[dtype],
            'traffic': traffic_per_launch(args, names[dom]),
            'alg_bytes_per_launch': stage_by[dom] / nlaunch,
            'flops_per_launch': stage_fl[dom] / nlaunch,
            'achieved_tflops': ach_f / 1e12, 'mfma_frac': ach_f / PEAK[dtype],
            'launch_ms': avg_ms, 'launches_per_step': nlaunch,
            'dispatches_per_step': ndisp, 'dispatch_avg_ms': stage_ms[dom] / ndisp,
            'sigma_alg_bytes': sigma_bytes,
            'sigma_hbm_floor_ms': sigma_bytes / HBM * 1e3,
            'intermediate_bytes': inter_bytes,
            'stage_names': names,
            'stage_ms': [round(x, 4) for x in stage_ms],
            'stage_gflop': [round(x / 1e9, 3) for x in stage_fl],
            'stage_alg_gb': [round(x / 1e9, 3) for x in stage_by]}
    # the same sigma in FP32 operands (the precision form's low-precision
    # phase), after the timed region: its time and its error against FP64
    prec = None
    if dtype == 0 and not args.no_f32:
        psi32 = dv.upload(psi_h, 1)
        sig32 = dv.empty(st.qtmp.size_allowed, 1)
        prog.apply(mpo, L, R, psi32, sig32, 1)
        dv.sync()
        t = time.time()
        for _ in range(args.steps):
            prog.apply(mpo, L, R, psi32, sig32, 1)
        dv.sync()
        ms32 = 1e3 * (time.time() - t) / args.steps
        prog.apply(mpo, L, R, psi, sig, 0)
        s64 = dv.download(sig)
        err = float(np.linalg.norm(dv.download(sig32).astype(np.float64) - s64) / np.linalg.norm(s64))
        _, st_ms32, _, _ = time_stages(prog, mpo, L, R, psi32, sig32, 1)
        prec = {'f32_ms_per_step': ms32, 'f32_gflops': prog.flops / ms32 / 1e6,
                'f32_speedup_vs_f64': ms_step / ms32, 'f32_rel_err': err,
                'f32_stage_ms': [round(x, 4) for x in st_ms32]}
        log('FP32 sigma: %.3f ms/step (%.2fx FP64), relative error %.2e, stages %s' % (
            ms32, ms_step / ms32, err, st_ms32))
        # BF16 operands, FP32 accumulation (BF16 MFMA), FP32 sigma
        psi16 = dv.upload(psi_h, 2)
        prog.apply(mpo, L, R, psi16, sig32, 2)
        dv.sync()
        t = time.time()
        for _ in range(args.steps):
            prog.apply(mpo, L, R, psi16, sig32, 2)
        dv.sync()
        ms16 = 1e3 * (time.time() - t) / args.steps
        err16 = float(np.linalg.norm(dv.download(sig32).astype(np.float64) - s64) / np.linalg.norm(s64))
        names16, st_ms16, _, _ = time_stages(prog, mpo, L, R, psi16, sig32, 2)
        prec.update({'bf16_ms_per_step': ms16, 'bf16_gflops': prog.flops / ms16 / 1e6,
                     'bf16_speedup_vs_f64': ms_step / ms16, 'bf16_rel_err': err16,
                     'bf16_stage_ms': [round(x, 4) for x in st_ms16]})
        log('BF16 sigma: %.3f ms/step (%.2fx FP64), relative error %.2e, stages %s' % (
            ms16, ms_step / ms16, err16, st_ms16))
    # full two-site steps (solve, decimation, renormalisation) in each
    # precision from the same state, after the timed region
    window = None
    if args.window > 0 and dtype == 0:
        prog._buffers.clear()
        window = {}
        for dt in (0, 1, 2):
            window[dt] = step_window(mpo, isite, L, R, extra, args.D, dt,
                                     int(args.budget_gb * 2 ** 30), args.window)
        # the first step's H_eff is the same in every run: its lowest
        # eigenvalue from an FP64 Davidson run to convergence (no iteration
        # cap) is the value every precision is measured against
        exact = step_window(mpo, isite, L, R, extra, args.D, 0, int(args.budget_gb * 2 ** 30),
                            1, maxcycle=400)['eigs'][0]
        base = window[0]
        for dt in (0, 1, 2):
            w = window[dt]
            w['dE_first_step'] = abs(w['eigs'][0] - exact)
            w['step_speedup'] = sum(base['step_s']) / sum(w['step_s'])
        nstep = 2 * (args.K - 1)
        # steps after the first start from the reference's guess, as sweep
        # steps do; the first starts from the lowest diagonal element
        warm = lambda w: w['step_s'][1:] if len(w['step_s']) > 1 else w['step_s']
        window = {'steps': args.window, 'tol': 1e-8, 'D': args.D, 'first_step_exact': exact,
                  'sweep_s_estimate': {str(dt): round(nstep * float(np.mean(warm(window[dt]))), 1)
                                       for dt in window},
                  'speedup_vs_f64': {str(dt): round(float(np.mean(warm(window[0])))
                                                    / float(np.mean(warm(window[dt]))), 3)
                                     for dt in window},
                  'how': 'mean of the window steps that start from the guess (middle bond, '
                         'the largest blocks) x 2(K-1) steps per sweep, an upper estimate; '
                         'dtype 0 FP64, 1 FP32 + FP64 correction, 2 BF16 then FP32 + FP64 '
                         'correction',
                  'runs': {str(k): v for k, v in window.items()}}
        log('window: %s' % json.dumps(window['sweep_s_estimate']))
    cpu = None
    if not args.no_cpu_baseline:
        try:
            fl, tcpu, sample = cpu_baseline(mpo, isite, st, L, R, args.cpu_ops, psi_h)
            cpu = {'value': fl / tcpu / 1e9, 'unit': 'GFLOP/s', 'cores': os.cpu_count(),
                   'kind': 'port',
                   'sample': 'oracle HVecQt (numpy FP64 per-block tensordot, plans rebuilt '
                             'per call as the reference) over %d of %d operators of the same '
                             'middle-bond sigma; %.1f GFLOP in %.1fs' % (
                                 len(sample), mpo.nops, fl / 1e9, tcpu)}
        except Exception as e:  # baseline is informative only
            log('cpu baseline failed: %r' % (e,))
    out = {
        'metric': 'sigma GFLOP/s (two-site Davidson H*psi, all 2K operators), K=%d D=%d' % (
            args.K, args.D),
        'value': gflops, 'unit': 'GFLOP/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': ms_step, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': args.dtype,
        'data': 'synthetic (seeded integrals; synthetic block-sparse MPS with the sector '
                'structure of a real K=36 D=1000 DMRG middle bond (%s), device-built envs)' % args.structure,
        'config': {'workload': 'middle-bond two-site sigma, K=%d N=%d Sz=0 D=%d, nops=%d' % (
            args.K, args.K, args.D, mpo.nops), 'ndim0': int(st.qtmp.size_allowed),
            'sigma_gflop': prog.flops / 1e9, 'chunks': len(prog.chunks),
            'launches': ' | '.join(names)},
        'roofline': roof, 'cpu_baseline': cpu, 'precision': prec, 'sweep_window': window,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
