"""Command-line DMRG run on the seeded integrals (one instance, one GPU).

    python -m zmpo_dmrg_amd.run --K 36 --D 1000 [--schedule normal|fixed]
        [--sweeps 9] [--dtype 0|1|2] [--tol 1e-7] [--seed 0] [--save path.npz]
        [--load path.npz]

Prints one line per two-site step and per sweep (energy, seconds, sigma
count, resident device memory) and, at the end, one JSON line with the
per-sweep energies and times.  --load starts from a checkpoint written by
--save (the MPS of the last sweep, all sites left-canonical), so runs in
different precisions can continue from the same state.
"""
import argparse
import json
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--K', type=int, default=24)
    ap.add_argument('--D', type=int, default=500)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--dtype', type=int, default=0)
    ap.add_argument('--schedule', choices=['normal', 'fixed'], default='normal')
    ap.add_argument('--sweeps', type=int, default=20)
    ap.add_argument('--tol', type=float, default=1e-7)
    ap.add_argument('--noise', type=float, default=0.0, help='noise of the fixed schedule')
    ap.add_argument('--budget-gb', type=float, default=64.0,
                    help='device bytes of the sigma intermediates per op chunk')
    ap.add_argument('--env-budget-gb', type=float, default=None,
                    help='device bytes the environments may hold; the rest wait in host '
                         'memory (default: all resident)')
    ap.add_argument('--host-budget-gb', type=float, default=200.0,
                    help='host bytes the spilled environments may hold')
    ap.add_argument('--save', default=None)
    ap.add_argument('--load', default=None)
    ap.add_argument('--start-m', type=int, default=None,
                    help='first bond dimension of the Normal2 ramp (default: the reference rule)')
    ap.add_argument('--compare', type=int, nargs='*', default=[],
                    help='after the schedule, one more full sweep (tolerance --tol, no noise) '
                         'in FP64 and then, from the same state, in each of these dtypes')
    args = ap.parse_args(argv)
    from . import device as dv
    from . import dmrg as pd
    from . import mpsio
    from .integrals import gen_spatial, spin_orbital
    from .schedule import Schedule
    dv.require()
    torch = dv.torch()
    K = args.K
    h, eri = gen_spatial(K, seed=args.seed)
    t0 = time.time()
    gb = lambda x: None if x is None else int(x * 2 ** 30)
    d = pd.MPODMRG(*spin_orbital(h, eri), K, 0.0, Dmax=args.D, tol=args.tol, dtype=args.dtype,
                   budget=gb(args.budget_gb), env_budget=gb(args.env_budget_gb),
                   host_budget=gb(args.host_budget_gb))
    if args.load:
        sites, qnums = mpsio.load(args.load)
        pd.restart_from_left_mps(d, sites, qnums)
    print('[run] K=%d D=%d dtype=%d set up in %.1fs' % (K, args.D, args.dtype, time.time() - t0),
          flush=True)
    sched = Schedule(maxM=args.D, tol=args.tol)
    sched.startM = args.start_m
    sched = sched.normal() if args.schedule == 'normal' else sched.fixed(noise=args.noise)

    def gib(peak=False):
        if not torch.cuda.is_available():
            return 0.0
        f = torch.cuda.max_memory_allocated if peak else torch.cuda.memory_allocated
        return f() / 2 ** 30

    def step_log(status, isite, eig, dt, nsig):
        ls = getattr(d, 'last_step', {})
        edev, ehost = pd.env_bytes(d)
        print('[run]   %s site %2d E=%.10f %.2fs sigma %d mem %.1f GiB (environments %.1f, '
              'host %.1f) ndim0 %s sigma %.1f GFLOP sectors %s dims %s' % (
                  status, isite, eig, dt, nsig, gib(), edev / 2 ** 30, ehost / 2 ** 30,
                  ls.get('ndim0'), ls.get('sigma_gflop', 0.0), ls.get('sectors'),
                  ls.get('dims')), flush=True)

    d.step_log = step_log
    out = {'energies': [], 'sweep_s': [], 'Dmax': [], 'nsigma': []}

    def sweep_log(isweep, eav, dt):
        out['energies'].append(eav)
        out['sweep_s'].append(dt)
        out['Dmax'].append(d.Dmax)
        out['nsigma'].append(d.nsigma)
        print('[run] sweep %d D=%d tol=%.0e noise=%.0e E=%.10f %.1fs sigma %d peak %.1f GiB' % (
            isweep, d.Dmax, d.crit_e, d.noise, eav, dt, d.nsigma, gib(True)), flush=True)
        print('[run]   timing so far: %s' % ' '.join('%s %.1f' % kv for kv in sorted(
            d.timing.items())), flush=True)
        from .kernelQt import SPILL
        if SPILL['download_bytes']:
            print('[run]   environment moves: %.1f GiB to host in %.1fs, %.1f GiB back in %.1fs' % (
                SPILL['download_bytes'] / 2 ** 30, SPILL['download_s'],
                SPILL['upload_bytes'] / 2 ** 30, SPILL['upload_s']), flush=True)
        if args.save:
            mpsio.save(args.save, {i: d.lmps[i] for i in range(K)},
                       [np.asarray(d.qnuml[i]) for i in range(K + 1)])

    pd.run(d, args.sweeps, schedule=sched, log=sweep_log)
    out.update({'K': K, 'D': args.D, 'dtype': args.dtype, 'schedule': args.schedule,
                'timing': dict(d.timing)})
    if args.compare:
        out['compare'] = compare_precisions(d, args, step_log)
    print(json.dumps(out), flush=True)
    return 0


def compare_precisions(d, args, step_log):
    """One full sweep (R then L, Davidson tolerance args.tol, no noise) from
    the current state in FP64, then the same sweep from the same state in
    each dtype of args.compare.  Per dtype: sweep seconds, the per-step
    eigenvalues and their difference from the FP64 sweep (the first step
    of each run solves the same H_eff; later steps follow each run's own
    truncations), sigma counts and the phase timing."""
    import os
    import tempfile
    from . import dmrg as pd
    from . import mpsio
    K = d.nsite
    path = os.path.join(tempfile.gettempdir(), 'zmpo_compare_%d.npz' % os.getpid())
    mpsio.save(path, {i: d.lmps[i] for i in range(K)},
               [np.asarray(d.qnuml[i]) for i in range(K + 1)])
    res = {}
    base = None
    for dt in [0] + list(args.compare):
        sites, qnums = mpsio.load(path)
        pd.restart_from_left_mps(d, sites, qnums)
        d.dtype = dt
        d.Dmax, d.crit_e, d.noise = args.D, args.tol, 0.0
        d.timing = {'solve': 0.0, 'renorm': 0.0}
        n0, n640 = d.nsigma, getattr(d, 'nsigma64', 0)
        t0 = time.time()
        r1 = pd.sweep(d, 'R')
        t1 = time.time()
        r2 = pd.sweep(d, 'L')
        dt_s = time.time() - t0
        steps = [float(e[0]) for e in r1[3]] + [float(e[0]) for e in r2[3]]
        ent = {'sweep_s': dt_s, 'half_s': [t1 - t0, dt_s - (t1 - t0)],
               'energy': min(r1[2], r2[2]), 'steps': steps,
               'nsigma_low': d.nsigma - n0 if dt else 0,
               'nsigma_f64': (getattr(d, 'nsigma64', 0) - n640) if dt else d.nsigma - n0,
               'timing': dict(d.timing)}
        if base is None:
            base = ent
        else:
            ent['dE_first_step'] = abs(steps[0] - base['steps'][0])
            ent['dE_sweep'] = abs(ent['energy'] - base['energy'])
            ent['speedup'] = base['sweep_s'] / dt_s
            ent['speedup_half'] = [a / b for a, b in zip(base['half_s'], ent['half_s'])]
        res[dt] = ent
        print('[run] compare dtype %d: sweep %.1fs (R %.1fs, L %.1fs) E=%.10f %s' % (
            dt, dt_s, ent['half_s'][0], ent['half_s'][1], ent['energy'],
            ' '.join('%s=%.3g' % (k, v) for k, v in ent.items()
                     if k in ('dE_first_step', 'dE_sweep', 'speedup', 'nsigma_low',
                              'nsigma_f64'))), flush=True)
        print('[run]   timing: %s' % ' '.join('%s %.1f' % kv for kv in sorted(d.timing.items())),
              flush=True)
    os.remove(path)
    return res


if __name__ == '__main__':
    sys.exit(main())
