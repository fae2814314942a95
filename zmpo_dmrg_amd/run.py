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
    ap.add_argument('--budget-gb', type=float, default=64.0)
    ap.add_argument('--save', default=None)
    ap.add_argument('--load', default=None)
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
    d = pd.MPODMRG(*spin_orbital(h, eri), K, 0.0, Dmax=args.D, tol=args.tol, dtype=args.dtype,
                   budget=int(args.budget_gb * 2 ** 30))
    if args.load:
        sites, qnums = mpsio.load(args.load)
        pd.restart_from_left_mps(d, sites, qnums)
    print('[run] K=%d D=%d dtype=%d set up in %.1fs' % (K, args.D, args.dtype, time.time() - t0),
          flush=True)
    sched = Schedule(maxM=args.D, tol=args.tol)
    sched = sched.normal() if args.schedule == 'normal' else sched.fixed(noise=args.noise)

    def gib(peak=False):
        if not torch.cuda.is_available():
            return 0.0
        f = torch.cuda.max_memory_allocated if peak else torch.cuda.memory_allocated
        return f() / 2 ** 30

    def step_log(status, isite, eig, dt, nsig):
        print('[run]   %s site %2d E=%.10f %.2fs sigma %d mem %.1f GiB' % (
            status, isite, eig, dt, nsig, gib()), flush=True)

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
        if args.save:
            mpsio.save(args.save, {i: d.lmps[i] for i in range(K)},
                       [np.asarray(d.qnuml[i]) for i in range(K + 1)])

    pd.run(d, args.sweeps, schedule=sched, log=sweep_log)
    out.update({'K': K, 'D': args.D, 'dtype': args.dtype, 'schedule': args.schedule,
                'timing': d.timing})
    print(json.dumps(out), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
