"""Batching form: many independent seeded Hamiltonians on one GPU.

Instances (K from {12, 16, 20, 24}, N = K, Sz = 0, D from {128, 256, 384,
512}, integrals from the seeded generator) are independent DMRG problems.
At these sizes a two-site step is dominated by host work (plan construction,
Davidson bookkeeping, decimation), so the batch runs instances concurrently
in worker processes that share the GPU: each worker owns its instances'
device buffers and streams, host work proceeds in parallel across workers,
and their kernels interleave on the device.  Kernels of different instances
are not fused into one launch.

    python -m zmpo_dmrg_amd.batch [--instances 48] [--workers 8] [--sweeps 2]
                                  [--dtype 1] [--seed 0]

prints one JSON line: aggregate sweeps/s, p50/p99 per-instance sweep time
and per-instance energies.
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

KS = (12, 16, 20, 24)
DS = (128, 256, 384, 512)


def instances(n=48, seed=0):
    """Deterministic instance list: (index, K, D, integral seed)."""
    rng = np.random.default_rng(seed)
    ks = rng.choice(KS, size=n)
    ds = rng.choice(DS, size=n)
    return [(i, int(k), int(d), int(seed) * 1000 + i) for i, (k, d) in enumerate(zip(ks, ds))]


def run_instance(spec, nsweeps, dtype, budget_gb=6.0):
    """Sweeps of one instance; returns its energies and per-sweep wall times.
    ``budget_gb`` caps the sigma intermediates per op chunk, so that several
    workers' instances fit in device memory together."""
    from . import device as dv
    from . import dmrg as pdmrg
    from .integrals import gen_integrals
    i, K, D, iseed = spec
    t0 = time.time()
    i1, i2 = gen_integrals(K, seed=iseed)
    d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=D, tol=1e-8, dtype=dtype,
                      budget=int(budget_gb * 2 ** 30))
    dv.sync()
    setup = time.time() - t0
    print('[batch]   instance %d K=%d D=%d set up in %.1fs' % (i, K, D, setup), flush=True)
    energies, times = [], []
    for isw in range(nsweeps):
        t = time.time()
        r1 = pdmrg.sweep(d, 'R')
        print('[batch]   instance %d sweep %d R E=%.8f' % (i, isw, r1[2]), flush=True)
        r2 = pdmrg.sweep(d, 'L')
        dv.sync()
        times.append(time.time() - t)
        energies.append(min(r1[2], r2[2]))
        print('[batch]   instance %d sweep %d L E=%.8f' % (i, isw, r2[2]), flush=True)
    nsigma = d.nsigma
    # <Psi|H|Psi> of the final MPS (left environment sweep) beside the last
    # eigenvalue of the final 'L' half-sweep: equal for a consistent MPS
    from .grad import expectation
    e_mps = expectation(d.fhop, d.lmps) if nsweeps else None
    e_last = float(r2[3][-1][0]) if nsweeps else None
    # hand the cached blocks back: workers share one device's memory
    del d
    dv.torch().cuda.empty_cache()
    return {'index': i, 'K': K, 'D': D, 'seed': iseed, 'energies': energies,
            'sweep_s': times, 'setup_s': setup, 'nsigma': nsigma, 'pid': os.getpid(),
            'e_mps': e_mps, 'e_last': e_last}


def _worker(args):
    spec, nsweeps, dtype, budget_gb = args
    r = run_instance(spec, nsweeps, dtype, budget_gb)
    print('[batch] instance %d K=%d D=%d E=%.8f sweep %.2fs (pid %d)' % (
        r['index'], r['K'], r['D'], r['energies'][-1], float(np.mean(r['sweep_s'])), r['pid']),
        flush=True)
    return r


def summarize(results, wall, nsweeps):
    """Aggregate sweeps/s over the batch wall time and p50/p99 of per-instance
    mean sweep times."""
    per = np.array([float(np.mean(r['sweep_s'])) for r in results])
    return {'instances': len(results), 'sweeps_per_instance': nsweeps,
            'wall_s': wall, 'sweeps_per_s': len(results) * nsweeps / wall if wall > 0 else None,
            'p50_sweep_s': float(np.percentile(per, 50)) if len(per) else None,
            'p99_sweep_s': float(np.percentile(per, 99)) if len(per) else None}


THREAD_VARS = ('OMP_NUM_THREADS', 'OPENBLAS_NUM_THREADS', 'MKL_NUM_THREADS')


def run_batch(specs, workers, nsweeps, dtype, cores=16, budget_gb=6.0, spawn=False):
    """Run the instances on ``workers`` processes (spawned: each initialises
    its own HIP context) with cores // workers host threads each, so the
    workers' BLAS and torch thread pools do not oversubscribe the ``cores``
    host cores; returns (results sorted by index, summary).  workers = 1
    runs in this process unless spawn is set.  Host BLAS results (the
    decimation's eigh and products) can depend on the thread count, so runs
    compared bit for bit need the same threads per worker."""
    t = time.time()
    if workers <= 1 and not spawn:
        results = [_worker((s, nsweeps, dtype, budget_gb)) for s in specs]
    else:
        saved = {k: os.environ.get(k) for k in THREAD_VARS}
        for k in THREAD_VARS:   # read by the spawned interpreters at start-up
            os.environ[k] = str(max(1, cores // workers))
        try:
            ctx = mp.get_context('spawn')
            with ctx.Pool(workers) as pool:
                # largest instances first: shorter tail
                order = sorted(specs, key=lambda s: -(s[1] * s[2] * s[2]))
                results = list(pool.imap_unordered(
                    _worker, [(s, nsweeps, dtype, budget_gb) for s in order]))
        finally:
            for k, v in saved.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
    wall = time.time() - t
    results.sort(key=lambda r: r['index'])
    return results, summarize(results, wall, nsweeps)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--instances', type=int, default=48)
    ap.add_argument('--workers', type=int, default=8)
    ap.add_argument('--sweeps', type=int, default=2)
    ap.add_argument('--dtype', type=int, default=1, help='0 = FP64, 1 = FP32 sigma + FP64 correction')
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--cores', type=int, default=16, help='host cores shared by the workers')
    ap.add_argument('--budget-gb', type=float, default=6.0,
                    help='sigma intermediates per op chunk and instance (GiB)')
    args = ap.parse_args(argv)
    if args.workers > 16:
        print('at most 16 worker processes share one GPU', file=sys.stderr)
        return 2
    specs = instances(args.instances, args.seed)
    results, summ = run_batch(specs, args.workers, args.sweeps, args.dtype, args.cores,
                              args.budget_gb)
    summ.update({'workers': args.workers, 'dtype': args.dtype,
                 'energies': {r['index']: r['energies'][-1] for r in results}})
    print(json.dumps(summ))
    return 0


if __name__ == '__main__':
    sys.exit(main())
