"""Measured tile choice for the grouped-GEMM launches of the two-site sigma.

The two outer launches of a sigma (L x psi into top1, top3 x R into sigma)
are grouped block GEMMs whose speed on MI355X depends on the tile in ways a
cost model does not capture: latency per unit, padding of the block edges,
register blocking, LDS per workgroup and occupancy all move with the block
sizes and the operand type.  A launch's *signature* bins what decides the
best tile: its role, the kernel dtype code, and half-octave bins of the
work-weighted median block rows M, columns N and summed K per output
block.  ``tune`` times every candidate (kernel, tile shape) of a program on
the device with HIP events on the launch stream -- each after one untimed
launch that builds its schedule -- and returns the fastest.

Modes (environment ZMPO_AUTOTUNE):
  'table' (default)  the choice comes from the table measured on MI355X
                     that ships with the package (tiles_mi355x.json, written
                     by ``python -m zmpo_dmrg_amd.autotune``): the nearest
                     signature of the same role and dtype; launches keep
                     their static choice when the table has none.  Lookups
                     are deterministic, so runs repeat bit for bit.
  'online'           as 'table', and a signature further than one bin from
                     every table entry is tuned on the device the first time
                     it is met (the result is kept for the process).
  'off'              static choices only.
"""
import json
import os
import sys
import time

import numpy as np

from . import device as dv

MODE = os.environ.get('ZMPO_AUTOTUNE', 'table')
TABLE_PATH = os.environ.get('ZMPO_TILE_TABLE',
                            os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                         'tiles_mi355x.json'))
REPS = 2
# host cost per schedule unit in ms per sigma application: building and
# uploading a unit costs ~40 ns (more on freshly faulted memory) and a
# program's schedule serves the ~20 sigma applications of one Davidson
# solve, so each unit adds ~2 ns to every application
HOST_MS_PER_UNIT = 2e-6

# candidates per role: (kernel, qtensor.SHAPES id)
CANDIDATES = {
    # L x psi: tall blocks (bra x MPO-bond rows), one pair per output block
    'sigma_l': [('shaped', 2), ('shaped', 3), ('shaped', 0), ('shaped', 1), ('shaped', 6),
                ('shaped', 7), ('shaped', 10), ('shaped', 11)],
    # top3 x R: square sigma blocks, hundreds of long pairs each (split-K)
    'sigma_r': [('stream', 2), ('shaped', 2), ('shaped', 3), ('shaped', 6), ('shaped', 8),
                ('shaped', 10), ('shaped', 11), ('shaped', 12)],
}
# the three launches of a renormalisation step (site x environment, MPO x
# that, x site into the new environment)
for _r in ('renorm1', 'renorm2', 'renorm3'):
    CANDIDATES[_r] = [('shaped', 0), ('shaped', 1), ('shaped', 2), ('shaped', 3), ('shaped', 4),
                      ('shaped', 5), ('shaped', 6), ('shaped', 7), ('shaped', 10), ('shaped', 11)]

_table = None
_online = {}


def _wmedian(x, w):
    """Weighted median of non-negative integer-valued x: the smallest value
    whose cumulative weight (values in ascending order) reaches half the
    total; by counting when the values are small, else by sorting."""
    if len(x) == 0 or w.sum() <= 0:
        return 1.0
    xi = x.astype(np.int64)
    if xi.max() <= 1 << 20:
        c = np.cumsum(np.bincount(xi, weights=w))
        return float(np.searchsorted(c, 0.5 * c[-1]))
    o = np.argsort(xi, kind='stable')
    c = np.cumsum(w[o])
    return float(xi[o][np.searchsorted(c, 0.5 * c[-1])])


def _bin(v):
    return round(2.0 * np.log2(max(float(v), 1.0))) / 2.0


def signature(prog, role, dtype):
    """(role, dtype, bins of M, N, summed K) of a grouped-GEMM program (the
    bins are computed once per program)."""
    bins = getattr(prog, '_sig_bins', None)
    if bins is None:
        bins = prog._sig_bins = _size_bins(prog)
    return (role, int(dtype)) + bins


def _size_bins(prog):
    t, p = prog.tasks_host, prog.pairs_host
    cnt = (t['p_end'] - t['p_begin']).astype(np.int64)
    rows = np.repeat(np.arange(prog.ntasks), cnt)
    M = t['M'][rows].astype(np.float64)
    N = t['N'][rows].astype(np.float64)
    K = p['K'].astype(np.float64)
    w = M * N * K
    ksum = np.bincount(rows, weights=K, minlength=prog.ntasks)
    wt = np.bincount(rows, weights=w, minlength=prog.ntasks)
    return (_bin(_wmedian(M, w)), _bin(_wmedian(N, w)), _bin(_wmedian(ksum, wt)))


def _load_table():
    global _table
    if _table is None:
        _table = []
        try:
            with open(TABLE_PATH) as f:
                for e in json.load(f).get('entries', []):
                    _table.append((tuple(e['signature']), (e['kernel'], int(e['shape']))))
        except (OSError, ValueError, KeyError):
            _table = []
    return _table


def lookup(sig, max_dist=None):
    """Choice of the nearest table (or tuned) signature of the same role and
    dtype: (kernel, shape) or None.  max_dist bounds the L1 distance in
    bins."""
    best, bd = None, None
    for s, choice in _load_table() + list(_online.items()):
        if s[0] != sig[0] or s[1] != sig[1]:
            continue
        d = sum(abs(a - b) for a, b in zip(s[2:], sig[2:]))
        if bd is None or d < bd:
            best, bd = choice, d
    if best is None or (max_dist is not None and bd > max_dist):
        return None
    return best


def tune(prog, role, dtype, launch, reps=REPS):
    """Time every candidate of ``role`` for prog (``launch()`` runs prog in
    kernel dtype code ``dtype`` with its current per-dtype choice) and leave
    the one with the lowest kernel time plus host schedule cost
    (HOST_MS_PER_UNIT per unit) set.  Returns ({(kernel, shape): ms}, best)."""
    torch = dv.torch()
    stream = torch.cuda.current_stream()
    times = {}
    score = {}
    for kernel, shape in CANDIDATES[role]:
        prog.tiles[dtype] = (kernel, shape)
        k_eff, s_eff = prog.launch_choice(dtype)
        if s_eff != shape:
            continue                  # the program cannot use this tile
        launch()                      # builds and uploads the schedule
        dv.sync()
        nunits = (prog._sched or {}).get((k_eff, shape), {}).get('nunits', 0)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(reps):
            launch()
        ev1.record(stream)
        ev1.synchronize()
        times[(kernel, shape)] = ev0.elapsed_time(ev1) / reps
        score[(kernel, shape)] = times[(kernel, shape)] + HOST_MS_PER_UNIT * nunits
        prog.__dict__.setdefault('_tune_units', {})[(kernel, shape)] = nunits
    best = min(score, key=score.get)
    prog.tiles[dtype] = best
    # keep only the chosen schedule on the device
    if prog._sched:
        keep = prog.launch_choice(dtype)
        for key in list(prog._sched):
            if key != keep:
                del prog._sched[key]
    return times, best


def choose(prog, role, dtype, launch):
    """Per-dtype tile of one launch by the current mode (sets prog.tiles)."""
    if MODE == 'off' or prog.ntasks == 0 or dtype in prog.tiles:
        return prog.tiles.get(dtype)
    sig = signature(prog, role, dtype)
    choice = lookup(sig, max_dist=1.0 if MODE == 'online' else None)
    if choice is None and MODE == 'online':
        _, choice = tune(prog, role, dtype, launch)
        _online[sig] = choice
    if choice is not None:
        prog.tiles[dtype] = choice
    return choice


def main():
    """Measure the tile table over middle-bond sigma workloads: for each
    (K, D) and each operand type, tune both outer launches of every op
    chunk and write the signatures with their fastest tiles."""
    import argparse
    from . import kernelQt as kq
    from . import synthetic
    ap = argparse.ArgumentParser()
    ap.add_argument('--sizes', default='36:1000,36:500,24:500,24:256,16:256,12:128',
                    help='comma-separated K:D workloads')
    ap.add_argument('--dtypes', default='0,1,2')
    ap.add_argument('--budget-gb', type=float, default=96.0)
    ap.add_argument('--roles', default='sigma,renorm')
    ap.add_argument('--out', default=TABLE_PATH)
    args = ap.parse_args()
    dv.require()
    entries = {}
    log = []
    for kd in args.sizes.split(','):
        K, D = (int(x) for x in kd.split(':'))
        extra = {}
        mpo, isite, st, L, R, prog = synthetic.middle_bond(
            K, D, 0, int(args.budget_gb * 2 ** 30), extra=extra,
            log=lambda m: print('[autotune] ' + m, file=sys.stderr, flush=True))
        n = st.qtmp.size_allowed
        psi_h = np.random.default_rng(1).standard_normal(n)
        for dt in (int(x) for x in args.dtypes.split(',')):
            psi = dv.upload(psi_h, dt)
            sig_out = dv.empty(n, kq.sigma_dtype(dt))
            t1, _, t3 = prog.buffers(dt)
            Lb, Rb = L.buffer(dt), R.buffer(dt)
            for st_ in prog.stages:
                for role, k, code, a, b, c in (('sigma_l', 0, dt, Lb, psi, t1),
                                               ('sigma_r', 3, kq.out_dtype(dt), t3, Rb, sig_out)):
                    pg = st_[k]
                    if pg.ntasks == 0:
                        continue
                    sig = signature(pg, role, code)
                    if sig in entries:
                        continue
                    from .qtensor import run_program
                    t = time.time()
                    times, best = tune(pg, role, code,
                                       lambda: run_program(pg, a, b, c, code, beta=0))
                    entries[sig] = best
                    rec = {'K': K, 'D': D, 'signature': list(sig), 'kernel': best[0],
                           'shape': best[1],
                           'ms': {'%s:%d' % kk: round(v, 3) for kk, v in times.items()},
                           'units': {'%s:%d' % kk: int(v) for kk, v in
                                     getattr(pg, '_tune_units', {}).items()}}
                    log.append(rec)
                    print('[autotune] K=%d D=%d %s dtype %d -> %s (%.1fs) %s' % (
                        K, D, role, code, best, time.time() - t, json.dumps(rec['ms'])),
                        flush=True)
            prog._buffers.clear()
            for env in (L, R):
                env.drop_low()
            if dt == 0 and 'renorm' in args.roles:
                # the renormalisation launches of the step that built L
                from . import kernelQt as kq2
                rp = kq2.RenormProgram(mpo, extra['lidx'], extra['lsite'], extra['Lprev'], 'L')
                sv = extra['lsite'].value
                sv = dv.upload(sv, 0) if isinstance(sv, np.ndarray) else sv
                E = extra['Lprev'].buffer(0)
                W = mpo.buffer(extra['lidx'], 0)
                r1 = dv.empty(rp.inter[0], 0)
                r2 = dv.empty(rp.inter[1], 0)
                ro = dv.zeros(rp.out_size, 0)
                from .qtensor import run_program as rpg
                for st_ in rp.stages:
                    for k, (pg, a, b, c) in enumerate(((st_[0], sv, E, r1), (st_[1], W, r1, r2),
                                                       (st_[2], r2, sv, ro))):
                        role = 'renorm%d' % (k + 1)
                        if pg.ntasks == 0:
                            continue
                        sig = signature(pg, role, 0)
                        if sig in entries:
                            continue
                        t = time.time()
                        times, best = tune(pg, role, 0,
                                           lambda pg=pg, a=a, b=b, c=c: rpg(pg, a, b, c, 0))
                        entries[sig] = best
                        rec = {'K': K, 'D': D, 'signature': list(sig), 'kernel': best[0],
                               'shape': best[1],
                               'ms': {'%s:%d' % kk: round(v, 3) for kk, v in times.items()},
                               'units': {'%s:%d' % kk: int(v) for kk, v in
                                         getattr(pg, '_tune_units', {}).items()}}
                        log.append(rec)
                        print('[autotune] K=%d D=%d %s -> %s (%.1fs) %s' % (
                            K, D, role, best, time.time() - t, json.dumps(rec['ms'])), flush=True)
                del r1, r2, ro, rp
        del mpo, L, R, prog
        dv.torch().cuda.empty_cache()
    out = {'device': 'MI355X (gfx950)',
           'how': 'python -m zmpo_dmrg_amd.autotune --sizes %s --dtypes %s' % (args.sizes,
                                                                              args.dtypes),
           'entries': [{'signature': list(s), 'kernel': c[0], 'shape': c[1]}
                       for s, c in entries.items()],
           'measurements': log}
    with open(args.out, 'w') as f:
        json.dump(out, f, indent=1)
    print('[autotune] wrote %d entries to %s' % (len(entries), args.out), flush=True)


if __name__ == '__main__':
    main()
