"""Helpers for measuring the sigma kernels with rocprofv3.

    python -m zmpo_dmrg_amd.profiling copy [--gib 2]
        a device copy of known size (3 passes) to calibrate the FETCH_SIZE /
        WRITE_SIZE counters: run it under ``rocprofv3 --pmc FETCH_SIZE`` (then
        WRITE_SIZE) and divide the known bytes by what the counter reports.

    python -m zmpo_dmrg_amd.profiling stats KERNEL_STATS_CSV
        markdown table of a rocprofv3 kernel statistics file.

    python -m zmpo_dmrg_amd.profiling mfma COUNTER_CSV TRACE_CSV [--cu 256]
        executed MFMA TFLOP/s and MFMA busy fraction per kernel from
        SQ_INSTS_VALU_MFMA_MOPS_* / SQ_VALU_MFMA_BUSY_CYCLES / GRBM_GUI_ACTIVE.

    python -m zmpo_dmrg_amd.profiling launches BENCH_JSON TRACE_CSV
            [--fetch CSV] [--write CSV] [--fetch-scale F] [--write-scale W]
        per-launch duration and HBM bytes of the sigma launches (L*psi, the
        chained W[i]*W[i+1] launch, *R) that bench.py times last (3
        repetitions x op chunks), from rocprofv3 ``-f csv`` kernel-trace /
        counter-collection files.  Prints a markdown table and a JSON line
        {"launches": {name: {"avg_ms", "bytes_per_launch", ...}}}.
"""
import argparse
import collections
import csv
import json
import sys

KERNELS = ('shaped_gemm', 'shaped_reduce', 'chain_gemm', 'stream_gemm', 'wave_gemm', 'wave_reduce',
           'skinny_gemm', 'dot_gemm', 'blockgemm', 'panel_gemm')
# executed FLOPs per SQ_INSTS_VALU_MFMA_MOPS_* count
FLOPS_PER_MOP = 512


def _copy(gib):
    import torch
    n = int(gib * 2 ** 30) // 8
    a = torch.randn(n, dtype=torch.float64, device='cuda')
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print('copied %d bytes x3 (read %d, write %d per pass)' % (n * 8, n * 8, n * 8), flush=True)


def _kind(name):
    for k in KERNELS:
        if k in name:
            return k
    return None


def _dispatches(path, counters=False):
    """Sigma-kernel dispatches in order: (kind, name, ms or counter dict)."""
    out = collections.OrderedDict()
    with open(path) as f:
        for r in csv.DictReader(f):
            k = _kind(r['Kernel_Name'])
            if k is None:
                continue
            if counters:
                d = out.setdefault(r['Dispatch_Id'], [k, r['Kernel_Name'], {}])
                d[2][r['Counter_Name']] = d[2].get(r['Counter_Name'], 0.0) + float(r['Counter_Value'])
            else:
                out[r['Dispatch_Id']] = [k, r['Kernel_Name'],
                                         (int(r['End_Timestamp']) - int(r['Start_Timestamp'])) * 1e-6]
    return list(out.values())


def _sigma_blocks(ds):
    """Group dispatches into sigma applications of the chained path: the
    L*psi dispatch before each run of chain_gemm dispatches, the run, then
    the *R dispatch (and its split-K reduction)."""
    blocks, i = [], 0
    while i < len(ds):
        if ds[i][0] != 'chain_gemm':
            i += 1
            continue
        j = i
        while j < len(ds) and ds[j][0] == 'chain_gemm':
            j += 1
        if i == 0 or j >= len(ds):
            break
        r_end = j + 1
        if r_end < len(ds) and ds[r_end][0] == 'shaped_reduce':
            r_end += 1
        blocks.append({'L*psi': [ds[i - 1]], 'W[i]*W[i+1] chained': ds[i:j], '*R': ds[j:r_end]})
        i = r_end
    return blocks


def _launches(args):
    with open(args.bench) as f:
        b = json.load(f)
    names = b['roofline']['stage_names']
    n = 3 * int(b['config']['chunks'])
    tr = _sigma_blocks(_dispatches(args.trace))[-n:]
    fe = _sigma_blocks(_dispatches(args.fetch, True))[-n:] if args.fetch else None
    wr = _sigma_blocks(_dispatches(args.write, True))[-n:] if args.write else None
    print('| launch | kernels | avg ms | FETCH GB | WRITE GB | HBM GB/s |')
    print('|---|---|---|---|---|---|')
    res = {}
    for nm in names:
        ms = sum(sum(d[2] for d in blk[nm]) for blk in tr) / len(tr)
        kn = sorted({d[1].replace('void ', '').replace('(anonymous namespace)::', '').split('(')[0]
                     for blk in tr for d in blk[nm]})
        fgb = (sum(sum(d[2].get('FETCH_SIZE', 0.0) for d in blk[nm]) for blk in fe) * 1024
               * args.fetch_scale / len(fe) / 1e9) if fe else None
        wgb = (sum(sum(d[2].get('WRITE_SIZE', 0.0) for d in blk[nm]) for blk in wr) * 1024
               * args.write_scale / len(wr) / 1e9) if wr else None
        tot = (fgb or 0.0) + (wgb or 0.0)
        res[nm] = {'avg_ms': ms, 'bytes_per_launch': tot * 1e9 if (fe and wr) else None,
                   'fetch_gb': fgb, 'write_gb': wgb}
        print('| %s | %s | %.3f | %s | %s | %s |' % (
            nm, ', '.join('`%s`' % k for k in kn), ms, '%.2f' % fgb if fe else '-',
            '%.2f' % wgb if wr else '-', '%.0f' % (tot / ms * 1e3) if (fe and wr) else '-'))
    print(json.dumps({'launches': res}))


def _short(name):
    n = name.replace('void ', '').replace('(anonymous namespace)::', '')
    return n.split('(zmpo')[0].split('(int,')[0].split('(float*')[0].split('(double*')[0]


def _stats(args):
    """Markdown table of a rocprofv3 ``*_kernel_stats.csv``."""
    with open(args.csv) as f:
        rows = list(csv.DictReader(f))
    print('| kernel | calls | total ms | avg ms | min ms | max ms | % |')
    print('|---|---|---|---|---|---|---|')
    for r in rows[:args.top]:
        print('| `%s` | %s | %.2f | %.3f | %.3f | %.3f | %.2f |' % (
            _short(r['Name']), r['Calls'], float(r['TotalDurationNs']) / 1e6,
            float(r['AverageNs']) / 1e6, float(r['MinNs']) / 1e6, float(r['MaxNs']) / 1e6,
            float(r['Percentage'])))


def _mfma(args):
    """Per-kernel MFMA work from a counter-collection CSV with
    SQ_INSTS_VALU_MFMA_MOPS_F64 (and _F32 / _BF16 when present),
    SQ_VALU_MFMA_BUSY_CYCLES and GRBM_GUI_ACTIVE, joined with the kernel
    durations of a kernel-trace CSV of the same command (dispatches in the
    same order).  Executed TFLOP/s = FLOPS_PER_MOP * MOPS / duration; busy =
    SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE * cu * 4 SIMDs)."""
    per = collections.OrderedDict()
    with open(args.counters) as f:
        for r in csv.DictReader(f):
            d = per.setdefault(r['Dispatch_Id'], {'name': _short(r['Kernel_Name'])})
            d[r['Counter_Name']] = d.get(r['Counter_Name'], 0.0) + float(r['Counter_Value'])
    dur = {}
    with open(args.trace) as f:
        for r in csv.DictReader(f):
            dur.setdefault(_short(r['Kernel_Name']), []).append(
                (int(r['End_Timestamp']) - int(r['Start_Timestamp'])) * 1e-9)
    agg = collections.OrderedDict()
    for d in per.values():
        a = agg.setdefault(d['name'], collections.Counter())
        a['n'] += 1
        for k, v in d.items():
            if k != 'name':
                a[k] += v
    print('| kernel | dispatches | executed MFMA TFLOP/s | MFMA busy |')
    print('|---|---|---|---|')
    for nm, a in agg.items():
        mops = sum(a.get(k, 0.0) for k in a if k.startswith('SQ_INSTS_VALU_MFMA_MOPS'))
        t = dur.get(nm, [])
        tsum = sum(t) * a['n'] / max(len(t), 1) if t else 0.0
        rate = FLOPS_PER_MOP * mops / tsum / 1e12 if tsum else float('nan')
        busy = (a.get('SQ_VALU_MFMA_BUSY_CYCLES', 0.0) / (a['GRBM_GUI_ACTIVE'] * args.cu * 4)
                if a.get('GRBM_GUI_ACTIVE') else float('nan'))
        print('| `%s` | %d | %.2f | %.3f |' % (nm, a['n'], rate, busy))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest='cmd', required=True)
    c = sub.add_parser('copy')
    c.add_argument('--gib', type=float, default=2.0)
    l_ = sub.add_parser('launches')
    l_.add_argument('bench')
    l_.add_argument('trace')
    l_.add_argument('--fetch')
    l_.add_argument('--write')
    l_.add_argument('--fetch-scale', type=float, default=1.0)
    l_.add_argument('--write-scale', type=float, default=1.0)
    st = sub.add_parser('stats')
    st.add_argument('csv')
    st.add_argument('--top', type=int, default=16)
    m = sub.add_parser('mfma')
    m.add_argument('counters')
    m.add_argument('trace')
    m.add_argument('--cu', type=int, default=256)
    args = ap.parse_args(argv)
    if args.cmd == 'copy':
        _copy(args.gib)
    elif args.cmd == 'stats':
        _stats(args)
    elif args.cmd == 'mfma':
        _mfma(args)
    else:
        _launches(args)


if __name__ == '__main__':
    sys.exit(main())
