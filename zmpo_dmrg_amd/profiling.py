"""Helpers for measuring the sigma kernels with rocprofv3.

    python -m zmpo_dmrg_amd.profiling copy [--gib 2]
        a device copy of known size (3 passes) to calibrate the FETCH_SIZE /
        WRITE_SIZE counters: run it under ``rocprofv3 --pmc FETCH_SIZE`` (then
        WRITE_SIZE) and divide the known bytes by what the counter reports.

    python -m zmpo_dmrg_amd.profiling launches BENCH_JSON TRACE_CSV
            [--fetch CSV] [--write CSV] [--fetch-scale F] [--write-scale W]
        per-launch duration and HBM bytes of the sigma launches that bench.py
        times last (3 repetitions x op chunks x launches, in that order), from
        rocprofv3 ``-f csv`` kernel-trace / counter-collection files.  Prints a
        markdown table and a JSON line {launch: bytes_per_launch}.
"""
import argparse
import collections
import csv
import json
import sys

KERNELS = ('shaped_gemm', 'chain_gemm', 'stream_gemm', 'wave_gemm', 'blockgemm')


def _copy(gib):
    import torch
    n = int(gib * 2 ** 30) // 8
    a = torch.randn(n, dtype=torch.float64, device='cuda')
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print('copied %d bytes x3 (read %d, write %d per pass)' % (n * 8, n * 8, n * 8), flush=True)


def _is_sigma_kernel(name):
    return any(k in name for k in KERNELS)


def _trace(path):
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if _is_sigma_kernel(r['Kernel_Name']):
                out.append((r['Kernel_Name'], (int(r['End_Timestamp']) - int(r['Start_Timestamp'])) * 1e-6))
    return out


def _counters(path):
    agg = collections.OrderedDict()
    with open(path) as f:
        for r in csv.DictReader(f):
            if not _is_sigma_kernel(r['Kernel_Name']):
                continue
            a = agg.setdefault(r['Dispatch_Id'], {})
            a[r['Counter_Name']] = a.get(r['Counter_Name'], 0.0) + float(r['Counter_Value'])
    return list(agg.values())


def _launches(args):
    with open(args.bench) as f:
        b = json.load(f)
    names = b['roofline']['stage_names']
    nch = int(b['config']['chunks'])
    per = len(names)
    n = 3 * nch * per
    tr = _trace(args.trace)[-n:]
    fe = _counters(args.fetch)[-n:] if args.fetch else None
    wr = _counters(args.write)[-n:] if args.write else None
    print('| launch | kernel | avg ms | FETCH GB | WRITE GB | HBM GB/s |')
    print('|---|---|---|---|---|---|')
    res = {}
    for s, nm in enumerate(names):
        idx = list(range(s, n, per))
        ms = sum(tr[i][1] for i in idx) / len(idx)
        fgb = sum(fe[i]['FETCH_SIZE'] for i in idx) * 1024 * args.fetch_scale / len(idx) / 1e9 if fe else None
        wgb = sum(wr[i]['WRITE_SIZE'] for i in idx) * 1024 * args.write_scale / len(idx) / 1e9 if wr else None
        tot = (fgb or 0) + (wgb or 0)
        res[nm] = tot * 1e9 if fe and wr else None
        kn = tr[idx[0]][0].split('(')[0].replace('void ', '').replace('(anonymous namespace)::', '')
        print('| %s | `%s` | %.3f | %s | %s | %s |' % (
            nm, kn, ms, '%.2f' % fgb if fe else '-', '%.2f' % wgb if wr else '-',
            '%.0f' % (tot / ms * 1e3) if fe and wr else '-'))
    print(json.dumps({'launches': res}))


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
    args = ap.parse_args(argv)
    if args.cmd == 'copy':
        _copy(args.gib)
    else:
        _launches(args)


if __name__ == '__main__':
    sys.exit(main())
