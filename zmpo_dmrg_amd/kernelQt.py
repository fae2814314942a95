"""Device two-site sigma, diagonal, renormalisation and environment build.

Drop-in counterparts of source/mpo_dmrg_kernelQt.py (HDiagQt :30-127, HVecQt
:153-319, renormQt_H00 :633-767) and source/mpo_dmrg_initQt.py (genBopsQt
:49-85, genHopsQt :88-271) with maxslc=1.  Operator-level work is the same
block-pair contraction as the reference, lowered to libzmpo grouped-GEMM
launches:

* sigma: 4 launches per op chunk (all ops of the chunk in each launch); the
  last launch sums hpwts[p] * top4_p over every op of the chunk directly
  into sigma, one owner per sigma block, so no per-op reduction pass exists;
* renormalisation: 3 launches per op chunk, new environments written in place
  into one device buffer per bond;
* diagonal: 2 launches (L-diagonal x MPO-diagonal, then x R-diagonal).

Ops of one ``case`` (qtensor_opers.py:72-90) share block structure, so one
symbolic plan per case is instantiated for every op by adding base offsets.

Device layouts (leg storage order inside each block):
  environments [bra, mpo, ket]; top1 [b, w, n1, n2, r]; top3 [b, wr2, r, u1, u2];
  every other tensor in the reference's row-major order.
"""
import collections
import hashlib
import os
import time

import numpy as np

from . import _lib
from . import autotune
from . import device as dv
from .qtensor import (ChainProgram, Program, QTensor, _workspace, chain_structure, gather,
                      run_chain, run_program, seq_add_chain, seq_add_program)

ENV_PERM = (1, 0, 2)
TOP1_PERM = (1, 0, 2, 3, 4)
TOP3_PERM = (0, 3, 1, 2, 4)
DEFAULT_BUDGET = 32 * 2 ** 30     # bytes of intermediates per op chunk
# sigma stages 2-3 as one chained launch (intermediate top2 kept in LDS)
FUSED = os.environ.get('ZMPO_FUSED', '1') != '0'
# row tile of the chained launch: 16*rb rows.  32-row tiles read each
# W[i] / W[i+1] fragment for twice the rows: at K = 36, D = 1000 the FP32
# launch takes 65 ms instead of 77, BF16 44 instead of 59, FP64 the same
CHAIN_RB = int(os.environ.get('ZMPO_CHAIN_RB', '2'))
# sigma op chunks: at least this many (1: as the budget allows); with two or
# more, L x psi of chunk c+1 runs on a second stream beside the chained
# launch of chunk c (both are latency-bound, so they share the GPU well)
SIGMA_CHUNKS = int(os.environ.get('ZMPO_CHUNKS', '1'))
OVERLAP = os.environ.get('ZMPO_OVERLAP', '1') != '0'
SIGMA_STAGE1 = os.environ.get('ZMPO_STAGE1', 'shaped')   # grouped-GEMM kernel of L x psi
# grouped-GEMM kernel of top3 x R: the pipelined stream kernel, one unit per
# workgroup, k-tile 16 (FP64) / 32 (FP32); bit-identical to 'shaped' and
# measured 5% (FP64) / 11% (FP32) faster at K = 36, D = 1000, 3% / 8% at
# K = 24, D = 500
SIGMA_STAGE4 = os.environ.get('ZMPO_STAGE4', 'stream')
STAGE4_KT = {0: 16, 1: 32}
STAGE4_KT.update({2: 32, 3: 32})
# tile shapes (qtensor.SHAPES ids) of the L x psi and x R launches ('' = cost
# model).  32x32 for both: a sigma block's right dimension (~26 at D = 1000)
# fits one tile width, and the small LDS footprint keeps 8 workgroups per CU
# (measured at K = 36, D = 1000: L x psi 15.6 ms vs 16.6 for 64x16, x R 13.9
# ms vs 17.5 for zmpo_dotgemm)
SIGMA_SHAPES = [os.environ.get('ZMPO_SHAPE_S1', '2'), os.environ.get('ZMPO_SHAPE_S4', '2')]
# issue a whole sigma through one recorded launch sequence (zmpo_seq_run)
# instead of one library call per launch
SIGMA_SEQ = os.environ.get('ZMPO_SEQ', '1') != '0'
# operand slots of a sigma sequence
S_L, S_R, S_PI, S_PJ, S_PSI, S_SIG, S_T1, S_T3, S_WORK = range(9)
# ZMPO_PLAN_CACHE=1: keep each site's sigma program (plan tables only; the
# intermediates are released between steps) and reuse it when a later sweep
# meets the same block structure at that site
PLAN_CACHE = os.environ.get('ZMPO_PLAN_CACHE', '0') != '0'


def case(porb, isite, ncsite=1):
    """Operator class relative to the dot (qtensor_opers.py:72-90)."""
    if porb // 2 < isite:
        return porb % 2
    if porb // 2 == isite:
        return 2 + porb % 2
    if ncsite == 1 or porb // 2 == isite + 1:
        return 4 + porb % 2
    return 6 + porb % 2


def op_case(mpo, p, isite):
    """Class of operator p at one site for the one-site contractions
    (renormalisation): ops of one class share their plans.  Stores of
    other operators than H (mpo.op_class) define their own classes."""
    f = getattr(mpo, 'op_class', None)
    return f(p, isite) if f is not None else case(mpo.opers[p], isite)


# ------------------------------------------------------------------ stores
# environment moves between device and host memory (EnvStore.spill / buffer)
SPILL = {'download_s': 0.0, 'download_bytes': 0, 'upload_s': 0.0, 'upload_bytes': 0}


class EnvStore:
    """Environments of one bond for every op: ``meta[p]`` (QTensor, ENV_PERM)
    and one device buffer ``buf`` with op p at ``off[p]`` (replaces the HDF5
    files lop_site_i / rop_site_i of mpo_dmrg_opt.py:196-238)."""

    def __init__(self, meta, buf, off):
        self.meta, self.buf, self.off = meta, buf, np.asarray(off, dtype=np.int64)
        self._lo = {}

    @property
    def nops(self):
        return len(self.meta)

    def buffer(self, dtype=0):
        """The values in one precision (0 FP64, 1 FP32, 2 BF16); lower
        precisions are converted once and kept until drop_low()."""
        if isinstance(self.buf, np.ndarray):
            t0 = time.time()
            self.buf = dv.upload(self.buf, 0)
            SPILL['upload_s'] += time.time() - t0
            SPILL['upload_bytes'] += self.nbytes()
        if dtype == 0:
            return self.buf
        if dtype not in self._lo:
            b = dv.empty(self.buf.numel(), dtype)
            _lib.check(_lib.lib().zmpo_convert(dtype, dv.ptr(b), 0, dv.ptr(self.buf),
                                               self.buf.numel(), dv.stream()))
            self._lo[dtype] = b
        return self._lo[dtype]

    def drop_low(self):
        """Release the low-precision copies."""
        self._lo = {}

    def nbytes(self):
        """Bytes of the FP64 values, wherever they are held."""
        b = self.buf
        return int(b.nbytes) if isinstance(b, np.ndarray) else int(b.numel() * b.element_size())

    def resident(self):
        """Whether the FP64 values are in device memory."""
        return not isinstance(self.buf, np.ndarray)

    def spill(self):
        """Move the FP64 values to host memory and release every device
        copy; the next buffer() call uploads them again."""
        if self.resident():
            t0 = time.time()
            self.buf = dv.download(self.buf)
            SPILL['download_s'] += time.time() - t0
            SPILL['download_bytes'] += self.nbytes()
        self._lo = {}

    def op_tensor(self, p):
        """QTensor of op p with a host copy of its values (tests, export)."""
        m = self.meta[p]
        qt = m.template()
        qt.idlst = m.idlst
        n = m.size_allowed
        b = self.buf
        seg = b[self.off[p]:self.off[p] + n]
        qt.value = (seg.copy() if isinstance(b, np.ndarray) else dv.download(seg)) if n \
            else np.zeros(0)
        return qt


def boundary_store(nops, nqnum=2):
    """1x1x1 boundary operators with zero QN (initQt.py:49-85)."""
    qt = QTensor([True, False, True], ENV_PERM)
    z = np.zeros((1, nqnum))
    qt.fromQsyms(3, [z, z, z], [np.ones(1, np.int64)] * 3, ifallocate=False,
                 idlst=[[[0]], [[0]], [[0]]])
    return EnvStore([qt] * nops, np.ones(nops), np.arange(nops))


def _instantiate_many(items):
    """Case plans instantiated for a list of ops, concatenated: items =
    [(lowered, a_base, b_base, c_base, scale)] adds the op's operand base
    offsets to a copy of its case plan, sets the pair scale and shifts the
    op's p_begin/p_end past the pairs of the ops before it."""
    if not items:
        return np.zeros(0, _lib.TASK_DTYPE), np.zeros(0, _lib.PAIR_DTYPE)
    # concatenated as raw bytes (numpy copies structured arrays field by field)
    t = np.concatenate([it[0]['tasks'].view(np.uint8) for it in items]).view(_lib.TASK_DTYPE)
    p = np.concatenate([it[0]['pairs'].view(np.uint8) for it in items]).view(_lib.PAIR_DTYPE)
    nt = np.array([len(it[0]['tasks']) for it in items], dtype=np.int64)
    npr = np.array([len(it[0]['pairs']) for it in items], dtype=np.int64)
    base = np.array([it[1:4] for it in items], dtype=np.int64)
    p['a_off'] += np.repeat(base[:, 0], npr)
    p['b_off'] += np.repeat(base[:, 1], npr)
    p['scale'] = np.repeat(np.array([it[4] for it in items], dtype=np.float64), npr)
    t['c_off'] += np.repeat(base[:, 2], nt)
    shift = np.repeat(np.concatenate([[0], np.cumsum(npr)[:-1]]), nt)
    t['p_begin'] += shift
    t['p_end'] += shift
    return t, p


def _merge_shared_many(items, nout):
    """Instantiate the ops of ``items`` (every op's tasks are the same nout
    output blocks) and group their pairs per output block; the stable sort
    keeps, inside every output block, the ops in chunk order (fixed
    summation order)."""
    t, p = _instantiate_many(items)
    nt = np.array([len(it[0]['tasks']) for it in items], dtype=np.int64)
    cnt = (t['p_end'] - t['p_begin']).astype(np.int64)
    local = np.arange(len(t), dtype=np.int64) - np.repeat(np.concatenate([[0], np.cumsum(nt)[:-1]]), nt)
    rows = np.repeat(local, cnt)
    order = np.argsort(rows, kind='stable')
    counts = np.bincount(rows, minlength=nout)
    tasks = t[:nt[0]].copy()
    tasks['p_begin'] = np.concatenate([[0], np.cumsum(counts)[:-1]])
    tasks['p_end'] = tasks['p_begin'] + counts
    return tasks, p[order]


def out_dtype(dtype):
    """Kernel dtype code of the sigma launch into sigma: BF16 operands
    accumulate into an FP32 sigma (ZMPO_BF16_F32)."""
    return 3 if dtype == 2 else dtype


def sigma_dtype(dtype):
    """Storage code of sigma for a sigma computed with operands of dtype."""
    return 1 if dtype == 2 else dtype


def _chunks(sizes, budget):
    """Split ops into consecutive chunks whose intermediate bytes fit the budget."""
    out, cur, acc = [], [], 0
    for p, s in enumerate(sizes):
        if cur and acc + s > budget:
            out.append(cur)
            cur, acc = [], 0
        cur.append(p)
        acc += s
    if cur:
        out.append(cur)
    return out


# ------------------------------------------------------------------ sigma
class SigmaProgram:
    """Compiled two-site sigma for one (site, psi template): per op chunk,
    four grouped-GEMM programs and the intermediate offsets."""

    def __init__(self, mpo, isite, qtmp, qkey, lenv, renv, hpwts, budget=DEFAULT_BUDGET,
                 fused=None):
        self.isite = isite
        self.fused = FUSED if fused is None else fused
        self.ndim0 = qtmp.size_allowed
        jsite = isite + 1
        nops = mpo.nops
        plans = {}
        self.flops_stage = np.zeros(4, dtype=np.int64)
        for p in range(nops):
            c = case(mpo.opers[p], isite, 2)
            if c in plans:
                continue
            lop = lenv.meta[p]
            cop = mpo.sites[isite][p]
            wop = mpo.sites[jsite][p]
            rop = renv.meta[p]
            t1 = QTensor()
            t1.tensordotSYM(lop, qtmp, ([2], [0]), perm=TOP1_PERM)
            t2 = QTensor()
            t2.tensordotSYM(t1, cop, ([0, 2], [0, 3]))
            t3 = QTensor()
            t3.tensordotSYM(t2, wop, ([1, 3], [3, 0]), perm=TOP3_PERM)
            rf = rop.template()
            rf.qsyms = [-rop.qsyms[0], qkey - rop.qsyms[1], qkey - rop.qsyms[2]]
            t4 = QTensor()
            t4.tensordotSYM(t3, rf, ([1, 3], [2, 0]))
            if t4.nblks_allowed and not (np.array_equal(t4.idx_allowed, qtmp.idx_allowed)
                                         and np.array_equal(t4.blk_dims, qtmp.blk_dims)):
                raise RuntimeError('sigma output structure differs from the psi template')
            plans[c] = [t.plan.lowered() for t in (t1, t2, t3, t4)] + [t1, t2, t3]
        self.plans = plans
        self.case_of = [case(mpo.opers[p], isite, 2) for p in range(nops)]
        kept = (4, 6) if self.fused else (4, 5, 6)
        sz = [sum(plans[c][k].size_allowed for k in kept) * 8 for c in self.case_of]
        if SIGMA_CHUNKS > 1:
            budget = min(budget, max(1, sum(sz) // SIGMA_CHUNKS + 1))
        self.chunks = _chunks(sz, budget)
        self.stages = []
        self.chains = []
        self.inter = [0, 0, 0]
        structs = {}
        if self.fused:
            _, poff_i, pmap_i = mpo.padded(isite)
            _, poff_j, pmap_j = mpo.padded(jsite)
            rep = {}
            for p in range(nops):
                rep.setdefault(case(mpo.opers[p], isite, 2), p)
            for c, pl in plans.items():
                t2, t3 = pl[5], pl[6]
                nr = 1 << 24
                mid_key = t2.blk_isyms[:, 0] * nr + t2.blk_isyms[:, 2]
                out_key = t3.blk_isyms[:, 0] * nr + t3.blk_isyms[:, 1]
                structs[c] = chain_structure(pl[1], pl[2], mid_key, out_key, t2.blk_off,
                                             pmap_i[rep[c]], pmap_j[rep[c]])
        for ch in self.chunks:
            offs = [0, 0, 0]
            parts = [[], [], [], []]
            inst = []
            for p in ch:
                pl = plans[self.case_of[p]]
                sizes = [pl[4].size_allowed, pl[5].size_allowed, pl[6].size_allowed]
                if pl[4].nblks_allowed == 0 or pl[3]['pairs'].size == 0:
                    continue
                lo = lenv.off[p]
                wo, vo = mpo.off[isite][p], mpo.off[jsite][p]
                ro = renv.off[p]
                parts[0].append((pl[0], lo, 0, offs[0], 1.0))
                parts[1].append((pl[1], offs[0], wo, offs[1], 1.0))
                parts[2].append((pl[2], offs[1], vo, offs[2], 1.0))
                parts[3].append((pl[3], offs[2], ro, 0, float(hpwts[p])))
                if self.fused:
                    inst.append((self.case_of[p], offs[0], poff_i[p], poff_j[p], offs[2]))
                for k in range(3):
                    offs[k] += sizes[k]
            progs = []
            for k in range(3):
                t, pr = _instantiate_many(parts[k])
                # L x psi: tall bra x MPO-bond rows, few columns
                progs.append(Program(t, pr, prefer=SIGMA_STAGE1 if k == 0 else None))
            t, pr = (_merge_shared_many(parts[3], qtmp.nblks_allowed) if parts[3]
                     else _instantiate_many([]))
            # sigma blocks with the same right sector read the same R blocks
            progs.append(Program(t, pr, reuse_key=qtmp.blk_isyms[:, 3] if len(t) else None,
                                 prefer=SIGMA_STAGE4))
            if SIGMA_STAGE4 == 'stream':
                progs[3].upw = 1
                progs[3].kt_dtype = STAGE4_KT
            for k, sh in ((0, SIGMA_SHAPES[0]), (3, SIGMA_SHAPES[1])):
                if sh:
                    progs[k].shape = int(sh)
            for k in range(4):
                self.flops_stage[k] += progs[k].flops
            self.stages.append(progs)
            self.chains.append(ChainProgram(structs, inst, rb=CHAIN_RB) if self.fused else None)
            self.inter = [max(a, b) for a, b in zip(self.inter, offs)]
        self.flops = int(self.flops_stage.sum())
        self._buffers = {}

    def sequence(self, dtype):
        """The recorded launch sequence of a fused sigma in dtype (built once
        per dtype; None when a launch cannot be recorded).  Returns (seq,
        workspace elements)."""
        seqs = self.__dict__.setdefault('_seqs', {})
        if dtype not in seqs:
            seq = _lib.Seq()
            work = 0
            first = True
            ok = self.fused
            for st, chain in zip(self.stages, self.chains):
                if not ok:
                    break
                w = seq_add_program(seq, st[0], dtype, S_L, S_PSI, S_T1, S_WORK)
                seq_add_chain(seq, chain, S_T1, S_PI, S_PJ, S_T3)
                w4 = 0
                if st[3].ntasks:
                    w4 = seq_add_program(seq, st[3], out_dtype(dtype), S_T3, S_R, S_SIG, S_WORK,
                                         beta=0 if first else 1, out_f32=True)
                    first = False
                if w is None or w4 is None:
                    ok = False
                else:
                    work = max(work, w, w4)
            seqs[dtype] = (seq, work) if ok and not first else None
        return seqs[dtype]

    def device_refs(self):
        """The uploaded plan tables and schedules the launches read (a
        captured graph keeps them alive)."""
        out = []
        for st in self.stages:
            for pg in st:
                if pg._dev is not None:
                    out += list(pg._dev)
                for sc in (pg._sched or {}).values():
                    out += [sc[k] for k in ('units', 'groups') if k in sc]
        for ch in self.chains:
            if ch is not None and ch._dev is not None:
                out += list(ch._dev[0]) + list(ch._dev[1])
        return out

    def release_device(self):
        """Drop the uploaded plan tables of every launch (host arrays stay;
        the next apply uploads them again) and the sequences recording them."""
        self.__dict__.pop('_seqs', None)
        for st in self.stages:
            for pg in st:
                pg.release_device()
        for ch in self.chains:
            if ch is not None:
                ch._dev = None

    def drop_buffers(self, dtype):
        """Forget the intermediate views of one operand type (their memory is
        the shared scratch buffer, kept for the next launches)."""
        self._buffers.pop(dtype, None)
        self._buffers.pop(('t1b', dtype), None)

    def _layout(self, mid):
        """Byte offsets of top1, the second top1 buffer, top3 and top2 in the
        shared intermediate buffer, sized for FP64 so that every operand
        type fits the same bytes."""
        seg = [((n * 8 + 255) // 256) * 256 for n in self.inter]
        off = [0, seg[0], 2 * seg[0], 2 * seg[0] + seg[2]]
        return off, off[3] + (seg[1] if mid else 0)

    def buffers(self, dtype, mid=None):
        """Intermediate buffers [top1, top2, top3] (views of the scratch
        buffer that the sigma and renormalisation launches of every
        operand type share); top2 only for the unfused path (mid=True,
        default: not self.fused)."""
        mid = (not self.fused) if mid is None else mid
        off, total = self._layout(True)
        buf = dv.scratch('inter', total if mid else off[3])
        b = self._buffers.get(dtype)
        if b is None or b[0].data_ptr() != buf[0:1].data_ptr():
            b = [dv.typed(buf, off[0], self.inter[0], dtype), None,
                 dv.typed(buf, off[2], self.inter[2], dtype)]
            self._buffers[dtype] = b
        if mid and b[1] is None:
            b[1] = dv.typed(buf, off[3], self.inter[1], dtype)
        return b

    def apply(self, mpo, lenv, renv, psi, sigma, dtype=0, fused=None, overlap=None):
        """sigma = H psi (device arrays of length ndim0, same dtype; for dtype
        2, BF16 operands with FP32 accumulation, psi is BF16 and sigma FP32).
        overlap=False keeps every launch on the current stream (as graph
        capture needs)."""
        fused = self.fused if fused is None else fused
        overlap = OVERLAP if overlap is None else overlap
        t1, t2, t3 = self.buffers(dtype, mid=not fused)
        L, R = lenv.buffer(dtype), renv.buffer(dtype)
        Wi, Wj = mpo.buffer(self.isite, dtype), mpo.buffer(self.isite + 1, dtype)
        if not self.stages or all(st[3].ntasks == 0 for st in self.stages):
            sigma[:] = 0
            return sigma
        tuned = self.__dict__.setdefault('_tuned', set())
        if dtype not in tuned and getattr(psi, 'is_cuda', False):
            self._choose_tiles(L, R, psi, sigma, dtype, t1, t3)
            tuned.add(dtype)
        if fused and not self.fused:
            raise ValueError('SigmaProgram built without the chained stages (fused=False)')
        if fused and overlap and len(self.stages) > 1 and not isinstance(psi, np.ndarray) \
                and getattr(psi, 'is_cuda', False):
            return self._apply_overlapped(mpo, L, R, psi, sigma, dtype, t1, t3)
        if fused and SIGMA_SEQ and getattr(psi, 'is_cuda', False):
            sq = self.sequence(dtype)
            if sq is not None:
                seq, nwork = sq
                work = _workspace(nwork, out_dtype(dtype)) if nwork else None
                # the sigma's split-K slots and L x psi's share one workspace;
                # they are FP32 in BF16 runs (L x psi never splits at these K)
                _lib.check(_lib.lib().zmpo_sigma_two_site(
                    seq._h, dtype, dv.ptr(L), dv.ptr(R),
                    dv.ptr(mpo.padded_buffer(self.isite, dtype)),
                    dv.ptr(mpo.padded_buffer(self.isite + 1, dtype)), dv.ptr(psi),
                    dv.ptr(sigma), dv.ptr(t1), dv.ptr(t3), dv.ptr(work) if work is not None
                    else None, dv.stream()))
                return sigma
        first = True
        for st, chain in zip(self.stages, self.chains):
            run_program(st[0], L, psi, t1, dtype)
            if fused:
                run_chain(chain, t1, mpo.padded_buffer(self.isite, dtype),
                          mpo.padded_buffer(self.isite + 1, dtype), t3, dtype)
            else:
                run_program(st[1], t1, Wi, t2, dtype)
                run_program(st[2], t2, Wj, t3, dtype)
            if st[3].ntasks:
                run_program(st[3], t3, R, sigma, out_dtype(dtype), beta=0 if first else 1)
                first = False
        return sigma

    def _choose_tiles(self, L, R, psi, sigma, dtype, t1, t3):
        """Tiles of the L x psi and x R launches in this dtype (autotune.py:
        the measured table, or tuning on the device in 'online' mode, which
        writes t1 and sigma before the sigma proper overwrites them)."""
        odt = out_dtype(dtype)
        for st in self.stages:
            autotune.choose(st[0], 'sigma_l', dtype,
                            lambda pg=st[0]: run_program(pg, L, psi, t1, dtype))
            if st[3].ntasks:
                autotune.choose(st[3], 'sigma_r', odt,
                                lambda pg=st[3]: run_program(pg, t3, R, sigma, odt, beta=0))

    def _apply_overlapped(self, mpo, L, R, psi, sigma, dtype, t1, t3):
        """Chunked sigma with L x psi of chunk c+1 on a side stream while the
        chained launch and x R of chunk c run on the current stream; top1 is
        double-buffered, x R stays in chunk order (fixed summation order)."""
        torch = dv.torch()
        main = torch.cuda.current_stream()
        if getattr(self, '_side', None) is None:
            self._side = torch.cuda.Stream()   # default priority
        side = self._side
        off, _ = self._layout(False)
        bufs = [t1, dv.typed(dv.scratch('inter', off[3]), off[1], self.inter[0], dtype)]
        free = [None, None]
        Pi, Pj = mpo.padded_buffer(self.isite, dtype), mpo.padded_buffer(self.isite + 1, dtype)
        side.wait_stream(main)   # psi and the buffers are ready
        first = True
        for c, (st, chain) in enumerate(zip(self.stages, self.chains)):
            buf = bufs[c % 2]
            if free[c % 2] is not None:
                side.wait_event(free[c % 2])   # chunk c-2 is done reading buf
            run_program(st[0], L, psi, buf, dtype, stream=side.cuda_stream)
            ready = torch.cuda.Event()
            ready.record(side)
            main.wait_event(ready)
            run_chain(chain, buf, Pi, Pj, t3, dtype, stream=main.cuda_stream)
            done = torch.cuda.Event()
            done.record(main)
            free[c % 2] = done
            if st[3].ntasks:
                run_program(st[3], t3, R, sigma, out_dtype(dtype), beta=0 if first else 1,
                            stream=main.cuda_stream)
                first = False
        if first:
            dv.clear(sigma)
        return sigma


class SigmaGraph:
    """One sigma application of a SigmaProgram captured as a HIP graph over
    fixed input/output vectors: every launch of ``SigmaProgram.apply`` (the
    grouped-GEMM, split-K reduce and chained dispatches) is re-issued by one
    host call, which removes the per-launch host cost that dominates small
    sigmas.  The graph holds references to every buffer it reads or writes
    (environments, MPO copies, intermediates, split-K workspace), so they
    stay allocated while it lives; replays run on the current stream.
    The captured launches contain no memset nodes."""

    def __init__(self, prog, mpo, lenv, renv, dtype=0, warmup=True):
        """warmup=False: the caller has already run ``prog.apply`` eagerly in
        this dtype with these environments (lazy allocations are done)."""
        torch = dv.torch()
        n = prog.ndim0
        self.prog, self.dtype = prog, dtype
        self.x = dv.zeros(n, dtype)
        self.y = dv.zeros(n, sigma_dtype(dtype))
        self.empty = not prog.stages or all(st[3].ntasks == 0 for st in prog.stages)
        if self.empty:
            self.graph = None
            return
        # eager run first: device tables, intermediates, converted
        # environments and the workspace are allocated outside the capture
        if warmup:
            prog.apply(mpo, lenv, renv, self.x, self.y, dtype, overlap=False)
        self._keep = [lenv.buffer(dtype), renv.buffer(dtype), dv._SCRATCH.get((str(dv.device()), 'splitk')),
                      mpo.buffer(prog.isite, dtype), mpo.buffer(prog.isite + 1, dtype)]
        if prog.fused:
            self._keep += [mpo.padded_buffer(prog.isite, dtype),
                           mpo.padded_buffer(prog.isite + 1, dtype)]
        self._keep += [b for b in prog.buffers(dtype) if b is not None]
        self._keep += prog.device_refs()
        main = torch.cuda.current_stream()
        cap = torch.cuda.Stream()   # default priority
        cap.wait_stream(main)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.stream(cap):
            self.graph.capture_begin()
            try:
                prog.apply(mpo, lenv, renv, self.x, self.y, dtype, overlap=False)
            finally:
                self.graph.capture_end()
        main.wait_stream(cap)

    def run(self, x, out):
        """out = H x through the captured launches (x, out: device vectors of
        any floating dtype; converted on the copies in and out)."""
        if self.graph is None:
            out.zero_()
            return out
        self.x.copy_(x)
        self.graph.replay()
        out.copy_(self.y)
        return out


# ------------------------------------------------------------------ diagonal
def _zero_sector(qsyms0):
    z = np.flatnonzero(np.all(np.abs(qsyms0) < 1e-10, axis=1))
    return int(z[0]) if len(z) else -1


def _diag_gather_index(meta):
    """Indices of env[b, w, b] for w in the zero MPO sector, as [w][b'] with b'
    running over the bra sectors in sector order.  Returns (idx, zdim, bdim)."""
    z = _zero_sector(meta.qsyms[0])
    bdim = int(meta.ndims[1].sum())
    if z < 0:
        return None, 0, bdim
    zd = int(meta.ndims[0][z])
    soff = np.concatenate([[0], np.cumsum(meta.ndims[1])[:-1]])
    idx = -np.ones((zd, bdim), dtype=np.int64)
    st = meta.blk_strides()
    for j in range(meta.nblks_allowed):
        w_s, b_s, k_s = meta.blk_isyms[j]
        if w_s != z or b_s != k_s:
            continue
        d = int(meta.ndims[1][b_s])
        bl = np.arange(d)
        wl = np.arange(zd)
        off = (meta.blk_off[j] + bl[None, :] * st[j, 1] + wl[:, None] * st[j, 0]
               + bl[None, :] * st[j, 2])
        idx[:, soff[b_s]:soff[b_s] + d] = off
    return idx, zd, bdim


def _diag_mpo_factor(mpo, isite, p):
    """P[n1 n2][w, wr2] = sum_wr Wi_p[w, wr, n1, n1] Wj_p[wr, wr2, n2, n2] over the
    zero-QN rows of W[i] and columns of W[i+1] (None when it vanishes).  Depends
    only on the MPO, so it is computed once per (site, op) and kept with it."""
    cache = mpo.__dict__.setdefault('_diag_factor', {})
    key = (isite, p)
    if key not in cache:
        cop, wop = mpo.sites[isite][p], mpo.sites[isite + 1][p]
        zc = _zero_sector(cop.qsyms[0])
        zw = _zero_sector(wop.qsyms[1])
        P = None
        if zc >= 0 and zw >= 0:
            ci = cop.idlst[0][zc]
            wi = wop.idlst[1][zw]
            cdz = np.einsum('abnn->abn', cop.toDenseTensor()[ci])      # [w, wr, n]
            wdz = np.einsum('abnn->abn', wop.toDenseTensor()[:, wi])   # [wr, wr2, n]
            P = np.einsum('awn,wbm->nmab', cdz, wdz).reshape(16, len(ci), len(wi))
            if not np.any(P):
                P = None
        cache[key] = P
    return cache[key]


class DiagProgram:
    """HDiagQt: diag[l,n1,n2,r] = sum_p hp sum_{w,wr,wr2} Ld_p[w,l] Wi_p[w,wr,n1,n1]
    Wj_p[wr,wr2,n2,n2] Rd_p[wr2,r], on the psi template's allowed blocks."""

    def __init__(self, mpo, isite, qtmp, lenv, renv, hpwts):
        nops = mpo.nops
        self.ndim0 = qtmp.size_allowed
        lgi, rgi, pmats = [], [], []
        lbase = rbase = xbase = 0
        ldim = int(qtmp.ndims[0].sum())
        rdim = int(qtmp.ndims[3].sum())
        loff = np.concatenate([[0], np.cumsum(qtmp.ndims[0])[:-1]])
        roff = np.concatenate([[0], np.cumsum(qtmp.ndims[3])[:-1]])
        h1_parts, h2_rows = [], []
        ops_info = []
        pbase = 0
        gidx = {}   # ops of one case share their environment template

        def gather_index(meta):
            if id(meta) not in gidx:
                gidx[id(meta)] = _diag_gather_index(meta)
            return gidx[id(meta)]

        for p in range(nops):
            li, z1, ld = gather_index(lenv.meta[p])
            ri, z3, rd = gather_index(renv.meta[p])
            if li is None or ri is None:
                continue
            assert ld == ldim and rd == rdim
            P = _diag_mpo_factor(mpo, isite, p)
            if P is None:
                continue
            lgi.append(li.reshape(-1) + np.where(li.reshape(-1) >= 0, lenv.off[p], 0))
            rgi.append(ri.reshape(-1) + np.where(ri.reshape(-1) >= 0, renv.off[p], 0))
            pmats.append(P.reshape(-1))
            ops_info.append((p, z1, z3, lbase, rbase, pbase, xbase))
            lbase += z1 * ldim
            rbase += z3 * rdim
            pbase += P.size
            xbase += 16 * ldim * z3
        self.nact = len(ops_info)
        if not self.nact:
            self.h1 = self.h2 = None
            return
        self.lidx = dv.upload(np.concatenate(lgi))
        self.ridx = dv.upload(np.concatenate(rgi))
        self.pbuf = dv.upload(np.concatenate(pmats), 0)
        self.nl, self.nr, self.nx = lbase, rbase, xbase
        # H1: X_p[nn][l, wr2] = sum_w Ld_p[w, l] P_p[nn][w, wr2]
        t1 = np.zeros(16 * self.nact, dtype=_lib.TASK_DTYPE)
        p1 = np.zeros(16 * self.nact, dtype=_lib.PAIR_DTYPE)
        k = 0
        for (p, z1, z3, lb, rb, pb, xb) in ops_info:
            for nn in range(16):
                t1[k] = (xb + nn * ldim * z3, 0, ldim, z3, ldim, 0, z3, 1, k, k + 1, 0, 0)
                p1[k] = (lb, pb + nn * z1 * z3, 1.0, z1, ldim, 0, 1, ldim, z3, 0, 1, z3, 0)
                k += 1
        self.h1 = Program(t1, p1)
        # H2: diag block j = sum_p hp X_p[n1n2][l in l_s, :] Rd_p[:, r in r_s]
        nb = qtmp.nblks_allowed
        iso = qtmp.blk_isyms
        t2 = np.zeros(nb, dtype=_lib.TASK_DTYPE)
        p2 = np.zeros(nb * self.nact, dtype=_lib.PAIR_DTYPE)
        k = 0
        for j in range(nb):
            ls, n1, n2, rs = iso[j]
            dl, dr = int(qtmp.ndims[0][ls]), int(qtmp.ndims[3][rs])
            nn = int(n1) * 4 + int(n2)
            beg = k
            for (p, z1, z3, lb, rb, pb, xb) in ops_info:
                p2[k] = (xb + nn * ldim * z3 + loff[ls] * z3, rb + roff[rs], float(hpwts[p]),
                         z3, dl, 0, z3, 1, dr, 0, 1, rdim, 0)
                k += 1
            t2[j] = (qtmp.blk_off[j], 0, dl, dr, dl, 0, dr, 1, beg, k, 0, 0)
        self.h2 = Program(t2, p2)
        self.flops = self.h1.flops + self.h2.flops

    def sequence(self):
        """The two environment gathers and two launches as one recorded
        sequence over the slots (L, R, MPO factors, ldiag, rdiag, x, diag,
        workspace), run by zmpo_hdiag; None when a launch cannot be
        recorded."""
        if '_seq' not in self.__dict__:
            seq = _lib.Seq()
            for idx, a, c in ((self.lidx, 0, 3), (self.ridx, 1, 4)):
                rec = np.zeros(1, _lib.SEQ_OP_DTYPE)
                rec['kind'] = _lib.SEQ_GATHER
                rec['tasks'], rec['nunits'] = dv.ptr(idx), idx.numel()
                rec['slot_a'], rec['slot_b'], rec['slot_b2'], rec['slot_c'] = a, a, -1, c
                rec['slot_work'] = -1
                seq.add(rec)
            w1 = seq_add_program(seq, self.h1, 0, 3, 2, 5, 7)
            w2 = seq_add_program(seq, self.h2, 0, 5, 4, 6, 7)
            self._seq = (seq, max(w1, w2)) if (w1 is not None and w2 is not None) else None
        return self._seq

    def apply(self, lenv, renv, out):
        if not self.nact:
            out[:] = 0
            return out
        ld, rd, x = dv.empty(self.nl, 0), dv.empty(self.nr, 0), dv.empty(self.nx, 0)
        sq = self.sequence() if SIGMA_SEQ and getattr(out, 'is_cuda', False) else None
        if sq is not None:
            seq, nwork = sq
            work = _workspace(nwork, 0) if nwork else None
            _lib.check(_lib.lib().zmpo_hdiag(
                seq._h, dv.ptr(lenv.buffer(0)), dv.ptr(renv.buffer(0)), dv.ptr(self.pbuf),
                dv.ptr(ld), dv.ptr(rd), dv.ptr(x), dv.ptr(out),
                dv.ptr(work) if work is not None else None, dv.stream()))
            return out
        gather(ld, lenv.buffer(0), self.lidx)
        gather(rd, renv.buffer(0), self.ridx)
        run_program(self.h1, ld, self.pbuf, x, 0)
        run_program(self.h2, x, rd, out, 0)
        return out


# ------------------------------------------------------------------ renorm
class RenormProgram:
    """Three-stage renormalisation of every op for one site (kernelQt.py:633-767)."""

    def __init__(self, mpo, site_idx, qsite, env_in, status, budget=DEFAULT_BUDGET):
        nops = mpo.nops
        self.site_idx = site_idx
        self.status = status
        plans = {}
        case_of = [op_case(mpo, p, site_idx) for p in range(nops)]
        for p in range(nops):
            c = case_of[p]
            if c in plans:
                continue
            cop = mpo.sites[site_idx][p]
            op_in = env_in.meta[p]
            t1, t2, t3 = QTensor(), QTensor(), QTensor()
            if status == 'L':
                t1.tensordotSYM(qsite, op_in, ([0], [1]))
                t2.tensordotSYM(cop, t1, ([0, 2], [2, 0]))
                t3.tensordotSYM(t2, qsite, ([1, 3], [1, 0]), perm=ENV_PERM)
                t3.idlst = [cop.idlst[1], qsite.idlst[2], qsite.idlst[2]]
            else:
                cf = cop.template()
                cf.qsyms = [-cop.qsyms[0], -cop.qsyms[1], cop.qsyms[2], cop.qsyms[3]]
                cf.status = cop.status.copy()
                cf.status[0] = not cf.status[0]
                cf.status[1] = not cf.status[1]
                cf.fromQsyms(4, cf.qsyms, cf.ndims, ifallocate=False, idlst=cop.idlst)
                t1.tensordotSYM(qsite, op_in, ([2], [1]))
                t2.tensordotSYM(cf, t1, ([1, 2], [2, 1]))
                t3.tensordotSYM(t2, qsite, ([1, 3], [1, 2]), perm=ENV_PERM)
                t3.idlst = [cop.idlst[0], qsite.idlst[0], qsite.idlst[0]]
            plans[c] = [t.plan.lowered() for t in (t1, t2, t3)] + [t1, t2, t3]
        self.plans = plans
        self.meta = [plans[c][5] for c in case_of]
        sizes = np.array([m.size_allowed for m in self.meta], dtype=np.int64)
        self.out_off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.out_size = int(sizes.sum())
        isz = [(plans[c][3].size_allowed + plans[c][4].size_allowed) * 8 for c in case_of]
        self.chunks = _chunks(isz, budget)
        self.stages = []
        self.inter = [0, 0]
        self.flops = 0
        for ch in self.chunks:
            offs = [0, 0]
            parts = [[], [], []]
            for p in ch:
                pl = plans[case_of[p]]
                # the same operand roles in both directions (site, env | MPO, t1 | t2, site)
                parts[0].append((pl[0], 0, env_in.off[p], offs[0], 1.0))
                parts[1].append((pl[1], mpo.off[site_idx][p], offs[0], offs[1], 1.0))
                parts[2].append((pl[2], offs[1], 0, self.out_off[p], 1.0))
                offs[0] += pl[3].size_allowed
                offs[1] += pl[4].size_allowed
            progs = [Program(*_instantiate_many(parts[k])) for k in range(3)]
            self.flops += sum(pg.flops for pg in progs)
            self.stages.append(progs)
            self.inter = [max(a, b) for a, b in zip(self.inter, offs)]

    def sequence(self, dtype):
        """The three launches per op chunk as one recorded sequence over the
        slots (site, environment, MPO, t1, t2, out, workspace); None when a
        launch cannot be recorded."""
        seqs = self.__dict__.setdefault('_seqs', {})
        if dtype not in seqs:
            seq = _lib.Seq()
            work, ok = 0, True
            for st in self.stages:
                for pg, a, b, c in ((st[0], 0, 1, 3), (st[1], 2, 3, 4), (st[2], 4, 0, 5)):
                    w = seq_add_program(seq, pg, dtype, a, b, c, 6)
                    if w is None:
                        ok = False
                    else:
                        work = max(work, w)
            seqs[dtype] = (seq, work) if ok else None
        return seqs[dtype]

    def apply(self, mpo, site_values, env_in, dtype=0):
        """Return the new EnvStore (values on the device)."""
        # intermediates in the scratch buffer the sigma launches also use
        s1 = ((self.inter[0] * 8 + 255) // 256) * 256
        buf = dv.scratch('inter', s1 + self.inter[1] * 8)
        t1 = dv.typed(buf, 0, self.inter[0], dtype)
        t2 = dv.typed(buf, s1, self.inter[1], dtype)
        out = dv.zeros(self.out_size, dtype)
        W = mpo.buffer(self.site_idx, dtype)
        E = env_in.buffer(dtype)
        if getattr(out, 'is_cuda', False):
            # tiles of the three launches (autotune.py: the measured table,
            # or tuning on the device in 'online' mode)
            for st in self.stages:
                for k, (pg, a, b, c) in enumerate(((st[0], site_values, E, t1),
                                                   (st[1], W, t1, t2), (st[2], t2, site_values, out))):
                    autotune.choose(pg, 'renorm%d' % (k + 1), dtype,
                                    lambda pg=pg, a=a, b=b, c=c: run_program(pg, a, b, c, dtype))
        sq = self.sequence(dtype) if SIGMA_SEQ and getattr(out, 'is_cuda', False) else None
        if sq is not None:
            seq, nwork = sq
            work = _workspace(nwork, dtype) if nwork else None
            _lib.check(_lib.lib().zmpo_renorm(
                seq._h, dtype, dv.ptr(site_values), dv.ptr(E), dv.ptr(W), dv.ptr(t1),
                dv.ptr(t2), dv.ptr(out), dv.ptr(work) if work is not None else None,
                dv.stream()))
        else:
            for st in self.stages:
                run_program(st[0], site_values, E, t1, dtype)
                run_program(st[1], W, t1, t2, dtype)
                run_program(st[2], t2, site_values, out, dtype)
        if dtype != 0:
            o64 = dv.empty(self.out_size, 0)
            _lib.check(_lib.lib().zmpo_convert(0, dv.ptr(o64), dtype, dv.ptr(out),
                                               self.out_size, dv.stream()))
            out = o64
        return EnvStore(self.meta, out, self.out_off)


def renorm(mpo, site_idx, qsite, env_in, status, dtype=0, budget=DEFAULT_BUDGET):
    """One renormalisation step; qsite.value may be host or device."""
    prog = RenormProgram(mpo, site_idx, qsite, env_in, status, budget)
    v = qsite.value
    v = dv.upload(v, dtype) if isinstance(v, np.ndarray) else (v if dtype == 0 else v.float())
    return prog.apply(mpo, v, env_in, dtype), prog


def genHops(mpo, sites, status, nqnum=2, fit=None):
    """All environments from an MPS (initQt.py:88-271).  Returns (envs, exphop).
    ``fit(envs, i)``, when given, runs after environment i is built (the
    sweep driver's device-memory policy; it keeps envs[i] resident)."""
    K = mpo.nsite
    envs = {}
    if status == 'L':
        envs[-1] = boundary_store(mpo.nops, nqnum)
        for i in range(K):
            envs[i], _ = renorm(mpo, i, sites[i], envs[i - 1], 'L')
            if fit is not None:
                fit(envs, i)
        last = envs[K - 1]
    else:
        envs[K] = boundary_store(mpo.nops, nqnum)
        for i in range(K - 1, -1, -1):
            envs[i], _ = renorm(mpo, i, sites[i], envs[i + 1], 'R')
            if fit is not None:
                fit(envs, i)
        last = envs[0]
    vals = last.buf if isinstance(last.buf, np.ndarray) else dv.download(last.buf)
    exphop = np.array([vals[last.off[p]] if last.meta[p].size_allowed else 0.0
                       for p in range(mpo.nops)])
    return envs, exphop


# ------------------------------------------------------------------ noise RDM
class PRDMProgram:
    """Noise term of the decimation RDM (pRDMQt, kernelQt.py:453-626):
    sum_p X_p X_p^T with, for status 'L', X_p[l, u, m] = sum L_p[w, l, l']
    W_p[i][w, wr, u, d] C[l', d, m] (m = (n2, r) with wr kept), and for 'R'
    X_p[u, r, m] = sum W_p[j][wl, wr, u, d] R_p[wr, r, r'] C[m, d, r']
    (m = (l, n1) with wl kept).  Three launches per op chunk; the Gram stage
    of the ops of one case is one launch whose pairs are grouped per output
    block (ops in a fixed order).  C is the 3-leg template of
    dotutil.setupQtmpPRDM (dotutil.py:562-588)."""

    def __init__(self, mpo, isite, qvec, env, status, qkey, budget=DEFAULT_BUDGET):
        nops = mpo.nops
        self.status = status
        site = isite if status == 'L' else isite + 1
        self.site = site
        case_of = [case(mpo.opers[p], site) for p in range(nops)]
        plans = {}
        for p in range(nops):
            c = case_of[p]
            if c in plans:
                continue
            op = env.meta[p]
            cop = mpo.sites[site][p]
            t1, t2, t3 = QTensor(), QTensor(), QTensor()
            if status == 'L':
                t1.tensordotSYM(op, qvec, ([2], [0]))              # [w, l, n1, m]
                t2.tensordotSYM(t1, cop, ([0, 2], [0, 3]))         # [l, m, wr, u]
                t3.tensordotSYM(t2, t2, ([1, 2], [1, 2]))          # [l, u, l', u']
            else:
                rf = op.template()
                rf.qsyms = [-op.qsyms[0], -(qkey - op.qsyms[1]), qkey - op.qsyms[2]]
                rf.status = op.status.copy()
                rf.status[1] = not rf.status[1]
                rf.fromQsyms(3, rf.qsyms, rf.ndims, ifallocate=False, idlst=op.idlst)
                if not np.array_equal(rf.idx_allowed, op.idx_allowed):
                    raise RuntimeError('relabelled right environment has other blocks')
                t1.tensordotSYM(qvec, rf, ([2], [2]))              # [m, n2, w, r]
                t2.tensordotSYM(cop, t1, ([1, 3], [2, 1]))         # [wl, u, m, r]
                t3.tensordotSYM(t2, t2, ([0, 2], [0, 2]))          # [u, r, u', r']
            plans[c] = [t.plan.lowered() for t in (t1, t2, t3)] + [t1, t2, t3]
        self.plans, self.case_of = plans, case_of
        isz = [(plans[c][3].size_allowed + plans[c][4].size_allowed) * 8 for c in case_of]
        self.chunks = _chunks(isz, budget)
        self.cases = sorted(plans)
        self.out_off = {}
        tot = 0
        for c in self.cases:
            self.out_off[c] = tot
            tot += plans[c][5].size_allowed
        self.out_size = tot
        self.stages, self.inter = [], [0, 0]
        for ch in self.chunks:
            offs = [0, 0]
            parts = [[], []]
            gram = {c: [] for c in self.cases}
            for p in ch:
                c = case_of[p]
                pl = plans[c]
                if status == 'L':
                    parts[0].append((pl[0], env.off[p], 0, offs[0], 1.0))
                    parts[1].append((pl[1], offs[0], mpo.off[site][p], offs[1], 1.0))
                else:
                    parts[0].append((pl[0], 0, env.off[p], offs[0], 1.0))
                    parts[1].append((pl[1], mpo.off[site][p], offs[0], offs[1], 1.0))
                gram[c].append((pl[2], offs[1], offs[1], self.out_off[c], 1.0))
                offs[0] += pl[3].size_allowed
                offs[1] += pl[4].size_allowed
            progs = [Program(*_instantiate_many(parts[k])) for k in range(2)]
            grams = [Program(*_merge_shared_many(gram[c], plans[c][5].nblks_allowed))
                     for c in self.cases if gram[c] and plans[c][5].nblks_allowed]
            self.stages.append((progs, grams))
            self.inter = [max(a, b) for a, b in zip(self.inter, offs)]

    def apply(self, mpo, env, cvals):
        """Device buffer of every case's Gram blocks (case c at out_off[c]);
        the intermediates live in the shared scratch buffer of the sigma
        and renormalisation launches."""
        s1 = int(self.inter[0]) * 8
        buf = dv.scratch('inter', s1 + int(self.inter[1]) * 8)
        t1 = dv.typed(buf, 0, self.inter[0], 0)
        t2 = dv.typed(buf, s1, self.inter[1], 0)
        out = dv.zeros(self.out_size, 0)
        E = env.buffer(0)
        W = mpo.buffer(self.site, 0)
        for progs, grams in self.stages:
            if self.status == 'L':
                run_program(progs[0], E, cvals, t1, 0)
                run_program(progs[1], t1, W, t2, 0)
            else:
                run_program(progs[0], cvals, E, t1, 0)
                run_program(progs[1], W, t1, t2, 0)
            for g in grams:   # each chunk adds its ops' share
                run_program(g, t2, t2, out, 0, beta=1)
        return out

    def dense(self, values, idlst):
        """Sum over cases of the Gram blocks as one dense matrix; idlst = the
        index lists of the two kept legs."""
        tot = None
        for c in self.cases:
            t3 = self.plans[c][5].template()
            o = self.out_off[c]
            t3.value = values[o:o + t3.size_allowed]
            d = t3.toDenseTensor([idlst[0], idlst[1], idlst[0], idlst[1]])
            tot = d if tot is None else tot + d
        n = tot.shape[0] * tot.shape[1]
        return tot.reshape(n, n)


def prdm_template(dmrg, key, status):
    """dotutil.setupQtmpPRDM (ncsite=2): the 3-leg C template and its index
    lists; returns (qvec, idlst)."""
    from .dot import dpt
    qkey = np.array(key, dtype=np.float64)
    if status == 'L':
        qnums = [dmrg.qlst[0], dmrg.qlst[1], qkey - dpt(dmrg.qlst[2], dmrg.qlst[3])]
    else:
        qnums = [dpt(dmrg.qlst[0], dmrg.qlst[1]), dmrg.qlst[2], qkey - dmrg.qlst[3]]
    qvec = QTensor([False, False, True])
    from .qtensor import fromQnums
    rank, qsyms, ndims, idlst = fromQnums(qnums)
    qvec.fromQsyms(rank, qsyms, ndims, ifallocate=False)
    qvec.idlst = [list(x) for x in idlst]
    return qvec, idlst


def prdm_noise(cimat, info):
    """sum over states and ops of X_p X_p^T (the operator term of pRDMQt,
    kernelQt.py:478-533 / 562-622), dense over the kept legs: (ldim*4)^2 for
    'L', (4*rdim)^2 for 'R'.  cimat: float64[neig, ldim, 4, 4*rdim] ('L') or
    [neig, ldim*4, 4, rdim] ('R'), dense product basis."""
    dmrg, isite, flst = _info_parts(info)
    status = info[4]
    env = flst[0][0] if status == 'L' else flst[0][1]
    keep = (0, 1) if status == 'L' else (1, 2)
    tm = dmrg.__dict__.setdefault('timing', {})

    def lap(key, t0):
        t = time.time()
        tm[key] = tm.get(key, 0.0) + t - t0
        return t
    out = None
    idx = 0
    for key, neig in dmrg.qsectors.items():
        t0 = time.time()
        qvec, idlst = prdm_template(dmrg, key, status)
        prog = PRDMProgram(dmrg.fhop, isite, qvec, env, status, np.array(key, dtype=np.float64),
                           getattr(dmrg, 'budget', DEFAULT_BUDGET))
        t0 = lap('noise_plan', t0)
        for _ in range(neig):
            q = qvec.template()
            q.fromDense(np.asarray(cimat[idx]), idlst)
            idx += 1
            vals = prog.apply(dmrg.fhop, env, dv.upload(q.value, 0))
            vals = dv.download(vals)
            t0 = lap('noise_apply', t0)
            d = prog.dense(vals, [idlst[k] for k in keep])
            out = d if out is None else out + d
            t0 = lap('noise_dense', t0)
    return out


def prdm_noise_device(C, neig, info, scale=1.0):
    """scale times the operator term of pRDMQt (prdm_noise) from the dense
    two-site vectors on the device (C: float64, neig rows of ldim*16*rdim
    in (l, n1, n2, r) order), as a device dense m x m matrix (m = ldim*4
    for 'L', 4*rdim for 'R'): the C template's values are gathered from C
    through its dense index map, the three launches of PRDMProgram run,
    and each case's Gram blocks are scattered into the dense matrix and
    summed -- no host copy of C or of the result.  One symmetry sector."""
    dmrg, isite, flst = _info_parts(info)
    status = info[4]
    env = flst[0][0] if status == 'L' else flst[0][1]
    keep = (0, 1) if status == 'L' else (1, 2)
    if len(dmrg.qsectors) != 1:
        raise NotImplementedError('prdm_noise_device: one symmetry sector')
    tm = dmrg.__dict__.setdefault('timing', {})

    def lap(key, t0):
        dv.sync()
        t = time.time()
        tm[key] = tm.get(key, 0.0) + t - t0
        return t
    t0 = time.time()
    key = next(iter(dmrg.qsectors))
    qvec, idlst = prdm_template(dmrg, key, status)
    prog = PRDMProgram(dmrg.fhop, isite, qvec, env, status, np.array(key, dtype=np.float64),
                       getattr(dmrg, 'budget', DEFAULT_BUDGET))
    kl = [idlst[k] for k in keep]
    m = int(qvec.shape[keep[0]]) * int(qvec.shape[keep[1]])
    ntot = int(np.prod([int(x) for x in qvec.shape]))
    cmap = dv.upload(qvec.prjmap(idlst).astype(np.int64))
    gmaps = {}
    for c in prog.cases:
        t3 = prog.plans[c][5].template()
        if t3.size_allowed:
            gmaps[c] = (t3.size_allowed, dv.upload(t3.prjmap([kl[0], kl[1], kl[0], kl[1]])
                                                    .astype(np.int64)))
    t0 = lap('noise_plan', t0)
    out = dv.zeros(m * m, 0)
    tmp = dv.empty(max([n for n, _ in gmaps.values()] + [1]), 0)
    coef = dv.upload(np.ones(1), 0)
    cv = dv.empty(qvec.size_allowed, 0)
    for e in range(neig):
        gather(cv, C[e * ntot:(e + 1) * ntot], cmap)
        vals = prog.apply(dmrg.fhop, env, cv)
        # out[map] += scale * blocks, case by case (a case's map has no
        # repeated position; cases overlap, hence one case at a time)
        for c, (n, gm) in gmaps.items():
            o = prog.out_off[c]
            gather(tmp[:n], out, gm)
            _lib.check(_lib.lib().zmpo_lincomb(0, dv.ptr(tmp), dv.ptr(vals[o:o + n]), n, 1,
                                               dv.ptr(coef), float(scale), 1.0, n, dv.stream()))
            _lib.check(_lib.lib().zmpo_scatter(0, dv.ptr(out), 0, dv.ptr(tmp), dv.ptr(gm), n,
                                               dv.stream()))
    lap('noise_apply', t0)
    return out


def pRDMQt(cimat, info):
    """Reduced density matrix of the kept side (kernelQt.py:453-626): sum
    over states of C C^T plus dmrg.noise times the operator term
    (prdm_noise) when dmrg.noise >= 1e-10; dense, shapes as prdm_noise."""
    dmrg = info[0]
    status = info[4]
    rdm = None
    for c in np.asarray(cimat):
        m = c.reshape(c.shape[0] * c.shape[1], -1) if status == 'L' else \
            c.reshape(c.shape[0], -1)
        plain = m @ m.T if status == 'L' else m.T @ m
        rdm = plain if rdm is None else rdm + plain
    noise = float(getattr(dmrg, 'noise', 0.0))
    if noise >= 1e-10:
        rdm = rdm + noise * prdm_noise(cimat, info)
    return rdm


# ------------------------------------------------------------------ drop-in API
def _info_parts(info):
    dmrg, isite, ncsite, flst, status, ifsym = info
    if ncsite != 2:
        raise NotImplementedError('the device path implements the two-site (ncsite=2) kernels')
    return dmrg, isite, flst


def _structure_key(dmrg, isite, qtmp, fL, fR):
    """Bytes that fix a sigma program's plans: the MPO store and its weights,
    the intermediate budget and the launch switches, the site, the psi
    template's legs, every distinct environment template's legs, which
    template each operator uses, and the per-operator offsets of both
    environments."""
    budget = getattr(dmrg, 'budget', DEFAULT_BUDGET)
    parts = [np.array([id(dmrg.fhop), budget, int(FUSED), SIGMA_CHUNKS, CHAIN_RB, isite],
                      np.int64).tobytes(),
             np.ascontiguousarray(dmrg.hpwts, np.float64).tobytes()]
    tid = {}
    for m in list(fL.meta) + list(fR.meta):
        tid.setdefault(id(m), (len(tid), m))
    for t in [qtmp] + [m for _, m in tid.values()]:
        parts.append(np.int64(len(t.qsyms)).tobytes())
        for q, d in zip(t.qsyms, t.ndims):
            parts += [np.ascontiguousarray(q).tobytes(), np.ascontiguousarray(d).tobytes(), b'|']
    parts += [np.array([tid[id(m)][0] for m in fL.meta], np.int64).tobytes(), b'#',
              np.array([tid[id(m)][0] for m in fR.meta], np.int64).tobytes(), b'#',
              np.asarray(fL.off, np.int64).tobytes(), b'/', np.asarray(fR.off, np.int64).tobytes()]
    return hashlib.sha1(b''.join(parts)).hexdigest()


def _sigma_program(dmrg, isite, fL, fR):
    key = (isite, id(dmrg.qtmp), id(fL), id(fR))
    cache = getattr(dmrg, '_sigma_cache', None)
    if cache is None or cache[0] != key:
        if cache is not None:
            # intermediates and the uploaded plan tables are re-created on
            # use: a cached program keeps only host arrays
            cache[1]._buffers.clear()
            cache[1].release_device()
        prog = None
        if PLAN_CACHE:
            store = dmrg.__dict__.setdefault('_plan_cache', collections.OrderedDict())
            skey = _structure_key(dmrg, isite, dmrg.qtmp, fL, fR)
            prog = store.get(skey)
            if prog is not None:
                store.move_to_end(skey)
                dmrg.plan_hits = getattr(dmrg, 'plan_hits', 0) + 1
        if prog is None:
            prog = SigmaProgram(dmrg.fhop, isite, dmrg.qtmp, dmrg.qkey, fL, fR, dmrg.hpwts,
                                getattr(dmrg, 'budget', DEFAULT_BUDGET))
            if PLAN_CACHE:
                store[skey] = prog
                while len(store) > 2 * dmrg.nsite:
                    store.popitem(last=False)
        dmrg._sigma_cache = (key, prog)
    return dmrg._sigma_cache[1]


def HVecQt(vec0, info, ndim0, prjmap, ifile=0, ioff=0):
    """sigma = sum_p hpwts[p] H_p psi for the two-site dot (kernelQt.py:153-319).

    vec0: float64[ndim0] in qtmp block order (numpy or device tensor); the
    result has the same kind.  flst[ifile][ioff:ioff+2] are the L/R EnvStores.
    """
    dmrg, isite, flst = _info_parts(info)
    fL, fR = flst[ifile][ioff:ioff + 2]
    prog = _sigma_program(dmrg, isite, fL, fR)
    dtype = getattr(dmrg, 'dtype', 0)
    host = isinstance(vec0, np.ndarray)
    psi = dv.upload(vec0, dtype) if host else vec0
    sig = dv.empty(ndim0, sigma_dtype(dtype))
    prog.apply(dmrg.fhop, fL, fR, psi, sig, dtype)
    return dv.download(sig).astype(np.float64) if host else sig


def HDiagQt(info, ndim0, prjmap):
    """Diagonal of H_eff in qtmp order (kernelQt.py:30-127); returns float64 numpy."""
    dmrg, isite, flst = _info_parts(info)
    fL, fR = flst[0][0:2]
    prog = DiagProgram(dmrg.fhop, isite, dmrg.qtmp, fL, fR, dmrg.hpwts)
    out = dv.zeros(ndim0, 0)
    prog.apply(fL, fR, out)
    return dv.download(out)


def renormQt_H00(dmrg, isite, ncsite, flst, flstN, site, status):
    """New environments for every op (kernelQt.py:633-767).

    The result EnvStore is stored into flstN[0][0] (the reference writes the
    HDF5 file opened there).  Returns 0.
    """
    if status == 'L':
        env_in = flst[0][0]
        idx = isite
    else:
        env_in = flst[0][1]
        idx = isite + ncsite - 1
    # environments stay FP64 in every precision form: the FP64 residual
    # correction of a low-precision Davidson solve needs an H_eff that is
    # symmetric and accurate to FP64 (FP32-rounded environments leave a
    # ~1e-7 relative asymmetry that bounds the corrected energy)
    new, _ = renorm(dmrg.fhop, idx, site, env_in, status, getattr(dmrg, 'renorm_dtype', 0),
                    getattr(dmrg, 'budget', DEFAULT_BUDGET))
    flstN[0][0] = new
    return 0
