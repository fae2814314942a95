"""Block-sparse quantum-number tensor with device-resident values.

Same block tables and value layout as the reference's qtensor
(source/qtensor/qtensor.py:108-346): sectors per leg, C-ordered blocks, an
allowed-block mask from the block rule (libzmpo ``zmpo_symAllowed``, the
replacement of ctypes/qsym.c), values concatenated over allowed blocks in
ascending block index.  Two differences, both internal to this package:

* ``perm`` -- the storage order of the legs inside every block (identity =
  the reference's row-major layout).  Environments and sigma intermediates
  use non-identity orders so that the dominant GEMMs read K-contiguous
  operands; ``toDenseTensor``/``fromDense`` honour it.
* ``value`` -- a device tensor (torch, float64/float32) or a host numpy
  array; ``tensordotCAL`` requires device values and runs every block pair
  of the plan in one libzmpo grouped-GEMM launch.

``tensordotSYM`` builds the same pair list as the reference (same output
block order, same internal-tuple order, qtensor.py:563-637) with integer
numpy, and lowers it to GEMM tasks (``lower_plan``).
"""
import collections
import os

import numpy as np

from . import _lib
from . import device as dv


def classification(qnums, ifclt=1):
    """Sectors in first-occurrence order (qtensor_util.py:34-51)."""
    rows = [tuple(float(x) for x in q) for q in qnums]
    if not ifclt:
        return (np.array([list(r) for r in rows], dtype=np.float64),
                np.ones(len(rows), dtype=np.int64), [[i] for i in range(len(rows))])
    order, groups = [], {}
    for i, r in enumerate(rows):
        if r not in groups:
            groups[r] = []
            order.append(r)
        groups[r].append(i)
    return (np.array([list(r) for r in order], dtype=np.float64),
            np.array([len(groups[r]) for r in order], dtype=np.int64), [groups[r] for r in order])


def fromQnums(qnums, ifcollect=None):
    """qtensor_util.py:20-31."""
    ifclt = [1] * len(qnums) if ifcollect is None else ifcollect
    cls = [classification(q, c) for q, c in zip(qnums, ifclt)]
    return len(qnums), [c[0] for c in cls], [c[1] for c in cls], [c[2] for c in cls]


# Allowed-block lists by (sector counts, leg directions, sector quantum
# numbers): a sweep step rebuilds the same tensor templates many times (about
# five calls per distinct table), so the block rule runs once per table.
_BLOCKS = collections.OrderedDict()
_BLOCKS_MAX = 4096


def _allowed_blocks(nsyms, status, qsyms):
    """Flat indices of the allowed blocks (ascending), via zmpo_symAllowed."""
    key = (np.asarray(nsyms, np.int64).tobytes(), bytes(int(bool(x)) for x in status),
           tuple(q.shape[1] for q in qsyms), b''.join(q.tobytes() for q in qsyms))
    idx = _BLOCKS.get(key)
    if idx is None:
        idx = np.flatnonzero(_lib.sym_allowed(nsyms, status, qsyms))
        _BLOCKS[key] = idx
        if len(_BLOCKS) > _BLOCKS_MAX:
            _BLOCKS.popitem(last=False)
    else:
        _BLOCKS.move_to_end(key)
    return idx.copy()


class QTensor:
    def __init__(self, status=None, perm=None):
        self.status = None if status is None else np.array(status, dtype=bool)
        self.perm = None if perm is None else tuple(perm)
        self.rank = None
        self.qsyms = None
        self.ndims = None
        self.idlst = None
        self.value = None
        self.plan = None

    # ---------------------------------------------------------- tables
    def fromQsyms(self, rank, qsyms, ndims, ifallocate=True, idlst=None):
        """qtensor.py:272-346 (mask from libzmpo)."""
        self.rank = rank
        self.qsyms = [np.array(q, dtype=np.float64).reshape(len(d), -1) for q, d in
                      zip(qsyms, ndims)]
        self.ndims = [np.array(d, dtype=np.int64) for d in ndims]
        if idlst is not None:
            self.idlst = [list(x) for x in idlst]
        if self.perm is None:
            self.perm = tuple(range(rank))
        self.shape = np.array([int(d.sum()) for d in self.ndims], dtype=np.int64)
        self.size = int(np.prod(self.shape))
        self.nsyms = np.array([len(d) for d in self.ndims], dtype=np.int32)
        self.nblks = int(np.prod(self.nsyms))
        self.idx_allowed = _allowed_blocks(self.nsyms, self.status, self.qsyms)
        self.nblks_allowed = len(self.idx_allowed)
        ib = np.zeros(self.nblks, dtype=np.int32)
        ib[self.idx_allowed] = self.idx_allowed + 1
        self.iblks_allowed = ib.reshape(self.nsyms)
        # per allowed block (position j): sector tuple, dims, size, offset
        self.blk_isyms = np.array(np.unravel_index(self.idx_allowed, self.nsyms),
                                  dtype=np.int64).T.reshape(-1, rank)
        self.blk_dims = np.stack([self.ndims[i][self.blk_isyms[:, i]] for i in range(rank)],
                                 axis=1) if self.nblks_allowed else np.zeros((0, rank), np.int64)
        self.blk_size = np.prod(self.blk_dims, axis=1).astype(np.int64)
        self.blk_off = np.concatenate([[0], np.cumsum(self.blk_size)[:-1]]).astype(np.int64) \
            if self.nblks_allowed else np.zeros(0, np.int64)
        self.size_allowed = int(self.blk_size.sum())
        self.value = np.zeros(self.size_allowed) if ifallocate else None
        return 0

    def blk_strides(self):
        """Element strides (nblks_allowed x rank) of each leg inside each block."""
        st = np.zeros_like(self.blk_dims)
        acc = np.ones(self.nblks_allowed, dtype=np.int64)
        for leg in reversed(self.perm):
            st[:, leg] = acc
            acc = acc * self.blk_dims[:, leg]
        return st

    def same_structure(self, other):
        return (self.rank == other.rank and np.array_equal(self.idx_allowed, other.idx_allowed)
                and all(np.array_equal(a, b) for a, b in zip(self.ndims, other.ndims)))

    def template(self, perm=None):
        """Metadata copy (no values) with an optional new storage order."""
        qt = QTensor(self.status, self.perm if perm is None else perm)
        qt.fromQsyms(self.rank, self.qsyms, self.ndims, ifallocate=False, idlst=self.idlst)
        return qt

    # ---------------------------------------------------------- host values
    def _host_value(self):
        v = self.value
        if v is None:
            return None
        if not isinstance(v, np.ndarray):
            v = dv.download(v)
        return v

    def _ix(self, j, idlst):
        isyms = self.blk_isyms[j]
        return np.ix_(*[idlst[i][isyms[i]] for i in range(self.rank)])

    def _dense_positions(self, idlst):
        """(dense flat indices, value positions) of every stored entry, vectorised."""
        shp = [int(x) for x in self.shape]
        sec, loc = [], []
        for i in range(self.rank):
            s_i = np.zeros(shp[i], np.int64)
            l_i = np.zeros(shp[i], np.int64)
            for k, ids in enumerate(idlst[i]):
                s_i[ids] = k
                l_i[ids] = np.arange(len(ids))
            sec.append(s_i)
            loc.append(l_i)
        bshape = [[1] * self.rank for _ in range(self.rank)]
        for i in range(self.rank):
            bshape[i][i] = shp[i]
        nsy = [int(x) for x in self.nsyms]
        flat = np.zeros(shp, np.int64)
        mult = 1
        for i in reversed(range(self.rank)):
            flat = flat + sec[i].reshape(bshape[i]) * mult
            mult *= nsy[i]
        j = self.iblks_allowed.reshape(-1)[flat] - 1
        mask = j >= 0
        jm = np.searchsorted(self.idx_allowed, j[mask])
        st = self.blk_strides()
        pos = self.blk_off[jm].copy()
        idx = np.nonzero(mask)
        for i in range(self.rank):
            pos += loc[i][idx[i]] * st[jm, i]
        return np.flatnonzero(mask.reshape(-1)), pos

    def fromDense(self, val, idlst):
        """qtensor.py:364-379 (honours perm)."""
        out = np.zeros(self.size_allowed)
        if self.size <= 20_000_000:
            d, pos = self._dense_positions(idlst)
            out[pos] = np.asarray(val).reshape(-1)[d]
        else:
            for j in range(self.nblks_allowed):
                blk = np.asarray(val[self._ix(j, idlst)]).transpose(self.perm)
                out[self.blk_off[j]:self.blk_off[j] + self.blk_size[j]] = blk.reshape(-1)
        self.value = out
        return 0

    def fromDenseTensor(self, val, qnums, ifcollect=None):
        """qtensor.py:356-362."""
        rank, qsyms, ndims, idlst = fromQnums(qnums, ifcollect)
        self.fromQsyms(rank, qsyms, ndims, ifallocate=False)
        self.fromDense(val, idlst)
        self.idlst = [list(x) for x in idlst]
        return 0

    def toDenseTensor(self, idlst=None):
        """qtensor.py:382-398 (honours perm)."""
        idlst = self.idlst if idlst is None else idlst
        v = self._host_value()
        val = np.zeros(self.shape)
        if self.size <= 20_000_000:
            d, pos = self._dense_positions(idlst)
            val.reshape(-1)[d] = v[pos]
            return val
        inv = np.argsort(self.perm)
        for j in range(self.nblks_allowed):
            dm = self.blk_dims[j]
            blk = v[self.blk_off[j]:self.blk_off[j] + self.blk_size[j]].reshape(
                dm[list(self.perm)]).transpose(inv)
            val[self._ix(j, idlst)] = blk
        return val

    def toDenseDiag(self, idlst=None):
        """Diagonal over the last two legs (qtensor.py:401-411)."""
        full = self.toDenseTensor(idlst)
        return np.einsum('...ii->...i', full)

    def prjmap(self, idlst):
        """Dense flat index of every stored entry, storage order (qtensor.py:442-453)."""
        out = []
        for j in range(self.nblks_allowed):
            isyms = self.blk_isyms[j]
            ind = [np.asarray(idlst[i][isyms[i]]) for i in range(self.rank)]
            grids = np.meshgrid(*[ind[i] for i in self.perm], indexing='ij')
            full = [None] * self.rank
            for k, leg in enumerate(self.perm):
                full[leg] = grids[k].reshape(-1)
            out.append(np.ravel_multi_index(full, self.shape))
        return np.concatenate(out) if out else np.zeros(0, dtype=np.int64)

    # ---------------------------------------------------------- contraction
    def tensordotSYM(self, qt1, qt2, axes, perm=None):
        """Symbolic plan with the reference's pair order (qtensor.py:563-637).

        ``perm`` sets the storage order of the output blocks.  Returns the
        number of allowed output blocks.
        """
        i1, i2 = list(axes[0]), list(axes[1])
        e1 = sorted(set(range(qt1.rank)) - set(i1))
        e2 = sorted(set(range(qt2.rank)) - set(i2))
        ne1, ne2 = len(e1), len(e2)
        for a, b in zip(i1, i2):
            if not np.array_equal(qt1.qsyms[a], qt2.qsyms[b]):
                raise ValueError('tensordotSYM: contracted legs carry different sectors')
        ints1 = [bool(qt1.status[i]) for i in i1]
        ints2 = [bool(qt2.status[i]) for i in i2]
        exts1 = [bool(qt1.status[i]) for i in e1]
        exts2 = [bool(qt2.status[i]) for i in e2]
        if i1:
            if ints1[0] == ints2[0]:
                ints2 = [not x for x in ints2]
                exts2 = [not x for x in exts2]
            if any(a == b for a, b in zip(ints1, ints2)):
                raise ValueError('tensordotSYM: contracted legs must pair in/out')
        self.status = np.array(exts1 + exts2, dtype=bool)
        self.perm = tuple(range(ne1 + ne2)) if perm is None else tuple(perm)
        self.fromQsyms(ne1 + ne2, [qt1.qsyms[i] for i in e1] + [qt2.qsyms[i] for i in e2],
                       [qt1.ndims[i] for i in e1] + [qt2.ndims[i] for i in e2],
                       ifallocate=False)
        ns1 = qt1.nsyms
        ns2 = qt2.nsyms
        nint = int(np.prod([ns1[i] for i in i1])) if i1 else 1
        ib1 = qt1.iblks_allowed.transpose(e1 + i1).reshape(-1, nint)
        ib2 = qt2.iblks_allowed.transpose(e2 + i2).reshape(-1, nint)
        iso = self.blk_isyms
        x1 = (np.ravel_multi_index(iso[:, :ne1].T, [ns1[i] for i in e1]) if ne1
              else np.zeros(len(iso), np.int64))
        x2 = (np.ravel_multi_index(iso[:, ne1:].T, [ns2[i] for i in e2]) if ne2
              else np.zeros(len(iso), np.int64))
        sel = (ib1[x1] > 0) & (ib2[x2] > 0)
        o, k = np.nonzero(sel)
        a_flat = ib1[x1[o], k].astype(np.int64) - 1
        b_flat = ib2[x2[o], k].astype(np.int64) - 1
        self.plan = Plan(qt1, qt2, self, (i1, i2), e1, e2, o.astype(np.int64),
                         np.searchsorted(qt1.idx_allowed, a_flat),
                         np.searchsorted(qt2.idx_allowed, b_flat))
        return self.nblks_allowed

    def tensordotCAL(self, qt1, qt2, ifc1=False, ifc2=False, thresh=1e-20, tsize=10000,
                     dtype=_lib.F64):
        """Run the plan on the device in one grouped-GEMM launch (qtensor.py:642-659).

        Real arithmetic: ifc1/ifc2 (complex conjugation) are no-ops.  The
        reference's skip of all-zero blocks (< thresh, > tsize entries) is
        not applied; such blocks contribute nothing.
        """
        if self.plan is None:
            raise ValueError('tensordotCAL before tensordotSYM')
        prog = self.plan.lowered()
        out = dv.zeros(self.size_allowed, dtype)
        a = _as_device(qt1.value, dtype)
        b = _as_device(qt2.value, dtype)
        run_program(prog, a, b, out, dtype=dtype, beta=0)
        self.value = out
        return 0


def _as_device(v, dtype):
    if isinstance(v, np.ndarray):
        return dv.upload(v, dtype)
    want = dv.tdtype(dtype)
    return v if v.dtype == want else v.to(want)


def _mergeable(qt, outer, inner, always_unit):
    """Legs outer, inner adjacent in qt's storage order (ignoring always-unit legs)."""
    p = list(qt.perm)
    a, b = p.index(outer), p.index(inner)
    if a >= b:
        return False
    return all(always_unit[x] for x in p[a + 1:b])


class Plan:
    """Pair list of one contraction: rows (out block j, qt1 block ja, qt2 block jb)."""

    def __init__(self, qt1, qt2, out, axes, e1, e2, jo, ja, jb):
        self.qt1, self.qt2, self.out = qt1, qt2, out
        self.axes, self.e1, self.e2 = axes, e1, e2
        self.jo, self.ja, self.jb = jo, ja, jb
        self._lowered = None

    def flops(self):
        """Algorithmic FLOPs: sum over pairs of 2 * free1 * contracted * free2."""
        lw = self.lowered()
        p = lw['pairs']
        t = lw['tasks']
        rows_task = np.repeat(np.arange(len(t)), (t['p_end'] - t['p_begin']).astype(np.int64))
        return int(np.sum(2 * t['M'][rows_task].astype(np.int64) * t['N'][rows_task]
                          * p['K'].astype(np.int64)))

    def lowered(self):
        if self._lowered is None:
            self._lowered = lower_plan(self)
        return self._lowered


def _unit_legs(qt):
    return [bool(np.all(qt.blk_dims[:, i] <= 1)) if qt.nblks_allowed else True
            for i in range(qt.rank)]


def _groups(legs_x, legs_c, qx, qc, ux, uc):
    """Split a free-leg list into runs that are storage-adjacent in both X and C."""
    groups = [[0]] if legs_x else []
    for j in range(1, len(legs_x)):
        if (_mergeable(qx, legs_x[j - 1], legs_x[j], ux)
                and _mergeable(qc, legs_c[j - 1], legs_c[j], uc)):
            groups[-1].append(j)
        else:
            groups.append([j])
    if len(groups) > 2:
        raise NotImplementedError('more than two independent free legs in one block GEMM')
    return groups


def _apply(groups, dims, strides):
    """(lo_dim, hi_stride, lo_stride, total) of a <=2-level index over merged legs."""
    nrow = dims.shape[0]
    one = np.ones(nrow, np.int64)
    zero = np.zeros(nrow, np.int64)
    if not groups:
        return one, zero, zero, one
    gd = [np.prod(dims[:, g], axis=1) for g in groups]
    gs = [strides[:, g[-1]] for g in groups]
    if len(groups) == 1:
        return gd[0], zero, gs[0], gd[0]
    return gd[1], gs[0], gs[1], gd[0] * gd[1]


def lower_plan(plan):
    """Lower a pair list to libzmpo tasks/pairs (offsets relative to each value array)."""
    qa, qb, qc = plan.qt1, plan.qt2, plan.out
    i1, i2 = plan.axes
    e1, e2 = plan.e1, plan.e2
    ua, ub, uc = _unit_legs(qa), _unit_legs(qb), _unit_legs(qc)
    sa, sb, sc = qa.blk_strides(), qb.blk_strides(), qc.blk_strides()
    jo, ja, jb = plan.jo, plan.ja, plan.jb
    da, db = qa.blk_dims[ja], qb.blk_dims[jb]
    # free legs of A <-> out legs 0..ne1-1 ; free legs of B <-> out legs ne1..
    mA = [l for l in e1 if not (ua[l] and uc[e1.index(l)])]
    mC = [e1.index(l) for l in mA]
    nB = [l for l in e2 if not (ub[l] and uc[len(e1) + e2.index(l)])]
    nC = [len(e1) + e2.index(l) for l in nB]
    gm = _groups(mA, mC, qa, qc, ua, uc)
    gn = _groups(nB, nC, qb, qc, ub, uc)
    # contracted legs (pairs), active ones only, merged into one K index
    kp = [(a, b) for a, b in zip(i1, i2) if not (ua[a] and ub[b])]
    if len(kp) > 1:
        order = None
        for cand in (kp, kp[::-1]):
            if all(_mergeable(qa, cand[j - 1][0], cand[j][0], ua)
                   and _mergeable(qb, cand[j - 1][1], cand[j][1], ub)
                   for j in range(1, len(cand))):
                order = cand
                break
        if order is None:
            raise NotImplementedError('contracted legs not mergeable in both operands')
        kp = order
    nrow = len(jo)
    if kp:
        K = np.prod(da[:, [a for a, _ in kp]], axis=1)
        a_sk = sa[ja][:, kp[-1][0]]
        b_sk = sb[jb][:, kp[-1][1]]
    else:
        K = np.ones(nrow, np.int64)
        a_sk = np.zeros(nrow, np.int64)
        b_sk = np.zeros(nrow, np.int64)
    a_mlo, a_hi, a_lo, _ = _apply(gm, da[:, mA], sa[ja][:, mA])
    b_nlo, b_hi, b_lo, _ = _apply(gn, db[:, nB], sb[jb][:, nB])
    # task (output block) geometry from the output tensor, same leg groups
    nout = qc.nblks_allowed
    dc = qc.blk_dims
    scs = qc.blk_strides()
    c_mlo, c_hi, c_lo, Mt = _apply(gm, dc[:, mC], scs[:, mC])
    n_lo, cn_hi, cn_lo, Nt = _apply(gn, dc[:, nC], scs[:, nC])
    if len(gn) == 2 and np.any((Nt > n_lo) & (cn_hi != cn_lo * n_lo)):
        raise NotImplementedError('output N index is not single-strided')
    pairs = np.zeros(nrow, dtype=_lib.PAIR_DTYPE)
    pairs['a_off'] = qa.blk_off[ja]
    pairs['b_off'] = qb.blk_off[jb]
    pairs['scale'] = 1.0
    pairs['K'] = K
    pairs['a_mlo'] = np.maximum(a_mlo, 1)
    pairs['a_smhi'] = a_hi
    pairs['a_smlo'] = a_lo
    pairs['a_sk'] = a_sk
    pairs['b_nlo'] = np.maximum(b_nlo, 1)
    pairs['b_snhi'] = b_hi
    pairs['b_snlo'] = b_lo
    pairs['b_sk'] = b_sk
    counts = np.bincount(jo, minlength=nout)
    pbeg = np.concatenate([[0], np.cumsum(counts)[:-1]]) if nout else np.zeros(0, np.int64)
    tasks = np.zeros(nout, dtype=_lib.TASK_DTYPE)
    tasks['c_off'] = qc.blk_off
    tasks['M'] = Mt
    tasks['N'] = Nt
    tasks['c_mlo'] = np.maximum(c_mlo, 1)
    tasks['c_smhi'] = c_hi
    tasks['c_smlo'] = c_lo
    tasks['c_sn'] = cn_lo if len(gn) else 0
    tasks['p_begin'] = pbeg
    tasks['p_end'] = pbeg + counts
    finish_tasks(tasks)
    return {'tasks': tasks, 'pairs': pairs}


def finish_tasks(tasks):
    """Fill tiles_n and tile_base (exclusive prefix sum of 64x64 tiles)."""
    T = _lib.TILE
    tm = (tasks['M'].astype(np.int64) + T - 1) // T
    tn = (tasks['N'].astype(np.int64) + T - 1) // T
    tasks['tiles_n'] = np.maximum(tn, 1)
    nt = np.maximum(tm, 1) * np.maximum(tn, 1)
    tasks['tile_base'] = np.concatenate([[0], np.cumsum(nt)[:-1]]) if len(tasks) else 0
    return int(nt.sum()) if len(tasks) else 0


def make_schedule(tasks, pairs, rb, kchunk=2048, reuse_key=None, tile=None):
    """Units (one output tile x one pair chunk) and split-K groups, built
    by the library's host-side schedule builder (zmpo_schedule_build; the
    same units, order, slots and groups as make_schedule_np, which states
    them in numpy).  Returns (units, groups, number of split-K slots)."""
    TMs, TNs = (16 * rb, 16) if tile is None else tile
    t = np.ascontiguousarray(tasks, dtype=_lib.TASK_DTYPE)
    p = np.ascontiguousarray(pairs, dtype=_lib.PAIR_DTYPE)
    key = None if reuse_key is None else np.ascontiguousarray(reuse_key, dtype=np.int64)
    if key is not None and len(key) != len(t):
        raise ValueError('reuse_key needs one value per task')
    L = _lib.lib()
    counts = np.zeros(3, np.int64)
    kp = key.ctypes.data if key is not None else None
    _lib.check(L.zmpo_schedule_build(len(t), t.ctypes.data, p.ctypes.data, TMs, TNs, int(kchunk),
                                     kp, None, 0, None, 0, counts.ctypes.data))
    units = np.zeros(int(counts[0]), dtype=_lib.UNIT_DTYPE)
    groups = np.zeros(int(counts[1]), dtype=_lib.RGROUP_DTYPE)
    _lib.check(L.zmpo_schedule_build(len(t), t.ctypes.data, p.ctypes.data, TMs, TNs, int(kchunk),
                                     kp, units.ctypes.data, len(units),
                                     groups.ctypes.data if len(groups) else None, len(groups),
                                     counts.ctypes.data))
    return units, groups, int(counts[2])


def make_schedule_np(tasks, pairs, rb, kchunk=2048, reuse_key=None, tile=None):
    """Units (one output tile x one pair chunk) and split-K groups.

    The tile is (16*rb) x 16 for the wave kernel, or tile=(TM, TN) for the
    shaped kernel.
    Slots of one split tile are consecutive (task, tile, chunk order); a task
    whose summed K exceeds kchunk is split between pairs.  With ``reuse_key``
    (one int per task: tasks with equal keys read the same B blocks in the
    same pair order) the execution order becomes (key, chunk, task, tile), so
    co-resident waves stream the same B slice through one L2.
    Pairs must be stored task by task (p_begin ascending), as lower_plan and
    the program builders produce them.
    """
    nt = len(tasks)
    M = tasks['M'].astype(np.int64)
    N = tasks['N'].astype(np.int64)
    pb = tasks['p_begin'].astype(np.int64)
    pe = tasks['p_end'].astype(np.int64)
    K = pairs['K'].astype(np.int64)
    TMs, TNs = (16 * rb, 16) if tile is None else tile
    tm = np.maximum((M + TMs - 1) // TMs, 1)
    tn = np.maximum((N + TNs - 1) // TNs, 1)
    cnt = pe - pb
    tp = np.repeat(np.arange(nt), cnt)
    # pair indices task by task: p_begin[t] + 0 .. cnt[t]-1 (vectorised ranges)
    tot = int(cnt.sum())
    pidx = np.repeat(pb, cnt) + (np.arange(tot, dtype=np.int64)
                                 - np.repeat(np.cumsum(cnt) - cnt, cnt))
    if len(pidx):
        kk = K[pidx]
        csum = np.cumsum(kk)
        first = np.concatenate([[0], np.cumsum(cnt)[:-1]])[tp]
        start = np.concatenate([[0], csum])[first]
        cid = (csum - kk - start) // kchunk
        new = np.ones(len(kk), bool)
        new[1:] = (tp[1:] != tp[:-1]) | (cid[1:] != cid[:-1])
        cst = np.flatnonzero(new)
        cen = np.append(cst[1:], len(kk))
        ctask = tp[cst]
        cb, ce = pidx[cst], pidx[cen - 1] + 1
    else:
        ctask = np.zeros(0, np.int64)
        cb = ce = np.zeros(0, np.int64)
    empty = np.flatnonzero(cnt == 0)
    ctask = np.concatenate([ctask, empty])
    cb = np.concatenate([cb, pb[empty]])
    ce = np.concatenate([ce, pb[empty]])
    order = np.argsort(ctask, kind='stable')
    ctask, cb, ce = ctask[order], cb[order], ce[order]
    nch = np.bincount(ctask, minlength=nt)
    cfirst = np.concatenate([[0], np.cumsum(nch)[:-1]])
    per = tm * tn * nch
    ut = np.repeat(np.arange(nt), per)
    within = np.arange(int(per.sum())) - np.repeat(np.concatenate([[0], np.cumsum(per)[:-1]]), per)
    # the divisions are skipped where every divisor is 1 (no split-K, one
    # tile column): the common case of the large launches
    if nt and nch.max() == 1:
        tilei, chunk = within, np.zeros(len(within), np.int64)
    else:
        nchu = nch[ut]
        tilei, chunk = np.divmod(within, nchu)
    units = np.zeros(len(ut), dtype=_lib.UNIT_DTYPE)
    units['task'] = ut
    if nt and tn.max() == 1:
        units['m0'] = tilei * TMs
    else:
        tr, tc = np.divmod(tilei, tn[ut])
        units['m0'] = tr * TMs
        units['n0'] = tc * TNs
    gi = cfirst[ut] + chunk
    units['p_begin'] = cb[gi]
    units['p_end'] = ce[gi]
    split = nch[ut] > 1
    slots = -np.ones(len(ut), np.int64)
    slots[split] = np.arange(int(split.sum()))
    units['slot'] = slots
    first = split & (chunk == 0)
    if reuse_key is not None:
        key = np.asarray(reuse_key, dtype=np.int64)[ut]
        order = np.lexsort((tilei, ut, chunk, key))
        units = units[order]
        first_units = units[np.flatnonzero(first[order])]
    else:
        first_units = units[np.flatnonzero(first)]
    groups = np.zeros(len(first_units), dtype=_lib.RGROUP_DTYPE)
    groups['task'] = first_units['task']
    groups['m0'] = first_units['m0']
    groups['n0'] = first_units['n0']
    groups['slot_begin'] = first_units['slot']
    groups['nslots'] = nch[first_units['task']]
    return units, groups, int(split.sum())


SHAPES = {0: (64, 16), 1: (128, 16), 2: (32, 32), 3: (64, 32), 4: (16, 16), 5: (64, 64),
          # register-blocked panel tiles (zmpo_shapedgemm ZMPO_PANEL_*): same
          # schedule, 2x2 waves, k-contiguous LDS, BF16 on the 16x16x32 MFMA
          6: (64, 64), 7: (128, 64), 8: (64, 128), 9: (128, 128), 10: (32, 64), 11: (64, 32),
          12: (32, 160)}
PANEL_SHAPES = (6, 7, 8, 9, 10, 11, 12)
# panel tiles whose row count does not divide 256: k-contiguous operands only
KMAJOR_SHAPES = (12,)
# summed K per unit of the panel tiles: each unit of a long reduction (sigma
# x R: hundreds of K ~ 2500 pairs per output block) runs ~10 pairs, which
# keeps the split-K workspace to a few GB at K = 36, D = 1000
PANEL_KCHUNK = int(os.environ.get('ZMPO_PANEL_KCHUNK', '32768'))
# shapes the static cost model chooses among (the panel tiles are chosen by
# measurement, autotune.py)
MODEL_SHAPES = (0, 1, 2, 3, 4, 5)


def choose_shape(tasks, pairs):
    """Tile shape for the shaped kernel: minimise the modelled cost
    sum over units of (TM*TN*Kpad + per-unit overhead), Kpad = K rounded to 16."""
    M = tasks['M'].astype(np.int64)
    N = tasks['N'].astype(np.int64)
    cnt = (tasks['p_end'] - tasks['p_begin']).astype(np.int64)
    rows = np.repeat(np.arange(len(tasks)), cnt)
    kpad = np.bincount(rows, weights=((pairs['K'].astype(np.int64) + 15) // 16) * 16,
                       minlength=len(tasks)) if len(pairs) else np.zeros(len(tasks))
    best, cost_best = 0, None
    for sid in MODEL_SHAPES:
        tm_, tn_ = SHAPES[sid]
        ntile = np.maximum((M + tm_ - 1) // tm_, 1) * np.maximum((N + tn_ - 1) // tn_, 1)
        cost = float(np.sum(ntile * (tm_ * tn_ * np.maximum(kpad, 16) + 4096)))
        if cost_best is None or cost < cost_best:
            best, cost_best = sid, cost
    return best


def choose_rb(tasks):
    """Row blocks per wave: the largest rb whose tile padding keeps at least
    85% of the best achievable useful fraction."""
    M = tasks['M'].astype(np.int64)
    N = tasks['N'].astype(np.int64)
    useful = float(np.sum(M * N)) or 1.0
    util = {}
    for rb in (1, 2, 4):
        tm = np.maximum((M + 16 * rb - 1) // (16 * rb), 1)
        tn = np.maximum((N + 15) // 16, 1)
        util[rb] = useful / float(np.sum(tm * 16 * rb * tn * 16))
    best = max(util.values())
    return max(rb for rb in util if util[rb] >= 0.85 * best)


STREAM_KT = int(os.environ.get('ZMPO_KT', '32'))
STREAM_UPW = int(os.environ.get('ZMPO_UPW', '0'))


SINGLE = 1 << 30   # "mlo" of a single-level index map (no index reaches it)


def pair_tasks(tasks, npairs):
    """Task index of every pair from the tasks' [p_begin, p_end) ranges (any task order)."""
    cnt = (tasks['p_end'] - tasks['p_begin']).astype(np.int64)
    out = np.full(npairs, -1, np.int64)
    if cnt.sum():
        pos = np.repeat(tasks['p_begin'].astype(np.int64), cnt) + (
            np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt))
        out[pos] = np.repeat(np.arange(len(tasks)), cnt)
    return out


def collapse_maps(tasks, pairs):
    """Copies of the arrays with every 2-level index map that is really one
    stride (hi stride == lo extent * lo stride) turned into a single-level
    map, which the kernels apply without a division."""
    t, p = tasks.copy(), pairs.copy()
    for arr, lo, hi, sl in ((p, 'a_mlo', 'a_smhi', 'a_smlo'), (p, 'b_nlo', 'b_snhi', 'b_snlo'),
                            (t, 'c_mlo', 'c_smhi', 'c_smlo')):
        if len(arr) == 0:
            continue
        one = arr[hi].astype(np.int64) == arr[lo].astype(np.int64) * arr[sl]
        arr[lo][one] = SINGLE
        arr[hi][one] = 0
    return t, p


def _max_row_offset(lo, shi, slo, extent):
    """Largest in-block row offset (r // lo) * shi + (r % lo) * slo, r < extent."""
    lo = lo.astype(np.int64)
    ext = extent.astype(np.int64)
    hi = (ext - 1) // lo
    last = np.minimum(ext, lo) - 1
    return int(np.max(hi * np.abs(shi.astype(np.int64)) + last * np.abs(slo.astype(np.int64))))


def stream_units(tasks, units):
    """zmpo_sunit records: the units with their task geometry folded in."""
    t = tasks[units['task']]
    su = np.zeros(len(units), dtype=_lib.SUNIT_DTYPE)
    for f in ('c_off', 'c_smhi', 'c_mlo', 'c_smlo', 'c_sn', 'M', 'N'):
        su[f] = t[f]
    for f in ('m0', 'n0', 'p_begin', 'p_end', 'slot'):
        su[f] = units[f]
    return su


def stream_upw(nunits):
    """Units per workgroup: long ranges for big launches (>= ~4096 workgroups
    stay in flight), one unit each for small ones."""
    return int(min(64, max(1, nunits // 4096)))


class Program:
    """Task/pair arrays of one grouped-GEMM launch (uploaded on first use)."""

    def __init__(self, tasks, pairs, reuse_key=None, prefer=None):
        self.reuse_key = reuse_key
        self.prefer = prefer   # kernel to use when eligible (None: the library default)
        self.ntasks = len(tasks)
        self.npairs = len(pairs)
        self._raw = (tasks, pairs)
        self._host = None
        self._dev = None
        self._sched = None
        self._rb = self._shape = self._small = self._kmajor = None
        rows = np.repeat(np.arange(self.ntasks), (tasks['p_end'] - tasks['p_begin']).astype(np.int64))
        self.flops = int(np.sum(2 * tasks['M'][rows].astype(np.int64) * tasks['N'][rows]
                                * pairs['K'].astype(np.int64))) if self.npairs else 0
        self.kt = None    # stream kernel k-tile (None: library default)
        self.kt_dtype = None   # stream kernel k-tile per dtype code (overrides kt)
        self.upw = None   # stream kernel units per workgroup (None: by launch size)
        self.kchunk = None   # summed K per split-K unit (None: by kernel)
        # per kernel dtype code: (kernel, tile shape) replacing (prefer, shape)
        # for launches in that type (set by the autotuner)
        self.tiles = {}

    # The device-ready arrays and the schedule choices are derived on first
    # use (a sigma plan builds programs that only the unfused path launches).
    def _prepare(self):
        if self._host is None:
            t, p = collapse_maps(*self._raw)
            self.ntiles = finish_tasks(t) if self.ntasks else 0
            self._host = (t, p)
            self._raw = None
        return self._host

    @property
    def tasks_host(self):
        return self._prepare()[0]

    @property
    def pairs_host(self):
        return self._prepare()[1]

    @property
    def rb(self):
        if self._rb is None:
            self._rb = choose_rb(self.tasks_host) if self.ntasks else 1
        return self._rb

    @rb.setter
    def rb(self, v):
        self._rb = v

    @property
    def shape(self):
        if self._shape is None:
            self._shape = choose_shape(self.tasks_host, self.pairs_host) if self.ntasks else 0
        return self._shape

    @shape.setter
    def shape(self, v):
        self._shape = v

    @property
    def small_blocks(self):
        """Whether in-block row offsets fit 32 bits (the stream kernel's)."""
        if self._small is None:
            t, p = self.tasks_host, self.pairs_host
            rows = np.repeat(np.arange(self.ntasks), (t['p_end'] - t['p_begin']).astype(np.int64))
            self._small = bool(self.npairs == 0 or max(
                _max_row_offset(p['a_mlo'], p['a_smhi'], p['a_smlo'], t['M'][rows]),
                _max_row_offset(p['b_nlo'], p['b_snhi'], p['b_snlo'], t['N'][rows])) < 2 ** 31 - 1)
        return self._small

    def kmajor_ok(self):
        """Whether every pair reads both operands k-contiguously (a_sk ==
        b_sk == 1), as the KMAJOR_SHAPES tiles need."""
        if self._kmajor is None:
            p = self.pairs_host
            self._kmajor = bool(self.npairs == 0 or (np.all(p['a_sk'] == 1)
                                                     and np.all(p['b_sk'] == 1)))
        return self._kmajor

    def skinny_ok(self):
        """Whether zmpo_skinnygemm applies: K-contiguous single-level A rows."""
        if self.npairs == 0:
            return False
        t, p = self.tasks_host, self.pairs_host
        rows = np.repeat(np.arange(self.ntasks), (t['p_end'] - t['p_begin']).astype(np.int64))
        return bool(np.all(p['a_sk'] == 1) and np.all(p['a_mlo'] >= t['M'][rows]))

    def skinny_units(self, rb):
        """zmpo_unit records of the skinny kernel: (64*rb)-row strips x 64-column
        chunks of every task, each over all of the task's pairs."""
        t = self.tasks_host
        tm = 64 * rb
        ns = np.maximum((t['M'].astype(np.int64) + tm - 1) // tm, 1)
        nc = np.maximum((t['N'].astype(np.int64) + 63) // 64, 1)
        per = ns * nc
        ti = np.repeat(np.arange(self.ntasks), per)
        w = np.arange(int(per.sum())) - np.repeat(np.cumsum(per) - per, per)
        u = np.zeros(len(ti), _lib.UNIT_DTYPE)
        u['task'] = ti
        u['m0'] = (w // nc[ti]) * tm
        u['n0'] = (w % nc[ti]) * 64
        u['p_begin'] = t['p_begin'][ti]
        u['p_end'] = t['p_end'][ti]
        u['slot'] = -1
        return u

    def alg_bytes(self, elsize=8, beta=0):
        """Algorithmic HBM bytes of one launch: every distinct A and B block
        read once, every output block written once (and read once if beta)."""
        if self.npairs == 0:
            return 0
        t, p = self.tasks_host, self.pairs_host
        rows = np.repeat(np.arange(self.ntasks), (t['p_end'] - t['p_begin']).astype(np.int64))
        a_sz = t['M'][rows].astype(np.int64) * p['K']
        b_sz = t['N'][rows].astype(np.int64) * p['K']
        _, ia = np.unique(p['a_off'], return_index=True)
        _, ib = np.unique(p['b_off'], return_index=True)
        c = int(np.sum(t['M'].astype(np.int64) * t['N']))
        return elsize * (int(a_sz[ia].sum()) + int(b_sz[ib].sum()) + c * (2 if beta else 1))

    def device(self):
        if self._dev is None:
            self._dev = (dv.upload_bytes(self.tasks_host), dv.upload_bytes(self.pairs_host))
        return self._dev

    def release_device(self):
        """Free the uploaded tables and schedules; host arrays are kept and
        uploaded again on the next launch."""
        self._dev = None
        for sc in (self._sched or {}).values():
            sc.pop('units', None)
            sc.pop('groups', None)

    def launch_choice(self, dtype, kernel=None):
        """(kernel, tile shape) of a launch in kernel dtype code ``dtype``: an
        explicit kernel keeps the program's shape; otherwise the per-dtype
        choice (autotuner) or the program's preference.  Kernels that cannot
        run the program fall back to the shaped one."""
        shape = self.shape
        if kernel is None:
            kernel, shape = self.tiles.get(dtype, (self.prefer or KERNEL, self.shape))
        if shape in KMAJOR_SHAPES and not self.kmajor_ok():
            shape = 2
        if kernel == 'stream' and (not self.small_blocks or shape in PANEL_SHAPES):
            kernel = 'shaped'
        if kernel == 'skinny' and not self.skinny_ok():
            kernel = 'shaped'
        return kernel, shape

    def schedule(self, kernel='wave', shape=None):
        """Device units / split-K groups for one kernel and tile shape (built
        once each; shape None: the program's shape)."""
        shape = self.shape if shape is None else shape
        if self._sched is None:
            self._sched = {}
        key = (kernel, shape)
        if kernel == 'skinny' and key not in self._sched:
            u = self.skinny_units(SKINNY_RB)
            self._sched[key] = {'units_dev_host': u, 'nunits': len(u), 'rb': SKINNY_RB}
        if key not in self._sched:
            tile = None if kernel == 'wave' else (32, 32) if kernel == 'dot' else SHAPES[shape]
            kchunk = self.kchunk or (PANEL_KCHUNK if (kernel == 'shaped' and
                                                      shape in PANEL_SHAPES) else 2048)
            units, groups, nslots = make_schedule(self.tasks_host, self.pairs_host, self.rb,
                                                  kchunk=kchunk, reuse_key=self.reuse_key,
                                                  tile=tile)
            size = self.rb * 256 if kernel == 'wave' else tile[0] * tile[1]
            sc = {'nunits': len(units), 'ngroups': len(groups), 'nslots': nslots,
                  'slot_size': size, 'units_host': units, 'groups_host': groups}
            if kernel == 'stream':
                su = stream_units(self.tasks_host, units)
                sc['units_dev_host'] = su
                sc['kt'] = self.kt or STREAM_KT
                sc['upw'] = self.upw or STREAM_UPW or stream_upw(len(su))
            else:
                sc['units_dev_host'] = units
            self._sched[key] = sc
        sc = self._sched[key]
        if 'units' not in sc:
            sc['units'] = dv.upload_bytes(sc['units_dev_host'])
        if 'groups' not in sc and 'groups_host' in sc:
            sc['groups'] = dv.upload_bytes(sc['groups_host'])
        return sc


def _workspace(n, dtype):
    """Split-K workspace of n elements of dtype: a view of one byte buffer
    shared by every launch and operand type (launches on one stream use it
    one after the other)."""
    return dv.typed(dv.scratch('splitk', int(max(n, 1)) * dv.ELSIZE[dtype]), 0, n, dtype)


KERNEL = os.environ.get('ZMPO_KERNEL', 'shaped')
SKINNY_RB = int(os.environ.get('ZMPO_SKINNY_RB', '1'))


def run_program(prog, A, B, C, dtype=_lib.F64, beta=0, stream=None, kernel=None):
    """Launch one grouped block GEMM: C (=|+=) sum over pairs of A_pair B_pair.

    kernel='stream' runs workgroups over ranges of block-shaped tiles as one
    software pipeline, 'shaped' one LDS-staged tile per workgroup, 'wave' one
    wave per small tile, all with deterministic split-K;
    'tile' runs the 64x64 LDS-tiled workgroup kernel."""
    if isinstance(prog, dict):
        prog = Program(prog['tasks'], prog['pairs'])
    kernel, shape = prog.launch_choice(dtype, kernel)
    if prog.ntasks == 0:
        return
    tasks, pairs = prog.device()
    L = _lib.lib()
    st = dv.stream() if stream is None else stream
    if kernel == 'tile':
        _lib.check(L.zmpo_blockgemm(dtype, dv.ptr(tasks), prog.ntasks, dv.ptr(pairs),
                                    prog.ntiles, dv.ptr(A), dv.ptr(B), dv.ptr(C), beta, st))
        return
    sc = prog.schedule(kernel, shape)
    work = _workspace(sc['nslots'] * sc['slot_size'], dtype) if sc.get('nslots') else None
    if kernel == 'dot':
        _lib.check(L.zmpo_dotgemm(dtype, dv.ptr(tasks), dv.ptr(sc['units']), sc['nunits'],
                                  dv.ptr(sc['groups']), sc['ngroups'], dv.ptr(pairs), dv.ptr(A),
                                  dv.ptr(B), dv.ptr(C), dv.ptr(work), beta, st))
        return
    if kernel == 'skinny':
        _lib.check(L.zmpo_skinnygemm(dtype, sc['rb'], dv.ptr(tasks), dv.ptr(sc['units']),
                                     sc['nunits'], dv.ptr(pairs), dv.ptr(A), dv.ptr(B), dv.ptr(C),
                                     beta, st))
        return
    if kernel == 'stream':
        kt = prog.kt_dtype.get(dtype, sc['kt']) if prog.kt_dtype else sc['kt']
        _lib.check(L.zmpo_streamgemm(dtype, shape, kt, dv.ptr(sc['units']),
                                     sc['nunits'], sc['upw'], dv.ptr(tasks), dv.ptr(sc['groups']),
                                     sc['ngroups'], dv.ptr(pairs), dv.ptr(A), dv.ptr(B),
                                     dv.ptr(C), dv.ptr(work), beta, st))
    elif kernel == 'wave':
        _lib.check(L.zmpo_wavegemm(dtype, prog.rb, dv.ptr(tasks), dv.ptr(sc['units']),
                                   sc['nunits'], dv.ptr(sc['groups']), sc['ngroups'],
                                   dv.ptr(pairs), dv.ptr(A), dv.ptr(B), dv.ptr(C), dv.ptr(work),
                                   beta, st))
    else:
        _lib.check(L.zmpo_shapedgemm(dtype, shape, dv.ptr(tasks), dv.ptr(sc['units']),
                                     sc['nunits'], dv.ptr(sc['groups']), sc['ngroups'],
                                     dv.ptr(pairs), dv.ptr(A), dv.ptr(B), dv.ptr(C),
                                     dv.ptr(work), beta, st))


def program_kernel(prog, kernel=None, dtype=None):
    """The grouped-GEMM kernel run_program would use for prog (in kernel
    dtype code ``dtype``)."""
    return prog.launch_choice(dtype, kernel)[0]


def seq_add_program(seq, prog, dtype, a, b, c, work, beta=0, out_f32=False):
    """Record prog (as run_program would launch it in dtype) into seq with
    operand slots a, b, c and split-K workspace slot work.  Returns the
    workspace elements it needs, or None if its kernel cannot be recorded
    (only the shaped and stream kernels can)."""
    if prog.ntasks == 0:
        return 0
    kernel, shape = prog.launch_choice(dtype)
    if kernel not in ('shaped', 'stream'):
        return None
    tasks, pairs = prog.device()
    sc = prog.schedule(kernel, shape)
    rec = np.zeros(1, _lib.SEQ_OP_DTYPE)
    rec['kind'] = _lib.SEQ_STREAM if kernel == 'stream' else _lib.SEQ_SHAPED
    rec['shape'] = shape
    if kernel == 'stream':
        kt = prog.kt_dtype.get(dtype, sc['kt']) if prog.kt_dtype else sc['kt']
        rec['kt'], rec['upw'] = kt, sc['upw']
    rec['tasks'], rec['pairs'] = dv.ptr(tasks), dv.ptr(pairs)
    rec['units'], rec['nunits'] = dv.ptr(sc['units']), sc['nunits']
    rec['groups'], rec['ngroups'] = dv.ptr(sc['groups']), sc['ngroups']
    rec['slot_a'], rec['slot_b'], rec['slot_b2'], rec['slot_c'] = a, b, -1, c
    rec['slot_work'] = work if sc.get('nslots') else -1
    rec['beta'], rec['out_f32'] = beta, int(bool(out_f32))
    seq.add(rec)
    return sc['nslots'] * sc['slot_size'] if sc.get('nslots') else 0


def seq_add_chain(seq, prog, a, b1, b2, c, beta=0):
    """Record every launch of a ChainProgram (run_chain) into seq."""
    if prog.nunits == 0:
        return
    (g, bl, t, p), units = prog.device()
    for l_, u in zip(prog.launches, units):
        rec = np.zeros(1, _lib.SEQ_OP_DTYPE)
        rec['kind'] = _lib.SEQ_CHAIN16 if l_.get('n16') else _lib.SEQ_CHAIN
        rec['rb'], rec['atile'], rec['kmax'], rec['nmax'] = prog.rb, l_['atile'], l_['kmax'], l_['nmax']
        rec['tasks'], rec['pairs'] = dv.ptr(t), dv.ptr(p)
        rec['units'], rec['nunits'] = dv.ptr(u), len(l_['units'])
        rec['groups'], rec['blocks'] = dv.ptr(g), dv.ptr(bl)
        rec['slot_a'], rec['slot_b'], rec['slot_b2'], rec['slot_c'] = a, b1, b2, c
        rec['slot_work'] = -1
        rec['beta'] = beta
        seq.add(rec)


def _map_lookup(offs, bmap):
    """(new offset, row stride) of every B block offset through a padded-copy map."""
    if len(offs) == 0:
        return np.zeros((0, 2), np.int64)
    ks = np.fromiter(bmap.keys(), np.int64, len(bmap))
    vs = np.array(list(bmap.values()), np.int64).reshape(-1, 2)
    o = np.argsort(ks)
    ks, vs = ks[o], vs[o]
    i = np.searchsorted(ks, offs)
    if np.any(i >= len(ks)) or np.any(ks[np.minimum(i, len(ks) - 1)] != offs):
        raise RuntimeError('chained plans: B block missing from the padded MPO map')
    return vs[i]


def _segments(sizes):
    """Exclusive prefix sum and, per element, its segment's start."""
    sizes = np.asarray(sizes, np.int64)
    starts = np.cumsum(sizes) - sizes
    return starts, np.repeat(starts, sizes)


def chain_structure(first, second, mid_key, out_key, mid_off, b1_map, b2_map):
    """Row-group structure of two chained plans (one operator case).

    ``first`` lowers X = A B1 (its tasks are the intermediate blocks, in
    block order, with c_off = mid_off), ``second`` lowers C = X B2 (pair
    a_off = an intermediate block offset).  mid_key / out_key give the row
    group of every intermediate / output block; the rows of blocks of one
    group coincide.  b1_map / b2_map send every B block offset to its
    zero-padded copy (offset, row stride): the padded B blocks are
    [ceil4(K)][ceil32(N)] row-major.  A-tile and intermediate columns are
    reserved in multiples of 4 (the padding columns hold zeros).  Returns
    (groups, tasks, pairs, blocks) for zmpo_chaingemm with offsets relative
    to each operand's base; groups without output blocks are dropped (their
    intermediate is never read).  Vectorised (numpy only).
    """
    t1, p1 = first['tasks'], first['pairs']
    t2, p2 = second['tasks'], second['pairs']
    mid_key = np.asarray(mid_key, np.int64)
    out_key = np.asarray(out_key, np.int64)
    mid_off = np.asarray(mid_off, np.int64)
    keys = np.unique(out_key)
    ng = len(keys)
    grp_of_mid = np.full(len(t1), -1, np.int64)
    gi = np.searchsorted(keys, mid_key)
    ok = (gi < ng) & (keys[np.minimum(gi, max(ng - 1, 0))] == mid_key) if ng else np.zeros(len(t1), bool)
    grp_of_mid[ok] = gi[ok]
    # intermediate column offsets: cumulative widths (multiples of 4) per group
    order1 = np.lexsort((np.arange(len(t1)), grp_of_mid))
    order1 = order1[grp_of_mid[order1] >= 0]
    g1 = grp_of_mid[order1]
    n1 = np.bincount(g1, minlength=ng).astype(np.int64)
    w1 = (t1['N'][order1].astype(np.int64) + 3) // 4 * 4
    col = np.zeros(len(t1), np.int64)
    csum = np.cumsum(w1) - w1
    _, gstart1 = _segments(n1)
    col[order1] = csum - csum[gstart1] if len(g1) else csum
    ncols = np.bincount(g1, weights=w1, minlength=ng).astype(np.int64)
    grp_of_out = np.searchsorted(keys, out_key)
    order2 = np.lexsort((np.arange(len(t2)), grp_of_out))
    g2 = grp_of_out[order2]
    n2 = np.bincount(g2, minlength=ng).astype(np.int64)
    # the intermediate block each second-stage pair reads
    pa = np.searchsorted(mid_off, p2['a_off'].astype(np.int64))
    if len(pa) and (np.any(pa >= len(mid_off)) or
                    np.any(mid_off[np.minimum(pa, len(mid_off) - 1)] != p2['a_off'])):
        raise RuntimeError('chained plans: a second-stage pair reads no intermediate block')
    rows2 = pair_tasks(t2, len(p2))
    if len(pa) and np.any(grp_of_mid[pa] != grp_of_out[rows2]):
        raise RuntimeError('chained plans: a second-stage pair reads another row group')
    rows1 = pair_tasks(t1, len(p1))
    if len(p1) and np.any(p1['b_nlo'] < t1['N'][rows1]) or \
            len(p2) and np.any(p2['b_nlo'] < t2['N'][rows2]):
        raise NotImplementedError('chained plans: B column map is not single-level')
    if np.any(p1['scale'] != 1.0) or np.any(p2['scale'] != 1.0):
        raise NotImplementedError('chained plans: pair scales must be 1')
    # pairs with B re-pointed at the padded MPO copies; second-stage A = columns
    P = np.concatenate([p1, p2])
    nb = np.concatenate([_map_lookup(p1['b_off'].astype(np.int64), b1_map),
                         _map_lookup(p2['b_off'].astype(np.int64), b2_map)])
    if len(P):
        P['b_off'] = nb[:, 0]
        P['b_sk'] = nb[:, 1]
        P['b_snlo'] = 1
        P['b_snhi'] = 0
        P['b_nlo'] = 1 << 30
        P['a_off'][len(p1):] = col[pa]
    # tasks: per group its first-stage tasks, then its second-stage tasks
    tb1 = np.cumsum(n1 + n2) - (n1 + n2)
    ntask = int((n1 + n2).sum())
    tasks = np.zeros(ntask, _lib.TASK_DTYPE)
    pos1 = tb1[g1] + (np.arange(len(g1)) - gstart1) if len(g1) else np.zeros(0, np.int64)
    _, gstart2 = _segments(n2)
    pos2 = tb1[g2] + n1[g2] + (np.arange(len(g2)) - gstart2) if len(g2) else np.zeros(0, np.int64)
    tasks[pos1] = t1[order1]
    tasks[pos2] = t2[order2]
    tasks['c_off'][pos1] = col[order1]
    src_pb = np.zeros(ntask, np.int64)
    src_pb[pos1] = t1['p_begin'][order1]
    src_pb[pos2] = t2['p_begin'][order2].astype(np.int64) + len(p1)
    cnt = (tasks['p_end'] - tasks['p_begin']).astype(np.int64)
    new_pb, seg = _segments(cnt)
    gather = np.repeat(src_pb, cnt) + (np.arange(int(cnt.sum())) - seg)
    pairs = P[gather]
    tasks['p_begin'] = new_pb
    tasks['p_end'] = new_pb + cnt
    stage_of_task = np.zeros(ntask, np.int64)
    stage_of_task[pos2] = 1
    group_of_task = np.zeros(ntask, np.int64)
    group_of_task[pos1] = g1
    group_of_task[pos2] = g2
    prow = np.repeat(np.arange(ntask), cnt)
    # distinct first-stage A blocks (each top1 block belongs to one group),
    # ordered by group then first use, at A-tile columns of width ceil4(K)
    s1 = np.flatnonzero(stage_of_task[prow] == 0)
    if len(s1):
        ao = pairs['a_off'][s1].astype(np.int64)
        uniq, first_i, inv = np.unique(ao, return_index=True, return_inverse=True)
        bgrp = group_of_task[prow[s1[first_i]]]
        border = np.lexsort((first_i, bgrp))
        rank = np.empty(len(uniq), np.int64)
        rank[border] = np.arange(len(uniq))
        q = s1[first_i[border]]
        blocks = np.zeros(len(uniq), _lib.CBLOCK_DTYPE)
        blocks['a_off'] = uniq[border]
        for f in ('a_smhi', 'a_smlo', 'a_sk', 'K', 'a_mlo'):
            blocks[f] = pairs[f][q]
        bg = bgrp[border]
        kw = (blocks['K'].astype(np.int64) + 3) // 4 * 4
        nblk = np.bincount(bg, minlength=ng).astype(np.int64)
        bstart, bseg = _segments(nblk)
        kc = np.cumsum(kw) - kw
        blocks['col'] = kc - kc[bseg]
        kcols = np.bincount(bg, weights=kw, minlength=ng).astype(np.int64)
        pairs['a_off'][s1] = blocks['col'][rank[inv]]
        pairs['pad'][s1] = rank[inv]
        ab = np.append(bstart, len(uniq))
    else:
        blocks = np.zeros(0, _lib.CBLOCK_DTYPE)
        kcols = np.zeros(ng, np.int64)
        ab = np.zeros(ng + 1, np.int64)
        bg = np.zeros(0, np.int64)
    groups = np.zeros(ng, _lib.CGROUP_DTYPE)
    Mg = np.zeros(ng, np.int64)
    Mg[g2] = t2['M'][order2]
    if len(g1) and np.any(t1['M'][order1] != Mg[g1]):
        raise RuntimeError('chained plans: blocks of one row group differ in rows')
    groups['M'] = Mg
    groups['ncols'] = ncols
    groups['t1_begin'] = tb1
    groups['t1_end'] = tb1 + n1
    groups['t2_begin'] = tb1 + n1
    groups['t2_end'] = tb1 + n1 + n2
    groups['kcols'] = kcols
    groups['ab_begin'] = ab[:-1]
    groups['ab_end'] = ab[1:]
    # every block and output task of a group splits its rows the same way:
    # 2-level maps (mlo < M) must agree; single-level ones (mlo >= M, e.g. a
    # bra sector of dimension 1) are rewritten to that split, which leaves
    # their offsets unchanged ((m / s) * s * lo + (m % s) * lo = m * lo)
    t2i = np.flatnonzero(stage_of_task == 1)
    vals = np.concatenate([blocks['a_mlo'].astype(np.int64), tasks['c_mlo'][t2i].astype(np.int64)])
    vgrp = np.concatenate([bg, group_of_task[t2i]])
    two = vals < Mg[vgrp] if len(vals) else np.zeros(0, bool)
    lo_ = np.full(ng, np.iinfo(np.int64).max)
    hi_ = np.full(ng, -1, np.int64)
    np.minimum.at(lo_, vgrp[two], vals[two])
    np.maximum.at(hi_, vgrp[two], vals[two])
    if np.any((hi_ >= 0) & (hi_ != lo_)):
        raise NotImplementedError('chained plans: row maps differ inside a row group')
    rsplit = np.where(hi_ >= 0, hi_, np.maximum(Mg, 1))
    groups['rsplit'] = rsplit
    if len(blocks):
        one = blocks['a_mlo'] >= Mg[bg]
        blocks['a_smhi'] = np.where(one, rsplit[bg] * blocks['a_smlo'], blocks['a_smhi'])
        blocks['a_mlo'] = np.where(one, rsplit[bg], blocks['a_mlo'])
        r = np.maximum(Mg[bg], 1) - 1
        big = (r // blocks['a_mlo']) * blocks['a_smhi'] + np.minimum(r, blocks['a_mlo'] - 1) * \
            blocks['a_smlo'] + (blocks['K'].astype(np.int64) - 1) * blocks['a_sk']
        if big.max() >= 2 ** 31:
            raise NotImplementedError('chained plans: an A block exceeds 2^31 elements')
    if len(t2i):
        tg = group_of_task[t2i]
        one = tasks['c_mlo'][t2i] >= Mg[tg]
        tasks['c_smhi'][t2i] = np.where(one, rsplit[tg] * tasks['c_smlo'][t2i], tasks['c_smhi'][t2i])
        tasks['c_mlo'][t2i] = np.where(one, rsplit[tg], tasks['c_mlo'][t2i])
    # inside each stage of each group, costly tasks first: the waves take
    # tasks round-robin, so this balances them (pair ranges are unchanged)
    kp = (pairs['K'].astype(np.int64) + 3) // 4
    cost = np.bincount(prow, weights=kp, minlength=ntask) * ((tasks['N'] + 31) // 32)
    segid = 2 * group_of_task + stage_of_task
    tasks = tasks[np.lexsort((np.arange(ntask), -cost, segid))]
    return groups, tasks, pairs, blocks


def balance_waves(groups, tasks, pairs):
    """Order of the chained launch's tasks that evens out its four waves.

    Wave w of a workgroup runs the tasks at positions lo + w, lo + w + 4, ...
    of each stage range [lo, hi) of its row group, so it owns ceil((n - w) / 4)
    of the n tasks.  Inside every range the tasks are taken by cost
    (k-steps x 32-column tiles, plus one flush per tile) in descending order,
    each to the least-loaded wave with a free position (ties: the wave with
    fewer positions).  Returns perm with new_tasks = tasks[perm]; every task
    stays in its range and keeps its pairs and output, so results do not
    change.  At K = 36, D = 1000 the summed per-unit maximum wave cost drops
    from 1.29x to 1.20x its mean."""
    n = len(tasks)
    perm = np.arange(n)
    if n == 0 or len(groups) == 0:
        return perm
    lo = np.concatenate([groups['t1_begin'], groups['t2_begin']]).astype(np.int64)
    hi = np.concatenate([groups['t1_end'], groups['t2_end']]).astype(np.int64)
    keep = hi - lo > 1
    lo, hi = lo[keep], hi[keep]
    if len(lo) == 0:
        return perm
    ks = (pairs['K'].astype(np.int64) + 3) // 4
    cum = np.concatenate([[0], np.cumsum(ks)])
    ntile = (tasks['N'].astype(np.int64) + 31) // 32
    cost = (cum[tasks['p_end']] - cum[tasks['p_begin']] + 1) * ntile
    ln = hi - lo
    nr = len(lo)
    rid = np.repeat(np.arange(nr), ln)
    first = np.concatenate([[0], np.cumsum(ln)[:-1]])
    idx = lo[rid] + np.arange(int(ln.sum())) - first[rid]
    order = idx[np.lexsort((-cost[idx], rid))]
    cap = (ln[:, None] - np.arange(4)[None, :] + 3) // 4
    cnt = np.zeros((nr, 4), np.int64)
    load = np.zeros((nr, 4), np.int64)
    big = np.iinfo(np.int64).max
    for j in range(int(ln.max())):
        act = np.flatnonzero(ln > j)
        t = order[first[act] + j]
        ld = np.where(cnt[act] < cap[act], load[act], big)
        w = 3 - np.argmin(ld[:, ::-1], axis=1)
        perm[lo[act] + 4 * cnt[act, w] + w] = t
        cnt[act, w] += 1
        load[act, w] += cost[t]
    return perm


# LDS size classes (A-tile + intermediate columns): one launch per class
CHAIN_CLASSES = tuple(int(x) for x in os.environ.get('ZMPO_CHAIN_CLASSES', '96,160,256').split(','))
CHAIN_ATILE = os.environ.get('ZMPO_CHAIN_ATILE', '0')   # '1', '0' or 'auto' (by class)
# 16-column output tiles (zmpo_chaingemm16) for row groups whose blocks are
# all <= 16 columns wide: 'auto', '1' (every group) or '0' (none)
CHAIN_N16 = os.environ.get('ZMPO_CHAIN_N16', 'auto')
# reorder each row group's tasks to balance the four waves (balance_waves)
CHAIN_BALANCE = os.environ.get('ZMPO_CHAIN_BALANCE', '1') != '0'


class ChainProgram:
    """zmpo_chaingemm launches: the row-group structures of several cases
    instantiated for many operators (units = row tiles of 16*rb rows).
    Units are split by the LDS their group needs (kcols + ncols, classes
    CHAIN_CLASSES), one launch per class, so small groups run at full
    occupancy; every output element is still written by exactly one unit."""

    def __init__(self, structures, instances, rb=1):
        """structures: {case: (groups, tasks, pairs, blocks)}; instances: list
        of (case, a_base, b1_base, b2_base, c_base)."""
        self.rb = rb
        self._structs, self._inst = structures, instances
        tm = 16 * rb
        gparts, tparts, pparts, bparts = [], [], [], []
        gbase, tbase, pbase, bbase = {}, 0, 0, 0
        self._gspan = {}
        ng = 0
        for c in sorted(structures):
            g, t, p, b = structures[c]
            g = g.copy()
            t = t.copy()
            for f in ('t1_begin', 't1_end', 't2_begin', 't2_end'):
                g[f] += tbase
            for f in ('ab_begin', 'ab_end'):
                g[f] += bbase
            t['p_begin'] += pbase
            t['p_end'] += pbase
            p = p.copy()
            p['pad'] += bbase
            gbase[c] = ng
            self._gspan[c] = (ng, len(g))
            ng += len(g)
            tbase += len(t)
            pbase += len(p)
            bbase += len(b)
            gparts.append(g)
            tparts.append(t)
            pparts.append(p)
            bparts.append(b)
        self.groups_host = np.concatenate(gparts) if gparts else np.zeros(0, _lib.CGROUP_DTYPE)
        self.tasks_host = np.concatenate(tparts) if tparts else np.zeros(0, _lib.TASK_DTYPE)
        self.pairs_host = np.concatenate(pparts) if pparts else np.zeros(0, _lib.PAIR_DTYPE)
        self.blocks_host = np.concatenate(bparts) if bparts else np.zeros(0, _lib.CBLOCK_DTYPE)
        if CHAIN_BALANCE:
            self.tasks_host = self.tasks_host[balance_waves(self.groups_host, self.tasks_host,
                                                            self.pairs_host)]
        uparts = []
        self.flops = 0
        t, p = self.tasks_host, self.pairs_host
        prow = pair_tasks(t, len(p))
        pfl = 2 * t['M'][prow].astype(np.int64) * t['N'][prow] * p['K'].astype(np.int64)
        tfl = np.bincount(prow, weights=pfl, minlength=len(t)) if len(p) else np.zeros(len(t))
        self.instance_cases = [c for c, *_ in instances]
        templ = {}   # per case: unit records without the operator's bases, flops
        bases = []
        for c, a0, b10, b20, c0 in instances:
            g = structures[c][0]
            if len(g) == 0:
                continue
            if c not in templ:
                ntile = np.maximum((g['M'].astype(np.int64) + tm - 1) // tm, 1)
                gi = np.repeat(np.arange(len(g)), ntile)
                m0 = (np.arange(int(ntile.sum())) - np.repeat(np.cumsum(ntile) - ntile, ntile)) * tm
                u = np.zeros(len(gi), _lib.CUNIT_DTYPE)
                u['group'] = gi + gbase[c]
                u['m0'] = m0
                gl = gbase[c]
                templ[c] = (u, int(tfl[self.groups_host['t1_begin'][gl]:
                                       self.groups_host['t2_end'][gl + len(g) - 1]].sum()))
            u0, fl = templ[c]
            uparts.append(u0)
            bases.append((len(u0), a0, b10, b20, c0))
            self.flops += fl
        # one copy of every operator's units, then its bases set per field
        self.units_host = np.concatenate(uparts) if uparts else np.zeros(0, _lib.CUNIT_DTYPE)
        if bases:
            bs = np.array(bases, dtype=np.int64)
            for k, f in enumerate(('a_base', 'b1_base', 'b2_base', 'c_base')):
                self.units_host[f] = np.repeat(bs[:, k + 1], bs[:, 0])
        self.nunits = len(self.units_host)
        # LDS classes, by the columns a group keeps in LDS: the intermediate,
        # plus the A tile when it is staged there (CHAIN_ATILE '1')
        g = self.groups_host
        width = g['ncols'].astype(np.int64) + (g['kcols'].astype(np.int64)
                                               if CHAIN_ATILE == '1' else 0)
        cls = np.searchsorted(np.asarray(CHAIN_CLASSES), width[self.units_host['group']]) \
            if self.nunits else np.zeros(0, np.int64)
        # narrow row groups: every block of both stages at most 16 columns wide
        tN = self.tasks_host['N'].astype(np.int64)
        wide = np.zeros(len(g), bool)
        for lo, hi in (('t1_begin', 't1_end'), ('t2_begin', 't2_end')):
            cnt = (g[hi] - g[lo]).astype(np.int64)
            if len(tN) and cnt.sum():
                gid = np.repeat(np.arange(len(g)), cnt)
                tid = np.repeat(g[lo].astype(np.int64), cnt) + (
                    np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt))
                wide |= np.bincount(gid, weights=(tN[tid] > 16), minlength=len(g)) > 0
        n16 = {'1': np.ones(len(g), bool), '0': np.zeros(len(g), bool)}.get(CHAIN_N16, ~wide)
        unit_n16 = n16[self.units_host['group']] if self.nunits else np.zeros(0, bool)
        self.launches = []
        for k in range(len(CHAIN_CLASSES) + 1):
            for nar in (False, True):
                sel = np.flatnonzero((cls == k) & (unit_n16 == nar))
                if len(sel) == 0:
                    continue
                us = self.units_host[sel]
                gsel = us['group']
                # A blocks read from global memory (default: the LDS then holds
                # only the intermediate, so more workgroups fit per CU; measured
                # faster at K=36, D=1000) or staged in LDS; 'auto' stages them
                # for every class but the widest
                atile = {'1': 1, '0': 0}.get(CHAIN_ATILE, int(k < len(CHAIN_CLASSES)))
                self.launches.append({'units': us, 'kmax': int(max(1, g['kcols'][gsel].max())),
                                      'nmax': int(max(1, g['ncols'][gsel].max())),
                                      'atile': atile, 'n16': bool(nar)})
        self._dev = None

    def with_rb(self, rb):
        """The same launches with 16*rb-row tiles."""
        return self if rb == self.rb else ChainProgram(self._structs, self._inst, rb)

    def alg_bytes(self, elsize=8):
        """Algorithmic HBM bytes: the first-stage A blocks and the output
        blocks moved once, and every operator's B1/B2 blocks read once; the
        intermediate stays on chip."""
        if self.nunits == 0:
            return 0
        t, p, g = self.tasks_host, self.pairs_host, self.groups_host
        cols = g['kcols'].astype(np.int64).copy()
        for gi, gg in enumerate(g):
            cols[gi] += int(t['N'][gg['t2_begin']:gg['t2_end']].sum())
        u = self.units_host
        rows = np.minimum(16 * self.rb, g['M'][u['group']].astype(np.int64) - u['m0'])
        tot = int(np.sum(rows * cols[u['group']]))
        prow = pair_tasks(t, len(p))
        bsz = t['N'][prow].astype(np.int64) * p['K']
        stage2 = np.zeros(len(p), np.int64)
        for gg in g:
            for ti in range(gg['t2_begin'], gg['t2_end']):
                stage2[t['p_begin'][ti]:t['p_end'][ti]] = 1
        for c in self.instance_cases:
            gl, gn = self._gspan[c]
            if gn == 0:
                continue
            tr = slice(g['t1_begin'][gl], g['t2_end'][gl + gn - 1])
            lo = int(t['p_begin'][tr].min())
            hi = int(t['p_end'][tr].max())
            key = p['b_off'][lo:hi] * 2 + stage2[lo:hi]
            _, ib = np.unique(key, return_index=True)
            tot += int(bsz[lo:hi][ib].sum())
        return elsize * tot

    def device(self):
        if self._dev is None:
            common = tuple(dv.upload_bytes(x) for x in (self.groups_host, self.blocks_host,
                                                        self.tasks_host, self.pairs_host))
            self._dev = (common, [dv.upload_bytes(l_['units']) for l_ in self.launches])
        return self._dev


def run_chain(prog, A, B1, B2, C, dtype=_lib.F64, beta=0, stream=None):
    """Launch zmpo_chaingemm (one launch per LDS class): C (=|+=) (A B1) B2."""
    if prog.nunits == 0:
        return
    (g, b, t, p), units = prog.device()
    st = dv.stream() if stream is None else stream
    L = _lib.lib()
    for l_, u in zip(prog.launches, units):
        fn = L.zmpo_chaingemm16 if l_.get('n16') else L.zmpo_chaingemm
        _lib.check(fn(dtype, prog.rb, l_['atile'], l_['kmax'], l_['nmax'], dv.ptr(u),
                      len(l_['units']), dv.ptr(g), dv.ptr(b), dv.ptr(t), dv.ptr(p),
                      dv.ptr(A), dv.ptr(B1), dv.ptr(B2), dv.ptr(C), beta, st))


def gather(dst, src, idx, dtype_dst=_lib.F64, dtype_src=_lib.F64):
    """dst[i] = src[idx[i]] on the device (idx < 0 gives 0)."""
    n = idx.numel()
    _lib.check(_lib.lib().zmpo_gather(dtype_dst, dv.ptr(dst), dtype_src, dv.ptr(src), dv.ptr(idx),
                                      n, dv.stream()))
    return dst


def _scalar_contract(qt1, qt2, i1, i2):
    """Full contraction of two block-sparse tensors (the scalar case of
    qtensor.py:727-750): the matching allowed blocks' elements gathered into
    two device vectors, then one fixed-order dot product."""
    from . import pdvdson
    r = qt1.rank
    pair = [i2[list(i1).index(leg)] for leg in range(r)]   # leg of qt2 matching leg of qt1
    for leg in range(r):
        if not np.array_equal(qt1.qsyms[leg], qt2.qsyms[pair[leg]]):
            raise ValueError('tensordot: contracted legs carry different sectors')
    ns = [int(x) for x in qt1.nsyms]
    k1 = np.ravel_multi_index(qt1.blk_isyms.T, ns) if qt1.nblks_allowed else np.zeros(0, np.int64)
    k2 = np.ravel_multi_index(qt2.blk_isyms[:, pair].T, ns) if qt2.nblks_allowed else \
        np.zeros(0, np.int64)
    common, j1s, j2s = np.intersect1d(k1, k2, return_indices=True)
    st1, st2 = qt1.blk_strides(), qt2.blk_strides()
    idx1, idx2 = [], []
    for j1, j2 in zip(j1s, j2s):
        dims = qt1.blk_dims[j1]
        grid = np.indices(dims).reshape(r, -1)
        idx1.append(qt1.blk_off[j1] + (st1[j1][:, None] * grid).sum(0))
        idx2.append(qt2.blk_off[j2] + (st2[j2][pair][:, None] * grid).sum(0))
    n = int(sum(len(x) for x in idx1))
    if n == 0:
        return 0.0
    a = gather(dv.empty(n, 0), _as_device(qt1.value, 0), dv.upload(np.concatenate(idx1)))
    b = gather(dv.empty(n, 0), _as_device(qt2.value, 0), dv.upload(np.concatenate(idx2)))
    return float(pdvdson._VecOps(n).dots(a.view(1, -1), 1, b)[0])


def tensordot(qt1, qt2, axes, perm=None, dtype=_lib.F64):
    """One-shot contraction on the device (qtensor.py:665-827): the general
    case (some legs contracted), the outer product (none: K = 1 block GEMMs)
    and the scalar case (all legs of both: returns a float).  The output
    keeps the legs' index lists when both inputs have them."""
    i1, i2 = list(axes[0]), list(axes[1])
    if len(i1) == qt1.rank and len(i2) == qt2.rank:
        return _scalar_contract(qt1, qt2, i1, i2)
    out = QTensor()
    out.tensordotSYM(qt1, qt2, axes, perm=perm)
    out.tensordotCAL(qt1, qt2, dtype=dtype)
    if qt1.idlst is not None and qt2.idlst is not None:
        e1 = [k for k in range(qt1.rank) if k not in i1]
        e2 = [k for k in range(qt2.rank) if k not in i2]
        out.idlst = [qt1.idlst[k] for k in e1] + [qt2.idlst[k] for k in e2]
    return out
