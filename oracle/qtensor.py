"""Block-sparse quantum-number tensor (oracle restatement).

Follows source/qtensor/qtensor.py (class qtensor, :108-827) and
source/qtensor/qtensor_util.py (:20-90).  Test infrastructure only.

Layout (identical to the reference, qtensor.py:272-346):
  * leg i has sectors qsyms[i] (nsyms_i x nqnum, float) of sizes ndims[i]
  * blocks are indexed in C order over nsyms; a block is allowed when
    sum_{status=False} q - sum_{status=True} q == 0 (L1 norm < 1e-10,
    ctypes/qsym.c:24-75)
  * iblks_allowed[isyms] = idx+1 for allowed blocks, 0 otherwise
  * value = concatenation of allowed blocks in ascending idx, each row-major
"""
import itertools

import numpy as np


def classification(qnums, ifclt=1):
    """Group an index QN list into sectors ordered by first occurrence.

    qtensor_util.py:34-51.
    """
    qnums = [tuple(float(x) for x in q) for q in qnums]
    if ifclt:
        order = []
        groups = {}
        for idx, q in enumerate(qnums):
            if q not in groups:
                groups[q] = []
                order.append(q)
            groups[q].append(idx)
        qsyms = np.array([list(q) for q in order], dtype=np.float64)
        idlst = [groups[q] for q in order]
        ndims = np.array([len(g) for g in idlst], dtype=np.int64)
    else:
        qsyms = np.array([list(q) for q in qnums], dtype=np.float64)
        idlst = [[i] for i in range(len(qnums))]
        ndims = np.ones(len(qnums), dtype=np.int64)
    return qsyms, ndims, idlst


def fromQnums(qnums, ifcollect=None):
    """qtensor_util.py:20-31."""
    rank = len(qnums)
    ifclt = [1] * rank if ifcollect is None else ifcollect
    cls = [classification(q, c) for q, c in zip(qnums, ifclt)]
    return rank, [c[0] for c in cls], [c[1] for c in cls], [c[2] for c in cls]


def blks_allowed(nsyms, status, qsyms):
    """Closed-form block rule, vectorised (qtensor_util.py:53-67, qsym.c:24-75).

    Returns the C-ordered int32 mask of allowed blocks.
    """
    rank = len(nsyms)
    nq = qsyms[0].shape[1]
    diff = np.zeros(tuple(nsyms) + (nq,))
    for i in range(rank):
        shp = [1] * rank + [nq]
        shp[i] = nsyms[i]
        q = np.asarray(qsyms[i], dtype=np.float64).reshape(shp)
        diff = diff - q if status[i] else diff + q
    return (np.abs(diff).sum(axis=-1) < 1e-10).astype(np.int32).reshape(-1)


def cartesian_prod(arrays):
    """qtensor_util.py:121-169 (row = one tuple, last array fastest)."""
    arrays = [np.asarray(a) for a in arrays]
    return np.array(list(itertools.product(*arrays)), dtype=arrays[0].dtype).reshape(
        -1, len(arrays))


class QTensor:
    """qtensor.py:108-148."""

    def __init__(self, status=None):
        self.status = None if status is None else np.array(status, dtype=bool)
        self.rank = None
        self.qsyms = None
        self.ndims = None
        self.idlst = None
        self.value = None
        self.clst = None
        self.axes = None

    # -- construction -------------------------------------------------
    def fromQsyms(self, rank, qsyms, ndims, ifallocate=True, idlst=None):
        """qtensor.py:272-346."""
        self.rank = rank
        self.qsyms = [np.array(q, dtype=np.float64) for q in qsyms]
        self.ndims = [np.array(d, dtype=np.int64) for d in ndims]
        if idlst is not None:
            self.idlst = [list(x) for x in idlst]
        self.nqnum = self.qsyms[0].shape[1]
        self.shape = np.array([int(d.sum()) for d in self.ndims])
        self.size = int(np.prod(self.shape))
        self.nsyms = np.array([q.shape[0] for q in self.qsyms], dtype=np.int32)
        self.nblks = int(np.prod(self.nsyms))
        assert self.status is not None
        mask = blks_allowed(self.nsyms, self.status, self.qsyms)
        self.nblks_allowed = int(mask.sum())
        self.idx_allowed = np.flatnonzero(mask)
        ib = np.zeros(self.nblks, dtype=np.int32)
        ib[self.idx_allowed] = self.idx_allowed + 1
        self.iblks_allowed = ib.reshape(self.nsyms)
        self.dims_allowed = cartesian_prod(self.ndims).astype(np.int32).reshape(
            self.nblks, self.rank)
        self.ndim_allowed = np.prod(self.dims_allowed.astype(np.int64), axis=1)
        self.ioff_allowed = np.zeros(self.nblks, dtype=np.int64)
        sizes = self.ndim_allowed[self.idx_allowed]
        offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]) if len(sizes) else np.zeros(0)
        self.ioff_allowed[self.idx_allowed] = offs
        self.size_allowed = int(sizes.sum())
        self.value = np.zeros(self.size_allowed) if ifallocate else None
        return 0

    def fromDenseTensor(self, val, qnums, ifcollect=None):
        """qtensor.py:356-362."""
        rank, qsyms, ndims, idlst = fromQnums(qnums, ifcollect)
        self.fromQsyms(rank, qsyms, ndims)
        self.fromDense(val, idlst)
        self.idlst = [list(x) for x in idlst]
        return 0

    def _block_index(self, idx, idlst):
        isyms = np.unravel_index(idx, self.nsyms)
        return np.ix_(*[idlst[i][isyms[i]] for i in range(self.rank)])

    def fromDense(self, val, idlst):
        """qtensor.py:364-379."""
        for idx in self.idx_allowed:
            o, n = self.ioff_allowed[idx], self.ndim_allowed[idx]
            self.value[o:o + n] = val[self._block_index(idx, idlst)].reshape(n)
        return 0

    def toDenseTensor(self, idlst=None):
        """qtensor.py:382-398."""
        idlst = self.idlst if idlst is None else idlst
        val = np.zeros(self.shape)
        for idx in self.idx_allowed:
            o, n = self.ioff_allowed[idx], self.ndim_allowed[idx]
            val[self._block_index(idx, idlst)] = self.value[o:o + n].reshape(
                self.dims_allowed[idx])
        return val

    def toDenseDiag(self, idlst=None):
        """Diagonal over the last two legs (qtensor.py:401-411)."""
        idlst = self.idlst if idlst is None else idlst
        val = np.zeros(self.shape[:-1])
        for idx in self.idx_allowed:
            isyms = np.unravel_index(idx, self.nsyms)
            if isyms[-1] != isyms[-2]:
                continue
            o, n = self.ioff_allowed[idx], self.ndim_allowed[idx]
            ind = np.ix_(*[idlst[i][isyms[i]] for i in range(self.rank - 1)])
            val[ind] = np.einsum('...ii->...i', self.value[o:o + n].reshape(
                self.dims_allowed[idx]))
        return val

    def prjmap(self, idlst):
        """Dense flat index of every stored entry (qtensor.py:442-453)."""
        out = []
        for idx in self.idx_allowed:
            isyms = np.unravel_index(idx, self.nsyms)
            ind = [np.asarray(idlst[i][isyms[i]]) for i in range(self.rank)]
            grids = np.meshgrid(*ind, indexing='ij')
            out.append(np.ravel_multi_index([g.reshape(-1) for g in grids], self.shape))
        return np.concatenate(out) if out else np.zeros(0, dtype=np.int64)

    def copy(self):
        other = QTensor(self.status)
        other.fromQsyms(self.rank, self.qsyms, self.ndims, ifallocate=False,
                        idlst=self.idlst)
        other.value = None if self.value is None else self.value.copy()
        return other

    def transpose(self, *args):
        """qtensor.py:501-522."""
        qt = QTensor([self.status[i] for i in args])
        qt.fromQsyms(self.rank, [self.qsyms[i] for i in args], [self.ndims[i] for i in args])
        for idx in self.idx_allowed:
            o, n = self.ioff_allowed[idx], self.ndim_allowed[idx]
            isyms = np.unravel_index(idx, self.nsyms)
            jdx = qt.iblks_allowed[tuple(isyms[i] for i in args)] - 1
            o2 = qt.ioff_allowed[jdx]
            qt.value[o2:o2 + n] = self.value[o:o + n].reshape(
                self.dims_allowed[idx]).transpose(*args).reshape(n)
        return qt

    # -- contraction --------------------------------------------------
    def tensordotSYM(self, qt1, qt2, axes):
        """Symbolic contraction plan (qtensor.py:563-637).

        clst rows: [ioff, ioff1, ioff2, ndim, ndim1, ndim2, dims, dims1, dims2]
        for every allowed output block (ascending idx) x every internal sector
        tuple allowed in both inputs (C order).
        """
        i1, i2 = list(axes[0]), list(axes[1])
        e1 = sorted(set(range(qt1.rank)) - set(i1))
        e2 = sorted(set(range(qt2.rank)) - set(i2))
        ne1 = len(e1)
        for a, b in zip(i1, i2):
            assert np.array_equal(qt1.qsyms[a], qt2.qsyms[b])
        ints1 = [qt1.status[i] for i in i1]
        ints2 = [qt2.status[i] for i in i2]
        exts1 = [qt1.status[i] for i in e1]
        exts2 = [qt2.status[i] for i in e2]
        if len(i1) != 0:
            if ints1[0] == ints2[0]:
                ints2 = [not x for x in ints2]
                exts2 = [not x for x in exts2]
            for s1, s2 in zip(ints1, ints2):
                assert s1 != s2
        self.status = np.array(exts1 + exts2, dtype=bool)
        self.fromQsyms(ne1 + len(e2), [qt1.qsyms[i] for i in e1] + [qt2.qsyms[i] for i in e2],
                       [qt1.ndims[i] for i in e1] + [qt2.ndims[i] for i in e2],
                       ifallocate=False)
        self.axes = (i1, i2)
        self.clst = []
        ib1 = qt1.iblks_allowed.transpose(e1 + i1)
        ib2 = qt2.iblks_allowed.transpose(e2 + i2)
        for idx in self.idx_allowed:
            isyms = np.unravel_index(idx, self.nsyms)
            nz1 = ib1[tuple(isyms[:ne1])]
            nz2 = ib2[tuple(isyms[ne1:])]
            for inz in zip(*np.nonzero(np.logical_and(nz1, nz2))):
                j1 = nz1[inz] - 1
                j2 = nz2[inz] - 1
                self.clst.append([self.ioff_allowed[idx], qt1.ioff_allowed[j1],
                                  qt2.ioff_allowed[j2], self.ndim_allowed[idx],
                                  qt1.ndim_allowed[j1], qt2.ndim_allowed[j2],
                                  self.dims_allowed[idx], qt1.dims_allowed[j1],
                                  qt2.dims_allowed[j2]])
        return len(self.idx_allowed)

    def tensordotCAL(self, qt1, qt2, ifc1=False, ifc2=False, thresh=1e-20, tsize=10000):
        """Numeric contraction over the plan (qtensor.py:642-659)."""
        self.value = np.zeros(self.size_allowed)
        for (o, o1, o2, n, n1, n2, dims, dims1, dims2) in self.clst:
            t1 = qt1.value[o1:o1 + n1].reshape(dims1)
            t2 = qt2.value[o2:o2 + n2].reshape(dims2)
            if n1 > tsize and np.max(np.abs(t1)) < thresh:
                continue
            if n2 > tsize and np.max(np.abs(t2)) < thresh:
                continue
            self.value[o:o + n] += np.tensordot(t1, t2, self.axes).reshape(n)
        return 0

    def flops(self):
        """Algorithmic FLOPs of the plan: sum over rows of 2*ndim*(contracted)."""
        tot = 0
        for row in self.clst:
            n, n1, n2 = int(row[3]), int(row[4]), int(row[5])
            # n = free1*free2 ; n1 = free1*k ; n2 = k*free2  => k = sqrt(n1*n2/n)
            k = int(round((n1 * n2 / n) ** 0.5)) if n else 0
            tot += 2 * n * k
        return tot


def tensordot(qt1, qt2, axes):
    """One-shot contraction returning a new QTensor (qtensor.py:665-827)."""
    out = QTensor()
    out.tensordotSYM(qt1, qt2, axes)
    out.tensordotCAL(qt1, qt2)
    return out
