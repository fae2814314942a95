"""MPO site tensors H_p[i] = a_p^dagger V_p and the device-resident MPO store.

Dense factors follow source/mpo_dmrg_opers.py:91-219 (genWfacSpin,
genHfacSpin, linkTwoOpers); quantum numbers and the block-sparse form follow
source/qtensor/qtensor_opers.py:95-193 and :358-408 (genHRfacSpatialQt with
maxslc=1).  The store replaces the HDF5 file ``hop`` written by
source/mpo_dmrg_io.py:168-206: every site's ops are concatenated into one
device buffer so a single grouped-GEMM launch can address all of them.
"""
import copy

import numpy as np

from . import device as dv
from .qtensor import QTensor

_S = {
    'sgn': np.array([[1., 0.], [0., -1.]]),
    'idn': np.eye(2),
    'idnt': np.array([[1., 0.], [0., -1.]]),
    'cre': np.array([[0., 0.], [1., 0.]]),
    'cret': np.array([[0., 0.], [1., 0.]]),
    'ann': np.array([[0., 1.], [0., 0.]]),
    'annt': np.array([[0., -1.], [0., 0.]]),
    'nii': np.array([[0., 0.], [0., 1.]]),
    'niit': np.array([[0., 0.], [0., -1.]]),
}
SPINS = (0.5, -0.5)


def wfac_spin(nsite, isite, hq, vqrs):
    """V_p factor on one spin orbital (mpo_dmrg_opers.py:91-194), vectorised.

    Bond layout: [0] | q1(right sites) | q2(right sites) | q3(left sites) |
    q4(left sites) | last; the first/last site keep row 0 / the last column.
    """
    S = _S
    nb = 2 * (nsite + 1)
    diml, dimr = isite, nsite - isite - 1
    rs = np.arange(isite + 1, nsite)
    ls = np.arange(diml)
    last = nb - 1
    w = np.zeros((nb, nb, 2, 2))
    w[0, 0] = S['idnt']
    w[0, 1 + dimr:1 + 2 * dimr] = vqrs[isite, isite, rs][:, None, None] * S['niit']
    w[0, 1 + 2 * dimr + diml] = -S['ann']
    w[0, 2 + 2 * dimr + 2 * diml] = S['cre']
    w[0, last] = hq[isite] * S['ann']
    w[1, last] = S['cre']
    w[2 + np.arange(dimr), 1 + np.arange(dimr)] = S['idnt']
    w[2 + dimr, last] = S['ann']
    w[3 + dimr + np.arange(dimr), 1 + dimr + np.arange(dimr)] = S['idnt']
    off = 3 + 2 * dimr
    if diml and dimr:
        a = vqrs[np.ix_(rs, ls, [isite])][:, :, 0].T            # [l, r] = v[r, l, i]
        b = vqrs[np.ix_([isite], ls, rs)][0]                     # [l, r] = v[i, l, r]
        c = vqrs[np.ix_(ls, [isite], rs)][:, 0, :]               # [l, r] = v[l, i, r]
        w[off:off + diml, 1:1 + dimr] = -a[:, :, None, None] * S['annt']
        w[off:off + diml, 1 + dimr:1 + 2 * dimr] = b[:, :, None, None] * S['cret']
        off4 = off + diml
        w[off4:off4 + diml, 1 + dimr:1 + 2 * dimr] = c[:, :, None, None] * S['annt']
    if diml:
        w[off + ls, 1 + 2 * dimr + ls] = S['idn']
        w[off:off + diml, last] = vqrs[isite, ls, isite][:, None, None] * S['nii']
        off4 = off + diml
        w[off4 + ls, 2 + 2 * dimr + diml + ls] = S['idn']
    w[last, last] = S['idn']
    if isite == 0:
        w = w[:1]
    if isite == nsite - 1:
        w = w[:, last:]
    return w


def wfac_spin_batch(nsite, isite, hq, vqrs):
    """wfac_spin for P operators at once: hq[P, n], vqrs[P, n, n, n] ->
    w[P, nb, nb, 2, 2] (same entries as wfac_spin of each operator)."""
    S = _S
    P = hq.shape[0]
    nb = 2 * (nsite + 1)
    diml, dimr = isite, nsite - isite - 1
    rs = np.arange(isite + 1, nsite)
    ls = np.arange(diml)
    last = nb - 1
    w = np.zeros((P, nb, nb, 2, 2))
    w[:, 0, 0] = S['idnt']
    w[:, 0, 1 + dimr:1 + 2 * dimr] = vqrs[:, isite, isite, rs][:, :, None, None] * S['niit']
    w[:, 0, 1 + 2 * dimr + diml] = -S['ann']
    w[:, 0, 2 + 2 * dimr + 2 * diml] = S['cre']
    w[:, 0, last] = hq[:, isite][:, None, None] * S['ann']
    w[:, 1, last] = S['cre']
    w[:, 2 + np.arange(dimr), 1 + np.arange(dimr)] = S['idnt']
    w[:, 2 + dimr, last] = S['ann']
    w[:, 3 + dimr + np.arange(dimr), 1 + dimr + np.arange(dimr)] = S['idnt']
    off = 3 + 2 * dimr
    if diml and dimr:
        a = vqrs[:, rs][:, :, ls, isite].transpose(0, 2, 1)      # [P, l, r] = v[r, l, i]
        b = vqrs[:, isite][:, ls][:, :, rs]                       # [P, l, r] = v[i, l, r]
        c = vqrs[:, ls, isite][:, :, rs]                          # [P, l, r] = v[l, i, r]
        w[:, off:off + diml, 1:1 + dimr] = -a[..., None, None] * S['annt']
        w[:, off:off + diml, 1 + dimr:1 + 2 * dimr] = b[..., None, None] * S['cret']
        off4 = off + diml
        w[:, off4:off4 + diml, 1 + dimr:1 + 2 * dimr] = c[..., None, None] * S['annt']
    if diml:
        w[:, off + ls, 1 + 2 * dimr + ls] = S['idn']
        w[:, off:off + diml, last] = vqrs[:, isite, ls, isite][..., None, None] * S['nii']
        off4 = off + diml
        w[:, off4 + ls, 2 + 2 * dimr + diml + ls] = S['idn']
    w[:, last, last] = S['idn']
    if isite == 0:
        w = w[:, :1]
    if isite == nsite - 1:
        w = w[:, :, last:]
    return w


def hfac_spin_batch(ops, nsite, isite, hq, vqrs):
    """hfac_spin for the operators ``ops`` (hq, vqrs rows in the same order)."""
    e = np.stack([_S['sgn'] if isite < p else (_S['cre'] if isite == p else _S['idn'])
                  for p in ops])
    return np.matmul(e[:, None, None], wfac_spin_batch(nsite, isite, hq, vqrs))


def hfac_spin(p, nsite, isite, hq, vqrs):
    """(mpo_dmrg_opers.py:56-66, 197-203): elem(a_p^dagger) applied on V_p."""
    e = _S['sgn'] if isite < p else (_S['cre'] if isite == p else _S['idn'])
    return np.einsum('ij,abjk->abik', e, wfac_spin(nsite, isite, hq, vqrs))


def link(op1, op2):
    """W1[a,b,i,j] W2[b,c,k,l] -> W[a,c,(ik),(jl)] (mpo_dmrg_opers.py:43-49)."""
    t = np.tensordot(op1, op2, axes=([1], [0])).transpose(0, 3, 1, 4, 2, 5)
    s = t.shape
    return t.reshape(s[0], s[1], s[2] * s[3], s[4] * s[5])


def qphys():
    """|0>, |b>, |a>, |ab> as (N, Sz) (qtensor_opers.py:95-109)."""
    return np.array([[0., 0.], [1., -0.5], [1., 0.5], [2., 0.]])


def _elem_q(p, isite):
    spin = 0.5 if p % 2 == 0 else -0.5
    if isite < p:
        return np.array([0., 0.]), np.array([0., 0.])
    if isite == p:
        return np.array([0., 0.]), np.array([1., spin])
    return np.array([1., spin]), np.array([1., spin])


def _right_q(nsite, isite, sz):
    q = [[0., 0.]]
    q += [[-2., -sz - SPINS[i % 2]] for i in range(isite + 1, nsite)]
    q += [[0., -sz + SPINS[i % 2]] for i in range(isite + 1, nsite)]
    q += [[-1., -SPINS[i % 2]] for i in range(isite + 1)]
    q += [[1., SPINS[i % 2]] for i in range(isite + 1)]
    q += [[-1., -sz]]
    return np.array(q)


def _w_q(nsite, isite, isz):
    sz = SPINS[isz]
    ql = np.array([[0., 0.]]) if isite == 0 else _right_q(nsite, isite - 1, sz)
    qr = np.array([[-1., -sz]]) if isite == nsite - 1 else _right_q(nsite, isite, sz)
    return ql, qr


_KEYS = {}


def _qkey(row):
    t = np.asarray(row, dtype=np.float64).tobytes()   # keeps -0.0 apart from 0.0
    k = _KEYS.get(t)
    if k is None:
        k = _KEYS[t] = str(np.asarray(row))
    return k


def _sort_key_order(q):
    """Stable order by the printed QN row (qtensor_opers.py:56-58)."""
    keys = [_qkey(r) for r in q]
    return np.array(sorted(range(len(q)), key=lambda k: keys[k]))


def site_order(p, nsite, isite):
    """Bond quantum numbers of H_p on spatial site isite and the row / column
    orders (il, ir) that sort them (qtensor_opers.py:389-404): the stored
    tensor is the dense link(...)[il][:, ir]."""
    isz = p % 2
    el, _ = _elem_q(p, 2 * isite)
    _, er = _elem_q(p, 2 * isite + 1)
    wl, _ = _w_q(nsite, 2 * isite, isz)
    _, wr = _w_q(nsite, 2 * isite + 1, isz)
    ql = wl + el[None, :]
    qr = wr + er[None, :]
    return ql, qr, _sort_key_order(ql), _sort_key_order(qr)


def site_tensor(p, nsite, isite, int1e, int2e):
    """Block-sparse H_p on spatial site isite, legs [wl, wr, u, d], status [F,T,F,T]."""
    hq, vqrs = int1e[p], int2e[p]
    w = link(hfac_spin(p, nsite, 2 * isite, hq, vqrs), hfac_spin(p, nsite, 2 * isite + 1, hq, vqrs))
    ql, qr, il, ir = site_order(p, nsite, isite)
    w = w[np.ix_(il, ir)]
    qt = QTensor([False, True, False, True])
    qp = qphys()
    qt.fromDenseTensor(w, [ql[il], qr[ir], qp, qp], ifcollect=[1, 1, 0, 0])
    return qt


def _rel(a, b):
    return 0 if a < b else (1 if a == b else 2)


def site_class(p, isite):
    """Operators of one class share the block structure of H_p[isite]: the
    spin of p and where p sits relative to the site's two spin orbitals."""
    return (p % 2, _rel(2 * isite, p), _rel(2 * isite + 1, p))


_TEMPLATES = {}


def _class_template(p, nsbas, isite):
    """Block tables of H_p[isite] (values unset) and, per stored value, its
    flat index in the unsorted dense tensor link(...)[wl, wr, 4, 4].  They
    depend only on (nsbas, isite, site_class(p, isite)), not on the
    integrals, so they are kept for later stores (read-only, shared)."""
    key = (nsbas, isite, site_class(p, isite))
    if key not in _TEMPLATES:
        if len(_TEMPLATES) > 4096:
            _TEMPLATES.clear()
        _TEMPLATES[key] = _build_class_template(p, nsbas, isite)
    return _TEMPLATES[key]


def _build_class_template(p, nsbas, isite):
    ql, qr, il, ir = site_order(p, nsbas, isite)
    qp = qphys()
    rank, qsyms, ndims, idlst = _fromQnums([ql[il], qr[ir], qp, qp], [1, 1, 0, 0])
    qt = QTensor([False, True, False, True])
    qt.fromQsyms(rank, qsyms, ndims, ifallocate=False)
    qt.idlst = [list(x) for x in idlst]
    d, pos = qt._dense_positions(idlst)
    nr = len(ir)
    x, y, uv = d // (nr * 16), (d // 16) % nr, d % 16
    gidx = np.zeros(qt.size_allowed, np.int64)
    gidx[pos] = (il[x] * nr + ir[y]) * 16 + uv
    return qt, gidx


def link_program(P, nl, nb, nr):
    """Grouped-GEMM plan of link(h0, h1) for P operators: h0 [P][nl][nb][2][2]
    and h1 [P][nb][nr][2][2] (row-major) give W [P][nl][nr][4][4] with
    W[a, c, 2i+k, 2j+l] = sum_b h0[a, b, i, j] h1[b, c, k, l] (one task per
    operator and (i, j, k, l))."""
    from . import _lib
    from .qtensor import Program
    n = 16 * P
    t = np.zeros(n, _lib.TASK_DTYPE)
    pr = np.zeros(n, _lib.PAIR_DTYPE)
    op = np.repeat(np.arange(P, dtype=np.int64), 16)
    ijkl = np.tile(np.arange(16), P)
    i, j, k, l = ijkl // 8, (ijkl // 4) % 2, (ijkl // 2) % 2, ijkl % 2
    t['c_off'] = op * (nl * nr * 16) + (2 * i + k) * 4 + 2 * j + l
    t['M'], t['N'] = nl, nr
    t['c_mlo'], t['c_smhi'], t['c_smlo'], t['c_sn'] = nl, 0, nr * 16, 16
    t['p_begin'] = np.arange(n)
    t['p_end'] = np.arange(n) + 1
    pr['a_off'] = op * (nl * nb * 4) + 2 * i + j
    pr['b_off'] = op * (nb * nr * 4) + 2 * k + l
    pr['scale'] = 1.0
    pr['K'] = nb
    pr['a_mlo'], pr['a_smhi'], pr['a_smlo'], pr['a_sk'] = nl, 0, nb * 4, 4
    pr['b_nlo'], pr['b_snhi'], pr['b_snlo'], pr['b_sk'] = nr, 0, 4, nr * 4
    return Program(t, pr)


def _fromQnums(qnums, ifcollect):
    from .qtensor import fromQnums
    return fromQnums(qnums, ifcollect)


class MPOStore:
    """All H_p[i] for i < K on the device ("hop" file of io.py:168-206).

    ``sites[i][p]`` is the host QTensor (metadata + values); ``buf[i]`` the
    device buffer with every op of site i concatenated at ``off[i][p]``.

    builder 'device' (default with upload): per site, the two spin-orbital
    factors of every operator are assembled on the host in one batch, their
    link product runs as one grouped-GEMM launch and one gather launch moves
    the allowed entries into the site buffer, with the block tables built
    once per operator class (site_class); 'host': one operator at a time in
    numpy (site_tensor).
    """

    def __init__(self, int1e, int2e, nops=None, upload=True, builder=None):
        self.sbas = int1e.shape[0]
        self.nsite = self.sbas // 2
        self.nops = self.sbas if nops is None else nops
        self.opers = list(range(self.nops))
        self.hpwts = np.ones(self.nops)
        self.sites = []
        self.off = []
        self.host = []
        self.buf = [None] * self.nsite
        self._buf32 = {}
        builder = builder or ('device' if upload else 'host')
        for i in range(self.nsite):
            if builder == 'device':
                ops, vals, dev = self._site_on_device(i, int1e, int2e)
                self.buf[i] = dev
            else:
                ops = [site_tensor(p, self.sbas, i, int1e, int2e) for p in self.opers]
                vals = np.concatenate([q.value for q in ops])
            sizes = np.array([q.size_allowed for q in ops], dtype=np.int64)
            offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
            self.sites.append(ops)
            self.off.append(offs)
            self.host.append(vals)
        if upload:
            self.to_device()

    @classmethod
    def from_sites(cls, sites, op_class=None, upload=True):
        """A store over given site tensors: sites[i] = [QTensor of op p at
        site i] (legs [wl, wr, u, d], status [F, T, F, T]).  op_class(p,
        isite) returns a key that is equal for operators whose site tensors
        have the same block tables (default: the Hamiltonian's classes)."""
        self = cls.__new__(cls)
        self.nsite = len(sites)
        self.nops = len(sites[0])
        self.sbas = 2 * self.nsite
        self.opers = list(range(self.nops))
        self.hpwts = np.ones(self.nops)
        self.sites = [list(s) for s in sites]
        self.off, self.host = [], []
        for ops in self.sites:
            sizes = np.array([q.size_allowed for q in ops], dtype=np.int64)
            self.off.append(np.concatenate([[0], np.cumsum(sizes)[:-1]]))
            self.host.append(np.concatenate([np.asarray(q.value, np.float64) for q in ops]))
        self.buf = [None] * self.nsite
        self._buf32 = {}
        if op_class is not None:
            self.op_class = op_class
        if upload:
            self.to_device()
        return self

    def _site_on_device(self, i, int1e, int2e):
        from .qtensor import gather, run_program
        ops = self.opers
        hq, vq = int1e[ops], int2e[ops]
        h0 = hfac_spin_batch(ops, self.sbas, 2 * i, hq, vq)
        h1 = hfac_spin_batch(ops, self.sbas, 2 * i + 1, hq, vq)
        P, nl, nb = h0.shape[:3]
        nr = h1.shape[2]
        dense = dv.empty(P * nl * nr * 16, 0)
        run_program(link_program(P, nl, nb, nr), dv.upload(h0.reshape(-1), 0),
                    dv.upload(h1.reshape(-1), 0), dense, 0)
        tmpl = {}
        parts, qts = [], []
        base = 0
        for k, p in enumerate(ops):
            c = site_class(p, i)
            if c not in tmpl:
                tmpl[c] = _class_template(p, self.sbas, i)
            qt, gidx = tmpl[c]
            parts.append(gidx + k * (nl * nr * 16))
            q = copy.copy(qt)
            qts.append(q)
            base += qt.size_allowed
        idx = dv.upload(np.concatenate(parts))
        vals = gather(dv.empty(base, 0), dense, idx)
        host = dv.download(vals).copy()
        o = 0
        for q in qts:
            q.value = host[o:o + q.size_allowed]
            o += q.size_allowed
        del dense, idx
        return qts, host, vals

    def to_device(self):
        self.buf = [b if b is not None else dv.upload(h, 0) for b, h in zip(self.buf, self.host)]
        return self

    def padded(self, i):
        """Site i re-laid for the chained sigma launch: every block (w_l, w_r,
        1, 1) of every op as a zero-padded row-major [ceil4(w_l)][ceil32(w_r)]
        matrix, so that MFMA operand loads need no bounds tests.  Returns
        (host values, op offsets, per-op {block offset: (padded offset, row
        stride)}); built once per site."""
        if not hasattr(self, '_pad'):
            self._pad = {}
        if i not in self._pad:
            parts, offs, maps = [], [], []
            tot = 0
            for p, q in enumerate(self.sites[i]):
                d = q.blk_dims
                if len(d) and (np.any(d[:, 2] != 1) or np.any(d[:, 3] != 1)):
                    raise NotImplementedError('padded MPO layout needs 1x1 physical blocks')
                K = d[:, 0].astype(np.int64) if len(d) else np.zeros(0, np.int64)
                N = d[:, 1].astype(np.int64) if len(d) else np.zeros(0, np.int64)
                Kp, Np = (K + 3) // 4 * 4, (N + 31) // 32 * 32
                sz = Kp * Np
                new_off = np.concatenate([[0], np.cumsum(sz)[:-1]]).astype(np.int64) if len(sz) else sz
                out = np.zeros(int(sz.sum()))
                v = self.host[i][self.off[i][p]:self.off[i][p] + q.size_allowed]
                for j in range(len(sz)):
                    blk = v[q.blk_off[j]:q.blk_off[j] + K[j] * N[j]].reshape(K[j], N[j])
                    out[new_off[j]:new_off[j] + sz[j]].reshape(Kp[j], Np[j])[:K[j], :N[j]] = blk
                maps.append({int(o): (int(n), int(w)) for o, n, w in zip(q.blk_off, new_off, Np)})
                offs.append(tot)
                tot += len(out)
                parts.append(out)
            host = np.concatenate(parts) if parts else np.zeros(1)
            self._pad[i] = {'host': host, 'off': np.array(offs, np.int64), 'map': maps, 'dev': {}}
        e = self._pad[i]
        return e['host'], e['off'], e['map']

    def padded_buffer(self, i, dtype=0):
        """Device copy of padded(i)[0] in the given precision."""
        self.padded(i)
        e = self._pad[i]
        if dtype not in e['dev']:
            e['dev'][dtype] = dv.upload(e['host'], dtype)
        return e['dev'][dtype]

    def buffer(self, i, dtype=0):
        if dtype == 0:
            if self.buf[i] is None:
                self.buf[i] = dv.upload(self.host[i], 0)
            return self.buf[i]
        if (i, dtype) not in self._buf32:
            self._buf32[(i, dtype)] = dv.upload(self.host[i], dtype)
        return self._buf32[(i, dtype)]
