"""Decimation and next-site guess on the device (counterpart of
source/mpo_dmrg_dotutil.py:288-495 decimation, source/mpo_dmrg_qparser.py:
216-293 rdm_blkdiag and dotutil.py:498-559 genGuess, ncsite = 2).

For one two-site step with the Davidson solutions on the device:

1. the dense two-site matrices C_e (kept side x other side: rows (l, n1)
   for a left-moving sweep 'L', columns (n2, r) for 'R') by one scatter of
   each solution through prjmap;
2. the reduced density matrix rho = sum_e C_e C_e^T (or C_e^T C_e) by one
   dense grouped-GEMM launch, and the blocks of its quantum-number sectors
   (first-occurrence order, as the reference) packed by one gather;
3. FP64 eigenpairs of every sector on the device (rocSOLVER dsyevd,
   zmpo_syevd_sectors);
4. the weights (one value per kept-side state) to the host, where the
   reference's normalisation, truncation order with its tie rule
   (dot.truncation_order), cut and QN re-sort run unchanged;
5. the kept eigenvectors gathered into the dense rotation on the device
   (a sector whose zero-weight vectors are kept gets the canonical
   null-space basis of dot.canonical_null on the host first);
6. srotR = rot^T C_e (or C_e rot), the new site tensor (block-sparse,
   values gathered from rot) and the next-site guess (srotR times the next
   site's dense tensor) by device launches; the guess stays on the device
   for the next step's Davidson start vector.

With the noise term (dmrg.noise >= 1e-10) the caller adds noise times the
operator term of pRDMQt (kernelQt.prdm_noise, dense over the kept side) to
rho before step 2's sector gather, as dotutil.py:312-333 adds it to the RDM.
The host copies of the reference's path (dense civecs, the RDM, eigh,
srotR and the guess in numpy) are gone from this path; with an all-zero
density matrix dot.decimation runs instead.
"""
import numpy as np

from . import _lib
from . import device as dv
from .qtensor import PANEL_SHAPES, Program, QTensor, gather, run_program


def _gemm(M, N, npair, K, a, b, c_sm, a_step=0, b_step=0):
    """One-task grouped GEMM C(M x N, row stride c_sm) = sum over npair
    pairs of A_e B_e, pair e reading A at a[0] + e * a_step with row stride
    a[1] and k stride a[2], B at b[0] + e * b_step with column stride b[1]
    and k stride b[2]."""
    t = np.zeros(1, _lib.TASK_DTYPE)
    t[0] = (0, 0, M, N, M, 0, c_sm, 1, 0, npair, 0, 0)
    p = np.zeros(npair, _lib.PAIR_DTYPE)
    for e in range(npair):
        p[e] = (a[0] + e * a_step, b[0] + e * b_step, 1.0, K, M, 0, a[1], a[2], N, 0, b[1],
                b[2], 0)
    prog = Program(t, p, prefer='shaped')
    prog.shape = PANEL_SHAPES[0]
    return prog


def dense_of(qt, idlst=None):
    """Device dense copy of a QTensor (values uploaded if on the host)."""
    idlst = qt.idlst if idlst is None else idlst
    d, pos = qt._dense_positions(idlst)
    v = dv.upload(qt.value, 0) if isinstance(qt.value, np.ndarray) else qt.value
    src = dv.empty(len(pos), 0)
    gather(src, v, dv.upload(pos.astype(np.int64)))
    out = dv.zeros(int(np.prod([int(x) for x in qt.shape])), 0)
    _lib.check(_lib.lib().zmpo_scatter(0, dv.ptr(out), 0, dv.ptr(src),
                                       dv.ptr(dv.upload(d.astype(np.int64))), len(d),
                                       dv.stream()))
    return out


class DeviceDecimation:
    """One decimation of a two-site step (see the module docstring).

    vecs: device float64 (neig, ndim0) in qtmp block order; prjmap: dense
    index of every block entry.  After run(): sigs, dwts, qred (host), site
    (QTensor with device values), srotR (device dense (neig, nres, 4, rdim)
    for 'L', (neig, ldim, 4, nres) for 'R'), rot (device dense)."""

    def __init__(self, dmrg, vecs, prjmap, isite, status):
        self.dmrg, self.isite, self.status = dmrg, isite, status
        ldim, _, rdim, ntot = dmrg.dims
        self.ldim, self.rdim = ldim, rdim
        neig = vecs.shape[0]
        self.neig = neig
        # 1. dense C_e: (neig, ldim*4, 4*rdim) row-major
        C = dv.zeros(neig * ntot, 0)
        idx = dv.upload(np.asarray(prjmap, dtype=np.int64))
        for e in range(neig):
            _lib.check(_lib.lib().zmpo_scatter(0, dv.ptr(C[e * ntot:]), 0, dv.ptr(vecs[e]),
                                               dv.ptr(idx), len(prjmap), dv.stream()))
        self.C = C
        self.m = ldim * 4 if status == 'L' else 4 * rdim          # kept-side dimension
        self.ncol = 4 * rdim if status == 'L' else ldim * 4       # other side

    def rdm(self):
        """Dense rho (m x m) on the device."""
        m, nc, ntot = self.m, self.ncol, self.m * self.ncol
        rho = dv.empty(m * m, 0)
        if self.status == 'L':
            # rho[i, j] = sum_e sum_k C_e[i, k] C_e[j, k]
            prog = _gemm(m, m, self.neig, nc, (0, nc, 1), (0, nc, 1), m, ntot, ntot)
        else:
            # rho[i, j] = sum_e sum_k C_e[k, i] C_e[k, j]
            prog = _gemm(m, m, self.neig, nc, (0, 1, m), (0, 1, m), m, ntot, ntot)
        run_program(prog, self.C, self.C, rho, 0)
        return rho

    def run(self, classes, thresh, Dcut, extra=None):
        """Steps 2-5; ``extra`` (device float64 m x m, or None) is added to
        rho.  False when rho is all zero (the host path handles it)."""
        from .dot import WZERO, canonical_null, truncation_order
        m = self.m
        keys, groups = [], {}
        for i, val in enumerate(classes):
            k = (float(val[0]), float(val[1]))
            if k not in groups:
                groups[k] = []
                keys.append(k)
            groups[k].append(i)
        rows = [np.array(groups[k], dtype=np.int64) for k in keys]
        ns = np.array([len(r) for r in rows], dtype=np.int64)
        off = np.concatenate([[0], np.cumsum(ns * ns)[:-1]]).astype(np.int64)
        woff = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
        # 2. rho and its sector blocks
        rho = self.rdm()
        if extra is not None:
            one = dv.upload(np.ones(1), 0)
            _lib.check(_lib.lib().zmpo_lincomb(0, dv.ptr(rho), dv.ptr(extra), m * m, 1,
                                               dv.ptr(one), 1.0, 1.0, m * m, dv.stream()))
        pidx = np.concatenate([(r[:, None] * m + r[None, :]).reshape(-1) for r in rows])
        A = dv.empty(len(pidx), 0)
        gather(A, rho, dv.upload(pidx))
        del rho
        # trace of rho: sum of the packed diagonals (one small gather)
        didx = np.concatenate([off[s] + np.arange(ns[s]) * (ns[s] + 1) for s in range(len(rows))])
        diag = dv.empty(len(didx), 0)
        gather(diag, A, dv.upload(didx.astype(np.int64)))
        tr = float(dv.download(diag).sum())
        if tr < 1e-12:
            return False
        # 3. eigenpairs per sector
        W = dv.empty(int(ns.sum()), 0)
        E = dv.empty(int(max(ns.max(), 1)), 0)
        info = dv.zeros(len(rows), 1)        # int32-sized scratch (FP32 zeros)
        _lib.check(_lib.lib().zmpo_syevd_sectors(
            len(rows), ns.ctypes.data, off.ctypes.data, woff.ctypes.data, dv.ptr(A), dv.ptr(W),
            dv.ptr(E), dv.ptr(info), dv.stream()))
        if np.any(dv.download(info).view(np.int32) != 0):
            raise RuntimeError('decimation: sector eigensolver did not converge')
        s2 = np.abs(dv.download(W)) / tr
        # 4. truncation on the host (the reference's order, tie rule and cut)
        qsym = []
        rot_cols = []
        for b, k in enumerate(keys):
            qsym += [k] * int(ns[b])
            rot_cols += [(b, c) for c in range(int(ns[b]))]
        sigs = s2 / np.sum(s2)
        indx = truncation_order(sigs, qsym)
        nres = len(sigs)
        for i in range(nres):
            if sigs[indx[i]] < thresh * 1.01:
                nres = i
                break
        nres = max(nres, 1)
        if Dcut > 0:
            nres = min(nres, Dcut)
        indx = indx[:nres]
        re = sorted(range(nres), key=lambda x: qsym[indx[x]])
        indx = indx[re]
        # 5. kept eigenvectors -> rot (m x nres); canonical null space where
        # zero-weight vectors are kept
        fixed = set()
        for i in indx:
            b, c = rot_cols[i]
            if b not in fixed and sigs[woff[b] + c] <= WZERO:
                n = int(ns[b])
                u = dv.download(A[off[b]:off[b] + n * n]).reshape(n, n).T.copy()  # columns
                u = canonical_null(u, sigs[woff[b]:woff[b] + n], WZERO)
                A[off[b]:off[b] + n * n] = dv.upload(np.ascontiguousarray(u.T).reshape(-1), 0)
                fixed.add(b)
        src, dst = [], []
        for cpos, i in enumerate(indx):
            b, c = rot_cols[i]
            n = int(ns[b])
            src.append(off[b] + c * n + np.arange(n))
            dst.append(rows[b] * nres + cpos)
        src = np.concatenate(src).astype(np.int64)
        dst = np.concatenate(dst).astype(np.int64)
        tmp = dv.empty(len(src), 0)
        gather(tmp, A, dv.upload(src))
        rot = dv.zeros(m * nres, 0)
        _lib.check(_lib.lib().zmpo_scatter(0, dv.ptr(rot), 0, dv.ptr(tmp), dv.ptr(dv.upload(dst)),
                                           len(dst), dv.stream()))
        sig_kept = sigs[indx]
        self.sigs = sig_kept
        self.dwts = 1.0 - float(np.sum(sig_kept))
        self.qred = np.array([qsym[i] for i in indx])
        self.rot, self.nres = rot, nres
        return True

    def site_and_srotR(self, qlst):
        """The new site tensor (QTensor, device values) and srotR."""
        nres, m, nc, ntot = self.nres, self.m, self.ncol, self.m * self.ncol
        ldim, rdim = self.ldim, self.rdim
        if self.status == 'L':
            q = QTensor([False, False, True])
            qnums = [qlst[0], qlst[1], self.qred]
            shape_d = (ldim, 4, nres)
        else:
            q = QTensor([True, False, False])
            qnums = [self.qred, qlst[-2], qlst[-1]]
            shape_d = (nres, 4, rdim)
        from .qtensor import fromQnums
        rank, qsyms, ndims, idlst = fromQnums(qnums)
        q.fromQsyms(rank, qsyms, ndims, ifallocate=False)
        q.idlst = [list(x) for x in idlst]
        d, pos = q._dense_positions(idlst)
        if self.status == 'L':
            dmap = d                                     # rot is (ldim*4, nres) = (l, n1, k)
        else:
            # site (k, n2, r) from rot (n2*rdim + r, k)
            kk, rest = np.divmod(d, 4 * rdim)
            dmap = rest * nres + kk
        src = np.empty(q.size_allowed, np.int64)
        src[pos] = dmap
        val = dv.empty(q.size_allowed, 0)
        gather(val, self.rot, dv.upload(src))
        q.value = val
        # srotR
        srot = dv.empty(self.neig * nres * nc, 0)
        if self.status == 'L':
            # srotR_e[k, j] = sum_i rot[i, k] C_e[i, j]   (nres x ncol)
            prog = _gemm(nres, nc, 1, m, (0, 1, nres), (0, 1, nc), nc)
        else:
            # srotR_e[i, k] = sum_j C_e[i, j] rot[j, k]   (ncol x nres)
            prog = _gemm(nc, nres, 1, m, (0, m, 1), (0, 1, nres), nres)
        for e in range(self.neig):
            if self.status == 'L':
                run_program(prog, self.rot, self.C[e * ntot:], srot[e * nres * nc:], 0)
            else:
                run_program(prog, self.C[e * ntot:], self.rot, srot[e * nres * nc:], 0)
        self.srotR = srot
        if self.status == 'L':
            return q, srot.view(self.neig, nres, 4, rdim)
        return q, srot.view(self.neig, ldim, 4, nres)

    def guess(self, nxt):
        """Next-site guess (device, (neig, ...) flattened): srotR times the
        dense next site (the reference's genGuess for ncsite = 2)."""
        nres, ldim, rdim = self.nres, self.ldim, self.rdim
        t = dense_of(nxt)
        sh = [int(x) for x in nxt.shape]
        outs = []
        if self.status == 'L':
            # (nres*4, rdim) x (rdim, 4*r')
            M, Kd, N = nres * 4, sh[0], sh[1] * sh[2]
            prog = _gemm(M, N, 1, Kd, (0, Kd, 1), (0, 1, N), N)
            per = nres * 4 * rdim
        else:
            # (l'*4, l) x (l, 4*nres)
            M, Kd, N = sh[0] * sh[1], sh[2], 4 * nres
            prog = _gemm(M, N, 1, Kd, (0, Kd, 1), (0, 1, N), N)
            per = ldim * 4 * nres
        out = dv.empty(self.neig * M * N, 0)
        for e in range(self.neig):
            if self.status == 'L':
                run_program(prog, self.srotR[e * per:], t, out[e * M * N:], 0)
            else:
                run_program(prog, t, self.srotR[e * per:], out[e * M * N:], 0)
        return out.view(self.neig, M * N)
