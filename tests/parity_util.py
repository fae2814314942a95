"""Shared set-up for parity tests: an oracle DMRG state plus the same MPS and
environments on the zmpo_dmrg_amd side."""
import numpy as np

from oracle import kernels as ok
from oracle import sweep as osw
from oracle.qtensor import QTensor as OQ
from oracle.qtensor import fromQnums as ofq
from zmpo_dmrg_amd import kernelQt as kq
from zmpo_dmrg_amd import mpo as pmpo
from zmpo_dmrg_amd import qtensor as pq
from zmpo_dmrg_amd.integrals import gen_integrals


def to_product(o):
    q = pq.QTensor(o.status)
    q.fromQsyms(o.rank, o.qsyms, o.ndims, idlst=o.idlst)
    q.value = o.value.copy()
    return q


class Case:
    """Oracle state after ``sweeps`` sweeps at (K, D); product MPO/MPS/envs mirror it."""

    def __init__(self, K=4, D=16, seed=3, sweeps=1, upload=False, nelec=None):
        self.K = K
        self.nelec = (K if K % 2 == 0 else K - 1) if nelec is None else nelec
        self.i1, self.i2 = gen_integrals(K, seed=seed)
        self.st = osw.State(self.i1, self.i2, self.nelec, 0.0, Dmax=D, tol=1e-10)
        for _ in range(sweeps):
            osw.sweep(self.st, 'R')
            osw.sweep(self.st, 'L')
        osw.sweep(self.st, 'R')
        self.mpo = pmpo.MPOStore(self.i1, self.i2, upload=upload)
        self.oL, _ = ok.genHops(self.st, self.st.lmps, 'L')
        self.oR, _ = ok.genHops(self.st, self.st.rmps, 'R')
        nops = 2 * K
        self.oL[-1] = [ok.boundary_env() for _ in range(nops)]
        self.oR[K] = [ok.boundary_env() for _ in range(nops)]

    def product_envs(self):
        lm = {i: to_product(self.st.lmps[i]) for i in range(self.K)}
        rm = {i: to_product(self.st.rmps[i]) for i in range(self.K)}
        L, _ = kq.genHops(self.mpo, lm, 'L')
        R, _ = kq.genHops(self.mpo, rm, 'R')
        return L, R

    def templates(self, isite):
        qkey = np.array([float(self.nelec), 0.0])
        qp = pmpo.qphys()
        qnums = [self.st.qnuml[isite], qp, qp, qkey - self.st.qnumr[isite + 2]]
        oq = OQ([False, False, False, True])
        r, qs, nd, il = ofq(qnums)
        oq.fromQsyms(r, qs, nd)
        pt = pq.QTensor([False, False, False, True])
        r, qs, nd, _ = pq.fromQnums(qnums)
        pt.fromQsyms(r, qs, nd)
        return qkey, oq, pt, il

    def oracle_sigma(self, isite, v):
        qkey, oq, _, _ = self.templates(isite)
        self.st.qtmp, self.st.qkey = oq, qkey
        return ok.HVecQt(v, self.st, self.oL[isite - 1], self.oR[isite + 2], isite)

    def oracle_diag(self, isite):
        qkey, oq, _, il = self.templates(isite)
        self.st.qtmp, self.st.qkey = oq, qkey
        lo, ro = self.oL[isite - 1], self.oR[isite + 2]
        ld, rd = int(lo[0].shape[2]), int(ro[0].shape[2])
        self.st.dims = (ld, 16, rd, ld * 16 * rd)
        return ok.HDiagQt(self.st, lo, ro, isite, oq.prjmap(il))
