"""Device-vs-oracle checks at the benchmark sizes (K = 24, D = 500 and
K = 36, D = 1000), on a subset of the operators so that the oracle's
per-block numpy loops finish in seconds.

The device side runs the full launches of every operator, with the sigma
weights of the operators outside the subset set to zero, through the
drop-in entry points HVecQt / HDiagQt (mpo_dmrg_kernelQt.py:30,153) called
with an ``info`` list as the reference's Davidson does; the renormalisation
compares every block of the subset's new environments.
"""
import types

import numpy as np

import bench
from oracle import kernels as ok
from oracle.qtensor import QTensor as OQ
from zmpo_dmrg_amd import kernelQt as kq


def one_op_per_case(mpo, isite, ncsite=2):
    """The first operator of every operator class (qtensor_opers.py:72-90)."""
    seen = {}
    for p in range(mpo.nops):
        seen.setdefault(kq.case(mpo.opers[p], isite, ncsite), p)
    return sorted(seen.values())


def dropin_info(mpo, isite, st, L, R, weights, budget):
    d = types.SimpleNamespace(fhop=mpo, hpwts=weights, qtmp=st.qtmp, qkey=st.qkey,
                              budget=budget, dtype=0, nsite=mpo.nsite)
    return [d, isite, 2, [[L, R]], 'R', True]


def check_sigma_and_diag(mpo, isite, st, L, R, budget, seed=0):
    """Returns (sigma relative error, diagonal relative error, sample)."""
    sample = one_op_per_case(mpo, isite)
    w = np.zeros(mpo.nops)
    w[sample] = 1.0
    info = dropin_info(mpo, isite, st, L, R, w, budget)
    n = st.qtmp.size_allowed
    v = np.random.default_rng(seed).standard_normal(n)
    got = kq.HVecQt(v, info, n, None)
    want, _, dt = bench.oracle_sigma(mpo, isite, st, L, R, sample, v)
    serr = float(np.abs(got - want).max() / np.abs(want).max())
    print('[fullsize] sigma over ops %s: max rel err %.2e (oracle %.1fs)' % (sample, serr, dt),
          flush=True)
    info[0]._sigma_cache[1]._buffers.clear()
    # diagonal: drop-in HDiagQt vs the oracle's dense HDiagQt on the subset
    oq = OQ(st.qtmp.status)
    oq.fromQsyms(st.qtmp.rank, st.qtmp.qsyms, st.qtmp.ndims)
    prjmap = oq.prjmap(st.idlstCI)
    dgot = kq.HDiagQt(info, n, prjmap)
    s = bench.oracle_state(mpo, (isite, isite + 1), sample)
    lenv = [bench.to_oracle(L.op_tensor(p)) for p in sample]
    renv = [bench.to_oracle(R.op_tensor(p)) for p in sample]
    ld, rd = int(lenv[0].shape[2]), int(renv[0].shape[2])
    s.dims = (ld, 16, rd, ld * 16 * rd)
    dwant = ok.HDiagQt(s, lenv, renv, isite, prjmap)
    derr = float(np.abs(dgot - dwant).max() / np.abs(dwant).max())
    print('[fullsize] diagonal: max rel err %.2e' % derr, flush=True)
    return serr, derr, sample


def check_renorm(mpo, env_prev, site, idx, env_new, status, sample):
    """Every block of env_new (device) for the ops in sample against the
    oracle's renorm_site of env_prev with the same site tensor."""
    s = bench.oracle_state(mpo, (idx,), sample)
    s.nsite = mpo.nsite
    prev = [bench.to_oracle(env_prev.op_tensor(p)) for p in sample]
    want = ok.renorm_site(s, prev, bench.to_oracle(site), idx, status)
    worst = 0.0
    for k, p in enumerate(sample):
        got = bench.to_oracle(env_new.op_tensor(p))
        w = want[k]
        assert np.array_equal(got.idx_allowed, w.idx_allowed), p
        scale = max(1.0, float(np.abs(w.value).max()))
        worst = max(worst, float(np.abs(got.value - w.value).max()) / scale)
    print('[fullsize] renorm %s at site %d, ops %s: max rel err %.2e' % (status, idx, sample, worst),
          flush=True)
    return worst
