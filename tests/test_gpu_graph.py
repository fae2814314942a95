"""Sigma replayed from a captured HIP graph (kernelQt.SigmaGraph) against the
oracle and against eager application.

Every replay is checked, with the program's own eager launches and host
synchronisations between the replays, and with a new input each time: a
graph whose nodes went stale (a freed buffer, a memset node, a pointer
captured from a temporary) would show up as a mismatch on some replay.
"""
import numpy as np
import pytest

from parity_util import Case
from zmpo_dmrg_amd import device as dv
from zmpo_dmrg_amd import dmrg as pdmrg
from zmpo_dmrg_amd import dot as pdot
from zmpo_dmrg_amd import kernelQt as kq
from zmpo_dmrg_amd.integrals import gen_integrals

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def case():
    c = Case(K=5, D=24, seed=4, sweeps=1, upload=True)
    c.L, c.R = c.product_envs()
    c.L[-1] = kq.boundary_store(2 * c.K)
    c.R[c.K] = kq.boundary_store(2 * c.K)
    return c


@pytest.mark.parametrize('chunks', [1, 3])
@pytest.mark.parametrize('dtype', [0, 1])
def test_graph_replays_match_eager_and_oracle(gpu, case, chunks, dtype, monkeypatch):
    monkeypatch.setattr(kq, 'SIGMA_CHUNKS', chunks)
    for isite in (1, case.K - 3):
        print('[graph] isite=%d chunks=%d dtype=%d' % (isite, chunks, dtype), flush=True)
        qkey, oq, pt, _ = case.templates(isite)
        L, R = case.L[isite - 1], case.R[isite + 2]
        prog = kq.SigmaProgram(case.mpo, isite, pt, qkey, L, R, case.mpo.hpwts)
        if chunks > 1:
            assert len(prog.chunks) >= 2
        g = kq.SigmaGraph(prog, case.mpo, L, R, dtype)
        assert g.graph is not None
        n = pt.size_allowed
        rng = np.random.default_rng(isite + 11)
        eager = dv.empty(n, dtype)
        out = dv.empty(n, 0)
        for rep in range(4):
            v = rng.standard_normal(n)
            # eager application (default stream layout, overlapped chunks)
            prog.apply(case.mpo, L, R, dv.upload(v, dtype), eager, dtype)
            want = dv.download(eager).astype(np.float64)
            dv.sync()
            g.run(dv.upload(v, 0), out)
            got = dv.download(out)
            assert np.array_equal(got, want), (isite, rep)
            if dtype == 0:
                ref = case.oracle_sigma(isite, v)
                assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_sweeps_with_and_without_graphs_agree(gpu, monkeypatch):
    """Davidson with graph replays after each step's first sigma gives the
    same energies and sigma counts as eager launches (bit-identical sigmas)."""
    K = 6
    i1, i2 = gen_integrals(K, seed=5)
    res = {}
    for on in (False, True):
        monkeypatch.setattr(pdot, 'GRAPH', on)
        for dtype in (0, 1):
            print('[graph] sweeps graph=%s dtype=%d' % (on, dtype), flush=True)
            d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=32, tol=1e-10, dtype=dtype)
            e = pdmrg.run(d, 2)
            res[on, dtype] = (e, d.nsigma, getattr(d, 'nsigma64', 0))
    for dtype in (0, 1):
        assert res[True, dtype] == res[False, dtype], (dtype, res[True, dtype], res[False, dtype])


def test_sweeps_with_plan_cache_agree(gpu, monkeypatch):
    """Sigma programs reused across sweeps for an unchanged block structure
    (ZMPO_PLAN_CACHE) give the same energies and sigma counts as programs
    rebuilt every step."""
    K = 6
    i1, i2 = gen_integrals(K, seed=7)
    res = {}
    for on in (False, True):
        monkeypatch.setattr(kq, 'PLAN_CACHE', on)
        print('[plan-cache] on=%s' % on, flush=True)
        d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=16, tol=1e-10)
        e = pdmrg.run(d, 4)
        res[on] = (e, d.nsigma)
        if on:
            assert getattr(d, 'plan_hits', 0) > 0
    assert res[True] == res[False], res
