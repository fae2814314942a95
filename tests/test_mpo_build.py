"""The batched MPO builder (one grouped-GEMM link launch and one gather per
site, block tables once per operator class) against the per-operator host
builder (mpo.site_tensor, itself bitwise equal to the oracle's MPO and
pinned by the Fock-space Hamiltonian in test_oracle.py)."""
import numpy as np
import pytest

from kmodel import host_model
from zmpo_dmrg_amd import mpo as pm
from zmpo_dmrg_amd.integrals import gen_integrals


def _compare(a, b, K, tol):
    worst = 0.0
    for i in range(K):
        assert np.array_equal(a.off[i], b.off[i])
        for p in range(2 * K):
            qa, qb = a.sites[i][p], b.sites[i][p]
            assert np.array_equal(qa.idx_allowed, qb.idx_allowed), (i, p)
            assert all(np.array_equal(x, y) for x, y in zip(qa.qsyms, qb.qsyms))
            assert all(np.array_equal(x, y) for x, y in zip(qa.ndims, qb.ndims))
            assert qa.idlst == qb.idlst
            assert np.array_equal(qa.value, b.host[i][b.off[i][p]:b.off[i][p] + qb.size_allowed])
        worst = max(worst, float(np.abs(a.host[i] - b.host[i]).max()))
    assert worst <= tol, worst
    return worst


@pytest.mark.parametrize('K', [2, 3, 5])
def test_batched_builder_matches_per_operator_builder_cpu_model(K):
    i1, i2 = gen_integrals(K, seed=K)
    a = pm.MPOStore(i1, i2, upload=False, builder='host')
    with host_model():
        b = pm.MPOStore(i1, i2, upload=False, builder='device')
    _compare(a, b, K, 0.0)


def test_site_classes_share_block_tables():
    """Operators of one site_class have identical block tables (the premise
    of building them once per class)."""
    K = 5
    i1, i2 = gen_integrals(K, seed=1)
    for i in range(K):
        seen = {}
        for p in range(2 * K):
            q = pm.site_tensor(p, 2 * K, i, i1, i2)
            c = pm.site_class(p, i)
            if c in seen:
                r = seen[c]
                assert np.array_equal(r.idx_allowed, q.idx_allowed)
                assert all(np.array_equal(x, y) for x, y in zip(r.qsyms, q.qsyms))
            else:
                seen[c] = q


@pytest.mark.gpu
def test_device_builder_matches_host_builder_k10(gpu):
    K = 10
    i1, i2 = gen_integrals(K, seed=4)
    a = pm.MPOStore(i1, i2, upload=False, builder='host')
    b = pm.MPOStore(i1, i2, upload=True, builder='device')
    _compare(a, b, K, 1e-14 * max(1.0, max(float(np.abs(h).max()) for h in a.host)))
    from zmpo_dmrg_amd import device as dv
    for i in range(K):
        assert np.array_equal(dv.download(b.buffer(i)), b.host[i])
