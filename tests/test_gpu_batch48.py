"""Batching form at its stated size: the 48 seeded instances (K from {12,
16, 20, 24}, D from {128, 256, 384, 512}, N = K, Sz = 0), FP32 sigma with
the FP64 correction, one sweep each, four worker processes sharing the GPU.
Every instance's <Psi|H|Psi> (environment sweep over its final MPS) equals
its last eigenvalue; the largest and the smallest instance run alone (one
worker process with the same host threads, 4) give the same energies bit
for bit."""
import pytest

from zmpo_dmrg_amd import batch

pytestmark = pytest.mark.gpu


def test_config4_48_instances(gpu):
    specs = batch.instances(48, seed=0)
    res, summ = batch.run_batch(specs, 4, 1, 1)
    print('[batch48] %s' % summ, flush=True)
    assert summ['instances'] == 48
    worst = max(abs(r['e_mps'] - r['e_last']) for r in res)
    print('[batch48] max |<Psi|H|Psi> - last eigenvalue| = %.2e' % worst, flush=True)
    assert worst < 1e-8
    big = max(specs, key=lambda s: s[1] * s[2] * s[2])
    small = min(specs, key=lambda s: s[1] * s[2] * s[2])
    seq, _ = batch.run_batch([big, small], 1, 1, 1, cores=4, spawn=True)
    for a in seq:
        assert a['energies'] == res[a['index']]['energies'], a['index']
