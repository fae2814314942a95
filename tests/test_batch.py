"""Batching form: instance list, summary statistics, and (GPU) concurrent
workers reproducing the sequential energies bit for bit."""
import numpy as np
import pytest

from zmpo_dmrg_amd import batch


def test_instances_are_deterministic_and_in_range():
    a = batch.instances(48, seed=0)
    b = batch.instances(48, seed=0)
    assert a == b and len(a) == 48
    assert all(k in batch.KS and d in batch.DS for _, k, d, _ in a)
    assert len({s for *_, s in a}) == 48
    assert batch.instances(48, seed=1) != a


def test_summary_statistics():
    res = [{'sweep_s': [float(t), float(t)]} for t in range(1, 101)]
    s = batch.summarize(res, wall=50.0, nsweeps=2)
    assert s['sweeps_per_s'] == pytest.approx(200 / 50.0)
    assert s['p50_sweep_s'] == pytest.approx(np.percentile(np.arange(1, 101), 50))
    assert s['p99_sweep_s'] == pytest.approx(np.percentile(np.arange(1, 101), 99))


@pytest.mark.gpu
def test_concurrent_workers_match_sequential(gpu):
    specs = [(0, 4, 16, 3), (1, 6, 32, 1), (2, 6, 16, 2)]
    seq, _ = batch.run_batch(specs, 1, 1, 1)
    par, summ = batch.run_batch(specs, 2, 1, 1)
    print('[batch] %s' % summ, flush=True)
    for a, b in zip(seq, par):
        assert a['index'] == b['index']
        assert a['energies'] == b['energies']
    assert summ['instances'] == 3 and summ['sweeps_per_s'] > 0
