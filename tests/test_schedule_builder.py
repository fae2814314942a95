"""The library's schedule builder (zmpo_schedule_build) against the numpy
statement of the schedule (qtensor.make_schedule_np): identical units,
order, split-K slots and reduction groups for random programs with empty
tasks, multi-pair tasks split at several chunk sizes, reuse keys and every
tile shape."""
import numpy as np
import pytest

import zmpo_dmrg_amd._lib as L
from zmpo_dmrg_amd import qtensor as pq


def _random(rng, nt):
    tasks = np.zeros(nt, L.TASK_DTYPE)
    pairs = []
    for i in range(nt):
        M, N = int(rng.integers(0, 300)), int(rng.integers(0, 200))
        npair = int(rng.choice([0, 1, 1, 3, 7]))
        tasks[i]['M'], tasks[i]['N'] = M, N
        tasks[i]['p_begin'], tasks[i]['p_end'] = len(pairs), len(pairs) + npair
        for _ in range(npair):
            pairs.append(int(rng.integers(1, 3000)))
    p = np.zeros(len(pairs), L.PAIR_DTYPE)
    p['K'] = pairs
    return tasks, p


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_builder_matches_numpy(seed):
    rng = np.random.default_rng(seed)
    tasks, pairs = _random(rng, 200)
    for tile in [None] + list(pq.SHAPES.values()):
        for rb in (1, 2):
            for kchunk in (500, 2048, 32768):
                for key in (None, rng.integers(0, 5, len(tasks))):
                    a = pq.make_schedule(tasks, pairs, rb, kchunk=kchunk, reuse_key=key, tile=tile)
                    b = pq.make_schedule_np(tasks, pairs, rb, kchunk=kchunk, reuse_key=key,
                                            tile=tile)
                    assert a[2] == b[2]
                    assert np.array_equal(a[0], b[0]), (tile, rb, kchunk, key is None)
                    assert np.array_equal(a[1], b[1]), (tile, rb, kchunk, key is None)


def test_builder_empty_program():
    u, g, n = pq.make_schedule(np.zeros(0, L.TASK_DTYPE), np.zeros(0, L.PAIR_DTYPE), 1,
                               tile=(32, 32))
    assert len(u) == 0 and len(g) == 0 and n == 0
