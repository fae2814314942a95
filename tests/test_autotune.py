"""Tile autotuner: signatures, nearest-entry lookup in the measured table
and the per-dtype launch choice (CPU; the device tuning is in
test_gpu_autotune.py)."""
import json

import numpy as np

import zmpo_dmrg_amd._lib as L
from zmpo_dmrg_amd import autotune as at
from zmpo_dmrg_amd import qtensor as pq


def _program(Ms, Ns, Ks):
    t = np.zeros(len(Ms), L.TASK_DTYPE)
    pairs = []
    for i, (M, N, ks) in enumerate(zip(Ms, Ns, Ks)):
        t[i] = (0, 0, M, N, M, 0, N, 1, len(pairs), len(pairs) + len(ks), 0, 0)
        for K in ks:
            pairs.append((0, 0, 1.0, K, M, 0, K, 1, N, 0, 1, N, 0))
    return pq.Program(t, np.array(pairs, dtype=L.PAIR_DTYPE))


def test_signature_bins_work_weighted_medians():
    # one big task dominates the work: its M, N and summed K set the bins
    prog = _program([2048, 8], [128, 8], [[100, 28], [4]])
    role, dt, bm, bn, bk = at.signature(prog, 'sigma_l', 0)
    assert (role, dt) == ('sigma_l', 0)
    assert bm == 11.0 and bn == 7.0 and bk == 7.0      # 2048, 128, 128 = 100 + 28
    # half-octave bins
    assert at._bin(181) == 7.5 and at._bin(1) == 0.0 and at._bin(0) == 0.0


def test_lookup_nearest_same_role_and_dtype(tmp_path, monkeypatch):
    table = {'entries': [
        {'signature': ['sigma_l', 0, 11.0, 7.0, 7.0], 'kernel': 'shaped', 'shape': 2},
        {'signature': ['sigma_l', 0, 8.0, 5.0, 5.0], 'kernel': 'shaped', 'shape': 3},
        {'signature': ['sigma_l', 2, 11.0, 7.0, 7.0], 'kernel': 'shaped', 'shape': 6},
        {'signature': ['sigma_r', 3, 7.0, 7.0, 20.0], 'kernel': 'shaped', 'shape': 10}]}
    path = tmp_path / 't.json'
    path.write_text(json.dumps(table))
    monkeypatch.setattr(at, 'TABLE_PATH', str(path))
    monkeypatch.setattr(at, '_table', None)
    monkeypatch.setattr(at, '_online', {})
    assert at.lookup(('sigma_l', 0, 10.5, 7.0, 7.5)) == ('shaped', 2)
    assert at.lookup(('sigma_l', 0, 8.5, 5.0, 5.0)) == ('shaped', 3)
    assert at.lookup(('sigma_l', 2, 3.0, 3.0, 3.0)) == ('shaped', 6)   # nearest of its dtype
    assert at.lookup(('sigma_l', 1, 11.0, 7.0, 7.0)) is None           # no FP32 entry
    assert at.lookup(('sigma_l', 2, 3.0, 3.0, 3.0), max_dist=1.0) is None
    # choose() sets the per-dtype tile; other dtypes keep the static choice
    prog = _program([2048], [128], [[128]])
    monkeypatch.setattr(at, 'MODE', 'table')
    assert at.choose(prog, 'sigma_l', 2, launch=None) == ('shaped', 6)
    assert prog.launch_choice(2) == ('shaped', 6)
    assert prog.launch_choice(0) == (prog.prefer or pq.KERNEL, prog.shape)
    monkeypatch.setattr(at, 'MODE', 'off')
    prog2 = _program([2048], [128], [[128]])
    assert at.choose(prog2, 'sigma_l', 2, launch=None) is None and prog2.tiles == {}


def test_launch_choice_falls_back_for_panel_tiles_on_stream():
    prog = _program([64], [64], [[64, 64]])
    prog.tiles[0] = ('stream', 6)        # the stream kernel has no panel tiles
    assert prog.launch_choice(0) == ('shaped', 6)
    prog.tiles[1] = ('stream', 2)
    assert prog.launch_choice(1) == ('stream', 2)
    # an explicit kernel keeps the program's own shape
    assert prog.launch_choice(1, 'shaped') == ('shaped', prog.shape)


def test_shipped_table_is_well_formed():
    import os
    if not os.path.exists(at.TABLE_PATH):
        return
    with open(at.TABLE_PATH) as f:
        d = json.load(f)
    for e in d['entries']:
        role, dt = e['signature'][:2]
        assert role in at.CANDIDATES and dt in (0, 1, 2, 3)
        assert (e['kernel'], e['shape']) in at.CANDIDATES[role]


def test_kmajor_tile_needs_k_contiguous_operands():
    kmaj = _program([64], [64], [[64, 32]])            # a_sk = 1, b_sk = N (row-major B)
    kmaj.tiles[0] = ('shaped', 12)
    assert kmaj.launch_choice(0) == ('shaped', 2)       # falls back: B is not k-contiguous
    t, p = kmaj.tasks_host.copy(), kmaj.pairs_host.copy()
    p['b_sk'], p['b_snlo'] = 1, p['K']
    ok = pq.Program(t, p)
    ok.tiles[0] = ('shaped', 12)
    assert ok.kmajor_ok() and ok.launch_choice(0) == ('shaped', 12)


def test_weighted_median_by_counting_equals_sorting():
    """The counting weighted median of the signature equals the sorted
    cumulative rule (the smallest value whose cumulative weight reaches half
    the total) for small and large value ranges."""
    rng = np.random.default_rng(0)

    def by_sorting(x, w):
        o = np.argsort(x, kind='stable')
        c = np.cumsum(w[o])
        return float(x[o][np.searchsorted(c, 0.5 * c[-1])])
    for hi in (3, 100, 5000, 1 << 22):
        for _ in range(200):
            n = int(rng.integers(1, 300))
            x = rng.integers(0, hi, n).astype(np.float64)
            w = rng.integers(1, 1000, n).astype(np.float64)
            assert at._wmedian(x, w) == by_sorting(x, w)
    assert at._wmedian(np.zeros(0), np.zeros(0)) == 1.0
