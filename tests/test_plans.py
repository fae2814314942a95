"""Plan lowering and engine wiring, executed by the CPU model of the block
GEMM (tests/kmodel.py): zmpo_dmrg_amd programs vs the oracle, no device."""
import numpy as np
import pytest

from kmodel import emulate, host_model
from oracle import mpo as ompo
from parity_util import Case, to_product
from zmpo_dmrg_amd import kernelQt as kq
from zmpo_dmrg_amd import mpo as pmpo
from zmpo_dmrg_amd import qtensor as pq
from zmpo_dmrg_amd.integrals import gen_integrals


@pytest.mark.parametrize('K', [3, 5])
def test_mpo_bitwise_equal_to_oracle(K):
    i1, i2 = gen_integrals(K, seed=7)
    ms = pmpo.MPOStore(i1, i2, upload=False)
    om = ompo.build_mpo(i1, i2)
    for i in range(K):
        for p in range(2 * K):
            a, b = ms.sites[i][p], om[i][p]
            assert np.array_equal(a.idx_allowed, b.idx_allowed)
            assert all(np.array_equal(x, y) for x, y in zip(a.qsyms, b.qsyms))
            assert np.array_equal(a.value, b.value)


@pytest.fixture(scope='module')
def case4():
    return Case(K=4, D=16, seed=3, sweeps=1)


def test_env_build_matches_oracle(case4):
    with host_model():
        L, R = case4.product_envs()
        for i in range(case4.K):
            for p in range(2 * case4.K):
                a = L[i].op_tensor(p).toDenseTensor()
                assert np.abs(a - case4.oL[i][p].toDenseTensor()).max() < 1e-13
                b = R[i].op_tensor(p).toDenseTensor()
                assert np.abs(b - case4.oR[i][p].toDenseTensor()).max() < 1e-13


@pytest.mark.parametrize('isite', [0, 1, 2])
def test_sigma_and_diag_match_oracle(case4, isite):
    with host_model():
        L, R = case4.product_envs()
        L[-1] = kq.boundary_store(2 * case4.K)
        R[case4.K] = kq.boundary_store(2 * case4.K)
        qkey, oq, pt, _ = case4.templates(isite)
        v = np.random.default_rng(isite).standard_normal(oq.size_allowed)
        want = case4.oracle_sigma(isite, v)
        prog = kq.SigmaProgram(case4.mpo, isite, pt, qkey, L[isite - 1], R[isite + 2],
                               case4.mpo.hpwts)
        got = np.zeros(pt.size_allowed)
        prog.apply(case4.mpo, L[isite - 1], R[isite + 2], v, got)
        assert np.abs(got - want).max() < 1e-12 * max(1.0, np.abs(want).max())
        d = kq.DiagProgram(case4.mpo, isite, pt, L[isite - 1], R[isite + 2], case4.mpo.hpwts)
        dg = np.zeros(pt.size_allowed)
        d.apply(L[isite - 1], R[isite + 2], dg)
        assert np.abs(dg - case4.oracle_diag(isite)).max() < 1e-12


def test_sigma_chunking_is_order_stable(case4):
    """Op chunks (bounded intermediates) give the same sigma as one chunk."""
    with host_model():
        L, R = case4.product_envs()
        isite = 1
        qkey, oq, pt, _ = case4.templates(isite)
        v = np.random.default_rng(0).standard_normal(oq.size_allowed)
        out = []
        for budget in (1 << 40, 4096):
            prog = kq.SigmaProgram(case4.mpo, isite, pt, qkey, L[isite - 1], R[isite + 2],
                                   case4.mpo.hpwts, budget=budget)
            got = np.zeros(pt.size_allowed)
            prog.apply(case4.mpo, L[isite - 1], R[isite + 2], v, got)
            out.append((len(prog.chunks), got))
        assert out[1][0] > 1
        assert np.abs(out[0][1] - out[1][1]).max() < 1e-12


def test_generic_tensordot_lowering():
    """Reference-layout tensordot for several axes patterns vs numpy on dense."""
    rng = np.random.default_rng(1)
    qp = np.array([[0., 0.], [1., -.5], [1., .5], [2., 0.]])
    qb = qp[rng.integers(0, 4, 7)]
    a = pq.QTensor([False, False, True])
    a.fromDenseTensor(rng.standard_normal((7, 4, 7)), [qb, qp, qb + qb[::-1] * 0])
    b = pq.QTensor([False, False, True])
    b.fromDenseTensor(rng.standard_normal((7, 4, 7)), [qb, qp, qb])
    with host_model():
        for axes in (([0], [0]), ([0, 1], [0, 1]), ([2], [0])):
            try:
                out = pq.QTensor()
                out.tensordotSYM(a, b, axes)
            except (AssertionError, ValueError):
                continue
            out.tensordotCAL(a, b)
            il = [a.idlst[i] for i in out.plan.e1] + [b.idlst[i] for i in out.plan.e2]
            ref = np.tensordot(a.toDenseTensor(), b.toDenseTensor(), axes)
            assert np.abs(out.toDenseTensor(il) - ref).max() < 1e-12


def test_wave_schedule_split_k_matches_direct():
    """Random programs with long pair lists: the wave schedule (tiles x pair
    chunks, split-K groups) reproduces the direct per-task sum for rb=1,2,4."""
    import zmpo_dmrg_amd._lib as L
    from kmodel import emulate_wave
    rng = np.random.default_rng(0)
    nt = 9
    tasks = np.zeros(nt, L.TASK_DTYPE)
    pairs, off = [], 0
    for i in range(nt):
        M, N = int(rng.integers(1, 70)), int(rng.integers(1, 40))
        npair = int(rng.integers(0, 5))
        tasks[i] = (off, 0, M, N, M, 0, N, 1, len(pairs), len(pairs) + npair, 0, 0)
        off += M * N
        for _ in range(npair):
            K = int(rng.integers(1, 3000))
            pairs.append((int(rng.integers(0, 1000)), int(rng.integers(0, 1000)),
                          float(rng.standard_normal()), K, M, 0, K, 1, N, 0, 1, N, 0))
    pairs = np.array(pairs, dtype=L.PAIR_DTYPE)
    prog = pq.Program(tasks, pairs)
    A, B = rng.standard_normal(300000), rng.standard_normal(300000)
    want = np.zeros(off)
    emulate(prog, A, B, want)
    configs = [('wave', rb, 0) for rb in (1, 2, 4)] + [('shaped', 1, s) for s in pq.SHAPES]
    for kernel, rb, shape in configs:
        for key in (None, rng.integers(0, 3, nt)):
            prog.rb, prog.shape, prog.reuse_key = rb, shape, key
            tile = None if kernel == 'wave' else pq.SHAPES[shape]
            _, groups, nslots = pq.make_schedule(tasks, pairs, rb, reuse_key=key, tile=tile)
            assert nslots > 0 and len(groups) > 0
            got = np.zeros(off)
            emulate_wave(prog, A, B, got, kernel=kernel)
            assert np.abs(got - want).max() < 1e-10


def test_stream_schedule_cursor_matches_direct():
    """The streaming kernel's item cursor (unit ranges of 1..64, k-tiles of 16
    and 32, empty tasks, split-K slots, beta=1) reproduces the direct sum."""
    import zmpo_dmrg_amd._lib as L
    from kmodel import emulate_stream
    rng = np.random.default_rng(5)
    nt = 14
    tasks = np.zeros(nt, L.TASK_DTYPE)
    pairs, off = [], 0
    for i in range(nt):
        M, N = int(rng.integers(1, 140)), int(rng.integers(1, 70))
        npair = int(rng.integers(0, 4)) if i % 5 else 0
        tasks[i] = (off, 0, M, N, M, 0, N, 1, len(pairs), len(pairs) + npair, 0, 0)
        off += M * N
        for _ in range(npair):
            K = int(rng.integers(1, 2600))
            if rng.random() < 0.5:   # K-contiguous A
                a = (M, 0, K, 1)
            else:                    # row-contiguous A
                a = (M, 0, 1, M)
            pairs.append((int(rng.integers(0, 1000)), int(rng.integers(0, 1000)),
                          float(rng.standard_normal()), K) + a + (N, 0, 1, N, 0))
    pairs = np.array(pairs, dtype=L.PAIR_DTYPE)
    prog = pq.Program(tasks, pairs)
    A, B = rng.standard_normal(400000), rng.standard_normal(400000)
    C0 = rng.standard_normal(off)
    want = C0.copy()
    emulate(prog, A, B, want, beta=1)
    for shape in pq.SHAPES:
        for kt in (16, 32):
            for upw in (1, 3, 64):
                prog.shape = shape
                got = C0.copy()
                emulate_stream(prog, A, B, got, kt, upw, beta=1)
                assert np.abs(got - want).max() < 1e-10, (shape, kt, upw)


@pytest.mark.parametrize('rb', [1, 2])
def test_chain_program_matches_two_stages(rb):
    """Sigma stages 2-3 as one chained program (row groups, intermediate
    columns) equal the two separate grouped GEMMs on a real K=5 plan."""
    from kmodel import emulate_chain
    case5 = Case(K=5, D=24, seed=4, sweeps=1, nelec=4)
    with host_model():
        Lh, Rh = case5.product_envs()
        isite = 2
        qkey, _, pt, _ = case5.templates(isite)
        sp = kq.SigmaProgram(case5.mpo, isite, pt, qkey, Lh[isite - 1], Rh[isite + 2],
                             case5.mpo.hpwts, fused=True)
    rng = np.random.default_rng(1)
    for st, ch in zip(sp.stages, sp.chains):
        ch = ch.with_rb(rb)
        n1 = int(st[1].pairs_host['a_off'].max() + 5000)
        n2 = int(st[1].tasks_host['c_off'].max() + 5000)
        n3 = int(st[2].tasks_host['c_off'].max() + 5000)
        A = rng.standard_normal(n1)
        Wi, Wj = case5.mpo.host[isite], case5.mpo.host[isite + 1]
        Pi, Pj = case5.mpo.padded(isite)[0], case5.mpo.padded(isite + 1)[0]
        mid = np.zeros(n2)
        emulate(st[1], A, Wi, mid)
        want = np.zeros(n3)
        emulate(st[2], mid, Wj, want)
        got = np.zeros(n3)
        emulate_chain(ch, A, Pi, Pj, got)
        assert np.abs(got - want).max() < 1e-11 * max(1.0, np.abs(want).max())
        assert ch.flops == st[1].flops + st[2].flops


def test_balance_waves_keeps_ranges_and_evens_waves():
    """balance_waves permutes tasks only inside each row group's stage range;
    on these seeded synthetic ranges, cost-sorted as chain_structure hands
    them over, its most loaded wave is never above the round-robin of that
    sorted order (what the kernel would otherwise run) and within 4/3 of a
    brute-force search over all stride-4 placements (ranges of at most 7
    tasks)."""
    import itertools
    rng = np.random.default_rng(3)
    # synthetic row groups: ranges of 1..9 tasks, pairs with K 1..40
    lens = rng.integers(1, 10, size=40)
    nt = int(lens.sum())
    groups = np.zeros(20, pq._lib.CGROUP_DTYPE)
    off = np.concatenate([[0], np.cumsum(lens)])
    groups['t1_begin'], groups['t1_end'] = off[0:40:2], off[1:41:2]
    groups['t2_begin'], groups['t2_end'] = off[1:41:2], off[2:41:2]
    tasks = np.zeros(nt, pq._lib.TASK_DTYPE)
    npair = rng.integers(1, 4, size=nt)
    tasks['p_begin'] = np.concatenate([[0], np.cumsum(npair)[:-1]])
    tasks['p_end'] = np.cumsum(npair)
    tasks['N'] = rng.integers(1, 70, size=nt)
    tasks['c_off'] = np.arange(nt) * 1000
    pairs = np.zeros(int(npair.sum()), pq._lib.PAIR_DTYPE)
    pairs['K'] = rng.integers(1, 41, size=len(pairs))
    # the production input is already cost-sorted inside each range by
    # chain_structure (k-steps x 32-column tiles, descending, stable)
    prow = np.repeat(np.arange(nt), npair)
    kp = np.bincount(prow, weights=(pairs['K'].astype(np.int64) + 3) // 4, minlength=nt)
    sort_cost = kp * ((tasks['N'].astype(np.int64) + 31) // 32)
    seg = np.repeat(np.arange(len(lens)), lens)
    tasks = tasks[np.lexsort((np.arange(nt), -sort_cost, seg))]
    perm = pq.balance_waves(groups, tasks, pairs)
    assert np.array_equal(np.sort(perm), np.arange(nt))
    ks = (pairs['K'].astype(np.int64) + 3) // 4
    cum = np.concatenate([[0], np.cumsum(ks)])
    cost = (cum[tasks['p_end']] - cum[tasks['p_begin']] + 1) * ((tasks['N'].astype(np.int64) + 31) // 32)

    def wave_max(c):
        return max(c[w::4].sum() for w in range(4))

    for lo, hi in zip(off[:-1], off[1:]):
        assert sorted(perm[lo:hi]) == list(range(lo, hi))          # stays in its range
        assert wave_max(cost[perm[lo:hi]]) <= wave_max(cost[lo:hi])
        if hi - lo <= 7:
            best = min(wave_max(cost[list(q)]) for q in itertools.permutations(range(lo, hi)))
            assert wave_max(cost[perm[lo:hi]]) <= best * 4 / 3 + 1e-9


def test_allowed_block_cache_matches_rule_and_is_private():
    """fromQsyms' cached allowed-block lists equal the block rule's result and
    each tensor gets its own copy."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        rank = int(rng.integers(2, 5))
        nsyms = rng.integers(1, 5, size=rank)
        qsyms = [rng.integers(-2, 3, size=(n, 2)).astype(np.float64) for n in nsyms]
        ndims = [np.ones(n, np.int64) for n in nsyms]
        status = [bool(x) for x in rng.integers(0, 2, size=rank)]
        a = pq.QTensor(status)
        a.fromQsyms(rank, qsyms, ndims)
        b = pq.QTensor(status)
        b.fromQsyms(rank, qsyms, ndims)                             # cache hit
        want = np.flatnonzero(pq._lib.sym_allowed(np.array(nsyms, np.int32), status,
                                                  [q.reshape(len(q), -1) for q in qsyms]))
        assert np.array_equal(a.idx_allowed, want) and np.array_equal(b.idx_allowed, want)
        if len(a.idx_allowed):
            a.idx_allowed[0] = -7
            assert b.idx_allowed[0] != -7
