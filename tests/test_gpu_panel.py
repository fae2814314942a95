"""Register-blocked panel tiles (zmpo_shapedgemm ZMPO_PANEL_*): dense block
products against numpy in every storage type, split-K programs with mixed
operand orders and scales against the CPU kernel model, and whole sigma
applications against the default tiles."""
import numpy as np
import pytest

import zmpo_dmrg_amd._lib as L
from kmodel import emulate
from parity_util import Case
from zmpo_dmrg_amd import device as dv
from zmpo_dmrg_amd import qtensor as pq

pytestmark = pytest.mark.gpu

SIZES = [(64, 64, 16), (37, 90, 33), (130, 7, 200), (300, 150, 140), (1, 1, 1), (5, 9, 5000)]


def bf16(x):
    """Round to BF16 (nearest even) and back to float64, on the host."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _dense_program(M, N, K, trans_a=False, trans_b=False, scale=1.0):
    t = np.zeros(1, L.TASK_DTYPE)
    t[0] = (0, 0, M, N, M, 0, N, 1, 0, 1, 0, 0)
    p = np.zeros(1, L.PAIR_DTYPE)
    a_sm, a_sk = (1, M) if trans_a else (K, 1)
    b_sk, b_sn = (1, K) if trans_b else (N, 1)
    p[0] = (0, 0, scale, K, M, 0, a_sm, a_sk, N, 0, b_sn, b_sk, 0)
    return pq.Program(t, p)


@pytest.mark.parametrize('shape', pq.PANEL_SHAPES)
@pytest.mark.parametrize('dtype', [L.F64, L.F32, L.ZMPO_BF16, L.ZMPO_BF16_F32])
def test_panel_dense(gpu, shape, dtype):
    store = {L.F64: 0, L.F32: 1, L.ZMPO_BF16: 2, L.ZMPO_BF16_F32: 2}[dtype]
    cstore = {L.F64: 0, L.F32: 1, L.ZMPO_BF16: 2, L.ZMPO_BF16_F32: 1}[dtype]
    for M, N, K in SIZES:
        for ta, tb in ((False, False), (True, False), (False, True), (True, True)):
            for scale in (1.0, -1.0, 0.75):
                rng = np.random.default_rng(M * N + K + 7 * ta + 3 * tb)
                A = rng.standard_normal(M * K)
                B = rng.standard_normal(K * N)
                if store == 2:
                    A, B = bf16(A), bf16(B)
                elif store == 1:
                    A, B = A.astype(np.float32), B.astype(np.float32)
                prog = _dense_program(M, N, K, ta, tb, scale)
                prog.shape = shape
                C = dv.zeros(M * N, cstore)
                pq.run_program(prog, dv.upload(A, store), dv.upload(B, store), C, dtype,
                               kernel='shaped')
                Am = (A.reshape(K, M).T if ta else A.reshape(M, K)).astype(np.float64)
                Bm = (B.reshape(N, K).T if tb else B.reshape(K, N)).astype(np.float64)
                ref = scale * (Am @ Bm)
                got = dv.download(C).astype(np.float64).reshape(M, N)
                mag = np.abs(Am) @ np.abs(Bm)
                if dtype == L.F64:
                    tol = 1e-15 * K + 1e-15
                elif dtype == L.F32 or dtype == L.ZMPO_BF16_F32:
                    # FP32 accumulation (the 0.75 scale rounds BF16 operands)
                    tol = 1e-6 * K + (2 ** -8 if (store == 2 and scale == 0.75) else 1e-6)
                else:
                    tol = 2 ** -8 * 3 + 1e-6 * K
                assert np.all(np.abs(got - ref) <= tol * mag + 1e-30), (shape, dtype, M, N, K,
                                                                        ta, tb, scale)


def _random_program(rng, nt=40, kmax=3000, twolevel=False):
    tasks = np.zeros(nt, L.TASK_DTYPE)
    pairs, off = [], 0
    for i in range(nt):
        M, N = int(rng.integers(1, 200)), int(rng.integers(1, 160))
        npair = int(rng.integers(0, 6))
        if twolevel and M % 2 == 0:
            # rows (m // 2, m % 2) with the low index the slower one
            tasks[i] = (off, 0, M, N, 2, N, M // 2 * N, 1, len(pairs), len(pairs) + npair, 0, 0)
        else:
            tasks[i] = (off, 0, M, N, M, 0, N, 1, len(pairs), len(pairs) + npair, 0, 0)
        off += M * N
        for j in range(npair):
            K = int(rng.integers(1, kmax))
            a = (M, 0, K, 1) if (i + j) % 2 else (M, 0, 1, M)
            b = (N, 0, 1, N) if (i + j) % 3 else (N, 0, K, 1)
            pairs.append((int(rng.integers(0, 1000)), int(rng.integers(0, 1000)),
                          float(rng.choice([1.0, -1.0, rng.standard_normal()])), K) + a + b + (0,))
    return tasks, np.array(pairs, dtype=L.PAIR_DTYPE), off


@pytest.mark.parametrize('twolevel', [False, True])
def test_panel_split_k_random_programs(gpu, twolevel):
    """Multi-pair tasks split over several units (small summed K per unit)
    and reduced in slot order: FP64 against the CPU model, BF16 with FP32
    outputs against the model on rounded operands."""
    rng = np.random.default_rng(5 + twolevel)
    tasks, pairs, off = _random_program(rng, twolevel=twolevel)
    A, B = rng.standard_normal(700000), rng.standard_normal(700000)
    want = np.zeros(off)
    emulate(pq.Program(tasks, pairs), A, B, want)
    for shape in pq.PANEL_SHAPES:
        for kchunk in (None, 700, 5000):
            prog = pq.Program(tasks, pairs)
            prog.shape, prog.kchunk = shape, kchunk
            C = dv.zeros(off, 0)
            pq.run_program(prog, dv.upload(A, 0), dv.upload(B, 0), C, 0, kernel='shaped')
            err = np.abs(dv.download(C) - want).max()
            assert err < 1e-10, (shape, kchunk, err)
    Ab, Bb = bf16(A), bf16(B)
    pb = pairs.copy()
    pb['scale'] = np.where(np.abs(pb['scale']) == 1.0, pb['scale'], 1.0)
    want = np.zeros(off)
    emulate(pq.Program(tasks, pb), Ab, Bb, want)
    mag = np.zeros(off)
    pa = pb.copy()
    pa['scale'] = 1.0
    emulate(pq.Program(tasks, pa), np.abs(Ab), np.abs(Bb), mag)
    for shape in pq.PANEL_SHAPES:
        prog = pq.Program(tasks, pb)
        prog.shape, prog.kchunk = shape, 700
        C = dv.zeros(off, 1)
        pq.run_program(prog, dv.upload(Ab, 2), dv.upload(Bb, 2), C, L.ZMPO_BF16_F32,
                       kernel='shaped')
        got = dv.download(C).astype(np.float64)
        assert np.all(np.abs(got - want) <= 1e-5 * mag + 1e-30), shape


def test_panel_beta_accumulates(gpu):
    M = N = K = 70
    rng = np.random.default_rng(2)
    A, B, C0 = (rng.standard_normal(n) for n in (M * K, K * N, M * N))
    for shape in pq.PANEL_SHAPES:
        C = dv.upload(C0, 0)
        prog = _dense_program(M, N, K)
        prog.shape = shape
        pq.run_program(prog, dv.upload(A, 0), dv.upload(B, 0), C, 0, beta=1, kernel='shaped')
        ref = C0.reshape(M, N) + A.reshape(M, K) @ B.reshape(K, N)
        assert np.abs(dv.download(C).reshape(M, N) - ref).max() < 1e-12, shape


@pytest.mark.parametrize('dtype', [0, 1, 2])
def test_panel_sigma_matches_default_tiles(gpu, dtype):
    """A whole sigma (K = 6, D = 32: two-level row maps, many small blocks)
    with the L x psi and x R launches on each panel tile against the default
    tiles: FP64 to rounding, FP32 and BF16 to their operand precision."""
    from zmpo_dmrg_amd import kernelQt as kq
    c = Case(K=6, D=32, seed=2, sweeps=1, upload=True)
    L_, R_ = c.product_envs()
    isite = 2
    qkey, _, pt, _ = c.templates(isite)
    n = pt.size_allowed
    v = np.random.default_rng(9).standard_normal(n)
    base = kq.SigmaProgram(c.mpo, isite, pt, qkey, L_[isite - 1], R_[isite + 2], c.mpo.hpwts)
    s0 = dv.empty(n, kq.sigma_dtype(dtype))
    base.apply(c.mpo, L_[isite - 1], R_[isite + 2], dv.upload(v, dtype), s0, dtype)
    want = dv.download(s0).astype(np.float64)
    tol = {0: 1e-12, 1: 1e-5, 2: 2e-2}[dtype] * max(1.0, np.abs(want).max())
    for shape in pq.PANEL_SHAPES:
        sp = kq.SigmaProgram(c.mpo, isite, pt, qkey, L_[isite - 1], R_[isite + 2], c.mpo.hpwts)
        for st in sp.stages:
            for k in (0, 3):
                st[k].shape = shape
                st[k].prefer = 'shaped'
        s1 = dv.empty(n, kq.sigma_dtype(dtype))
        sp.apply(c.mpo, L_[isite - 1], R_[isite + 2], dv.upload(v, dtype), s1, dtype)
        err = np.abs(dv.download(s1).astype(np.float64) - want).max()
        assert err < tol, (shape, dtype, err)


@pytest.mark.parametrize('dtype', [L.F64, L.F32, L.ZMPO_BF16_F32])
def test_panel_32x160_on_k_contiguous_operands(gpu, dtype):
    """The 32 x 160 tile (k-contiguous operands only) on single-block and
    split-K programs against numpy; other programs fall back to 32 x 32."""
    store = {L.F64: 0, L.F32: 1, L.ZMPO_BF16_F32: 2}[dtype]
    cstore = {L.F64: 0, L.F32: 1, L.ZMPO_BF16_F32: 1}[dtype]
    for M, N, K in [(134, 134, 2500), (37, 152, 300), (1, 1, 1), (200, 170, 33)]:
        rng = np.random.default_rng(M + N + K)
        A = rng.standard_normal(M * K)
        B = rng.standard_normal(K * N)
        if store == 2:
            A, B = bf16(A), bf16(B)
        prog = _dense_program(M, N, K, False, True)     # A (M, K), B stored (N, K)
        assert prog.kmajor_ok()
        prog.tiles[dtype] = ('shaped', 12)
        prog.kchunk = 700                                 # split K over several units
        assert prog.launch_choice(dtype) == ('shaped', 12)
        C = dv.zeros(M * N, cstore)
        pq.run_program(prog, dv.upload(A, store), dv.upload(B, store), C, dtype)
        Am, Bm = A.reshape(M, K), B.reshape(N, K).T
        ref = Am @ Bm
        mag = np.abs(Am) @ np.abs(Bm)
        got = dv.download(C).astype(np.float64).reshape(M, N)
        tol = 1e-15 * K + 1e-15 if dtype == L.F64 else 1e-6 * K + 1e-6
        assert np.all(np.abs(got - ref) <= tol * mag + 1e-30), (dtype, M, N, K)
