"""libzmpo kernels on the device vs numpy / the CPU kernel model."""
import numpy as np
import pytest

from kmodel import emulate
from parity_util import Case
import zmpo_dmrg_amd._lib as L
from zmpo_dmrg_amd import device as dv
from zmpo_dmrg_amd import qtensor as pq

pytestmark = pytest.mark.gpu


def _dense_program(M, N, K, trans_a=False, trans_b=False):
    t = np.zeros(1, L.TASK_DTYPE)
    t[0] = (0, 0, M, N, M, 0, N, 1, 0, 1, 0, 0)
    p = np.zeros(1, L.PAIR_DTYPE)
    a_sm, a_sk = (1, M) if trans_a else (K, 1)
    b_sk, b_sn = (1, K) if trans_b else (N, 1)
    p[0] = (0, 0, 1.0, K, M, 0, a_sm, a_sk, N, 0, b_sn, b_sk, 0)
    return pq.Program(t, p)


@pytest.mark.parametrize('shape', [(64, 64, 16), (37, 90, 33), (130, 7, 200), (1, 1, 1),
                                   (5, 9, 5000)])
@pytest.mark.parametrize('ta,tb', [(False, False), (True, False), (False, True)])
@pytest.mark.parametrize('kernel', ['wave', 'tile', 'shaped', 'stream', 'skinny', 'dot'])
def test_blockgemm_dense_f64(gpu, shape, ta, tb, kernel):
    M, N, K = shape
    rng = np.random.default_rng(M * N + K)
    A = rng.standard_normal(M * K)
    B = rng.standard_normal(K * N)
    prog = _dense_program(M, N, K, ta, tb)
    C = dv.zeros(M * N, 0)
    pq.run_program(prog, dv.upload(A, 0), dv.upload(B, 0), C, 0, kernel=kernel)
    Am = A.reshape(K, M).T if ta else A.reshape(M, K)
    Bm = B.reshape(N, K).T if tb else B.reshape(K, N)
    ref = Am @ Bm
    got = dv.download(C).reshape(M, N)
    assert np.abs(got - ref).max() < 1e-13 * max(1.0, np.abs(ref).max()) * K


@pytest.mark.parametrize('shape', [(64, 64, 16), (37, 90, 33)])
def test_blockgemm_dense_f32(gpu, shape):
    M, N, K = shape
    rng = np.random.default_rng(1)
    A = rng.standard_normal(M * K).astype(np.float32)
    B = rng.standard_normal(K * N).astype(np.float32)
    C = dv.zeros(M * N, 1)
    pq.run_program(_dense_program(M, N, K), dv.upload(A, 1), dv.upload(B, 1), C, 1)
    ref = A.reshape(M, K).astype(np.float64) @ B.reshape(K, N).astype(np.float64)
    assert np.abs(dv.download(C).reshape(M, N) - ref).max() < 1e-5 * K


def test_wavegemm_split_k_random_programs(gpu):
    """Long pair lists (split-K with workspace reduction) vs the CPU model, rb=1,2,4."""
    rng = np.random.default_rng(0)
    nt = 40
    tasks = np.zeros(nt, L.TASK_DTYPE)
    pairs, off = [], 0
    for i in range(nt):
        M, N = int(rng.integers(1, 90)), int(rng.integers(1, 40))
        npair = int(rng.integers(0, 6))
        tasks[i] = (off, 0, M, N, M, 0, N, 1, len(pairs), len(pairs) + npair, 0, 0)
        off += M * N
        for j in range(npair):
            K = int(rng.integers(1, 3000))
            # consecutive pairs alternate between K-contiguous and
            # row-contiguous operands (A and B independently)
            a = (M, 0, K, 1) if (i + j) % 2 else (M, 0, 1, M)
            b = (N, 0, 1, N) if (i + j) % 3 else (N, 0, K, 1)
            pairs.append((int(rng.integers(0, 1000)), int(rng.integers(0, 1000)),
                          float(rng.standard_normal()), K) + a + b + (0,))
    pairs = np.array(pairs, dtype=L.PAIR_DTYPE)
    A, B = rng.standard_normal(400000), rng.standard_normal(400000)
    want = np.zeros(off)
    prog = pq.Program(tasks, pairs)
    emulate(prog, A, B, want)
    configs = [('wave', r, 0, 0, 0) for r in (1, 2, 4)] + [('shaped', 1, s, 0, 0) for s in pq.SHAPES]
    configs += [('stream', 1, s, kt, upw) for s in pq.SHAPES for kt in (16, 32) for upw in (1, 3, 64)]
    configs += [('dot', 1, 2, 0, 0)]
    for kernel, rb, shape, kt, upw in configs:
        prog = pq.Program(tasks, pairs)
        prog.rb, prog.shape, prog.kt, prog.upw = rb, shape, kt or None, upw or None
        C = dv.zeros(off, 0)
        pq.run_program(prog, dv.upload(A, 0), dv.upload(B, 0), C, 0, kernel=kernel)
        assert np.abs(dv.download(C) - want).max() < 1e-10, (kernel, rb, shape, kt, upw)


@pytest.mark.parametrize('kernel', ['wave', 'tile', 'shaped', 'stream', 'skinny', 'dot'])
def test_blockgemm_beta_accumulates(gpu, kernel):
    M = N = K = 20
    rng = np.random.default_rng(2)
    A, B, C0 = (rng.standard_normal(n) for n in (M * K, K * N, M * N))
    C = dv.upload(C0, 0)
    pq.run_program(_dense_program(M, N, K), dv.upload(A, 0), dv.upload(B, 0), C, 0, beta=1,
                   kernel=kernel)
    ref = C0.reshape(M, N) + A.reshape(M, K) @ B.reshape(K, N)
    assert np.abs(dv.download(C).reshape(M, N) - ref).max() < 1e-12


@pytest.fixture(scope='module')
def case4():
    return Case(K=4, D=16, seed=3, sweeps=1)


def test_blockgemm_real_plans_match_model(gpu, case4):
    """Every stage program of one sigma: device == CPU model (FP64 rounding)."""
    from kmodel import host_model
    from zmpo_dmrg_amd import kernelQt as kq
    with host_model():
        Lh, Rh = case4.product_envs()
    isite = 1
    qkey, _, pt, _ = case4.templates(isite)
    prog = kq.SigmaProgram(case4.mpo, isite, pt, qkey, Lh[isite - 1], Rh[isite + 2],
                           case4.mpo.hpwts)
    rng = np.random.default_rng(0)
    for st in prog.stages:
        for k, pg in enumerate(st):
            if pg.ntasks == 0:
                continue
            na = int(max(pg.pairs_host['a_off'].max() + 10000, 1))
            nb = int(max(pg.pairs_host['b_off'].max() + 10000, 1))
            nc = int(pg.tasks_host['c_off'].max() + 10000)
            A, B = rng.standard_normal(na), rng.standard_normal(nb)
            want = np.zeros(nc)
            emulate(pg, A, B, want)
            for kernel, upw in (('shaped', None), ('stream', None), ('stream', 64), ('skinny', None),
                                ('dot', None)):
                pg.upw = upw
                pg._sched = None
                C = dv.zeros(nc, 0)
                pq.run_program(pg, dv.upload(A, 0), dv.upload(B, 0), C, 0, kernel=kernel)
                err = np.abs(dv.download(C) - want).max()
                assert err < 1e-12 * max(1.0, np.abs(want).max()), (kernel, upw, k)


def test_vector_kernels(gpu):
    lib = L.lib()
    n = 100003
    rng = np.random.default_rng(3)
    X = rng.standard_normal((3, n))
    y = rng.standard_normal(n)
    Xd, yd = dv.upload(X, 0).view(3, n), dv.upload(y, 0)
    out = dv.zeros(3, 0)
    work = dv.empty(lib.zmpo_dots_workspace(3), 0)
    L.check(lib.zmpo_dots(0, dv.ptr(Xd), n, 3, dv.ptr(yd), n, dv.ptr(out), dv.ptr(work), dv.stream()))
    assert np.abs(dv.download(out) - X @ y).max() < 1e-10
    c = np.array([0.5, -2.0, 3.0])
    z = rng.standard_normal(n)
    zd = dv.upload(z, 0)
    L.check(lib.zmpo_lincomb(0, dv.ptr(zd), dv.ptr(Xd), n, 3, dv.ptr(dv.upload(c, 0)), -1.5, 2.0,
                             n, dv.stream()))
    assert np.abs(dv.download(zd) - (2.0 * z - 1.5 * (c @ X))).max() < 1e-12
    diag = rng.standard_normal(n)
    r = rng.standard_normal(n)
    rd = dv.upload(r, 0)
    L.check(lib.zmpo_precond(0, dv.ptr(rd), dv.ptr(dv.upload(diag, 0)), 0.3, 1e-4, n, dv.stream()))
    assert np.abs(dv.download(rd) - r / (np.abs(diag - 0.3) + 1e-4)).max() < 1e-9
    idx = rng.permutation(n).astype(np.int64)
    g = dv.zeros(n, 0)
    pq.gather(g, dv.upload(y, 0), dv.upload(idx))
    assert np.array_equal(dv.download(g), y[idx])
    # deterministic reduction: repeated dots are bitwise identical
    vals = []
    for _ in range(3):
        L.check(lib.zmpo_dots(0, dv.ptr(Xd), n, 3, dv.ptr(yd), n, dv.ptr(out), dv.ptr(work),
                              dv.stream()))
        vals.append(dv.download(out).copy())
    assert all(np.array_equal(vals[0], v) for v in vals)


def test_errors_surface(gpu):
    lib = L.lib()
    rc = lib.zmpo_blockgemm(7, None, 1, None, 1, None, None, None, 0, None)
    assert rc != 0
    assert lib.zmpo_last_error()


@pytest.mark.parametrize('n16', [0, 1])
@pytest.mark.parametrize('atile', [1, 0])
@pytest.mark.parametrize('dtype', [0, 1])
@pytest.mark.parametrize('rb', [1, 2])
def test_chaingemm_real_plan_matches_two_stages(gpu, rb, dtype, atile, n16):
    """zmpo_chaingemm / zmpo_chaingemm16 (sigma stages 2-3 with the
    intermediate in LDS; 32- or 16-column tiles, every launch forced to one
    of them) vs the CPU model of the two separate stages, on a real K=5 sigma
    plan."""
    from kmodel import host_model
    from zmpo_dmrg_amd import kernelQt as kq
    case5 = Case(K=5, D=24, seed=4, sweeps=1, nelec=4)
    with host_model():
        Lh, Rh = case5.product_envs()
    isite = 2
    qkey, _, pt, _ = case5.templates(isite)
    sp = kq.SigmaProgram(case5.mpo, isite, pt, qkey, Lh[isite - 1], Rh[isite + 2],
                         case5.mpo.hpwts, fused=True)
    rng = np.random.default_rng(rb)
    for st, ch in zip(sp.stages, sp.chains):
        ch = ch.with_rb(rb)
        for l_ in ch.launches:   # A staged in LDS (1) or read from global memory (0)
            l_['atile'] = atile
            l_['n16'] = bool(n16)
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
        C = dv.zeros(n3, dtype)
        pq.run_chain(ch, dv.upload(A, dtype), dv.upload(Pi, dtype), dv.upload(Pj, dtype), C, dtype)
        tol = 1e-12 if dtype == 0 else 2e-5
        err = np.abs(dv.download(C).astype(np.float64) - want).max()
        assert err < tol * max(1.0, np.abs(want).max()), (rb, dtype, err)


def test_chaingemm16_bitwise_equal_to_chaingemm(gpu):
    """The 16-column tiles change which MFMA block holds a column, not the
    per-element summation order: FP64 results are bit-identical (K=6 plan)."""
    from kmodel import host_model
    from zmpo_dmrg_amd import kernelQt as kq
    case6 = Case(K=6, D=32, seed=2, sweeps=1)
    with host_model():
        Lh, Rh = case6.product_envs()
    isite = 2
    qkey, _, pt, _ = case6.templates(isite)
    sp = kq.SigmaProgram(case6.mpo, isite, pt, qkey, Lh[isite - 1], Rh[isite + 2],
                         case6.mpo.hpwts, fused=True)
    rng = np.random.default_rng(7)
    for st, ch in zip(sp.stages, sp.chains):
        n1 = int(st[1].pairs_host['a_off'].max() + 5000)
        n3 = int(st[2].tasks_host['c_off'].max() + 5000)
        A = dv.upload(rng.standard_normal(n1), 0)
        Pi = dv.upload(case6.mpo.padded(isite)[0], 0)
        Pj = dv.upload(case6.mpo.padded(isite + 1)[0], 0)
        out = []
        for flag in (False, True):
            for l_ in ch.launches:
                l_['n16'] = flag
            C = dv.zeros(n3, 0)
            pq.run_chain(ch, A, Pi, Pj, C, 0)
            out.append(dv.download(C))
        assert np.array_equal(out[0], out[1])
        assert np.abs(out[0]).max() > 0
