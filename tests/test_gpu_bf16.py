"""BF16 operands with FP32 accumulation on the gfx950 BF16 MFMA
(ZMPO_BF16 / ZMPO_BF16_F32): the grouped-GEMM kernels of the sigma launches
against numpy on BF16-representable inputs (so that only the FP32
accumulation and the output rounding differ), the chained launch against a
host model that rounds its intermediate like the kernel, and the whole BF16
sigma against FP64."""
import numpy as np
import pytest

import zmpo_dmrg_amd._lib as L
from kmodel import emulate
from parity_util import Case
from zmpo_dmrg_amd import device as dv
from zmpo_dmrg_amd import qtensor as pq

pytestmark = pytest.mark.gpu


def bf16(x):
    """Round to BF16 (nearest even) and back to float64, on the host."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _dense_program(M, N, K, trans_a=False, trans_b=False):
    t = np.zeros(1, L.TASK_DTYPE)
    t[0] = (0, 0, M, N, M, 0, N, 1, 0, 1, 0, 0)
    p = np.zeros(1, L.PAIR_DTYPE)
    a_sm, a_sk = (1, M) if trans_a else (K, 1)
    b_sk, b_sn = (1, K) if trans_b else (N, 1)
    p[0] = (0, 0, 1.0, K, M, 0, a_sm, a_sk, N, 0, b_sn, b_sk, 0)
    return pq.Program(t, p)


def test_host_rounding_is_nearest_even():
    assert bf16(1.0 + 2 ** -8) == 1.0                  # tie -> even
    assert bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6    # tie -> even (up)
    assert bf16(1.0 + 2 ** -7) == 1.0 + 2 ** -7        # representable


@pytest.mark.parametrize('shape', [(64, 64, 16), (37, 90, 33), (130, 7, 200), (5, 9, 5000),
                                   (1, 1, 1)])
@pytest.mark.parametrize('ta,tb', [(False, False), (True, False), (False, True)])
@pytest.mark.parametrize('kernel', ['shaped', 'stream'])
@pytest.mark.parametrize('dtype', [L.ZMPO_BF16, L.ZMPO_BF16_F32])
def test_bf16_grouped_gemm_dense(gpu, shape, ta, tb, kernel, dtype):
    M, N, K = shape
    rng = np.random.default_rng(M * N + K)
    A = bf16(rng.standard_normal(M * K))
    B = bf16(rng.standard_normal(K * N))
    prog = _dense_program(M, N, K, ta, tb)
    C = dv.zeros(M * N, 2 if dtype == L.ZMPO_BF16 else 1)
    pq.run_program(prog, dv.upload(A, 2), dv.upload(B, 2), C, dtype, kernel=kernel)
    Am = A.reshape(K, M).T if ta else A.reshape(M, K)
    Bm = B.reshape(N, K).T if tb else B.reshape(K, N)
    ref = Am @ Bm
    got = dv.download(C).astype(np.float64).reshape(M, N)
    scale = np.abs(Am) @ np.abs(Bm)
    if dtype == L.ZMPO_BF16_F32:
        tol = 1e-6 * K + 1e-6     # FP32 accumulation of exact products
    else:
        # output (and split-K partials) rounded to BF16
        tol = 2 ** -8 * 3 + 1e-6 * K
    assert np.all(np.abs(got - ref) <= tol * scale + 1e-30), (kernel, dtype, shape)


def test_bf16_chained_launch_matches_rounded_model(gpu):
    from kmodel import host_model
    from zmpo_dmrg_amd import kernelQt as kq
    case5 = Case(K=5, D=24, seed=4, sweeps=1, nelec=4)
    with host_model():
        Lh, Rh = case5.product_envs()
    isite = 2
    qkey, _, pt, _ = case5.templates(isite)
    sp = kq.SigmaProgram(case5.mpo, isite, pt, qkey, Lh[isite - 1], Rh[isite + 2],
                         case5.mpo.hpwts, fused=True)
    rng = np.random.default_rng(3)
    for st, ch in zip(sp.stages, sp.chains):
        for n16 in (False, True):
            for l_ in ch.launches:
                l_['n16'] = n16
            n1 = int(st[1].pairs_host['a_off'].max() + 5000)
            n2 = int(st[1].tasks_host['c_off'].max() + 5000)
            n3 = int(st[2].tasks_host['c_off'].max() + 5000)
            A = bf16(rng.standard_normal(n1))
            Wi, Wj = bf16(case5.mpo.host[isite]), bf16(case5.mpo.host[isite + 1])
            Pi = bf16(case5.mpo.padded(isite)[0])
            Pj = bf16(case5.mpo.padded(isite + 1)[0])
            mid = np.zeros(n2)
            emulate(st[1], A, Wi, mid)
            mid = bf16(mid)                 # the kernel keeps top2 in LDS as BF16
            want = np.zeros(n3)
            emulate(st[2], mid, Wj, want)
            C = dv.zeros(n3, 2)
            pq.run_chain(ch, dv.upload(A, 2), dv.upload(Pi, 2), dv.upload(Pj, 2), C, 2)
            err = np.abs(dv.download(C) - want).max()
            assert err < 2e-2 * max(1.0, np.abs(want).max()), (n16, err)


def test_bf16_sigma_close_to_fp64(gpu):
    from zmpo_dmrg_amd import kernelQt as kq
    c = Case(K=6, D=32, seed=2, sweeps=1, upload=True)
    L_, R_ = c.product_envs()
    for isite in (1, 2, 3):
        qkey, _, pt, _ = c.templates(isite)
        sp = kq.SigmaProgram(c.mpo, isite, pt, qkey, L_[isite - 1], R_[isite + 2], c.mpo.hpwts)
        v = np.random.default_rng(isite).standard_normal(pt.size_allowed)
        s64 = dv.empty(pt.size_allowed, 0)
        sp.apply(c.mpo, L_[isite - 1], R_[isite + 2], dv.upload(v, 0), s64, 0)
        s16 = dv.empty(pt.size_allowed, 1)
        sp.apply(c.mpo, L_[isite - 1], R_[isite + 2], dv.upload(v, 2), s16, 2)
        a, b = dv.download(s64), dv.download(s16).astype(np.float64)
        rel = np.linalg.norm(a - b) / np.linalg.norm(a)
        print('[bf16] K=6 site %d sigma relative error %.2e' % (isite, rel), flush=True)
        assert rel < 2e-2
