"""The sigma and diagonal launches at the full K=36, D=1000 size (72 operators,
ndim0 = 278932; the oracle cannot run this size), checked through properties
that hold at any size:

* H_eff is symmetric: u.(H v) = v.(H u);
* sigma is linear: H(a u + b v) = a H u + b H v;
* repeated applications are bit-identical (fixed summation order, no atomics);
* the diagonal launch agrees with the sigma launch: diag[i] = e_i.(H e_i);
* the FP32 sigma stays within FP32 rounding of the FP64 one.
"""
import numpy as np
import pytest

from zmpo_dmrg_amd import device as dv
from zmpo_dmrg_amd import kernelQt as kq

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def full(gpu):
    import bench
    mpo, isite, st, L, R, prog = bench.build_workload(36, 1000, 0, 0, 96 * 2 ** 30, fused=True)
    print('[fullsize] K=36 D=1000 ndim0=%d' % st.qtmp.size_allowed, flush=True)
    return mpo, isite, st, L, R, prog


def _sigma(full, v, dtype=0):
    mpo, isite, st, L, R, prog = full
    x = dv.upload(v, dtype)
    out = dv.empty(st.qtmp.size_allowed, dtype)
    prog.apply(mpo, L, R, x, out, dtype)
    return dv.download(out).astype(np.float64)


def test_fullsize_sigma_symmetric_linear_deterministic(full):
    n = full[2].qtmp.size_allowed
    rng = np.random.default_rng(5)
    u = rng.standard_normal(n)
    v = rng.standard_normal(n)
    Hu, Hv = _sigma(full, u), _sigma(full, v)
    scale = np.linalg.norm(Hu) * np.linalg.norm(v)
    assert abs(v @ Hu - u @ Hv) < 1e-12 * scale
    a, b = 0.75, -1.25
    Hw = _sigma(full, a * u + b * v)
    assert np.abs(Hw - (a * Hu + b * Hv)).max() < 1e-12 * np.abs(Hu).max() * 4
    assert np.array_equal(_sigma(full, u), Hu)


def test_fullsize_diagonal_matches_sigma(full):
    mpo, isite, st, L, R, prog = full
    n = st.qtmp.size_allowed
    d = kq.DiagProgram(mpo, isite, st.qtmp, L, R, mpo.hpwts)
    dg = dv.zeros(n, 0)
    d.apply(L, R, dg)
    diag = dv.download(dg)
    rng = np.random.default_rng(9)
    for i in rng.choice(n, size=4, replace=False):
        e = np.zeros(n)
        e[i] = 1.0
        hi = _sigma(full, e)[i]
        assert abs(hi - diag[i]) < 1e-11 * max(1.0, abs(hi)), (i, hi, diag[i])


def test_fullsize_fp32_sigma_within_rounding(full):
    n = full[2].qtmp.size_allowed
    u = np.random.default_rng(3).standard_normal(n)
    h64 = _sigma(full, u)
    h32 = _sigma(full, u, dtype=1)
    assert np.linalg.norm(h32 - h64) < 1e-5 * np.linalg.norm(h64)
