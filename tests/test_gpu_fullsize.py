"""The sigma and diagonal launches at the full K=36, D=1000 size (72 operators,
ndim0 = 278932).  Against the oracle on one operator of each class (the
oracle cannot run all 72 in reasonable time): sigma and the diagonal through
the drop-in HVecQt / HDiagQt, and every block of the renormalised left and
right environments of those operators.  Through properties that hold at any
size, on all operators:

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


EXTRA = {}


@pytest.fixture(scope='module')
def full(gpu):
    import bench
    print('[fullsize] building K=36 D=1000', flush=True)
    mpo, isite, st, L, R, prog = bench.build_workload(36, 1000, 0, 0, 96 * 2 ** 30, fused=True,
                                                      extra=EXTRA, structure='wide')
    print('[fullsize] K=36 D=1000 ndim0=%d' % st.qtmp.size_allowed, flush=True)
    yield mpo, isite, st, L, R, prog
    EXTRA.clear()


def test_fullsize_sigma_and_diagonal_match_oracle_on_each_case(full):
    from fullsize_util import check_sigma_and_diag
    mpo, isite, st, L, R, prog = full
    serr, derr, sample = check_sigma_and_diag(mpo, isite, st, L, R, 96 * 2 ** 30)
    assert len(sample) == 8
    assert serr < 1e-12 and derr < 1e-12


@pytest.mark.parametrize('side', ['L', 'R'])
def test_fullsize_renorm_matches_oracle(full, side):
    from fullsize_util import check_renorm, one_op_per_case
    mpo, isite, st, L, R, prog = full
    if side == 'L':
        prev, site, idx, new = EXTRA['Lprev'], EXTRA['lsite'], EXTRA['lidx'], L
    else:
        prev, site, idx, new = EXTRA['Rprev'], EXTRA['rsite'], EXTRA['ridx'], R
    sample = one_op_per_case(mpo, idx, 1)[:3]
    assert check_renorm(mpo, prev, site, idx, new, side, sample) < 1e-12


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
