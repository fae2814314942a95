"""Device vs oracle at K = 24, D = 500 (the FP64 drop-in size): sigma and
diagonal through the drop-in HVecQt / HDiagQt with one operator of each
class weighted (all 48 operators' launches run), and every block of the
renormalised environments of three operators in both directions.  The
environments are built on the device from a synthetic MPS with a DMRG-like
(N, Sz) sector structure."""
import pytest

pytestmark = pytest.mark.gpu

EXTRA = {}


@pytest.fixture(scope='module')
def k24(gpu):
    import bench
    print('[k24] building K=24 D=500', flush=True)
    w = bench.build_workload(24, 500, 1, 0, 32 * 2 ** 30, fused=True, extra=EXTRA,
                             structure='wide')
    yield w
    EXTRA.clear()


def test_k24_sigma_and_diagonal_match_oracle(k24):
    from fullsize_util import check_sigma_and_diag
    mpo, isite, st, L, R, prog = k24
    serr, derr, sample = check_sigma_and_diag(mpo, isite, st, L, R, 32 * 2 ** 30, seed=2)
    assert len(sample) == 8
    assert serr < 1e-12 and derr < 1e-12


@pytest.mark.parametrize('side', ['L', 'R'])
def test_k24_renorm_matches_oracle(k24, side):
    from fullsize_util import check_renorm, one_op_per_case
    mpo, isite, st, L, R, prog = k24
    if side == 'L':
        prev, site, idx, new = EXTRA['Lprev'], EXTRA['lsite'], EXTRA['lidx'], L
    else:
        prev, site, idx, new = EXTRA['Rprev'], EXTRA['rsite'], EXTRA['ridx'], R
    sample = one_op_per_case(mpo, idx, 1)
    assert check_renorm(mpo, prev, site, idx, new, side, sample) < 1e-12
