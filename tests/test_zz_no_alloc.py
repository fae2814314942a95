"""zmpo_streamgemm_static (the allocation-free entry point) against
zmpo_streamgemm, its counterpart with a caller-supplied workspace, through
the C-ABI: bit-identical outputs on two calls in a row, the second with a
smaller input, on random programs whose long pair lists use split-K slots,
and on a real sigma top3 x R program; sizes beyond the static scratch are
rejected before any launch."""
import numpy as np
import pytest

import zmpo_dmrg_amd._lib as L
from zmpo_dmrg_amd import device as dv
from zmpo_dmrg_amd import qtensor as pq

pytestmark = pytest.mark.gpu


def _random_program(rng, nt, kmax):
    tasks = np.zeros(nt, L.TASK_DTYPE)
    pairs, off = [], 0
    for i in range(nt):
        M, N = int(rng.integers(1, 70)), int(rng.integers(1, 40))
        npair = int(rng.integers(1, 6))
        tasks[i] = (off, 0, M, N, M, 0, N, 1, len(pairs), len(pairs) + npair, 0, 0)
        off += M * N
        for j in range(npair):
            K = int(rng.integers(1, kmax))
            a = (M, 0, K, 1) if (i + j) % 2 else (M, 0, 1, M)
            b = (N, 0, 1, N) if (i + j) % 3 else (N, 0, K, 1)
            pairs.append((int(rng.integers(0, 1000)), int(rng.integers(0, 1000)),
                          float(rng.standard_normal()), K) + a + b + (0,))
    return pq.Program(tasks, np.array(pairs, dtype=L.PAIR_DTYPE)), off


def _both(prog, A, B, nc, dtype):
    """C from zmpo_streamgemm (library workspace) and zmpo_streamgemm_static."""
    lib = L.lib()
    prog.shape = 2
    tasks, pairs = prog.device()
    sc = prog.schedule('stream')
    kt = 16 if dtype == 0 else 32
    welem = sc['nslots'] * sc['slot_size']
    work = pq._workspace(welem, dtype)
    c1 = dv.zeros(nc, dtype)
    c2 = dv.zeros(nc, dtype)
    args = (dtype, prog.shape, kt, dv.ptr(sc['units']), sc['nunits'], sc['upw'], dv.ptr(tasks),
            dv.ptr(sc['groups']), sc['ngroups'], dv.ptr(pairs), dv.ptr(A), dv.ptr(B))
    L.check(lib.zmpo_streamgemm(*args, dv.ptr(c1), dv.ptr(work), 0, dv.stream()))
    L.check(lib.zmpo_streamgemm_static(*args, dv.ptr(c2), welem, 0, dv.stream()))
    dv.sync()
    return dv.download(c1), dv.download(c2), sc['ngroups']


@pytest.mark.parametrize('dtype', [0, 1])
def test_static_scratch_matches_workspace_twice_shrinking(gpu, dtype):
    rng = np.random.default_rng(7 + dtype)
    ngroups_seen = []
    for nt, kmax in ((60, 4000), (9, 2500)):        # second call: smaller input
        prog, nc = _random_program(rng, nt, kmax)
        A = dv.upload(rng.standard_normal(300000), dtype)
        B = dv.upload(rng.standard_normal(300000), dtype)
        c1, c2, ng = _both(prog, A, B, nc, dtype)
        ngroups_seen.append(ng)
        assert np.array_equal(c1, c2), (nt, dtype)
        assert np.abs(c1).max() > 0
    assert min(ngroups_seen) > 0          # both calls went through split-K slots


def test_static_scratch_on_sigma_stage4(gpu):
    """The top3 x R program of a real two-site sigma (K = 6, D = 32)."""
    from parity_util import Case
    from zmpo_dmrg_amd import kernelQt as kq
    c = Case(K=6, D=32, seed=2, sweeps=1, upload=True)
    L_, R_ = c.product_envs()
    isite = 2
    qkey, _, pt, _ = c.templates(isite)
    sp = kq.SigmaProgram(c.mpo, isite, pt, qkey, L_[isite - 1], R_[isite + 2], c.mpo.hpwts)
    st4 = sp.stages[0][3]
    rng = np.random.default_rng(3)
    A = dv.upload(rng.standard_normal(max(sp.inter[2], 1)), 0)
    c1, c2, _ = _both(st4, A, R_[isite + 2].buffer(0), pt.size_allowed, 0)
    assert np.array_equal(c1, c2)


def test_static_scratch_rejects_oversize(gpu):
    lib = L.lib()
    cap = lib.zmpo_scratch_bytes()
    assert cap == 64 * 2 ** 20
    rc = lib.zmpo_streamgemm_static(0, 2, 16, None, 0, 1, None, None, 0, None, None, None, None,
                                    cap // 8 + 1, 0, None)
    assert rc == L.ZMPO_ERR_ARG
    assert b'64 MiB' in lib.zmpo_last_error()
