"""Device sigma / diagonal / renormalisation / sweeps vs the FP64 oracle."""
import os

import numpy as np
import pytest

from parity_util import Case
from zmpo_dmrg_amd import device as dv
from zmpo_dmrg_amd import dmrg as pdmrg
from zmpo_dmrg_amd import kernelQt as kq
from zmpo_dmrg_amd.integrals import gen_integrals, gen_spatial, spin_orbital

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


@pytest.fixture(scope='module', params=[(4, 16, 3), (5, 24, 4)])
def case(request):
    K, D, seed = request.param
    c = Case(K=K, D=D, seed=seed, sweeps=1, upload=True)
    c.L, c.R = c.product_envs()
    c.L[-1] = kq.boundary_store(2 * K)
    c.R[K] = kq.boundary_store(2 * K)
    return c


def test_envs_on_device_match_oracle(gpu, case):
    for i in range(case.K):
        for p in range(2 * case.K):
            a = case.L[i].op_tensor(p).toDenseTensor()
            assert np.abs(a - case.oL[i][p].toDenseTensor()).max() < 1e-12
            b = case.R[i].op_tensor(p).toDenseTensor()
            assert np.abs(b - case.oR[i][p].toDenseTensor()).max() < 1e-12


def test_sigma_and_diag_on_device_match_oracle(gpu, case):
    for isite in range(case.K - 1):
        qkey, oq, pt, _ = case.templates(isite)
        v = np.random.default_rng(isite).standard_normal(oq.size_allowed)
        want = case.oracle_sigma(isite, v)
        prog = kq.SigmaProgram(case.mpo, isite, pt, qkey, case.L[isite - 1], case.R[isite + 2],
                               case.mpo.hpwts)
        sig = dv.empty(pt.size_allowed, 0)
        prog.apply(case.mpo, case.L[isite - 1], case.R[isite + 2], dv.upload(v, 0), sig)
        got = dv.download(sig)
        assert np.abs(got - want).max() < 1e-12 * max(1.0, np.abs(want).max()), isite
        d = kq.DiagProgram(case.mpo, isite, pt, case.L[isite - 1], case.R[isite + 2],
                           case.mpo.hpwts)
        dg = dv.zeros(pt.size_allowed, 0)
        d.apply(case.L[isite - 1], case.R[isite + 2], dg)
        assert np.abs(dv.download(dg) - case.oracle_diag(isite)).max() < 1e-12


@pytest.mark.parametrize('overlap', [True, False])
def test_sigma_chunked_two_streams_matches_oracle(gpu, case, overlap, monkeypatch):
    """Three op chunks: L x psi of the next chunk on a side stream beside the
    chained launch of the current one (double-buffered top1), or all on one
    stream; sigma equals the oracle either way, repeated applications agree."""
    monkeypatch.setattr(kq, 'SIGMA_CHUNKS', 3)
    monkeypatch.setattr(kq, 'OVERLAP', overlap)
    for isite in (1, case.K - 3):
        qkey, oq, pt, _ = case.templates(isite)
        v = np.random.default_rng(isite + 7).standard_normal(oq.size_allowed)
        want = case.oracle_sigma(isite, v)
        prog = kq.SigmaProgram(case.mpo, isite, pt, qkey, case.L[isite - 1], case.R[isite + 2],
                               case.mpo.hpwts)
        assert len(prog.chunks) >= 2
        sig = dv.empty(pt.size_allowed, 0)
        vd = dv.upload(v, 0)
        outs = []
        for _ in range(3):
            prog.apply(case.mpo, case.L[isite - 1], case.R[isite + 2], vd, sig)
            outs.append(dv.download(sig).copy())
        assert np.abs(outs[0] - want).max() < 1e-12 * max(1.0, np.abs(want).max()), isite
        assert all(np.array_equal(outs[0], o) for o in outs[1:])


def test_chain_wave_balance_changes_no_bit(gpu, case, monkeypatch):
    """The chained launch's per-wave task order (qtensor.balance_waves) only
    moves tasks between the four waves of a workgroup: sigma is bit-identical
    with and without it, and equals the oracle."""
    from zmpo_dmrg_amd import qtensor as pq
    for isite in (1, case.K - 3):
        qkey, oq, pt, _ = case.templates(isite)
        v = np.random.default_rng(isite + 3).standard_normal(oq.size_allowed)
        outs = []
        for bal in (False, True):
            monkeypatch.setattr(pq, 'CHAIN_BALANCE', bal)
            prog = kq.SigmaProgram(case.mpo, isite, pt, qkey, case.L[isite - 1],
                                   case.R[isite + 2], case.mpo.hpwts)
            sig = dv.empty(pt.size_allowed, 0)
            prog.apply(case.mpo, case.L[isite - 1], case.R[isite + 2], dv.upload(v, 0), sig)
            outs.append(dv.download(sig))
        assert np.array_equal(outs[0], outs[1]), isite
        want = case.oracle_sigma(isite, v)
        assert np.abs(outs[1] - want).max() < 1e-12 * max(1.0, np.abs(want).max())


def test_sigma_fp32_within_tolerance(gpu, case):
    isite = 1
    qkey, oq, pt, _ = case.templates(isite)
    v = np.random.default_rng(9).standard_normal(oq.size_allowed)
    want = case.oracle_sigma(isite, v)
    prog = kq.SigmaProgram(case.mpo, isite, pt, qkey, case.L[isite - 1], case.R[isite + 2],
                           case.mpo.hpwts)
    sig = dv.empty(pt.size_allowed, 1)
    prog.apply(case.mpo, case.L[isite - 1], case.R[isite + 2], dv.upload(v, 1), sig, dtype=1)
    got = dv.download(sig).astype(np.float64)
    assert np.abs(got - want).max() < 1e-5 * max(1.0, np.abs(want).max())


def test_device_dmrg_equals_fci(gpu):
    from oracle import fci
    K = 4
    i1, i2 = gen_integrals(K, seed=1)
    h, eri = gen_spatial(K, seed=1)
    d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=100, tol=1e-10)
    e = pdmrg.run(d, 4)
    assert abs(e[-1] - fci.fci_energy(h, eri, K, 0.0)[0]) < 1e-8


def test_device_dmrg_h6_reference_energy(gpu):
    """H6/STO-3G from the reference's example: FCI -5.1491130015 (README:2)."""
    z = np.load(os.path.join(GOLDEN, 'h6_sto3g.npz'))
    i1, i2 = spin_orbital(z['h'], z['eri'])
    d = pdmrg.MPODMRG(i1, i2, 6, 0.0, Dmax=64, tol=1e-10, occun=np.array([1., 0., 0., 1.] * 3))
    e = pdmrg.run(d, 6)
    assert abs(e[-1] - (-5.1491130015)) < 2e-10


def test_device_dmrg_matches_oracle_sweep_by_sweep(gpu):
    """Truncated run (D below the full space): energies per half sweep and the
    bond QN structure after truncation identical to the oracle."""
    from oracle import sweep as osw
    K, D, N = 5, 12, 4
    i1, i2 = gen_integrals(K, seed=5)
    st = osw.State(i1, i2, N, 0.0, Dmax=D, tol=1e-10)
    d = pdmrg.MPODMRG(i1, i2, N, 0.0, Dmax=D, tol=1e-10)
    for _ in range(2):
        for status in ('R', 'L'):
            ro = osw.sweep(st, status)
            rp = pdmrg.sweep(d, status)
            eo = np.array([x[0] for x in ro[3]])
            ep = np.array([x[0] for x in rp[3]])
            assert np.abs(eo - ep).max() < 1e-9
            for i in range(K + 1):
                assert np.array_equal(np.asarray(st.qnuml[i]), np.asarray(d.qnuml[i]))
                assert np.array_equal(np.asarray(st.qnumr[i]), np.asarray(d.qnumr[i]))


@pytest.mark.parametrize('name', ['k6d20', 'cfg1'])
def test_device_dmrg_matches_oracle_fixture(gpu, name):
    """Full device sweeps vs the stored oracle run: every step's eigenvalue to
    1e-8 Ha and every bond's (N, Sz) list after truncation identical."""
    path = os.path.join(GOLDEN, 'dmrg_%s.npz' % name)
    if not os.path.exists(path):
        pytest.skip('fixture %s not generated' % name)
    z = np.load(path)
    K, D, seed = int(z['K']), int(z['D']), int(z['seed'])
    i1, i2 = gen_integrals(K, seed=seed)
    d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=D, tol=1e-8)
    eig = z['eig']
    k = 0
    for _ in range(eig.shape[0] // 2):
        for status in ('R', 'L'):
            r = pdmrg.sweep(d, status)
            ep = np.array([x[0] for x in r[3]])
            assert np.abs(ep - eig[k]).max() < 1e-8, (status, k, np.abs(ep - eig[k]).max())
            k += 1
    for i in range(K + 1):
        assert np.array_equal(np.asarray(d.qnuml[i]), z['qnuml%d' % i]), i
        assert np.array_equal(np.asarray(d.qnumr[i]), z['qnumr%d' % i]), i


@pytest.mark.parametrize('mode,lo', [('rq', 1), ('rq', 2), ('correction', 1), ('correction', 2),
                                     ('restart', 1)])
def test_fp32_sigma_with_fp64_correction_matches_fp64(gpu, mode, lo):
    """Precision form: low-precision sigma (FP32 or BF16 operands) with FP64
    correction -- 'correction': outer FP64 sigmas around anchored inner
    low-precision Davidson solves; 'restart': a low-precision Davidson
    restarted from its solution with FP64 sigma.  K=6, N=6 at D=64 keeps
    every bond untruncated, so the converged energy is the FCI energy: the
    precision run must match the FP64 run and FCI to 1e-6 Ha (the target;
    the reported value is an FP64 Rayleigh quotient), with both the
    low-precision and the FP64 sigma taking part.  (Per-step comparisons
    against a truncated FP64 run are not meaningful: RDM weight ties at the
    1e-9 level can pick a different truncated space.)"""
    from oracle import fci
    from zmpo_dmrg_amd import dot as pdot
    K = 6
    i1, i2 = gen_integrals(K, seed=2)
    h, eri = gen_spatial(K, seed=2)
    efci = fci.fci_energy(h, eri, K, 0.0)[0]
    out = {}
    saved = pdot.PRECISION_MODE
    pdot.PRECISION_MODE = mode
    try:
        for dtype in (0, lo):
            d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=64, tol=1e-10, dtype=dtype)
            out[dtype] = (pdmrg.run(d, 3)[-1], d)
    finally:
        pdot.PRECISION_MODE = saved
    e64, e32 = out[0][0], out[lo][0]
    print('[prec] %s dtype %d: |E - E64| = %.2e, sigma low %d / FP64 %d' % (
        mode, lo, abs(e32 - e64), out[lo][1].nsigma, getattr(out[lo][1], 'nsigma64', 0)),
        flush=True)
    assert abs(e64 - efci) < 1e-8, (e64, efci)
    assert abs(e32 - e64) < 1e-6, (e32, e64)
    d = out[lo][1]
    n64 = getattr(d, 'nsigma64', 0)
    assert d.nsigma > 0 and n64 > 0, (d.nsigma, n64)   # both phases ran
