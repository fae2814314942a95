"""Online tile tuning on the device: every candidate of both outer sigma
launches is timed, the fastest is kept per dtype, and the sigma computed
with the tuned tiles equals the one with the static tiles."""
import numpy as np
import pytest

from parity_util import Case
from zmpo_dmrg_amd import autotune as at
from zmpo_dmrg_amd import device as dv
from zmpo_dmrg_amd import kernelQt as kq

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('dtype', [0, 2])
def test_online_tuning_keeps_sigma(gpu, dtype, monkeypatch):
    c = Case(K=6, D=32, seed=2, sweeps=1, upload=True)
    L_, R_ = c.product_envs()
    isite = 2
    qkey, _, pt, _ = c.templates(isite)
    n = pt.size_allowed
    v = np.random.default_rng(4).standard_normal(n)
    monkeypatch.setattr(at, 'MODE', 'off')
    base = kq.SigmaProgram(c.mpo, isite, pt, qkey, L_[isite - 1], R_[isite + 2], c.mpo.hpwts)
    s0 = dv.empty(n, kq.sigma_dtype(dtype))
    base.apply(c.mpo, L_[isite - 1], R_[isite + 2], dv.upload(v, dtype), s0, dtype)
    want = dv.download(s0).astype(np.float64)
    monkeypatch.setattr(at, 'MODE', 'online')
    monkeypatch.setattr(at, '_table', [])
    monkeypatch.setattr(at, '_online', {})
    sp = kq.SigmaProgram(c.mpo, isite, pt, qkey, L_[isite - 1], R_[isite + 2], c.mpo.hpwts)
    s1 = dv.empty(n, kq.sigma_dtype(dtype))
    for _ in range(2):      # the first application tunes, the second reuses
        sp.apply(c.mpo, L_[isite - 1], R_[isite + 2], dv.upload(v, dtype), s1, dtype)
        got = dv.download(s1).astype(np.float64)
        tol = {0: 1e-12, 2: 2e-2}[dtype] * max(1.0, np.abs(want).max())
        assert np.abs(got - want).max() < tol
    st = sp.stages[0]
    assert dtype in st[0].tiles and st[0].tiles[dtype] in at.CANDIDATES['sigma_l']
    assert kq.out_dtype(dtype) in st[3].tiles
    assert len(at._online) >= 2
