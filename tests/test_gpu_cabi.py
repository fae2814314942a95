"""The named C-ABI entry points of the path (zmpo_sigma_two_site,
zmpo_renorm, zmpo_hdiag over recorded sequences; zmpo_davidson_residual)
against the per-launch calls they replace, bit for bit."""
import numpy as np
import pytest

import zmpo_dmrg_amd._lib as L
from parity_util import Case
from zmpo_dmrg_amd import device as dv
from zmpo_dmrg_amd import kernelQt as kq

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def case6():
    c = Case(K=6, D=32, seed=2, sweeps=1, upload=True)
    c.L, c.R = c.product_envs()
    return c


def test_sigma_two_site_and_hdiag_equal_per_launch(gpu, case6):
    c = case6
    isite = 2
    qkey, _, pt, _ = c.templates(isite)
    Le, Re = c.L[isite - 1], c.R[isite + 2]
    n = pt.size_allowed
    v = dv.upload(np.random.default_rng(1).standard_normal(n), 0)
    sp = kq.SigmaProgram(c.mpo, isite, pt, qkey, Le, Re, c.mpo.hpwts)
    s_seq = dv.empty(n, 0)
    sp.apply(c.mpo, Le, Re, v, s_seq, 0)          # zmpo_sigma_two_site
    saved = kq.SIGMA_SEQ
    kq.SIGMA_SEQ = False
    try:
        s_loop = dv.empty(n, 0)
        sp.apply(c.mpo, Le, Re, v, s_loop, 0)
        dp = kq.DiagProgram(c.mpo, isite, pt, Le, Re, c.mpo.hpwts)
        d_loop = dp.apply(Le, Re, dv.empty(n, 0))
    finally:
        kq.SIGMA_SEQ = saved
    assert np.array_equal(dv.download(s_seq), dv.download(s_loop))
    d_seq = kq.DiagProgram(c.mpo, isite, pt, Le, Re, c.mpo.hpwts).apply(Le, Re, dv.empty(n, 0))
    assert np.array_equal(dv.download(d_seq), dv.download(d_loop))


def test_renorm_entry_equals_per_launch(gpu, case6):
    from parity_util import to_product
    c = case6
    site = to_product(c.st.lmps[1])
    out_seq, _ = kq.renorm(c.mpo, 1, site, c.L[0], 'L')
    saved = kq.SIGMA_SEQ
    kq.SIGMA_SEQ = False
    try:
        out_loop, _ = kq.renorm(c.mpo, 1, site, c.L[0], 'L')
    finally:
        kq.SIGMA_SEQ = saved
    assert np.array_equal(dv.download(out_seq.buf), dv.download(out_loop.buf))


def test_davidson_residual(gpu):
    rng = np.random.default_rng(3)
    n, nb = 100003, 7
    W = rng.standard_normal((nb + 2, n))
    V = rng.standard_normal((nb + 2, n))
    c = rng.standard_normal(nb)
    diag = rng.standard_normal(n)
    theta, shift = -1.25, 1e-4
    Wd, Vd, cd, dd = (dv.upload(x, 0) for x in (W, V, c, diag))
    r = dv.empty(n, 0)
    lib = L.lib()
    L.check(lib.zmpo_davidson_residual(n, dv.ptr(Wd), dv.ptr(Vd), n, nb, dv.ptr(cd), theta,
                                       dv.ptr(dd), shift, dv.ptr(r), dv.stream()))
    # the same launches one by one
    r2 = dv.empty(n, 0)
    L.check(lib.zmpo_lincomb(0, dv.ptr(r2), dv.ptr(Wd), n, nb, dv.ptr(cd), 1.0, 0.0, n,
                             dv.stream()))
    L.check(lib.zmpo_lincomb(0, dv.ptr(r2), dv.ptr(Vd), n, nb, dv.ptr(cd), -theta, 1.0, n,
                             dv.stream()))
    L.check(lib.zmpo_precond(0, dv.ptr(r2), dv.ptr(dd), theta, shift, n, dv.stream()))
    assert np.array_equal(dv.download(r), dv.download(r2))
    want = (c @ W[:nb] - theta * (c @ V[:nb])) / (np.abs(diag - theta) + shift)
    assert np.abs(dv.download(r) - want).max() < 1e-10 * max(1.0, np.abs(want).max())
