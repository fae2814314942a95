"""Decimation on the device (decim.py) against the host path (dot.decimation
/ genGuess, which restates dotutil.py:288-559 and qparser.py:216-293 in
numpy): the same two-site vectors give the same kept quantum numbers and
weights, and the same site tensor, srotR and next-site guess up to the sign
of each kept eigenvector; whole sweeps give the same energies and bond
quantum numbers."""
import numpy as np
import pytest

from zmpo_dmrg_amd import device as dv
from zmpo_dmrg_amd import dmrg as pdmrg
from zmpo_dmrg_amd import dot as pdot
from zmpo_dmrg_amd.decim import DeviceDecimation
from zmpo_dmrg_amd.integrals import gen_integrals

pytestmark = pytest.mark.gpu


def _state(K, D, seed):
    i1, i2 = gen_integrals(K, seed=seed)
    d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=D, tol=1e-9)
    saved = pdot.DECIM
    pdot.DECIM = 'host'
    try:
        pdmrg.run(d, 1)
    finally:
        pdot.DECIM = saved
    return d


def _prepare(d, isite, status):
    """The structures ci_solver sets up for a two-site step at isite."""
    ldim = len(d.qnuml[isite])
    rdim = len(d.qnumr[isite + 2])
    d.dims = (ldim, 16, rdim, ldim * 16 * rdim)
    pdot.setupQlst(d, isite)
    key = list(d.qsectors.keys())[0]
    pdot.setupQtmpCI(d, key)
    return d.qtmp.prjmap(d.idlstCI), d.qtmp.size_allowed


@pytest.mark.parametrize('status', ['L', 'R'])
@pytest.mark.parametrize('neig', [1, 2])
def test_device_decimation_matches_host(gpu, status, neig):
    K, D = 6, 12
    d = _state(K, D, 3)
    isite = 2
    prjmap, n = _prepare(d, isite, status)
    rng = np.random.default_rng(7 + neig)
    vec = rng.standard_normal((neig, n))
    vec /= np.linalg.norm(vec, axis=1, keepdims=True)
    # host path
    civecs = np.zeros((neig, d.dims[3]))
    civecs[:, prjmap] = vec
    sigs, dwts, qred, site, srotR = pdot.decimation(d, civecs, isite, status)
    actlst = list(range(K - 1))
    pdot.genGuess(d, isite, actlst, civecs, srotR, status)
    psi_host = d.psi0
    # device path
    dd = DeviceDecimation(d, dv.upload(vec, 0).view(neig, n), prjmap, isite, status)
    if status == 'L':
        classes = pdot.dpt(d.qnuml[isite], d.qphys[isite])
    else:
        classes = pdot.dpt(d.qphys[isite + 1], d.qnumr[isite + 2])
    assert dd.run(classes, d.thresh, d.Dmax)
    q, srot = dd.site_and_srotR(d.qlst)
    nxt = d.lmps[isite - 1] if status == 'R' else d.rmps[isite + 2]
    psi_dev = dv.download(dd.guess(nxt))
    assert np.array_equal(np.asarray(dd.qred), np.asarray(qred))
    assert np.abs(dd.sigs - sigs).max() < 1e-12
    assert abs(dd.dwts - dwts) < 1e-12
    # site: kept eigenvector k is the last leg ('L') or the first ('R')
    sd = q.toDenseTensor()
    nres = len(qred)
    if status == 'L':
        hc, dc = site.reshape(-1, nres), sd.reshape(-1, nres)
    else:
        hc, dc = site.reshape(nres, -1).T, sd.reshape(nres, -1).T
    sign = np.sign(np.sum(hc * dc, axis=0))
    assert np.all(sign != 0)
    assert np.abs(hc - dc * sign).max() < 1e-10
    # srotR and the guess carry the kept index on the same leg as the site
    sr = dv.download(srot).reshape(srotR.shape)
    if status == 'L':
        assert np.abs(srotR - sr * sign[None, :, None, None]).max() < 1e-10
        ph = psi_host.reshape(neig, nres, -1)
        pd_ = psi_dev.reshape(neig, nres, -1)
        assert np.abs(ph - pd_ * sign[None, :, None]).max() < 1e-10
    else:
        assert np.abs(srotR - sr * sign[None, None, None, :]).max() < 1e-10
        ph = psi_host.reshape(neig, -1, nres)
        pd_ = psi_dev.reshape(neig, -1, nres)
        assert np.abs(ph - pd_ * sign[None, None, :]).max() < 1e-10


@pytest.mark.parametrize('noise', [0.0, 1e-4])
def test_device_decimation_sweeps_match_host(gpu, noise):
    """With the noise term too (the pRDMQt operator term added to the
    device density matrix)."""
    K, D = 6, 12
    out = {}
    for mode in ('host', 'device'):
        i1, i2 = gen_integrals(K, seed=5)
        d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=D, tol=1e-10)
        d.noise = noise
        saved = pdot.DECIM
        pdot.DECIM = mode
        try:
            e = pdmrg.run(d, 3)
        finally:
            pdot.DECIM = saved
        out[mode] = (e, [np.asarray(d.qnuml[i]) for i in range(K + 1)])
    eh, ed = out['host'][0], out['device'][0]
    assert len(eh) == len(ed)
    assert np.abs(np.array(eh) - np.array(ed)).max() < 1e-9, (eh, ed)
    for a, b in zip(out['host'][1], out['device'][1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize('status', ['L', 'R'])
def test_device_noise_term_matches_host(gpu, status):
    """prdm_noise_device (C template gathered from the dense vectors, Gram
    blocks scattered into the dense matrix on the device) against the host
    restatement prdm_noise on the same two-site vectors."""
    from zmpo_dmrg_amd import kernelQt as kq
    K, D = 6, 12
    i1, i2 = gen_integrals(K, seed=3)
    d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=D, tol=1e-9, keep_envs=True)
    d.noise = 1e-4
    pdmrg.run(d, 1)
    isite = 2
    prjmap, n = _prepare(d, isite, status)
    neig = 2
    key = list(d.qsectors.keys())[0]
    d.qsectors = {key: neig}          # both paths sum over the sector's states
    rng = np.random.default_rng(11)
    vec = rng.standard_normal((neig, n))
    civecs = np.zeros((neig, d.dims[3]))
    civecs[:, prjmap] = vec
    ldim, _, rdim, _ = d.dims
    flst = [[d.envL[isite - 1], d.envR[isite + 2]]]
    info = [d, isite, 2, flst, status, True]
    shp = (neig, ldim, 4, 4 * rdim) if status == 'L' else (neig, ldim * 4, 4, rdim)
    want = 0.5 * kq.prdm_noise(civecs.reshape(shp), info)
    got = dv.download(kq.prdm_noise_device(dv.upload(civecs.reshape(-1), 0), neig, info, 0.5))
    assert got.size == want.size
    assert np.abs(got.reshape(want.shape) - want).max() <= 1e-12 * max(1.0, np.abs(want).max())
