"""Noise term of the decimation RDM (pRDMQt, kernelQt.py:453-626) vs the same
sum written densely: for 'L', sum_p X_p X_p^T with X_p[l, u; wr, m] =
sum L_p[w, l, l'] W_p[i][w, wr, u, d] C[l', d, m]; for 'R', with
Y_p[u, r; wl, m] = sum W_p[i+1][wl, wr, u, d] R_p[wr, r, r'] C[m, d, r']."""
import numpy as np
import pytest

from zmpo_dmrg_amd import dmrg as pdmrg
from zmpo_dmrg_amd import kernelQt as kq
from zmpo_dmrg_amd.dot import setupQlst, setupQtmpCI
from zmpo_dmrg_amd.integrals import gen_spatial, spin_orbital


def _state(K=4, D=16, seed=3):
    h, eri = gen_spatial(K, seed=seed)
    i1, i2 = spin_orbital(h, eri)
    d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=D, tol=1e-10, keep_envs=True)
    pdmrg.run(d, 1)
    pdmrg.sweep(d, 'R')     # every bond has both environments
    return d


def _dense_noise(d, isite, C, status):
    K = d.nsite
    out = 0.0
    for p in range(d.nops):
        if status == 'L':
            L = d.envL[isite - 1].op_tensor(p).toDenseTensor()
            W = d.fhop.sites[isite][p].toDenseTensor()
            X = np.einsum('wab,wxud,bdm->auxm', L, W, C)
            X = X.reshape(X.shape[0] * X.shape[1], -1)
        else:
            R = d.envR[isite + 2].op_tensor(p).toDenseTensor()
            W = d.fhop.sites[isite + 1][p].toDenseTensor()
            X = np.einsum('wxud,xab,mdb->uawm', W, R, C)
            X = X.reshape(X.shape[0] * X.shape[1], -1)
        out = out + X @ X.T
    return out


def _check(d, isite, status):
    flst = [[d.envL[isite - 1], d.envR[isite + 2]]]
    ldim = int(flst[0][0].meta[0].shape[2])
    rdim = int(flst[0][1].meta[0].shape[2])
    d.dims = (ldim, 16, rdim, ldim * 16 * rdim)
    setupQlst(d, isite)
    info = [d, isite, 2, flst, status, True]
    # a random two-site state with the (N, Sz) block structure of the solver
    setupQtmpCI(d, list(d.qsectors)[0])
    q = d.qtmp.template()
    q.value = np.random.default_rng(isite).standard_normal(q.size_allowed)
    C = q.toDenseTensor(d.idlstCI)
    shp = (1, ldim, 4, 4 * rdim) if status == 'L' else (1, ldim * 4, 4, rdim)
    got = kq.prdm_noise(C.reshape(shp), info)
    want = _dense_noise(d, isite, C.reshape(shp[1:]), status)
    assert got.shape == want.shape
    assert np.abs(got - want).max() < 1e-10 * max(1.0, np.abs(want).max()), (isite, status)
    d.noise = 1e-3
    full = kq.pRDMQt(C.reshape(shp), info)
    m = C.reshape(ldim * 4, 4 * rdim)
    plain = m @ m.T if status == 'L' else m.T @ m
    assert np.abs(full - plain - 1e-3 * want).max() < 1e-10 * max(1.0, np.abs(plain).max())
    d.noise = 0.0


def test_prdm_noise_host_model_matches_dense():
    from kmodel import host_model
    with host_model():
        d = _state()
        for isite in (1, 2):
            for status in ('L', 'R'):
                _check(d, isite, status)


@pytest.mark.gpu
def test_prdm_noise_on_device_matches_dense(gpu):
    d = _state()
    for isite in (1, 2):
        for status in ('L', 'R'):
            _check(d, isite, status)


@pytest.mark.gpu
def test_sweeps_with_noise_reach_fci(gpu):
    """With noise in the decimation RDM the sweeps still converge to the
    exact energy (K=4, D=16: no truncation)."""
    from oracle.fci import fci_energy
    K = 4
    h, eri = gen_spatial(K, seed=3)
    i1, i2 = spin_orbital(h, eri)
    d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=16, tol=1e-12, crit_vec=1e-9)
    d.noise = 1e-4
    pdmrg.run(d, 2)
    d.noise = 0.0
    es = pdmrg.run(d, 2)
    assert abs(es[-1] - fci_energy(h, eri, K)[0]) < 1e-8


def _device_maps_check(d, isite, status):
    """The pipeline of kernelQt.prdm_noise_device, with its gather and
    scatter-add written in numpy: C template values gathered from the
    dense vector through qvec.prjmap, the PRDMProgram launches, each case's
    Gram blocks added at t3.prjmap positions -- against the dense sum."""
    flst = [[d.envL[isite - 1], d.envR[isite + 2]]]
    ldim = int(flst[0][0].meta[0].shape[2])
    rdim = int(flst[0][1].meta[0].shape[2])
    d.dims = (ldim, 16, rdim, ldim * 16 * rdim)
    setupQlst(d, isite)
    setupQtmpCI(d, list(d.qsectors)[0])
    q = d.qtmp.template()
    q.value = np.random.default_rng(isite + 7).standard_normal(q.size_allowed)
    C = q.toDenseTensor(d.idlstCI)
    key = list(d.qsectors)[0]
    qvec, idlst = kq.prdm_template(d, key, status)
    cmap = qvec.prjmap(idlst)
    ref = qvec.template()
    ref.fromDense(C.reshape([int(x) for x in qvec.shape]), idlst)
    assert np.array_equal(C.reshape(-1)[cmap], ref.value)
    env = flst[0][0] if status == 'L' else flst[0][1]
    prog = kq.PRDMProgram(d.fhop, isite, qvec, env, status, np.array(key), d.budget)
    from zmpo_dmrg_amd import device as dv
    vals = dv.download(prog.apply(d.fhop, env, dv.upload(C.reshape(-1)[cmap], 0)))
    keep = (0, 1) if status == 'L' else (1, 2)
    kl = [idlst[k] for k in keep]
    m = int(qvec.shape[keep[0]]) * int(qvec.shape[keep[1]])
    out = np.zeros(m * m)
    for c in prog.cases:
        t3 = prog.plans[c][5].template()
        if t3.size_allowed:
            gm = t3.prjmap([kl[0], kl[1], kl[0], kl[1]])
            assert len(np.unique(gm)) == len(gm)        # no repeated position in a case
            o = prog.out_off[c]
            out[gm] += vals[o:o + t3.size_allowed]
    want = _dense_noise(d, isite, C.reshape(qvec.shape), status)
    assert np.abs(out.reshape(m, m) - want).max() < 1e-10 * max(1.0, np.abs(want).max())


def test_device_noise_maps_host_model():
    from kmodel import host_model
    with host_model():
        d = _state()
        for isite in (1, 2):
            for status in ('L', 'R'):
                _device_maps_check(d, isite, status)
