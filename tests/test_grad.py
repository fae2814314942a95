"""dE/dh (1-RDM) and dE/d(pq|rs) (2-RDM / 2) through the hand-written backward.

Pinned against the exact FCI vector of the same integrals (oracle.fci): its
RDMs, and central finite differences of the FCI energy (Hellmann-Feynman),
plus size-independent identities of any N-electron state at truncated sizes.
"""
import numpy as np
import pytest

from zmpo_dmrg_amd import dmrg as pdmrg
from zmpo_dmrg_amd import grad
from zmpo_dmrg_amd import mpo as pmpo
from zmpo_dmrg_amd.integrals import gen_spatial, spin_orbital
from oracle.fci import fci_energy, fci_rdms


def _orbit_size(K):
    idx = np.indices((K,) * 4)
    keys = np.stack([np.ravel_multi_index(tuple(idx[list(pm)]), (K,) * 4) for pm in grad._PERM8])
    keys = np.sort(keys, axis=0)
    return 1 + (np.diff(keys, axis=0) != 0).sum(axis=0)


def _check_identities(E, gamma, Gamma, h, eri, nelec, tol):
    """Euler (E is linear in the integrals), trace and partial trace of the RDMs."""
    assert abs(np.sum(h * gamma) + 0.5 * np.sum(eri * Gamma) - E) < tol * max(1.0, abs(E))
    assert abs(np.trace(gamma) - nelec) < tol
    assert np.abs(gamma - gamma.T).max() < tol
    pt = np.einsum('prqq->pr', Gamma)
    assert np.abs(pt - (nelec - 1) * gamma).max() < tol * nelec


# ------------------------------------------------------------------ CPU
def test_torch_mpo_builder_matches_numpy():
    torch = pytest.importorskip('torch')
    K = 3
    h, eri = gen_spatial(K, seed=5)
    i1, i2 = spin_orbital(h, eri)
    t1, t2 = grad.spin_orbital_t(torch.from_numpy(h), torch.from_numpy(eri))
    assert np.array_equal(t1.numpy(), i1) and np.array_equal(t2.numpy(), i2)
    n = 2 * K
    for isite in range(K):
        w = grad.site_dense_t(list(range(n)), n, isite, t1, t2).numpy()
        for p in range(n):
            ref = pmpo.link(pmpo.hfac_spin(p, n, 2 * isite, i1[p], i2[p]),
                            pmpo.hfac_spin(p, n, 2 * isite + 1, i1[p], i2[p]))
            assert np.array_equal(w[p], ref), (isite, p)
            _, _, il, ir = pmpo.site_order(p, n, isite)
            st = pmpo.site_tensor(p, n, isite, i1, i2).toDenseTensor()
            assert np.array_equal(ref[np.ix_(il, ir)], st)


def test_fci_rdm_oracle_pins():
    """The oracle's RDMs: energy from RDMs, traces, and central differences of
    the FCI energy along h_pr and along 8-fold symmetric (pr|qs) perturbations."""
    K, N = 4, 4
    h, eri = gen_spatial(K, seed=3)
    E, g, G = fci_rdms(h, eri, N)
    assert abs(E - fci_energy(h, eri, N)[0]) < 1e-12
    _check_identities(E, g, grad.sym8(G), h, eri, N, 1e-10)
    rng = np.random.default_rng(0)
    step = 1e-5
    osz = _orbit_size(K)
    for _ in range(4):
        p, r = rng.integers(K, size=2)
        d = np.zeros((K, K))
        d[p, r] += step
        d[r, p] += step if p != r else 0.0
        fd = (fci_energy(h + d, eri, N)[0] - fci_energy(h - d, eri, N)[0]) / (2 * step)
        assert abs(fd - (g[p, r] + (g[r, p] if p != r else 0.0))) < 1e-6
        t = tuple(rng.integers(K, size=4))
        e = np.zeros((K,) * 4)
        idx = np.array(t)
        for pm in grad._PERM8:
            e[tuple(idx[list(pm)])] = step
        fd = (fci_energy(h, eri + e, N)[0] - fci_energy(h, eri - e, N)[0]) / (2 * step)
        assert abs(fd - 0.5 * osz[t] * grad.sym8(G)[t]) < 1e-6


def test_rdm_gradients_host_model_match_fci():
    """The backward pass (plans, three-launch dE/dW program, torch MPO VJP) on
    the CPU kernel model at K=4 with an exact MPS."""
    from kmodel import host_model
    K = 4
    h, eri = gen_spatial(K, seed=3)
    i1, i2 = spin_orbital(h, eri)
    with host_model():
        d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=64, tol=1e-10)
        pdmrg.run(d, 3)
        E, gamma, Gamma = grad.rdms(d.fhop, d.lmps, np.array([float(K), 0.0]), h, eri)
        site_e = list(grad.w_gradients.last_energies)
    Ef, gf, Gf = fci_rdms(h, eri, K)
    assert abs(E - Ef) < 1e-9
    assert np.ptp(site_e) < 1e-11
    assert np.abs(gamma - gf).max() < 1e-8
    assert np.abs(Gamma - grad.sym8(Gf)).max() < 1e-8
    _check_identities(E, gamma, Gamma, h, eri, K, 1e-9)


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize('K,D,seed', [(4, 16, 3), (6, 64, 1)])
def test_rdm_gradients_on_device_match_fci(gpu, K, D, seed):
    h, eri = gen_spatial(K, seed=seed)
    i1, i2 = spin_orbital(h, eri)
    d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=D, tol=1e-13, crit_vec=1e-10)
    pdmrg.run(d, 4)
    E, gamma, Gamma = grad.rdms(d.fhop, d.lmps, np.array([float(K), 0.0]), h, eri)
    print('[grad] K=%d D=%d E=%.12f site-energy spread %.2e' % (
        K, D, E, np.ptp(grad.w_gradients.last_energies)), flush=True)
    Ef, gf, Gf = fci_rdms(h, eri, K)
    assert abs(E - Ef) < 1e-8
    assert np.ptp(grad.w_gradients.last_energies) < 1e-10 * K
    assert np.abs(gamma - gf).max() < 1e-6
    assert np.abs(Gamma - grad.sym8(Gf)).max() < 1e-6
    _check_identities(E, gamma, Gamma, h, eri, K, 1e-9)


@pytest.mark.gpu
def test_energy_function_autograd_matches_finite_differences(gpu):
    """torch.autograd through EnergyFunction vs central differences (step 1e-5)
    of the FCI energy on 20 random integral entries (symmetric perturbations)."""
    import torch
    K = 4
    h, eri = gen_spatial(K, seed=7)
    i1, i2 = spin_orbital(h, eri)
    d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=16, tol=1e-13, crit_vec=1e-10)
    pdmrg.run(d, 4)
    ht = torch.from_numpy(h.copy()).requires_grad_(True)
    et = torch.from_numpy(eri.copy()).requires_grad_(True)
    E = grad.EnergyFunction.apply(ht, et, d.fhop, d.lmps, np.array([float(K), 0.0]))
    (2.0 * E).backward()
    gh, ge = ht.grad.numpy() / 2.0, et.grad.numpy() / 2.0
    assert abs(float(E) - fci_energy(h, eri, K)[0]) < 1e-8
    rng = np.random.default_rng(11)
    step = 1e-5
    osz = _orbit_size(K)
    for k in range(20):
        if k % 2 == 0:
            p, r = rng.integers(K, size=2)
            dh = np.zeros((K, K))
            dh[p, r] += step
            if p != r:
                dh[r, p] += step
            fd = (fci_energy(h + dh, eri, K)[0] - fci_energy(h - dh, eri, K)[0]) / (2 * step)
            an = gh[p, r] + (gh[r, p] if p != r else 0.0)
        else:
            t = np.array(rng.integers(K, size=4))
            de = np.zeros((K,) * 4)
            for pm in grad._PERM8:
                de[tuple(t[list(pm)])] = step
            fd = (fci_energy(h, eri + de, K)[0] - fci_energy(h, eri - de, K)[0]) / (2 * step)
            an = osz[tuple(t)] * grad.sym8(ge)[tuple(t)]
        assert abs(fd - an) < 1e-6, (k, fd, an)


@pytest.mark.gpu
def test_rdm_identities_truncated_k10(gpu):
    """Truncated MPS (K=10, D=48): every site gives the same energy, and the
    RDMs satisfy Euler's identity, the trace and the partial-trace rule."""
    K = 10
    h, eri = gen_spatial(K, seed=0)
    i1, i2 = spin_orbital(h, eri)
    d = pdmrg.MPODMRG(i1, i2, K, 0.0, Dmax=48, tol=1e-8)
    pdmrg.run(d, 2)
    E, gamma, Gamma = grad.rdms(d.fhop, d.lmps, np.array([float(K), 0.0]), h, eri)
    print('[grad] K=10 D=48 E=%.10f' % E, flush=True)
    assert abs(E - grad.expectation(d.fhop, d.lmps)) < 1e-10 * abs(E)
    assert np.ptp(grad.w_gradients.last_energies) < 1e-10 * abs(E)
    _check_identities(E, gamma, Gamma, h, eri, K, 1e-8)
