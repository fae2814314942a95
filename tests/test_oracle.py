"""Pins the CPU oracle: brute-force Fock space, exact FCI, and the reference's
recorded H6/STO-3G FCI energies (samps/example_conversion/0.h6_nqt/README:2-5)."""
import os

import numpy as np
import pytest

from oracle import davidson, fci, mpo, qtensor, sweep
from zmpo_dmrg_amd.integrals import gen_integrals, gen_spatial, spin_orbital

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


@pytest.mark.parametrize('K', [2, 3])
def test_mpo_equals_fock_hamiltonian(K):
    """Sum over ops of the contracted MPO == second-quantised H from the spatial integrals."""
    h, eri = gen_spatial(K, seed=11)
    H, _ = fci.fock_hamiltonian(h, eri)
    Hm = fci.mpo_dense(mpo.build_mpo(*spin_orbital(h, eri)), K, 2 * K)
    assert np.abs(H.toarray() - Hm).max() < 1e-12


def test_mpo_planted_sign_error_is_caught():
    """A flipped sign in one W entry must break the Fock-space identity."""
    K = 3
    h, eri = gen_spatial(K, seed=11)
    H, _ = fci.fock_hamiltonian(h, eri)
    m = mpo.build_mpo(*spin_orbital(h, eri))
    m[1][2].value[3] *= -1.0
    assert np.abs(H.toarray() - fci.mpo_dense(m, K, 2 * K)).max() > 1e-6


def test_elem_reference_matrices():
    """Hand-written a_0^+/a_0 matrices of qtensor_opers.py:535-562."""
    cre = mpo.link_two_opers(mpo.elem_spin_mat(0, 0).reshape(1, 1, 2, 2),
                             mpo.elem_spin_mat(0, 1).reshape(1, 1, 2, 2))
    ref = np.zeros((1, 1, 4, 4))
    ref[0, 0, 2, 0] = ref[0, 0, 3, 1] = 1.0
    assert np.abs(cre - ref).max() == 0.0


def test_h6_fixture_fci_matches_reference_readme():
    d = np.load(os.path.join(GOLDEN, 'h6_sto3g.npz'))
    e_sz0 = fci.fci_energy(d['h'], d['eri'], 6, 0.0, nroots=4)
    assert np.abs(e_sz0 - d['efci_readme']).max() < 2e-10


def test_fci_brute_force_small():
    """Two electrons in two orbitals: closed form of the 2x2 singlet/ionic problem."""
    h = np.array([[-1.0, -0.2], [-0.2, -0.5]])
    eri = np.zeros((2, 2, 2, 2))
    eri[0, 0, 0, 0] = 0.7
    eri[1, 1, 1, 1] = 0.6
    e = fci.fci_energy(h, eri, 2, 0.0, nroots=1)[0]
    # singlet basis |aa>, |bb>, (|ab>+|ba>)/sqrt2 ; triplet decouples
    Hs = np.array([[2 * -1.0 + 0.7, 0.0, np.sqrt(2) * -0.2],
                   [0.0, 2 * -0.5 + 0.6, np.sqrt(2) * -0.2],
                   [np.sqrt(2) * -0.2, np.sqrt(2) * -0.2, -1.5]])
    assert abs(e - np.linalg.eigvalsh(Hs)[0]) < 1e-12


def test_tensordot_equals_dense():
    rng = np.random.default_rng(0)
    q = [np.array([[0., 0.], [1., .5], [1., -.5], [2., 0.]])[rng.integers(0, 4, 6)] for _ in range(3)]
    a = qtensor.QTensor([False, False, True])
    a.fromDenseTensor(rng.standard_normal((6, 6, 6)), [q[0], q[1], q[0] + q[1] * 0 + q[1]])
    b = qtensor.QTensor([False, True, False])
    b.fromDenseTensor(rng.standard_normal((6, 6, 6)), [a.qsyms[2][np.repeat(np.arange(len(a.ndims[2])), a.ndims[2])] if False else (q[0] + q[1]), q[2], q[2]])
    da, db = a.toDenseTensor(), b.toDenseTensor()
    c = qtensor.tensordot(a, b, ([2], [0]))
    assert np.abs(c.toDenseTensor([a.idlst[0], a.idlst[1], b.idlst[1], b.idlst[2]])
                  - np.tensordot(da, db, ([2], [0]))).max() < 1e-12


def test_davidson_vs_eigh():
    """pdvdson.py:393-429 pattern: random symmetric matrix."""
    rng = np.random.default_rng(0)
    A = rng.uniform(-1, 1, (100, 100))
    A = A + A.T + np.diag(np.arange(100.0))
    e, v, nmvp, _ = davidson.solve_iter(lambda x: A @ x, np.diag(A).copy(), neig=2,
                                        crit_e=1e-12, crit_vec=1e-9, maxcycle=60)
    assert np.abs(e - np.linalg.eigvalsh(A)[:2]).max() < 1e-8


def test_dmrg_equals_fci_k4():
    K = 4
    i1, i2 = gen_integrals(K, seed=1)
    h, eri = gen_spatial(K, seed=1)
    st = sweep.State(i1, i2, K, 0.0, Dmax=100, tol=1e-10)
    e = sweep.run_dmrg(st, 4)
    assert abs(e[-1] - fci.fci_energy(h, eri, K, 0.0)[0]) < 1e-8


def test_checkmps_energy_equals_davidson():
    """<psi|H|psi> from a full environment sweep (class.py:470-558) == last eigenvalue."""
    K = 4
    i1, i2 = gen_integrals(K, seed=2)
    st = sweep.State(i1, i2, K, 0.0, Dmax=100, tol=1e-11)
    e = sweep.run_dmrg(st, 3)
    from oracle.kernels import genHops
    _, exphop = genHops(st, st.lmps, 'L')
    assert abs(exphop.sum() - e[-1]) < 1e-8
