"""The properties path: brute-force <E_pq> (eval_rdm1BF,
properties/mpo_dmrg_propsItrf.py:204-263) from products of elementary
operators, against the exact FCI 1-RDM (K = 4, the MPS is exact there) and
against the 1-RDM of the hand-written energy gradient dE/dh."""
import numpy as np
import pytest

from kmodel import host_model
from oracle.fci import fci_rdms
from zmpo_dmrg_amd import dmrg as pd
from zmpo_dmrg_amd import props
from zmpo_dmrg_amd.integrals import gen_spatial, spin_orbital


def test_elementary_products_match_reference_matrices():
    """a+_0 on spatial site 0 is |a><0| + |ab><b| in the (0, b, a, ab) basis
    (qtensor_opers.py:535-562); the number operator a+_0 a_0 is diag(0,0,1,1);
    a sign string acts left of the operator (site 0 for a+_2)."""
    assert np.array_equal(props.elem_product_site([(0, 1)], 0),
                          np.array([[0, 0, 0, 0], [0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0.]]))
    assert np.array_equal(props.elem_product_site([(0, 1), (0, 0)], 0), np.diag([0, 0, 1, 1.]))
    assert np.array_equal(props.elem_product_site([(2, 1)], 0), np.diag([1, -1, -1, 1.]))
    ql, qr = props.elem_product_q([(0, 1), (3, 0)], 0)
    assert np.array_equal(ql, [[0, 0]]) and np.array_equal(qr, [[1, 0.5]])
    ql, qr = props.elem_product_q([(0, 1), (3, 0)], 1)
    assert np.array_equal(ql, [[1, 0.5]]) and np.array_equal(qr, [[0, 1.0]])


def test_rdm1_bf_equals_fci_rdm_cpu_model():
    K = 4
    h, eri = gen_spatial(K, seed=3)
    i1, i2 = spin_orbital(h, eri)
    with host_model():
        d = pd.MPODMRG(i1, i2, K, 0.0, Dmax=16, tol=1e-12, crit_vec=1e-10)
        pd.run(d, 3)
        g, gs = props.eval_rdm1BF(d.lmps, K)
    _, gf, _ = fci_rdms(h, eri, K)
    assert np.abs(g - gf).max() < 1e-8
    assert abs(np.trace(g) - K) < 1e-10
    # singlet ground state: no spin density beyond the residual spin
    # contamination of the converged MPS
    assert np.abs(gs).max() < 1e-6
    assert np.abs(g - g.T).max() < 1e-12


@pytest.mark.gpu
def test_rdm1_bf_equals_gradient_rdm_k10(gpu):
    from zmpo_dmrg_amd import grad
    K = 10
    h, eri = gen_spatial(K, seed=0)
    i1, i2 = spin_orbital(h, eri)
    d = pd.MPODMRG(i1, i2, K, 0.0, Dmax=48, tol=1e-8)
    pd.run(d, 2)
    E, gamma, _ = grad.rdms(d.fhop, d.lmps, np.array([float(K), 0.0]), h, eri)
    g, _ = props.eval_rdm1BF(d.lmps, K)
    print('[props] K=10 max|gamma_BF - gamma_grad| = %.2e' % np.abs(g - gamma).max(), flush=True)
    assert np.abs(g - gamma).max() < 1e-10
