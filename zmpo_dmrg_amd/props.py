"""Expectation values of operator products on the device (the properties
path): the brute-force one-particle density matrix eval_rdm1BF
(source/properties/mpo_dmrg_propsItrf.py:204-263).

For every p <= q the reference builds the two-term MPO {a+_pa a_qa,
a+_pb a_qb} (genMPO_Epq, mpo_dmrg_propsMPO.py:115-141), each term a
bond-dimension-1 product of elementary operators (genElemProductSpatialQt,
qtensor_opers.py:228-240, with its quantum numbers from
genElemProductQnums, :145-160), and reads <Psi|op|Psi> off a left
environment sweep (evalProps, mpo_dmrg_props.py:131-234).  Here all 2 x
K(K+1)/2 terms are one operator store, so one environment sweep (one
grouped-GEMM launch per stage and site for all terms) evaluates every
element; operators with equal bond quantum numbers at a site share their
plans.
"""
import numpy as np

from . import kernelQt as kq
from .mpo import MPOStore, qphys
from .qtensor import QTensor

_SGN = np.array([[1., 0.], [0., -1.]])
_IDN = np.eye(2)
_CRE = np.array([[0., 0.], [1., 0.]])
_ANN = np.array([[0., 1.], [0., 0.]])


def elem_spin_mat(p, so, iop):
    """a+_p (iop 1) or a_p (iop 0) seen from spin orbital so: the Jordan-
    Wigner sign left of p, the operator at p, identity right of it
    (mpo_dmrg_opers.py:56-66)."""
    if so < p:
        return _SGN
    if so == p:
        return _CRE if iop == 1 else _ANN
    return _IDN


def elem_product_site(oplst, isite):
    """The 4x4 matrix of the product of oplst = [(p, iop), ...] on spatial
    site isite (mpo_dmrg_opers.py:68-88): kron over its two spin orbitals,
    the factors multiplied in list order."""
    mat = None
    for p, iop in oplst:
        m = np.kron(elem_spin_mat(p, 2 * isite, iop), elem_spin_mat(p, 2 * isite + 1, iop))
        mat = m if mat is None else mat @ m
    return mat


def _elem_q(p, so, iop):
    """(left, right) bond QN of one elementary operator at spin orbital so
    (qtensor_opers.py:113-142, status 'L'): the change (+-1, +-spin) of the
    operator lies on every bond right of p."""
    spin = 0.5 if p % 2 == 0 else -0.5
    sgn = 1.0 if iop == 1 else -1.0
    dq = np.array([sgn, sgn * spin])
    z = np.zeros(2)
    if so < p:
        return z, z
    if so == p:
        return z, dq
    return dq, dq


def elem_product_q(oplst, isite):
    """Bond QNs (1 x 2 each) of the product on spatial site isite
    (qtensor_opers.py:145-160)."""
    ql, qr = np.zeros(2), np.zeros(2)
    for p, iop in oplst:
        ql = ql + _elem_q(p, 2 * isite, iop)[0]
        qr = qr + _elem_q(p, 2 * isite + 1, iop)[1]
    return ql.reshape(1, 2), qr.reshape(1, 2)


def product_site_tensor(oplst, isite):
    """genElemProductSpatialQt (qtensor_opers.py:228-240): legs [wl, wr, u,
    d], status [F, T, F, T], bond dimension 1."""
    ql, qr = elem_product_q(oplst, isite)
    qt = QTensor([False, True, False, True])
    qp = qphys()
    qt.fromDenseTensor(elem_product_site(oplst, isite).reshape(1, 1, 4, 4), [ql, qr, qp, qp])
    return qt


def product_store(nsite, oplists, upload=True):
    """One operator store whose operator k is the product oplists[k]; the
    plan class of operator k at a site is its pair of bond QNs there."""
    sites = [[product_site_tensor(ol, i) for ol in oplists] for i in range(nsite)]
    keys = [[tuple(np.concatenate(elem_product_q(ol, i)[:2]).reshape(-1)) for ol in oplists]
            for i in range(nsite)]
    return MPOStore.from_sites(sites, op_class=lambda p, i: keys[i][p], upload=upload)


def epq_terms(nsite, pairs):
    """[a+_pa a_qa, a+_pb a_qb] for every (p, q) (genMPO_Epq)."""
    out = []
    for p, q in pairs:
        for s in (0, 1):
            out.append([(2 * p + s, 1), (2 * q + s, 0)])
    return out


def expectations(store, sites):
    """<Psi|op_k|Psi> for every operator of the store, from one left
    environment sweep over the MPS ``sites`` (left-canonical, the last
    site holding the normalised coefficients)."""
    _, exphop = kq.genHops(store, sites, 'L')
    return exphop


def eval_rdm1BF(sites, nsite, pairs=None):
    """(rdm1t, rdm1s): spin-summed <E_pq> and 0.5 <E_pq^a - E_pq^b>, both
    symmetric (propsItrf.py:204-263, without spin projection).  pairs
    (default: every p <= q) limits the elements computed; the others stay 0."""
    if pairs is None:
        pairs = [(p, q) for p in range(nsite) for q in range(p, nsite)]
    store = product_store(nsite, epq_terms(nsite, pairs))
    ex = expectations(store, sites).reshape(len(pairs), 2)
    rdm1 = np.zeros((2, nsite, nsite))
    for (p, q), e in zip(pairs, ex):
        rdm1[:, p, q] = e
        rdm1[:, q, p] = e
    return rdm1[0] + rdm1[1], 0.5 * (rdm1[0] - rdm1[1])
