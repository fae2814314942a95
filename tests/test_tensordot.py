"""One-shot qtensor.tensordot (qtensor.py:665-827): general, outer-product and
scalar cases vs numpy.tensordot of the dense tensors (the reference's own
check pattern, qtensor_examples.py:75-90)."""
import numpy as np
import pytest

from zmpo_dmrg_amd import device as dv
from zmpo_dmrg_amd.qtensor import QTensor, tensordot


def _rand(status, dims_q, rng):
    """Random block-sparse tensor: per leg a QN list (1-component, values in
    0..2), dense values zeroed outside the allowed blocks by fromDense."""
    qn = [np.array([[float(v)] for v in rng.integers(0, 3, size=n)]) for n in dims_q]
    q = QTensor(status)
    q.fromDenseTensor(rng.standard_normal(tuple(dims_q)), qn)
    return q, qn


def _cases(rng):
    a, qa = _rand([False, False, True], [4, 3, 5], rng)
    # same leg sectors for the contracted legs, opposite direction
    b = QTensor([False, True, True])
    b.fromDenseTensor(rng.standard_normal((5, 3, 4)), [qa[2], qa[1], qa[0]])
    c, _ = _rand([False, True], [3, 3], rng)
    return a, b, c


def _run():
    rng = np.random.default_rng(2)
    a, b, c = _cases(rng)
    da, db, dc = a.toDenseTensor(), b.toDenseTensor(), c.toDenseTensor()
    # scalar: every leg of both, legs in a permuted order
    s = tensordot(a, b, ([0, 1, 2], [2, 1, 0]))
    assert abs(s) > 1e-6
    assert abs(s - float(np.tensordot(da, db, ([0, 1, 2], [2, 1, 0])))) < 1e-12 * max(1.0, abs(s))
    # general: one leg
    g = tensordot(a, b, ([2], [0]))
    assert np.abs(g.toDenseTensor() - np.tensordot(da, db, ([2], [0]))).max() < 1e-12
    # outer product
    o = tensordot(a, c, ([], []))
    assert np.abs(o.toDenseTensor() - np.multiply.outer(da, dc)).max() < 1e-12


def test_tensordot_cases_host_model():
    from kmodel import host_model
    with host_model():
        _run()


@pytest.mark.gpu
def test_tensordot_cases_on_device(gpu):
    _run()
