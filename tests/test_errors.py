"""Error exits of the reference, forced from the host (no GPU work).

The reference prints and calls exit(1) (mpo_dmrg_dotutil.py:185-188 empty
symmetry sector, :348-350 negative RDM trace; tools/pdvdson.py:131-157
non-orthonormal basis / non-Hermitian projected H, :300-302 no new
direction); here they raise typed exceptions.  The C-ABI returns
ZMPO_ERR_ARG with a message for bad arguments before any device call.
"""
import numpy as np
import pytest

from kmodel import host_model
from zmpo_dmrg_amd import _lib
from zmpo_dmrg_amd import dot as pdot
from zmpo_dmrg_amd.pdvdson import DavidsonError, EigenSolver


def _rc_and_msg(rc):
    return rc, _lib.lib().zmpo_last_error().decode()


def test_cabi_argument_errors():
    L = _lib.lib()
    nd = np.ones(9, np.int32)
    st = np.zeros(9, np.int32)
    out = np.zeros(1, np.int32)
    q = np.zeros((1, 2))
    rc, msg = _rc_and_msg(L.zmpo_symAllowed(9, 2, nd.ctypes.data, st.ctypes.data,
                                            out.ctypes.data, *([q.ctypes.data] * 8)))
    assert rc == -1 and 'rank' in msg
    rc, msg = _rc_and_msg(L.zmpo_symAllowed(2, 2, nd.ctypes.data, st.ctypes.data,
                                            out.ctypes.data, q.ctypes.data, None,
                                            *([None] * 6)))
    assert rc == -1 and 'null QN' in msg
    rc, msg = _rc_and_msg(L.zmpo_shapedgemm(0, 2, None, None, 4, None, 0, None, None, None,
                                            None, None, 0, None))
    assert rc == -1 and 'null pointer' in msg
    rc, msg = _rc_and_msg(L.zmpo_shapedgemm(7, 2, None, None, 0, None, 0, None, None, None,
                                            None, None, 0, None))
    assert rc == -1 and 'dtype' in msg
    rc, msg = _rc_and_msg(L.zmpo_streamgemm(0, 2, 16, None, 0, 99, None, None, 0, None, None,
                                            None, None, None, 0, None))
    assert rc == -1 and 'upw' in msg
    rc, msg = _rc_and_msg(L.zmpo_chaingemm(0, 1, 0, 4, 5, None, 1, None, None, None, None,
                                           None, None, None, None, 0, None))
    assert rc == -1 and 'null pointer' in msg
    dummy = np.zeros(64).ctypes.data   # never dereferenced: the check comes first
    rc, msg = _rc_and_msg(L.zmpo_chaingemm(0, 1, 0, 0, 5, *([dummy] * 1), 1, *([dummy] * 8),
                                           0, None))
    assert rc == -1 and 'kcols_max' in msg
    with pytest.raises(_lib.ZmpoError):
        _lib.check(L.zmpo_symAllowed(1, 2, nd.ctypes.data, st.ctypes.data, out.ctypes.data,
                                     *([q.ctypes.data] * 8)))
    assert isinstance(L.zmpo_abi_version(), int)


def _solver(H, neig=1, crit=1e-10):
    import torch
    s = EigenSolver()
    s.ndim, s.neig = H.shape[0], neig
    s.crit_e = s.crit_vec = crit
    s.diag = torch.from_numpy(np.diag(H).copy())

    def hvec(x, out):
        out.copy_(torch.from_numpy(H @ x.numpy()))
    s.HVec = hvec
    return s


def test_davidson_non_hermitian_projected_h_raises():
    rng = np.random.default_rng(0)
    H = rng.standard_normal((12, 12))
    with host_model():
        with pytest.raises(DavidsonError, match='Hermitian'):
            _solver(H).solve_iter()


def test_davidson_neig_above_dimension_raises():
    with host_model():
        with pytest.raises(DavidsonError, match='neig'):
            _solver(np.eye(2), neig=3).solve_iter()


def test_davidson_dependent_start_vectors_raise():
    H = np.diag(np.arange(6.0))
    v0 = np.zeros((2, 6))
    v0[0, 1] = v0[1, 1] = 1.0
    with host_model():
        with pytest.raises(DavidsonError, match='orthonormal'):
            _solver(H, neig=2).solve_iter(v0)


def test_davidson_no_new_direction_raises():
    """Once the subspace spans the whole space the preconditioned residual
    has no component left; with the convergence tests switched off
    (criteria < 0) the solver must stop with the reference's error."""
    rng = np.random.default_rng(1)
    A = rng.standard_normal((2, 2))
    with host_model():
        with pytest.raises(DavidsonError, match='new direction'):
            _solver(A + A.T, crit=-1.0).solve_iter()


def test_negative_rdm_trace_raises():
    rng = np.random.default_rng(2)
    psi = rng.standard_normal((6, 5))
    ql = np.zeros((6, 2))
    with pytest.raises(RuntimeError, match='trace'):
        pdot.rdm_blkdiag([psi], ql, -1e-14, 4, 'L', extra=-10.0 * np.eye(6))


def test_empty_symmetry_sector_raises():
    """A target (N, Sz) no two-site basis state reaches (dotutil.py:185-188)."""
    from parity_util import Case
    from zmpo_dmrg_amd import dmrg as pd
    c = Case(K=4, D=16, seed=3, sweeps=0)
    with host_model():
        d = pd.MPODMRG(c.i1, c.i2, 4, Dmax=16, mpo=c.mpo)
        d.qsectors = {(40.0, 0.0): 1}
        with pytest.raises(RuntimeError, match='no basis'):
            pd.sweep(d, 'R')
