"""Pins the decimation / truncation of the oracle and of the package's host
decimation by what linear algebra fixes, independently of either
implementation (mpo_dmrg_qparser.py:101-180, mpo_dmrg_dotutil.py:288-495):

* Ky Fan: the kept weight equals the sum of the D largest eigenvalues of the
  whole trace-normalised RDM (the RDM is block diagonal in the kept side's
  (N, Sz) sectors, so the per-sector eigh must select exactly those);
* Eckart-Young: the discarded weight equals ||psi - P psi||^2 / ||psi||^2
  for the projector P onto the kept states, the best rank-D error, i.e. the
  sum of the discarded squared singular values of psi;
* the kept count is min(D, n) (zero weights are kept, thresh = -1e-14), the
  rotation is orthonormal, each kept vector lies in one sector whose QN is
  the returned one, and the kept states are sorted by (N, Sz);
* the diagonal of the explicit two-site H_eff (sigma applied to unit
  vectors) equals HDiagQt (the checkDiagQt pattern, kernelQt.py:129-150).

Inputs are random, so no weight tie straddles the cut.  A planted mistake
(ascending instead of descending order) fails the Ky Fan check.
"""
import numpy as np
import pytest

from oracle import sweep as osw
from zmpo_dmrg_amd import dot as pdot


def _sector_psi(rng, nl, nr, rank=None):
    """Random psi (rows = kept side) with (N, Sz) sector structure: row r in
    sector ql[r], column c in sector qr[c], nonzero only where they add up
    to the target (3, 0.5).  rank: optional cap on every block's rank."""
    lsec = [(0.0, 0.0), (1.0, 0.5), (1.0, -0.5), (2.0, 0.0), (3.0, 0.5)]
    rsec = [(3.0, 0.5), (2.0, 0.0), (2.0, 1.0), (1.0, 0.5), (0.0, 0.0)]
    ql = np.array([lsec[i] for i in rng.integers(0, len(lsec), nl)])
    qr = np.array([rsec[i] for i in rng.integers(0, len(rsec), nr)])
    psi = np.zeros((nl, nr))
    for a in lsec:
        rows = np.flatnonzero(np.all(ql == a, axis=1))
        cols = np.flatnonzero(np.all(np.abs(qr + np.array(a) - np.array([3.0, 0.5])) < 1e-12,
                                     axis=1))
        if len(rows) and len(cols):
            blk = rng.standard_normal((len(rows), len(cols)))
            if rank is not None and min(blk.shape) > rank:
                blk = rng.standard_normal((len(rows), rank)) @ rng.standard_normal((rank, len(cols)))
            psi[np.ix_(rows, cols)] = blk
    return psi / np.linalg.norm(psi), ql


def _check_truncation(psi, ql, D, dwts, qred, rot, sigs):
    n = psi.shape[0]
    rho = psi @ psi.T
    rho = rho / np.trace(rho)
    ev = np.sort(np.linalg.eigvalsh(rho))[::-1]
    nkeep = min(D, n)
    assert rot.shape == (n, nkeep)
    assert len(qred) == nkeep and len(sigs) == nkeep
    # Ky Fan: the D largest eigenvalues of the whole RDM
    assert abs(np.sum(sigs) - np.sum(ev[:nkeep])) < 1e-12
    assert abs(dwts - np.sum(ev[nkeep:])) < 1e-12
    # Eckart-Young: the projection error is the discarded weight
    P = rot @ rot.T
    err = np.linalg.norm(psi - P @ psi) ** 2 / np.linalg.norm(psi) ** 2
    sv = np.linalg.svd(psi, compute_uv=False) ** 2
    assert abs(err - dwts) < 1e-12
    assert abs(err - np.sum(np.sort(sv)[::-1][nkeep:]) / np.sum(sv)) < 1e-12
    # orthonormal, one sector per kept vector, QNs returned sorted
    assert np.abs(rot.T @ rot - np.eye(nkeep)).max() < 1e-12
    for c in range(nkeep):
        rows = np.flatnonzero(np.abs(rot[:, c]) > 1e-14)
        assert np.all(np.abs(ql[rows] - qred[c]) < 1e-12)
    keys = [tuple(q) for q in qred]
    assert keys == sorted(keys)
    # each kept vector is an eigenvector of the RDM with its weight
    for c in range(nkeep):
        assert np.linalg.norm(rho @ rot[:, c] - sigs[c] * rot[:, c]) < 1e-12


@pytest.mark.parametrize('seed,nl,nr,D', [(0, 40, 30, 7), (1, 64, 64, 20), (2, 25, 80, 24),
                                           (3, 50, 20, 33)])
def test_oracle_truncation_is_ky_fan_and_eckart_young(seed, nl, nr, D):
    rng = np.random.default_rng(seed)
    psi, ql = _sector_psi(rng, nl, nr)
    rho = psi @ psi.T
    dwts, qred, rot, sigs = osw.rdm_blkdiag(rho, ql, -1e-14, D)
    _check_truncation(psi, ql, D, dwts, qred, rot, sigs)


@pytest.mark.parametrize('seed,nl,nr,D', [(0, 40, 30, 7), (1, 64, 64, 20), (2, 25, 80, 24),
                                           (3, 50, 20, 33)])
def test_device_path_truncation_is_ky_fan_and_eckart_young(seed, nl, nr, D):
    """The host decimation of the device path (dot.rdm_blkdiag, given the
    psi matrices instead of the RDM), same checks, both sides."""
    rng = np.random.default_rng(seed)
    psi, ql = _sector_psi(rng, nl, nr)
    dwts, qred, rot, sigs = pdot.rdm_blkdiag([psi], ql, -1e-14, D, 'L')
    _check_truncation(psi, ql, D, dwts, qred, rot, sigs)
    # the same state seen from the other side: columns are the kept index
    dwts2, qred2, rot2, sigs2 = pdot.rdm_blkdiag([psi.T], ql, -1e-14, D, 'R')
    _check_truncation(psi, ql, D, dwts2, qred2, rot2, sigs2)
    assert np.abs(rot2 - rot).max() < 1e-12 and abs(dwts2 - dwts) < 1e-14


def test_oracle_and_device_truncation_agree_on_rank_deficient_sectors():
    """Rank-deficient sectors with D above the rank: the kept zero-weight
    vectors come from the fixed null-space basis in both implementations."""
    rng = np.random.default_rng(5)
    psi, ql = _sector_psi(rng, 48, 48, rank=2)
    D = 30
    a = osw.rdm_blkdiag(psi @ psi.T, ql, -1e-14, D)
    b = pdot.rdm_blkdiag([psi], ql, -1e-14, D, 'L')
    assert abs(a[0] - b[0]) < 1e-14
    assert np.array_equal(a[1], b[1])
    assert np.abs(a[2] - b[2]).max() < 1e-10
    _check_truncation(psi, ql, D, *b)


def test_planted_ascending_order_fails_ky_fan():
    """A truncation that keeps the smallest weights must fail the check."""
    rng = np.random.default_rng(0)
    psi, ql = _sector_psi(rng, 40, 30)
    rho = psi @ psi.T
    saved = osw.truncation_order
    try:
        osw.truncation_order = lambda s, q: np.argsort(s, kind='stable')
        bad = osw.rdm_blkdiag(rho, ql, -1e-14, 7)
    finally:
        osw.truncation_order = saved
    with pytest.raises(AssertionError):
        _check_truncation(psi, ql, 7, *bad)


def test_lazy_null_basis_matches_eager():
    """dot.rdm_blkdiag builds a sector's canonical null basis only when one
    of its zero-weight vectors is kept; applying canonical_null to every
    sector first gives the same rotation, weights and QNs."""
    rng = np.random.default_rng(11)
    for trial in range(20):
        nl, nr = int(rng.integers(10, 60)), int(rng.integers(5, 40))
        psi, ql = _sector_psi(rng, nl, nr, rank=int(rng.integers(1, 4)))
        D = int(rng.integers(1, nl + 5))
        lazy = pdot.rdm_blkdiag([psi], ql, -1e-14, D, 'L')
        # the oracle fixes the null basis of every sector before truncating
        ref = osw.rdm_blkdiag(psi @ psi.T, ql, -1e-14, D)
        assert abs(lazy[0] - ref[0]) < 1e-13, trial
        assert np.array_equal(lazy[1], ref[1]), trial
        assert np.abs(lazy[2] - ref[2]).max() < 1e-10, trial
        assert np.abs(lazy[3] - ref[3]).max() < 1e-13, trial


def test_heff_diagonal_equals_hdiag_k4():
    """diag(H_eff) from the oracle sigma on unit vectors equals the oracle's
    HDiagQt, and H_eff is symmetric (kernelQt.py:129-150)."""
    from parity_util import Case
    c = Case(K=4, D=16, seed=3, sweeps=1)
    for isite in (0, 1, 2):
        _, oq, _, _ = c.templates(isite)
        n = oq.size_allowed
        H = np.zeros((n, n))
        for j in range(n):
            e = np.zeros(n)
            e[j] = 1.0
            H[:, j] = c.oracle_sigma(isite, e)
        assert np.abs(H - H.T).max() < 1e-12
        d = c.oracle_diag(isite)
        assert np.abs(np.diag(H) - d).max() < 1e-12, isite
