"""The reference's stored Qt MPS (itools/fromFCIDUMP/lmpsQ0, extracted by
tests/golden/make_lmps_fixture.py) read into QTensors: block tables equal
the stored ones, and its energy through the environment sweep equals the
closed-form determinant energy of its occupation; checkpoint round trip."""
import os

import numpy as np
import pytest

from zmpo_dmrg_amd import mpsio
from zmpo_dmrg_amd.integrals import spin_orbital

G = os.path.join(os.path.dirname(__file__), 'golden')


def _reference_mps():
    d = np.load(os.path.join(G, 'lmps_fcidump_q0.npz'))
    K = int(d['nsite'])
    sites = {}
    for i in range(K):
        sites[i] = mpsio.qtensor_from_dump(
            d['s%d_rank' % i][0], d['s%d_nsyms' % i], d['s%d_status' % i], d['s%d_qsyms' % i],
            d['s%d_ndims' % i], d['s%d_idlst' % i], d['s%d_value' % i], d['s%d_idx_allowed' % i])
    qnums = [d['qnum%d' % i] for i in range(K + 1)]
    return sites, qnums


def _det_energy(h, eri, occ):
    """<D|H|D> for a determinant (spin-orbital occupation list, 2i alpha, 2i+1 beta)."""
    so = [p for p, o in enumerate(occ) if o]
    e = sum(h[p // 2, p // 2] for p in so)
    for p in so:
        for q in so:
            i, j = p // 2, q // 2
            e += 0.5 * eri[i, i, j, j]
            if p % 2 == q % 2:
                e -= 0.5 * eri[i, j, j, i]
    return e


def _energy(sites, K):
    from zmpo_dmrg_amd import grad, mpo as pmpo
    f = np.load(os.path.join(G, 'fcidump_k10.npz'))
    o = np.arange(K)[::-1]            # the example's integrals are in reversed order
    h, eri = f['h'][np.ix_(o, o)], f['eri'][np.ix_(o, o, o, o)]
    i1, i2 = spin_orbital(h, eri)
    m = pmpo.MPOStore(i1, i2)
    e = grad.expectation(m, sites)
    occ = [1, 1] + [1, 0] * 8 + [0, 0]     # main0.py:33
    return e, _det_energy(h, eri, occ)


def test_reference_mps_tables_and_qnums():
    sites, qnums = _reference_mps()
    assert qnums[-1].tolist() == [[10.0, 4.0]]
    for i, qt in sites.items():
        assert np.allclose(qt.qsyms[0], qnums[i]) and np.allclose(qt.qsyms[2], qnums[i + 1])


def test_reference_mps_energy_host_model():
    from kmodel import host_model
    sites, _ = _reference_mps()
    with host_model():
        e, want = _energy(sites, len(sites))
    assert abs(e - want) < 1e-10 * max(1.0, abs(want))


@pytest.mark.gpu
def test_reference_mps_energy_on_device(gpu):
    sites, _ = _reference_mps()
    e, want = _energy(sites, len(sites))
    assert abs(e - want) < 1e-10 * max(1.0, abs(want))


def test_checkpoint_round_trip(tmp_path):
    sites, qnums = _reference_mps()
    p = str(tmp_path / 'mps.npz')
    mpsio.save(p, sites, qnums)
    s2, q2 = mpsio.load(p)
    for i in sites:
        assert np.array_equal(s2[i].idx_allowed, sites[i].idx_allowed)
        assert np.array_equal(s2[i].value, sites[i].value)
        assert s2[i].idlst == sites[i].idlst
    assert all(np.array_equal(a, b) for a, b in zip(q2, qnums))
