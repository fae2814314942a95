"""FCIDUMP input and the reference's recorded BLOCK energy (golden value).

tests/golden/fcidump_k10.npz holds the integrals of the reference's
itools/fromFCIDUMP/FCIDUMP (K=10, N=10, MS2=8; parsed by
zmpo_dmrg_amd.fcidump.read, see make_fcidump_fixture.py) and its eBLOCK:1
total energy.
"""
import os

import numpy as np
import pytest

from zmpo_dmrg_amd import fcidump
from zmpo_dmrg_amd.integrals import gen_spatial
from oracle.fci import fci_energy, sector_fci_energy

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden', 'fcidump_k10.npz')


def _fixture():
    d = np.load(GOLDEN)
    return {k: d[k] for k in d.files}


def test_sector_fci_matches_fock_space_fci():
    for seed in (1, 3):
        h, eri = gen_spatial(4, seed=seed)
        for sz in (0.0, 1.0):
            a = fci_energy(h, eri, 4, sz)[0]
            b = sector_fci_energy(h, eri, 4, sz)[0]
            assert abs(a - b) < 1e-10, (seed, sz, a, b)


def test_fcidump_round_trip(tmp_path):
    d = _fixture()
    path = str(tmp_path / 'FCIDUMP')
    fcidump.write(path, d['h'], d['eri'], float(d['ecore']), int(d['nelec']), int(d['ms2']))
    r = fcidump.read(path)
    assert (r['norb'], r['nelec'], r['ms2']) == (10, 10, 8)
    assert np.array_equal(r['h'], d['h']) and np.array_equal(r['eri'], d['eri'])
    assert r['ecore'] == float(d['ecore'])
    # 8-fold symmetry restored by the reader
    e = r['eri']
    assert np.array_equal(e, e.transpose(1, 0, 2, 3)) and np.array_equal(e, e.transpose(2, 3, 0, 1))


def test_fcidump_fci_reproduces_block_energy():
    """The oracle's FCI on the reference's FCIDUMP reproduces the BLOCK total
    energy the reference records (eBLOCK:1) to its convergence (~1e-8)."""
    d = _fixture()
    e = sector_fci_energy(d['h'], d['eri'], int(d['nelec']), float(d['ms2']) / 2)[0]
    assert abs(e + float(d['ecore']) - float(d['eblock'])) < 5e-8


@pytest.mark.gpu
def test_device_dmrg_on_fcidump_reaches_block_energy(gpu):
    from zmpo_dmrg_amd import dmrg as pdmrg
    from zmpo_dmrg_amd.integrals import spin_orbital
    d = _fixture()
    i1, i2 = spin_orbital(d['h'], d['eri'])
    sz = float(d['ms2']) / 2
    m = pdmrg.MPODMRG(i1, i2, int(d['nelec']), sz, Dmax=100, tol=1e-12, crit_vec=1e-9)
    m.noise = 1e-4      # perturbative noise in the first sweeps (pRDMQt)
    pdmrg.run(m, 2)
    m.noise = 0.0
    es = pdmrg.run(m, 6)
    etot = es[-1] + float(d['ecore'])
    efci = sector_fci_energy(d['h'], d['eri'], int(d['nelec']), sz)[0] + float(d['ecore'])
    print('[fcidump] DMRG %.10f FCI %.10f BLOCK %.10f' % (etot, efci, float(d['eblock'])), flush=True)
    assert abs(etot - efci) < 1e-8
    assert abs(etot - float(d['eblock'])) < 5e-8


def test_orbital_energy_lines_are_skipped_and_bad_rows_rejected(tmp_path):
    """'e i 0 0 0' orbital-energy lines leave h untouched; a row such as
    'v i j k 0' is malformed and raises."""
    h, eri = gen_spatial(3, seed=2)
    path = str(tmp_path / 'FCIDUMP')
    fcidump.write(path, h, eri, 0.5, 4, 0)
    with open(path, 'a') as f:
        f.write('-1.25  1  0  0  0\n-0.75  3  0  0  0\n')
    r = fcidump.read(path)
    assert np.array_equal(r['h'], fcidump.read(path)['h'])
    ref = str(tmp_path / 'FCIDUMP0')
    fcidump.write(ref, h, eri, 0.5, 4, 0)
    r0 = fcidump.read(ref)
    assert np.array_equal(r['h'], r0['h']) and np.array_equal(r['eri'], r0['eri'])
    assert r['ecore'] == 0.5
    with open(path, 'a') as f:
        f.write('0.1  1  2  3  0\n')
    with pytest.raises(ValueError):
        fcidump.read(path)
