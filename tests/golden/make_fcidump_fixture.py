"""Integrals of the reference's FCIDUMP test case (K=10, NELEC=10, MS2=8)
and the BLOCK energy it records, as a numeric fixture.

Source: itools/fromFCIDUMP/FCIDUMP (parsed with zmpo_dmrg_amd.fcidump.read)
and itools/fromFCIDUMP/eBLOCK:1 in the reference (the BLOCK total energy,
core energy included; its convergence level is not recorded).  Run once in a
container that has the reference mounted; writes tests/golden/fcidump_k10.npz.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from zmpo_dmrg_amd import fcidump  # noqa: E402

SRC = '/root/reference/source/itools/fromFCIDUMP'


def main():
    d = fcidump.read(os.path.join(SRC, 'FCIDUMP'))
    with open(os.path.join(SRC, 'eBLOCK')) as f:
        eblock = float(f.read().split()[0])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'fcidump_k10.npz')
    np.savez_compressed(out, h=d['h'], eri=d['eri'], ecore=d['ecore'], norb=d['norb'],
                        nelec=d['nelec'], ms2=d['ms2'], eblock=eblock)
    print('wrote', out, 'norb', d['norb'], 'nelec', d['nelec'], 'ms2', d['ms2'],
          'ecore', d['ecore'], 'eBLOCK', eblock)


if __name__ == '__main__':
    main()
