"""The reference's stored Qt MPS of its FCIDUMP example as a numeric fixture.

Source: itools/fromFCIDUMP/lmpsQ0 in the reference -- the occupation
product state [1,1]+[1,0]*8+[0,0] (itools/fromFCIDUMP/main0.py:31-60) written
by qtensor_api.fmpsQt in the on-disk layout of qtensor.dump
(qtensor/qtensor.py:188-261): per site a group with attributes rank, nsyms,
status and datasets qsyms (sectors of all legs stacked), ndims, idlst,
idx_allowed, value.  h5py is not installed, so the groups are read with the
h5dump tool; writes tests/golden/lmps_fcidump_q0.npz.  Run once in a
container that has the reference mounted.
"""
import os
import re
import subprocess

import numpy as np

SRC = '/root/reference/source/itools/fromFCIDUMP/lmpsQ0'
H5DUMP = '/opt/conda/bin/h5dump'


def _numbers(text):
    body = text[text.index('DATA {') + 6:]
    body = body[:body.index('}')]
    toks = re.findall(r'TRUE|FALSE|[-+0-9.eE]+', body)
    return [1.0 if t == 'TRUE' else 0.0 if t == 'FALSE' else float(t) for t in toks]


def _read(flag, path):
    out = subprocess.run([H5DUMP, '-y', '-w', '100000', flag, path, SRC], capture_output=True,
                         text=True, check=True).stdout
    return np.array(_numbers(out))


def main():
    nsite = int(_read('-d', '/nsite')[0])
    arrs = {'nsite': nsite}
    for i in range(nsite + 1):
        arrs['qnum%d' % i] = _read('-d', '/qnum%d' % i).reshape(-1, 2)
    for i in range(nsite):
        g = '/site%d' % i
        for a in ('rank', 'nsyms', 'status'):
            arrs['s%d_%s' % (i, a)] = _read('-a', g + '/' + a)
        for d in ('qsyms', 'ndims', 'idlst', 'idx_allowed', 'value'):
            arrs['s%d_%s' % (i, d)] = _read('-d', g + '/' + d)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'lmps_fcidump_q0.npz')
    np.savez_compressed(out, **arrs)
    print('wrote', out, 'nsite', nsite)


if __name__ == '__main__':
    main()
