"""FCIDUMP reader: integrals for the sweep driver from the standard text format.

Counterpart of itools/fromFCIDUMP/tools_io.py:19-52 (loadERIs): the header
gives NORB, NELEC, MS2; every following line is "value i j k l" (1-based)
with (ij|kl) for four non-zero indices (8-fold symmetric, restored here),
h_ij for k = l = 0 (symmetric) and the core energy for i = j = k = l = 0.
Orbital-energy lines (i > 0, j = k = l = 0) are skipped; any other index
pattern with zeros raises ValueError.
Returns spatial h[K,K] and chemists' eri[K,K,K,K]; integrals.spin_orbital
maps them to the MPO convention (tools_itrf.py:22-35).
"""
import re

import numpy as np


def _header_int(text, key):
    m = re.search(r'\b%s\s*=\s*(-?\d+)' % key, text, re.IGNORECASE)
    if m is None:
        raise ValueError('FCIDUMP header lacks %s' % key)
    return int(m.group(1))


def read(path, ordering=None):
    """Returns dict(h, eri, ecore, norb, nelec, ms2).  ``ordering`` (a
    permutation of the orbitals) reorders both integral sets, as
    tools_itrf.dump does before the MPO is built."""
    with open(path) as f:
        text = f.read()
    end = re.search(r'(&END|/)\s*\n', text, re.IGNORECASE)
    if end is None:
        raise ValueError('FCIDUMP header not terminated')
    head, body = text[:end.end()], text[end.end():]
    n = _header_int(head, 'NORB')
    nelec = _header_int(head, 'NELEC')
    ms2 = _header_int(head, 'MS2')
    rows = np.array([line.split() for line in body.splitlines() if line.strip()], dtype=np.float64)
    val = rows[:, 0]
    idx = rows[:, 1:].astype(np.int64) - 1
    h = np.zeros((n, n))
    eri = np.zeros((n, n, n, n))
    ecore = 0.0
    pos = idx >= 0
    two = pos.all(axis=1)
    one = pos[:, 0] & pos[:, 1] & ~pos[:, 2] & ~pos[:, 3]
    core = ~pos.any(axis=1)
    # 'e i 0 0 0' lines carry orbital energies, which some producers write
    # after the integrals; they are not part of the Hamiltonian (the
    # reference's loader would store them as h[i, K-1], tools_io.py:36-40)
    orbe = pos[:, 0] & ~pos[:, 1] & ~pos[:, 2] & ~pos[:, 3]
    bad = ~(two | one | core | orbe)
    if np.any(bad):
        raise ValueError('FCIDUMP line with an invalid index pattern: %s' % (
            ' '.join('%d' % x for x in idx[np.flatnonzero(bad)[0]] + 1),))
    if np.any(core):
        ecore = float(val[core][-1])
    i, j = idx[one, 0], idx[one, 1]
    h[i, j] = val[one]
    h[j, i] = val[one]
    p, q, r, s = (idx[two, k] for k in range(4))
    v = val[two]
    for a, b, c, d in ((p, q, r, s), (q, p, r, s), (p, q, s, r), (q, p, s, r),
                       (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)):
        eri[a, b, c, d] = v
    if ordering is not None:
        o = np.asarray(ordering)
        h = h[np.ix_(o, o)]
        eri = eri[np.ix_(o, o, o, o)]
    return {'h': h, 'eri': eri, 'ecore': ecore, 'norb': n, 'nelec': nelec, 'ms2': ms2}


def write(path, h, eri, ecore=0.0, nelec=0, ms2=0, tol=1e-14):
    """Write integrals in the format ``read`` parses (one entry per 8-fold
    orbit of eri with i >= j, k >= l, (ij) >= (kl); h_ij with i >= j)."""
    n = h.shape[0]
    lines = [' &FCI NORB=%4d,NELEC=%2d,MS2=%d,' % (n, nelec, ms2),
             '  ORBSYM=' + '1,' * n, '  ISYM=1,', ' &END']
    for i in range(n):
        for j in range(i + 1):
            ij = i * (i + 1) // 2 + j
            for k in range(n):
                for l in range(k + 1):
                    if k * (k + 1) // 2 + l > ij:
                        continue
                    v = eri[i, j, k, l]
                    if abs(v) > tol:
                        lines.append('%.17g %4d %4d %4d %4d' % (v, i + 1, j + 1, k + 1, l + 1))
    for i in range(n):
        for j in range(i + 1):
            if abs(h[i, j]) > tol:
                lines.append('%.17g %4d %4d  0  0' % (h[i, j], i + 1, j + 1))
    lines.append('%.17g  0  0  0  0' % ecore)
    with open(path, 'w') as f:
        f.write('\n'.join(lines) + '\n')
