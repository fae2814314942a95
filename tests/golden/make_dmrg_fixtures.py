"""Oracle DMRG runs stored as fixtures for the device parity tests.

Each run: seeded integrals (zmpo_dmrg_amd.integrals), aufbau product-state
start, Fixed2 schedule, noise 0.  Stored: every half sweep's per-step
eigenvalues and discarded weights, the final energy, and every bond's QN list.
Usage: python tests/golden/make_dmrg_fixtures.py NAME K D SWEEPS SEED
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import sweep as osw  # noqa: E402
from zmpo_dmrg_amd.integrals import gen_integrals  # noqa: E402


def main(name, K, D, nsweeps, seed):
    K, D, nsweeps, seed = int(K), int(D), int(nsweeps), int(seed)
    i1, i2 = gen_integrals(K, seed=seed)
    st = osw.State(i1, i2, K, 0.0, Dmax=D, tol=1e-8)
    t0 = time.time()
    eig, dw = [], []
    for _ in range(nsweeps):
        for status in ('R', 'L'):
            r = osw.sweep(st, status)
            eig.append(np.array([x[0] for x in r[3]]))
            dw.append(np.array(r[4]))
            print(name, status, 'emin %.12f' % r[2], '%.1fs' % (time.time() - t0), flush=True)
    qn = {}
    for i in range(K + 1):
        qn['qnuml%d' % i] = np.asarray(st.qnuml[i])
        qn['qnumr%d' % i] = np.asarray(st.qnumr[i])
    np.savez(os.path.join(HERE, 'dmrg_%s.npz' % name), K=K, D=D, seed=seed,
             eig=np.array(eig), dwts=np.array(dw), **qn)


if __name__ == '__main__':
    main(*sys.argv[1:])
