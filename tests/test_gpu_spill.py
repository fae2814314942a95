"""Environments moved to host memory under a device budget (dmrg.fit_envs,
EnvStore.spill) on the device path: sweeps give the same energies bit for
bit as with every environment resident, and the moved environments come
back with the same values."""
import numpy as np
import pytest

from zmpo_dmrg_amd import dmrg as pd
from zmpo_dmrg_amd.integrals import gen_integrals

pytestmark = pytest.mark.gpu


def test_spilled_sweeps_match_resident(gpu):
    K, D = 6, 16
    i1, i2 = gen_integrals(K, seed=4)
    ref = pd.MPODMRG(i1, i2, K, 0.0, Dmax=D, tol=1e-9)
    e_ref = pd.run(ref, 2)
    big = max(e.nbytes() for e in ref.envL.values())
    d = pd.MPODMRG(i1, i2, K, 0.0, Dmax=D, tol=1e-9, env_budget=big)
    e = pd.run(d, 2)
    assert e == e_ref
    dev, host = pd.env_bytes(d)
    assert host > 0 and d.timing.get('spill', 0.0) > 0.0
    # a spilled environment uploads the values it held
    j = next(j for j, s in d.envL.items() if not s.resident())
    vals = d.envL[j].buf.copy()
    got = d.envL[j].buffer(0).cpu().numpy()
    assert np.array_equal(got, vals)
