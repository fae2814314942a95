"""Sweep schedules (source/mpo_dmrg_schedule.py) and the release of dead
environments during half-sweeps, on the CPU kernel model."""
import numpy as np

from kmodel import host_model
from zmpo_dmrg_amd import dmrg as pd
from zmpo_dmrg_amd.schedule import Schedule


def test_normal2_parameters_follow_the_reference_rule():
    """maxM = 1000, tol = 1e-7: startM = int(1.5 sqrt(1000)) = 47 doubled to
    752, then four sweeps at 1000 with Davidson tolerance 1e-4 .. 1e-7;
    noise 1e-4 up to and including the first sweep at maxM
    (schedule.py:126-168); change = 9 + 2."""
    s = Schedule(maxM=1000, tol=1e-7).normal()
    assert s.maxMs == [47, 94, 188, 376, 752, 1000, 1000, 1000, 1000]
    assert np.allclose(s.tols, [1e-4] * 6 + [1e-5, 1e-6, 1e-7])
    assert s.noises == [1e-4] * 6 + [0.0] * 3
    assert s.change == 11
    assert s.parameters(20) == (1000, s.tols[-1], 0.0)
    assert not s.converged(11, 1e-9) and s.converged(12, 1e-9)
    f = Schedule(maxM=50, tol=1e-8).fixed(noise=0.0)
    assert f.parameters(7) == (50, 1e-8, 0.0) and f.converged(0, 1e-9)
    small = Schedule(maxM=16, tol=1e-6).normal()
    assert small.maxMs == [16, 16, 16]          # startM = min(25, maxM)


def _run(keep, K=5, D=8, sweeps=1):
    from parity_util import Case
    c = Case(K=K, D=D, seed=5, sweeps=0)
    d = pd.MPODMRG(c.i1, c.i2, c.nelec, Dmax=D, tol=1e-9, mpo=c.mpo, keep_envs=keep)
    live = []
    orig = pd.release_dead_envs

    def spy(dmrg, isite, status, after):
        orig(dmrg, isite, status, after)
        if after:
            live.append(len(dmrg.envL) + len(dmrg.envR))
    pd.release_dead_envs = spy
    try:
        es = pd.run(d, sweeps)
    finally:
        pd.release_dead_envs = orig
    return es, live, d


def test_released_environments_change_no_energy():
    with host_model():
        e_keep, live_keep, _ = _run(True)
        e_rel, live, d = _run(False)
    assert e_rel == e_keep                     # bit for bit
    assert live_keep == []                     # keep_envs: nothing released
    # at most one environment per bond plus the two boundaries
    assert max(live) <= d.nsite + 2
    assert len(d.envR) == 1 and len(d.envL) == d.nsite   # after an 'L' half-sweep


def test_normal2_run_reaches_fci_on_small_case():
    """The Normal2 ramp (with its noise) ends at the FCI energy when maxM
    covers the whole space (K = 4)."""
    from oracle.fci import fci_energy
    from parity_util import Case
    c = Case(K=4, D=16, seed=3, sweeps=0)
    with host_model():
        d = pd.MPODMRG(c.i1, c.i2, 4, Dmax=16, tol=1e-10, mpo=c.mpo)
        sched = Schedule(maxM=16, tol=1e-7).normal()
        es = pd.run(d, 20, schedule=sched)
    assert len(es) >= sched.change
    from zmpo_dmrg_amd.integrals import gen_spatial
    h, eri = gen_spatial(4, seed=3)
    efci = fci_energy(h, eri, 4)[0]
    assert abs(es[-1] - efci) < 1e-8


def test_spilled_environments_change_no_energy():
    """An environment budget below one sweep's environments moves the ones
    read farthest ahead to host memory and back: the energies stay bit for
    bit, the steps' two environments are resident when read, and the
    resident bytes stay within the budget plus those two."""
    from parity_util import Case
    with host_model():
        c = Case(K=5, D=8, seed=5, sweeps=0)
        ref = pd.MPODMRG(c.i1, c.i2, c.nelec, Dmax=8, tol=1e-9, mpo=c.mpo)
        e_ref = pd.run(ref, 1)
        big = max(e.nbytes() for e in ref.envL.values())
        budget = big
        d = pd.MPODMRG(c.i1, c.i2, c.nelec, Dmax=8, tol=1e-9, mpo=c.mpo, env_budget=budget)
        seen = []
        orig = pd.fit_envs

        def spy(dmrg, isite, status, envL=None):
            orig(dmrg, isite, status, envL)
            if envL is None:
                dev, host = pd.env_bytes(dmrg)
                need = [dmrg.envL[isite - 1], dmrg.envR[isite + 2]]
                pending = sum(e.nbytes() for e in need if not e.resident())
                seen.append((dev + pending, host))
        pd.fit_envs = spy
        try:
            e = pd.run(d, 1)
        finally:
            pd.fit_envs = orig
    assert e == e_ref
    assert max(s[0] for s in seen) <= budget + 2 * big
    assert max(s[1] for s in seen) > 0                  # something was spilled
    assert d.timing.get('spill', 0.0) > 0.0


def test_host_budget_overflow_is_an_error():
    import pytest
    from parity_util import Case
    with host_model():
        c = Case(K=5, D=8, seed=5, sweeps=0)
        with pytest.raises(MemoryError):
            pd.MPODMRG(c.i1, c.i2, c.nelec, Dmax=8, tol=1e-9, mpo=c.mpo, env_budget=0,
                       host_budget=1)


class _FakeEnv:
    def __init__(self, n):
        self.n, self.on = n, True

    def nbytes(self):
        return self.n

    def resident(self):
        return self.on

    def spill(self):
        self.on = False


class _FakeDmrg:
    def __init__(self, budget, host=None):
        self.env_budget, self.host_budget, self.timing = budget, host, {}
        self.envL = {j: _FakeEnv(10) for j in range(-1, 5)}
        self.envR = {j: _FakeEnv(10) for j in range(4, 9)}


def test_fit_envs_moves_the_farthest_read_first():
    """'L' half-sweep at site 3: left environments go first, smallest bond
    first (read last in the next 'R' half-sweep), then right ones from the
    far end; the two the step reads (envL[2], envR[5]) never move."""
    d = _FakeDmrg(budget=60)
    pd.fit_envs(d, 3, 'L')
    moved = [('L', j) for j, e in d.envL.items() if not e.on] + \
        [('R', j) for j, e in d.envR.items() if not e.on]
    # 11 stores of 10 bytes, budget 60: five move
    assert moved == [('L', -1), ('L', 0), ('L', 1), ('L', 3), ('L', 4)]
    assert d.envL[2].on and d.envR[5].on
    # 'R' half-sweep at site 3 (reads envL[2], envR[5]): right ones from the
    # far end first, then left ones from the smallest bond
    d = _FakeDmrg(budget=60)
    pd.fit_envs(d, 3, 'R')
    moved = [('R', j) for j, e in d.envR.items() if not e.on] + \
        [('L', j) for j, e in d.envL.items() if not e.on]
    assert moved == [('R', 4), ('R', 6), ('R', 7), ('R', 8), ('L', -1)]
    # no budget: nothing moves; a host budget too small raises first
    d = _FakeDmrg(budget=None)
    pd.fit_envs(d, 3, 'L')
    assert all(e.on for e in list(d.envL.values()) + list(d.envR.values()))
    d = _FakeDmrg(budget=60, host=25)
    import pytest
    with pytest.raises(MemoryError):
        pd.fit_envs(d, 3, 'L')
