"""The FP64 drop-in and the differentiation form at their stated size:
K = 24, N = 24, Sz = 0, D = 500, seeded integrals.

Config 2: six FP64 sweeps (Fixed2, noise 0, Davidson tolerance 1e-8) from
the occupation product state.  Checked: the per-sweep energies never rise;
<Psi|H|Psi> of the final MPS from a full environment sweep (checkMPS,
mpo_dmrg_class.py:470-558) equals the last eigenvalue to 1e-8; every bond
keeps at most D states, its (N, Sz) list is sorted, holds sectors a block of
that many sites can hold and matches the site tensors.  (The reference's own energies
here; the oracle cannot run this size.)

Config 5, on that state: E(h, eri) = <Psi|H(h, eri)|Psi> at the fixed MPS.
The hand-written backward (EnergyFunction) against central differences
(step 1e-5) on 20 random symmetric integral entries, each energy from a
device-built MPO and an environment sweep, to 1e-6; the RDM identities
(Euler, trace, partial trace); gamma against the restated eval_rdm1BF
(<E_pq> from elementary-operator products) on 24 elements; and the FP32
variant of the gradient pass (every GEMM with FP32 operands) against FP64
and the finite differences.
"""
import time

import numpy as np
import pytest

from zmpo_dmrg_amd import dmrg as pd
from zmpo_dmrg_amd import grad, props
from zmpo_dmrg_amd.integrals import gen_spatial, spin_orbital
from zmpo_dmrg_amd.mpo import MPOStore

pytestmark = pytest.mark.gpu

K, D = 24, 500
FD_STEP = 1e-5


@pytest.fixture(scope='module')
def cfg2(gpu):
    h, eri = gen_spatial(K, seed=0)
    i1, i2 = spin_orbital(h, eri)
    t = time.time()
    d = pd.MPODMRG(i1, i2, K, 0.0, Dmax=D, tol=1e-8)
    print('[cfg2] K=%d D=%d set up in %.1fs' % (K, D, time.time() - t), flush=True)
    es, last = [], None
    for isw in range(6):
        t0 = time.time()
        r1 = pd.sweep(d, 'R')
        r2 = pd.sweep(d, 'L')
        es.append(min(r1[2], r2[2]))
        last = float(r2[3][-1][0])
        print('[cfg2] sweep %d E=%.10f (%.1fs)' % (isw, es[-1], time.time() - t0), flush=True)
    yield {'h': h, 'eri': eri, 'd': d, 'es': es, 'last': last}


def test_config2_six_fp64_sweeps(cfg2):
    d, es = cfg2['d'], cfg2['es']
    assert all(b <= a + 1e-8 for a, b in zip(es, es[1:])), es
    e_mps = grad.expectation(d.fhop, d.lmps)
    print('[cfg2] <Psi|H|Psi> = %.10f, last eigenvalue %.10f' % (e_mps, cfg2['last']), flush=True)
    assert abs(e_mps - cfg2['last']) < 1e-8
    for i in range(1, K):
        q = np.asarray(d.qnuml[i])
        assert 1 <= len(q) <= D, i
        n, sz = q[:, 0], q[:, 1]
        # states of any sector a block of i sites holds (zero-weight states
        # of sectors the right block cannot complete are kept when D exceeds
        # the rank, as the reference does with thresh = -1e-14)
        assert np.all(n >= 0) and np.all(n <= 2 * i)
        assert np.all(np.abs(sz) <= 0.5 * np.minimum(n, 2 * i - n) + 1e-12)
        keys = [tuple(x) for x in q]
        assert keys == sorted(keys)
        assert int(d.lmps[i - 1].shape[2]) == len(q) == int(d.lmps[i].shape[0])


def _energy(h, eri, sites):
    return grad.expectation(MPOStore(*spin_orbital(h, eri)), sites)


def _fd_directions(rng, n):
    """(kind, index) of n random entries, alternating one- and two-electron."""
    out = []
    for k in range(n):
        if k % 2 == 0:
            out.append(('h', tuple(int(x) for x in rng.integers(K, size=2))))
        else:
            out.append(('eri', tuple(int(x) for x in rng.integers(K, size=4))))
    return out


def _perturb(h, eri, kind, t, eps):
    h2, e2 = h.copy(), eri.copy()
    if kind == 'h':
        p, r = t
        h2[p, r] += eps
        if p != r:
            h2[r, p] += eps
    else:
        seen = set()
        for pm in grad._PERM8:
            u = tuple(t[i] for i in pm)
            if u not in seen:
                seen.add(u)
                e2[u] += eps
    return h2, e2


def _analytic(gh, ge, kind, t):
    if kind == 'h':
        p, r = t
        return gh[p, r] + (gh[r, p] if p != r else 0.0)
    orbit = {tuple(t[i] for i in pm) for pm in grad._PERM8}
    return sum(ge[u] for u in orbit)


@pytest.fixture(scope='module')
def fd(cfg2):
    """Central differences of the fixed-MPS energy on 20 entries."""
    h, eri, d = cfg2['h'], cfg2['eri'], cfg2['d']
    dirs = _fd_directions(np.random.default_rng(11), 20)
    t = time.time()
    vals = []
    for kind, tt in dirs:
        hp, ep = _perturb(h, eri, kind, tt, FD_STEP)
        hm, em = _perturb(h, eri, kind, tt, -FD_STEP)
        vals.append((_energy(hp, ep, d.lmps) - _energy(hm, em, d.lmps)) / (2 * FD_STEP))
    print('[cfg5] 20 central differences in %.1fs' % (time.time() - t), flush=True)
    return dirs, np.array(vals)


def test_config5_fp64_backward_matches_finite_differences(cfg2, fd):
    import torch
    h, eri, d = cfg2['h'], cfg2['eri'], cfg2['d']
    ht = torch.from_numpy(h.copy()).requires_grad_(True)
    et = torch.from_numpy(eri.copy()).requires_grad_(True)
    t = time.time()
    E = grad.EnergyFunction.apply(ht, et, d.fhop, d.lmps, np.array([float(K), 0.0]))
    E.backward()
    print('[cfg5] FP64 energy + backward in %.1fs' % (time.time() - t), flush=True)
    gh, ge = ht.grad.numpy(), et.grad.numpy()
    assert abs(float(E.detach()) - cfg2['last']) < 1e-8
    dirs, want = fd
    got = np.array([_analytic(gh, ge, k, tt) for k, tt in dirs])
    err = np.abs(got - want).max()
    print('[cfg5] FP64 max |backward - FD| = %.2e' % err, flush=True)
    assert err < 1e-6


def test_config5_rdm_identities_and_rdm1_bf(cfg2):
    h, eri, d = cfg2['h'], cfg2['eri'], cfg2['d']
    E, gamma, Gamma = grad.rdms(d.fhop, d.lmps, np.array([float(K), 0.0]), h, eri)
    assert abs(E - cfg2['last']) < 1e-8
    assert abs(np.einsum('pq,pq', h, gamma) + 0.5 * np.einsum('pqrs,pqrs', eri, Gamma) - E) \
        < 1e-8 * abs(E)
    assert abs(np.trace(gamma) - K) < 1e-9
    assert np.abs(np.einsum('prqq->pr', Gamma) - (K - 1) * gamma).max() < 1e-8
    rng = np.random.default_rng(5)
    pairs = [(p, p) for p in range(0, K, 3)] + \
        [tuple(sorted(int(x) for x in rng.choice(K, 2, replace=False))) for _ in range(16)]
    g_bf, _ = props.eval_rdm1BF(d.lmps, K, pairs)
    err = max(abs(g_bf[p, q] - gamma[p, q]) for p, q in pairs)
    print('[cfg5] max |eval_rdm1BF - gamma| over %d elements = %.2e' % (len(pairs), err),
          flush=True)
    assert err < 1e-9


def test_config5_fp32_gradient_pass(cfg2, fd):
    h, eri, d = cfg2['h'], cfg2['eri'], cfg2['d']
    qkey = np.array([float(K), 0.0])
    E64, gh64, ge64 = grad.integral_gradients(d.fhop, d.lmps, qkey, h, eri)
    t = time.time()
    E32, gh32, ge32 = grad.integral_gradients(d.fhop, d.lmps, qkey, h, eri, dtype=1)
    print('[cfg5] FP32 gradient pass in %.1fs' % (time.time() - t), flush=True)
    dh = np.abs(gh32 - gh64).max()
    de = np.abs(ge32 - ge64).max()
    dirs, want = fd
    got = np.array([_analytic(gh32, ge32, k, tt) for k, tt in dirs])
    err = np.abs(got - want).max()
    print('[cfg5] FP32: |dE| %.2e, max|dh| %.2e, max|deri| %.2e, max |grad - FD| %.2e' % (
        abs(E32 - E64), dh, de, err), flush=True)
    # FP32 operands: rounding ~6e-8 relative on environment entries of the
    # size of the energy (~1e2 Ha); measured |dE| 2.0e-4 Ha (1.5e-6
    # relative), gradients 8e-6, against the differences 1.2e-6
    assert abs(E32 - E64) < 1e-5 * abs(E64)
    assert dh < 1e-4 and de < 1e-4 and err < 1e-5
