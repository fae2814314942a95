"""Energy gradients with respect to the integrals: dE/dh = 1-RDM, dE/d(pq|rs) = 2-RDM / 2.

For a converged MPS |Psi> (all sites left-canonical, the last one holding
the normalised coefficients, as a sweep ending with an 'L' half-sweep leaves
it), E(h, eri) = <Psi|H(h, eri)|Psi> = sum_p w_p <Psi|H_p|Psi> is a chain
contraction over sites of  A_j^* . W_p[j](h, eri) . A_j.  Its reverse-mode
derivative is written out by hand:

* backward through the renormalisation chain: the cotangent of a left
  environment L_p[j-1] is the right environment R_p[j+1] of the same MPS,
  built by the right renormalisation (renormQt_H00 'R', kernelQt.py:633-767)
  over the sites relabelled to right-site quantum numbers;
* backward of one site: dE/dW_p[j] = w_p L_p[j-1] (x) A_j^* (x) A_j (x) R_p[j+1],
  three grouped-GEMM launches per op chunk, planned by the same block-pair
  tables (tensordotSYM, qtensor.py:563-637) as the sigma and renormalisation
  launches;
* left environments are recomputed site by site during the gradient pass
  (one live at a time) instead of being kept from the forward sweep;
* dW -> (int1e, int2e) -> (h, eri): torch autograd through a torch restatement
  of the MPO construction (genWfacSpin / genHfacSpin / linkTwoOpers,
  mpo_dmrg_opers.py:43-203) and of the spin-orbital mapping
  (itools/fromFCIDUMP/tools_itrf.py:22-35).

With H = sum h_pr E_pr + 1/2 sum (pr|qs) (E_pr E_qs - delta_qr E_ps) the
results are gamma_pr = dE/dh_pr and Gamma_prqs = 2 dE/d(pr|qs) summed over
the 8 index permutations that leave (pr|qs) unchanged (the MPO reads each
two-electron integral through one representative of that orbit), which is
what the reference's property code evaluates as <E_pq> and <a+a+aa> MPO
expectation values (properties/mpo_dmrg_propsItrf.py:204-263,
mpo_dmrg_propsMPO.py:73-141).  By the Hellmann-Feynman theorem these are the
gradients of the variational energy; ``EnergyFunction`` exposes them as a
torch.autograd.Function.
"""
import numpy as np

from . import _lib
from . import device as dv
from . import kernelQt as kq
from . import mpo as pmpo
from . import qtensor as qq
from .qtensor import Program, QTensor


# ------------------------------------------------------------------ torch MPO builder
def _mats(torch):
    return {k: torch.tensor(v, dtype=torch.float64) for k, v in pmpo._S.items()}


def wfac_spin_t(nsite, isite, hq, vqrs):
    """Torch form of mpo.wfac_spin batched over ops: hq[P, n], vqrs[P, n, n, n]
    -> w[P, nb, nb, 2, 2] (mpo_dmrg_opers.py:91-194)."""
    torch = dv.torch()
    S = _mats(torch)
    P = hq.shape[0]
    nb = 2 * (nsite + 1)
    diml, dimr = isite, nsite - isite - 1
    rs = torch.arange(isite + 1, nsite)
    ls = torch.arange(diml)
    ar = torch.arange(dimr)
    last = nb - 1
    w = torch.zeros((P, nb, nb, 2, 2), dtype=torch.float64)
    w[:, 0, 0] = S['idnt']
    w[:, 0, 1 + dimr:1 + 2 * dimr] = vqrs[:, isite, isite][:, rs][:, :, None, None] * S['niit']
    w[:, 0, 1 + 2 * dimr + diml] = -S['ann']
    w[:, 0, 2 + 2 * dimr + 2 * diml] = S['cre']
    w[:, 0, last] = hq[:, isite][:, None, None] * S['ann']
    w[:, 1, last] = S['cre']
    w[:, 2 + ar, 1 + ar] = S['idnt']
    w[:, 2 + dimr, last] = S['ann']
    w[:, 3 + dimr + ar, 1 + dimr + ar] = S['idnt']
    off = 3 + 2 * dimr
    if diml and dimr:
        a = vqrs[:, rs][:, :, ls][:, :, :, isite].transpose(1, 2)    # [P, l, r] = v[r, l, i]
        b = vqrs[:, isite][:, ls][:, :, rs]                           # [P, l, r] = v[i, l, r]
        c = vqrs[:, ls][:, :, isite][:, :, rs]                        # [P, l, r] = v[l, i, r]
        w[:, off:off + diml, 1:1 + dimr] = -a[..., None, None] * S['annt']
        w[:, off:off + diml, 1 + dimr:1 + 2 * dimr] = b[..., None, None] * S['cret']
        off4 = off + diml
        w[:, off4:off4 + diml, 1 + dimr:1 + 2 * dimr] = c[..., None, None] * S['annt']
    if diml:
        w[:, off + ls, 1 + 2 * dimr + ls] = S['idn']
        w[:, off:off + diml, last] = vqrs[:, isite][:, ls][:, :, isite][..., None, None] * S['nii']
        off4 = off + diml
        w[:, off4 + ls, 2 + 2 * dimr + diml + ls] = S['idn']
    w[:, last, last] = S['idn']
    if isite == 0:
        w = w[:, :1]
    if isite == nsite - 1:
        w = w[:, :, last:]
    return w


def site_dense_t(ops, nsite, isite, int1e, int2e):
    """Dense H_p[isite] (unsorted bond order) for the ops in ``ops``:
    [P, wl, wr, 4, 4], differentiable in int1e / int2e (torch tensors)."""
    torch = dv.torch()
    S = _mats(torch)
    idx = torch.as_tensor(np.asarray(ops, dtype=np.int64))
    hq, vq = int1e[idx], int2e[idx]
    out = []
    for so in (2 * isite, 2 * isite + 1):
        e = torch.stack([S['sgn'] if so < p else (S['cre'] if so == p else S['idn']) for p in ops])
        w = wfac_spin_t(nsite, so, hq, vq)
        out.append(torch.einsum('pij,pabjk->pabik', e, w))
    t = torch.einsum('pabij,pbckl->pacikjl', out[0], out[1])
    s = t.shape
    return t.reshape(s[0], s[1], s[2], s[3] * s[4], s[5] * s[6])


def spin_orbital_t(h, eri):
    """Torch form of integrals.spin_orbital (differentiable)."""
    torch = dv.torch()
    K = h.shape[0]
    n = 2 * K
    h1 = torch.zeros((n, n), dtype=torch.float64)
    h1[0::2, 0::2] = h
    h1[1::2, 1::2] = h
    g = torch.zeros((n, n, n, n), dtype=torch.float64)
    g[0::2, 0::2, 0::2, 0::2] = eri
    g[1::2, 1::2, 1::2, 1::2] = eri
    g[0::2, 0::2, 1::2, 1::2] = eri
    g[1::2, 1::2, 0::2, 0::2] = eri
    g = g.permute(0, 2, 1, 3)
    g = -0.5 * (g - g.permute(0, 1, 3, 2))
    return h1, g


# ------------------------------------------------------------------ device dE/dW
def as_right_site(qt, qkey):
    """A left-canonical site [l, n, r] (status F,F,T, left-counted quantum
    numbers) relabelled as a right site (status T,F,F, quantum numbers
    counted from the right end: q -> qkey - q).  Same blocks, same values."""
    r = QTensor([True, False, False])
    r.fromQsyms(3, [qkey - qt.qsyms[0], qt.qsyms[1], qkey - qt.qsyms[2]], qt.ndims,
                ifallocate=False, idlst=qt.idlst)
    if not (np.array_equal(r.idx_allowed, qt.idx_allowed)
            and np.array_equal(r.blk_off, qt.blk_off)):
        raise RuntimeError('relabelled site has a different block structure')
    r.value = qt.value
    return r


class WGradProgram:
    """dE/dW_p[j] = w_p L_p[j-1] (x) A_j^* (x) A_j (x) R_p[j+1] for every op of one site.

    Three launches per op chunk:  x = L_p . A (ket) -> [w_l, l_b, d, r_k],
    y = R_p . A (bra) -> [w_r, r_k, l_b, u],  z = x . y over (l_b, r_k)
    -> z[w_l, d, w_r, u], the gradient of the MPO block W[w_l, w_r, u, d]
    (u: bra, d: ket physical index)."""

    def __init__(self, mpo, site_idx, qsite, envL, envR, qkey, hpwts, budget=kq.DEFAULT_BUDGET):
        nops = mpo.nops
        self.site_idx = site_idx
        case_of = [kq.case(mpo.opers[p], site_idx) for p in range(nops)]
        plans = {}
        for p in range(nops):
            c = case_of[p]
            if c in plans:
                continue
            lop, rop = envL.meta[p], envR.meta[p]
            rf = rop.template()
            rf.qsyms = [-rop.qsyms[0], qkey - rop.qsyms[1], qkey - rop.qsyms[2]]
            x, y, z = QTensor(), QTensor(), QTensor()
            x.tensordotSYM(lop, qsite, ([2], [0]))
            # y[w_r, r_k, l_b, u] stored as [w_r, u, l_b, r_k], so that the legs
            # z contracts (l_b, r_k) are adjacent in x and in y (one K index)
            y.tensordotSYM(rf, qsite, ([1], [2]), perm=(0, 3, 2, 1))
            z.tensordotSYM(x, y, ([1, 3], [2, 1]))
            cop = mpo.sites[site_idx][p]
            if not (np.array_equal(z.ndims[0], cop.ndims[0]) and np.array_equal(z.ndims[2], cop.ndims[1])
                    and np.allclose(np.abs(z.qsyms[0]), np.abs(cop.qsyms[0]))
                    and np.allclose(np.abs(z.qsyms[2]), np.abs(cop.qsyms[1]))):
                raise RuntimeError('gradient bond sectors differ from the MPO site tensor')
            plans[c] = [x.plan.lowered(), y.plan.lowered(), z.plan.lowered(), x, y, z]
        self.plans = plans
        self.case_of = case_of
        self.meta = [plans[c][5] for c in case_of]
        sizes = np.array([m.size_allowed for m in self.meta], dtype=np.int64)
        self.out_off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.out_size = int(sizes.sum())
        isz = [(plans[c][3].size_allowed + plans[c][4].size_allowed) * 8 for c in case_of]
        self.chunks = kq._chunks(isz, budget)
        self.stages, self.inter, self.flops = [], [0, 0], 0
        for ch in self.chunks:
            offs = [0, 0]
            parts = [[], [], []]
            for p in ch:
                pl = plans[case_of[p]]
                parts[0].append((pl[0], envL.off[p], 0, offs[0], 1.0))
                parts[1].append((pl[1], envR.off[p], 0, offs[1], 1.0))
                parts[2].append((pl[2], offs[0], offs[1], self.out_off[p], float(hpwts[p])))
                offs[0] += pl[3].size_allowed
                offs[1] += pl[4].size_allowed
            progs = [Program(*kq._instantiate_many(parts[k])) for k in range(3)]
            self.flops += sum(pg.flops for pg in progs)
            self.stages.append(progs)
            self.inter = [max(a, b) for a, b in zip(self.inter, offs)]

    def apply(self, site_values, envL, envR, dtype=0):
        """Device buffer (FP64) with z of op p at out_off[p]; dtype 1 runs the
        three launches with FP32 operands (site_values then FP32 too)."""
        x = dv.empty(self.inter[0], dtype)
        y = dv.empty(self.inter[1], dtype)
        out = dv.zeros(self.out_size, dtype)
        EL, ER = envL.buffer(dtype), envR.buffer(dtype)
        for st in self.stages:
            qq.run_program(st[0], EL, site_values, x, dtype)
            qq.run_program(st[1], ER, site_values, y, dtype)
            qq.run_program(st[2], x, y, out, dtype)
        if dtype != 0:
            o64 = dv.empty(self.out_size, 0)
            _lib.check(_lib.lib().zmpo_convert(0, dv.ptr(o64), dtype, dv.ptr(out),
                                               self.out_size, dv.stream()))
            out = o64
        return out

    def dense(self, mpo, values):
        """Host dE/dW_p[j] for every op, dense [w_l, w_r, 4, 4] in the MPO's
        stored (sorted) bond order."""
        out = []
        for p in range(mpo.nops):
            cop = mpo.sites[self.site_idx][p]
            z = self.meta[p].template()
            z.value = values[self.out_off[p]:self.out_off[p] + z.size_allowed]
            d = z.toDenseTensor([cop.idlst[0], cop.idlst[3], cop.idlst[1], cop.idlst[2]])
            out.append(d.transpose(0, 2, 3, 1))
        return out


def right_environments(mpo, sites, qkey, dtype=0):
    """R_p[j] for j = K .. 1 over the left-canonical sites relabelled as right
    sites (the cotangents of the left environments); dtype 1: FP32 GEMMs,
    environments kept in FP64."""
    K = mpo.nsite
    envs = {K: kq.boundary_store(mpo.nops)}
    for j in range(K - 1, 0, -1):
        envs[j], _ = kq.renorm(mpo, j, as_right_site(sites[j], qkey), envs[j + 1], 'R', dtype)
    return envs


def w_gradients(mpo, sites, qkey, hpwts=None, dtype=0):
    """Yield (j, [dE/dW_p[j] dense sorted, for each op]) for j = 0 .. K-1.

    Also checks the invariant  sum_p <dE/dW_p[j], W_p[j]> = E  at every site
    (E is linear in each W[j]); the per-site values are collected in
    ``w_gradients.last_energies``.  dtype 1: every GEMM of the pass (the
    environment recomputation and the three dE/dW launches) with FP32
    operands; environments and results are held in FP64."""
    K = mpo.nsite
    hpwts = mpo.hpwts if hpwts is None else hpwts
    envR = right_environments(mpo, sites, qkey, dtype)
    envL = kq.boundary_store(mpo.nops)
    energies = []
    for j in range(K):
        s = sites[j]
        sv = dv.upload(s.value, dtype) if isinstance(s.value, np.ndarray) else \
            (s.value if dtype == 0 else s.value.float())
        prog = WGradProgram(mpo, j, s, envL, envR[j + 1], qkey, hpwts)
        vals = dv.download(prog.apply(sv, envL, envR[j + 1], dtype))
        g = prog.dense(mpo, vals)
        e = 0.0
        for p in range(mpo.nops):
            cop = mpo.sites[j][p]
            e += float(np.vdot(g[p], cop.toDenseTensor()))
        energies.append(e)
        yield j, g
        if j < K - 1:
            envL, _ = kq.renorm(mpo, j, s, envL, 'L', dtype)
    w_gradients.last_energies = energies


w_gradients.last_energies = []


def integral_gradients(mpo, sites, qkey, h, eri, dtype=0):
    """(E, dE/dh[K,K], dE/deri[K,K,K,K]) of <Psi|H(h, eri)|Psi> for the MPS
    ``sites`` (left-canonical, normalised), raw (unsymmetrised) gradients;
    dtype 1 runs the device pass with FP32 GEMM operands."""
    torch = dv.torch()
    # also called from a torch.autograd.Function backward, where grad mode is off
    with torch.enable_grad():
        return _integral_gradients(torch, mpo, sites, qkey, h, eri, dtype)


def _integral_gradients(torch, mpo, sites, qkey, h, eri, dtype=0):
    ht = torch.as_tensor(np.asarray(h, dtype=np.float64)).clone().requires_grad_(True)
    et = torch.as_tensor(np.asarray(eri, dtype=np.float64)).clone().requires_grad_(True)
    int1e, int2e = spin_orbital_t(ht, et)
    i1 = int1e.detach().clone().requires_grad_(True)
    i2 = int2e.detach().clone().requires_grad_(True)
    nsite = 2 * mpo.nsite
    ops = list(range(mpo.nops))
    for j, g in w_gradients(mpo, sites, qkey, dtype=dtype):
        w = site_dense_t(ops, nsite, j, i1, i2)
        G = np.zeros(tuple(w.shape))
        for p in ops:
            _, _, il, ir = pmpo.site_order(p, nsite, j)
            G[p][np.ix_(il, ir)] = g[p]
        (w * torch.from_numpy(G)).sum().backward()
    E = float(np.mean(w_gradients.last_energies))
    torch.autograd.backward([int1e, int2e], [i1.grad, i2.grad])
    return E, ht.grad.numpy(), et.grad.numpy()


_PERM8 = ((0, 1, 2, 3), (1, 0, 2, 3), (0, 1, 3, 2), (1, 0, 3, 2),
          (2, 3, 0, 1), (3, 2, 0, 1), (2, 3, 1, 0), (3, 2, 1, 0))


def sym8(g):
    """Mean of g over the 8 index permutations that fix (pq|rs) for real orbitals."""
    return sum(g.transpose(pm) for pm in _PERM8) / 8.0


def rdms(mpo, sites, qkey, h, eri, dtype=0):
    """(E, gamma[K,K], Gamma[K,K,K,K]): the spin-summed 1-RDM gamma_pr = <E_pr>
    and the 8-fold symmetrised 2-RDM, so that E = sum h gamma + 1/2 sum (pr|qs)
    Gamma for 8-fold symmetric eri.  Along a perturbation of eri that keeps
    the symmetry, dE/d(pr|qs) summed over the orbit of (pr|qs) is
    1/2 sum_orbit Gamma = |orbit| sym8(dE/deri)."""
    E, gh, ge = integral_gradients(mpo, sites, qkey, h, eri, dtype)
    return E, gh, 2.0 * sym8(ge)


class EnergyFunction:
    """torch.autograd.Function E(h, eri) = <Psi|H(h, eri)|Psi> at a fixed MPS
    (built on first use so that importing this module does not import torch)."""
    _fn = None

    @classmethod
    def apply(cls, h, eri, mpo, sites, qkey):
        if cls._fn is None:
            torch = dv.torch()

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(ctx, h_, eri_, state):
                    ctx.state = state
                    ctx.save_for_backward(h_, eri_)
                    E = expectation(state[0], state[1])
                    return h_.new_tensor(E)

                @staticmethod
                def backward(ctx, gE):
                    h_, eri_ = ctx.saved_tensors
                    mpo_, sites_, qkey_ = ctx.state
                    _, gh, ge = integral_gradients(mpo_, sites_, qkey_, h_.detach().numpy(),
                                                   eri_.detach().numpy())
                    t = dv.torch()
                    return gE * t.from_numpy(gh), gE * t.from_numpy(ge), None

            cls._fn = _F
        return cls._fn.apply(h, eri, (mpo, sites, qkey))


def expectation(mpo, sites):
    """<Psi|H|Psi> = sum_p w_p <H_p> over the left environments (genHopsQt, initQt.py:88-271)."""
    _, exphop = kq.genHops(mpo, sites, 'L')
    return float(np.dot(mpo.hpwts, exphop))
