"""Seeded synthetic molecular integrals and the spin-orbital convention.

Generator (fixed draw order, numpy.random.default_rng(seed)):
  h_pp ~ U[-2, -0.5]; h_pq ~ N(0, (0.1 e^{-|p-q|/2})^2), symmetric;
  L^x (x < 2K) symmetric, L^x_pq ~ N(0, (0.3 e^{-|p-q|/3})^2);
  (pq|rs) = sum_x L^x_pq L^x_rs  (positive semidefinite, 8-fold symmetric).

Spin-orbital mapping follows itools/fromFCIDUMP/tools_itrf.py:22-35 of the
reference: h on the alpha-alpha / beta-beta blocks; the two-electron tensor
is placed in all four spin blocks, transposed to <pq|rs> = [pr|qs] and
antisymmetrised as V[pqrs] = -1/2 (<pq|rs> - <pq|sr>).
"""
import numpy as np


def gen_spatial(K, seed=0):
    """Return (h[K,K], eri[K,K,K,K]) in chemists' notation."""
    rng = np.random.default_rng(seed)
    idx = np.arange(K)
    dist = np.abs(idx[:, None] - idx[None, :])
    diag = rng.uniform(-2.0, -0.5, size=K)
    off = rng.normal(size=(K, K)) * (0.1 * np.exp(-dist / 2.0))
    h = np.triu(off, 1)
    h = h + h.T + np.diag(diag)
    Lx = rng.normal(size=(2 * K, K, K)) * (0.3 * np.exp(-dist / 3.0))[None]
    Lx = np.triu(Lx) + np.transpose(np.triu(Lx, 1), (0, 2, 1))
    eri = np.einsum('xpq,xrs->pqrs', Lx, Lx, optimize=True)
    return h, eri


def spin_orbital(h, eri):
    """(int1e[2K,2K], int2e[2K,2K,2K,2K]) in the reference's MPO convention."""
    K = h.shape[0]
    n = 2 * K
    h1 = np.zeros((n, n))
    h1[0::2, 0::2] = h
    h1[1::2, 1::2] = h
    g = np.zeros((n, n, n, n))
    g[0::2, 0::2, 0::2, 0::2] = eri
    g[1::2, 1::2, 1::2, 1::2] = eri
    g[0::2, 0::2, 1::2, 1::2] = eri
    g[1::2, 1::2, 0::2, 0::2] = eri
    g = g.transpose(0, 2, 1, 3)
    g = -0.5 * (g - g.transpose(0, 1, 3, 2))
    return h1, np.ascontiguousarray(g)


def gen_integrals(K, seed=0):
    """Spin-orbital integrals of the synthetic generator."""
    return spin_orbital(*gen_spatial(K, seed))
