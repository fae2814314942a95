"""Brute-force Fock-space Hamiltonian and exact FCI (independent check).

Basis: K spatial sites, local state s = 2*n_alpha + n_beta (the order of the
reference's local basis, qtensor_opers.py:104-106), site 0 most significant,
Jordan-Wigner order over spin orbitals 2i (alpha), 2i+1 (beta).
H = sum_pr h_pr E_pr + 1/2 sum_prqs (pr|qs) (E_pr E_qs - delta_qr E_ps)
with spin-summed E_pr.  Test infrastructure only.
"""
import numpy as np
import scipy.sparse as sp


def _annihilators(K):
    n = 2 * K
    dim = 4 ** K
    states = np.arange(dim)
    # occupation bit of spin orbital p: site p//2, alpha (bit 1) or beta (bit 0)
    occ = np.zeros((n, dim), dtype=np.int64)
    for p in range(n):
        site, spin = divmod(p, 2)
        s = (states // 4 ** (K - 1 - site)) % 4
        occ[p] = (s >> 1) & 1 if spin == 0 else s & 1
    ops = []
    for p in range(n):
        site, spin = divmod(p, 2)
        mask = occ[p] == 1
        src = states[mask]
        delta = (2 if spin == 0 else 1) * 4 ** (K - 1 - site)
        dst = src - delta
        sign = (-1.0) ** occ[:p, mask].sum(axis=0)
        ops.append(sp.csr_matrix((sign, (dst, src)), shape=(dim, dim)))
    return ops, occ


def fock_hamiltonian(h, eri):
    """Sparse H (4^K x 4^K) from spatial h[K,K] and chemists' eri[K,K,K,K]."""
    K = h.shape[0]
    a, occ = _annihilators(K)
    E = [[None] * K for _ in range(K)]
    for p in range(K):
        for r in range(K):
            E[p][r] = (a[2 * p].T @ a[2 * r] + a[2 * p + 1].T @ a[2 * r + 1]).tocsr()
    dim = 4 ** K
    H = sp.csr_matrix((dim, dim))
    for p in range(K):
        for r in range(K):
            H = H + h[p, r] * E[p][r]
    for p in range(K):
        for r in range(K):
            for q in range(K):
                for s in range(K):
                    v = eri[p, r, q, s]
                    if v == 0.0:
                        continue
                    t = E[p][r] @ E[q][s]
                    if q == r:
                        t = t - E[p][s]
                    H = H + 0.5 * v * t
    return H.tocsr(), occ


def sector_states(K, nelec, sz):
    dim = 4 ** K
    states = np.arange(dim)
    na = np.zeros(dim, dtype=np.int64)
    nb = np.zeros(dim, dtype=np.int64)
    for site in range(K):
        s = (states // 4 ** (K - 1 - site)) % 4
        na += (s >> 1) & 1
        nb += s & 1
    return np.flatnonzero((na + nb == nelec) & (np.abs(0.5 * (na - nb) - sz) < 1e-12))


def fci_energy(h, eri, nelec, sz=0.0, nroots=1):
    """Lowest eigenvalues of H in the (N, Sz) sector."""
    K = h.shape[0]
    H, _ = fock_hamiltonian(h, eri)
    idx = sector_states(K, nelec, sz)
    Hs = H[idx][:, idx].toarray()
    return np.linalg.eigvalsh(Hs)[:nroots]


def mpo_dense(mpo_sites, K, nops):
    """Sum over ops of the fully contracted MPO, as a 4^K x 4^K matrix."""
    H = np.zeros((4 ** K, 4 ** K))
    for p in range(nops):
        t = mpo_sites[0][p].toDenseTensor()          # [1, wr, u, d]
        t = t[0]                                     # [wr, u, d]
        for i in range(1, K):
            w = mpo_sites[i][p].toDenseTensor()      # [wl, wr, u, d]
            t = np.tensordot(t, w, axes=([0], [0]))  # [U, D, wr, u, d]
            t = t.transpose(2, 0, 3, 1, 4)           # [wr, U, u, D, d]
            s = t.shape
            t = t.reshape(s[0], s[1] * s[2], s[3] * s[4])
        H += t[0]
    return H


def fci_rdms(h, eri, nelec, sz=0.0):
    """Ground state energy and spin-summed RDMs in the (N, Sz) sector:
    gamma_pr = <E_pr>, Gamma_prqs = <E_pr E_qs> - delta_qr <E_ps>, so that
    E = sum h gamma + 1/2 sum (pr|qs) Gamma (the Hamiltonian of fock_hamiltonian).
    The quantities of properties/mpo_dmrg_propsItrf.py:204-263 (<E_pq>) and
    mpo_dmrg_propsMPO.py:73-141 (<a+a+aa>), here from the exact vector."""
    K = h.shape[0]
    H, _ = fock_hamiltonian(h, eri)
    a, _ = _annihilators(K)
    idx = sector_states(K, nelec, sz)
    w, v = np.linalg.eigh(H[idx][:, idx].toarray())
    c = np.zeros(4 ** K)
    c[idx] = v[:, 0]
    V = np.zeros((K, K, 4 ** K))
    for q in range(K):
        for s in range(K):
            Eqs = a[2 * q].T @ a[2 * s] + a[2 * q + 1].T @ a[2 * s + 1]
            V[q, s] = Eqs @ c
    gamma = np.einsum('d,qsd->qs', c, V)
    # <E_pr E_qs> = <E_rp c, E_qs c> for a real vector
    G = np.einsum('rpd,qsd->prqs', V, V)
    G -= np.einsum('qr,ps->prqs', np.eye(K), gamma)
    return float(w[0]), gamma, G


def sector_fci_energy(h, eri, nelec, sz=0.0, nroots=1):
    """Lowest eigenvalues in one (N, Sz) sector, built on the sector's
    determinants only (no 4^K Fock space): E_pr restricted to the sector by
    applying a+_p a_r to every occupation string (Jordan-Wigner signs over
    the spin orbitals in the order 2i (alpha), 2i+1 (beta)).  Same
    Hamiltonian as fock_hamiltonian; usable where 4^K is too large."""
    import itertools
    K = h.shape[0]
    na = int(round(nelec / 2 + sz))
    nb = int(round(nelec / 2 - sz))
    dets = []
    for ao in itertools.combinations(range(K), na):
        for bo in itertools.combinations(range(K), nb):
            occ = 0
            for i in ao:
                occ |= 1 << (2 * i)
            for i in bo:
                occ |= 1 << (2 * i + 1)
            dets.append(occ)
    index = {d: k for k, d in enumerate(dets)}
    n = len(dets)

    def excite(occ, p, r):
        """a+_p a_r |occ> as (sign, new occ) or None (spin orbitals p, r)."""
        if not (occ >> r) & 1:
            return None
        s = (-1) ** bin(occ & ((1 << r) - 1)).count('1')
        occ2 = occ & ~(1 << r)
        if (occ2 >> p) & 1:
            return None
        s *= (-1) ** bin(occ2 & ((1 << p) - 1)).count('1')
        return s, occ2 | (1 << p)

    E = np.zeros((K, K, n, n))
    for col, occ in enumerate(dets):
        for p in range(K):
            for r in range(K):
                for spin in (0, 1):
                    res = excite(occ, 2 * p + spin, 2 * r + spin)
                    if res is not None:
                        E[p, r, index[res[1]], col] += res[0]
    H = np.einsum('pr,prij->ij', h, E)
    EE = np.einsum('prij,qsjk->prqsik', E, E, optimize=True)
    H += 0.5 * np.einsum('prqs,prqsik->ik', eri, EE, optimize=True)
    H -= 0.5 * np.einsum('prrs,psik->ik', eri, E, optimize=True)
    return np.linalg.eigvalsh(H)[:nroots]
