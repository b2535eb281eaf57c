"""Exact results used to pin the oracle and the GPU path (test infrastructure).

Only tests/, the smoke check and the bench's cpu_baseline leg may import this.

* kaufman_lnZ / kaufman_energy: B. Kaufman, Phys. Rev. 76, 1232 (1949), the
  exact partition function of the 2D Ising model on an m x n torus
  (ferromagnet, E = -sum_<ij> s_i s_j, coupling K = beta):
      Z = 1/2 (2 sinh 2K)^{mn/2} [ P1 + P2 + P3 + P4 ],
      P1 = prod_{r=0}^{n-1} 2 cosh(m g_{2r+1}/2),  P2 = prod 2 sinh(m g_{2r+1}/2),
      P3 = prod_{r=0}^{n-1} 2 cosh(m g_{2r}/2),    P4 = prod 2 sinh(m g_{2r}/2),
      cosh g_l = cosh 2K coth 2K - cos(pi l / n) for l >= 1,
      g_0 = 2K + ln tanh K  (signed: negative above K_c).
  Pinned in tests against brute-force enumeration of small tori, where it must
  agree to round-off.
* enumerate_*: brute-force Boltzmann averages over every state of tiny
  periodic lattices, with the reference's neighbour order and energy
  convention E = J * sum_i sum_{slots} b s_i s_j / 2
  (src/hamiltonian/util/observables.h:173-231).
"""
import itertools

import numpy as np


def kaufman_lnZ(K, m, n):
    """ln Z at the current mpmath precision."""
    import mpmath as mp
    K = mp.mpf(K)
    c2k = mp.cosh(2 * K)
    coth2k = mp.coth(2 * K)

    def gam(l):
        if l == 0:
            return 2 * K + mp.log(mp.tanh(K))
        return mp.acosh(c2k * coth2k - mp.cos(mp.pi * l / n))

    P1 = mp.mpf(1)
    P2 = mp.mpf(1)
    P3 = mp.mpf(1)
    P4 = mp.mpf(1)
    for r in range(n):
        go = gam(2 * r + 1)
        ge = gam(2 * r)
        P1 *= 2 * mp.cosh(m * go / 2)
        P2 *= 2 * mp.sinh(m * go / 2)
        P3 *= 2 * mp.cosh(m * ge / 2)
        P4 *= 2 * mp.sinh(m * ge / 2)
    return mp.log(mp.mpf(1) / 2) + (m * n / mp.mpf(2)) * mp.log(2 * mp.sinh(2 * K)) + mp.log(P1 + P2 + P3 + P4)


def kaufman_energy(K, m, n, dps=60):
    """<E>/N for the ferromagnet E = -sum_<ij> s_i s_j on an m x n torus."""
    import mpmath as mp
    with mp.workdps(dps):
        d = mp.diff(lambda k: kaufman_lnZ(k, m, n), mp.mpf(K))
        return float(-d / (m * n))


def _nbr(dim, L, i, slot):
    d, sign = slot >> 1, (-1 if slot & 1 else 1)
    pw = L ** d
    x = (i // pw) % L
    xn = (x + sign) % L
    return i + (xn - x) * pw


def neighbour_table(dim, L):
    N = L ** dim
    return np.array([[_nbr(dim, L, i, s) for s in range(2 * dim)] for i in range(N)], np.int64)


def bond_table(dim, L, bonds):
    """b for every (site, slot): +x_d slots own bonds[i*dim+d], -x_d slots use the neighbour's."""
    N = L ** dim
    nb = neighbour_table(dim, L)
    B = np.ones((N, 2 * dim), np.int64)
    if bonds is None:
        return B
    bonds = np.asarray(bonds).reshape(N, dim)
    for i in range(N):
        for s in range(2 * dim):
            d = s >> 1
            B[i, s] = bonds[i, d] if (s & 1) == 0 else bonds[nb[i, s], d]
    return B


def enumerate_ising(dim, L, beta, J=-1.0, bonds=None):
    """Exact <E/N>, <m^2>, <m^4>, <|m|> with E = J sum_i sum_slots b s_i s_j / 2."""
    N = L ** dim
    nb = neighbour_table(dim, L)
    B = bond_table(dim, L, bonds)
    states = ((np.arange(2 ** N)[:, None] >> np.arange(N)[None, :]) & 1).astype(np.int64) * 2 - 1
    e = np.zeros(len(states))
    for s in range(2 * dim):
        e += (states * states[:, nb[:, s]] * B[:, s][None, :]).sum(1)
    E = J * e / 2.0
    M = states.sum(1) / N
    w = np.exp(-beta * (E - E.min()))
    Z = w.sum()
    return {"e": float((w * E).sum() / Z / N), "m2": float((w * M ** 2).sum() / Z),
            "m4": float((w * M ** 4).sum() / Z), "mabs": float((w * np.abs(M)).sum() / Z)}


def enumerate_potts(dim, L, Q, beta, k=-1.0):
    """Exact <E/N> (E = k * #equal neighbour pairs) and <m_Potts> (Potts.h:40-65)."""
    N = L ** dim
    nb = neighbour_table(dim, L)
    states = np.array(list(itertools.product(range(Q), repeat=N)), np.int64)
    eq = np.zeros(len(states))
    for s in range(2 * dim):
        eq += (states == states[:, nb[:, s]]).sum(1)
    E = k * eq / 2.0
    counts = np.stack([(states == q).sum(1) for q in range(Q)], 1)
    m = (Q * counts.max(1) - N) / (N * (Q - 1))
    w = np.exp(-beta * (E - E.min()))
    Z = w.sum()
    return {"e": float((w * E).sum() / Z / N), "m": float((w * m).sum() / Z)}


def enumerate_graph(rp, ci, beta, J=-1.0, Q=0, bonds=None):
    """Exact <E/N> and <m^2> (Ising, Q = 0: E = J sum_edges b s_i s_j) or <E/N> and
    <m_Potts> (Q >= 2: E = J * #equal edges) on a small graph given as symmetric CSR;
    each edge appears in both rows and is counted once."""
    rp = np.asarray(rp, np.int64)
    ci = np.asarray(ci, np.int64)
    N = len(rp) - 1
    src = np.repeat(np.arange(N), np.diff(rp))
    keep = src < ci
    a, b = src[keep], ci[keep]
    w_b = np.ones(len(a)) if bonds is None else np.asarray(bonds, np.float64)[keep]
    if Q == 0:
        states = ((np.arange(2 ** N)[:, None] >> np.arange(N)[None, :]) & 1).astype(np.int64) * 2 - 1
        E = J * (states[:, a] * states[:, b] * w_b[None, :]).sum(1)
        M = states.sum(1) / N
        w = np.exp(-beta * (E - E.min()))
        Z = w.sum()
        return {"e": float((w * E).sum() / Z / N), "m2": float((w * M ** 2).sum() / Z)}
    states = np.array(list(itertools.product(range(Q), repeat=N)), np.int64)
    E = J * (states[:, a] == states[:, b]).sum(1)
    counts = np.stack([(states == q).sum(1) for q in range(Q)], 1)
    m = (Q * counts.max(1) - N) / (N * (Q - 1))
    w = np.exp(-beta * (E - E.min()))
    Z = w.sum()
    return {"e": float((w * E).sum() / Z / N), "m": float((w * m).sum() / Z)}
