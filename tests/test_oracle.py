"""Pins the oracle itself (CPU only).

- Philox4x32-10 against the Random123 known-answer vectors and against torch's
  independent implementation (ATen/core/PhiloxRNGEngine.h) via a committed fixture.
- std::mt19937_64 against the value the C++ standard mandates.
- The reference-faithful CPU restatement against the reference's own compiled
  Metropolis move (oracle/_ref), bit for bit.
- The CPU replay of the GPU schedule and the restatement against exact
  enumeration and Kaufman's exact torus energy, with planted mistakes showing
  that the checks have the power to catch a wrong sign or a dropped term.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_bind as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import exact  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")

# Random123 1.09 kat_vectors, philox4x32 10 rounds
R123_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,out", R123_KAT)
def test_philox_random123_kat(ctr, key, out):
    assert tuple(ob.philox(ctr, key)) == out


def test_philox_matches_torch_engine_fixture():
    with open(os.path.join(GOLDEN, "philox_torch.json")) as f:
        fx = json.load(f)
    assert len(fx["cases"]) >= 50
    for c in fx["cases"]:
        assert ob.philox(c["ctr"], c["key"]) == c["out"], c


def test_mt19937_64_standard_value():
    # C++ [rand.predef]: the 10000th invocation of a default-constructed
    # mt19937_64 (seed 5489) produces 9981545732273789042.
    assert ob.lib().ref_mt19937_64_nth(5489, 10000) == 9981545732273789042


def test_rngcache_mapping():
    """real() = (x >> 11) * 2^-53 and integer(r) = x % r (rngcache.h:179,204)."""
    n = 3000
    raw = np.zeros(n, np.uint64)
    ob.lib().ref_rngcache_draws(0, 0, 0, n, ob.p(raw))
    reals = np.zeros(n, np.uint64)
    ob.lib().ref_rngcache_draws(0, 1, 0, n, ob.p(reals))
    ints = np.zeros(n, np.uint64)
    ob.lib().ref_rngcache_draws(0, 2, 7, n, ob.p(ints))
    r = reals.view(np.float64)
    assert np.array_equal(r, (raw >> np.uint64(11)).astype(np.float64) * 2.0 ** -53)
    assert np.array_equal(ints, raw % np.uint64(7))
    # the cache refills every 1024 draws from the same engine stream
    assert ob.lib().ref_mt19937_64_nth(0, 1025) == int(raw[1024])


def _ref_metropolis(model, dim, L, beta, J, nsweeps, seed):
    out = subprocess.run([ob.REF_METRO, model, str(dim), str(L), str(beta), str(J), str(nsweeps), str(seed)],
                         capture_output=True, text=True, check=True).stdout
    return np.array([[float(x) for x in ln.split()] for ln in out.strip().split("\n")])


@pytest.mark.skipif(not os.path.exists(ob.REF_METRO), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("dim,L,beta,J,ns", [(3, 4, 0.8, -1.0, 20), (3, 6, 1.2, -1.0, 20), (2, 8, 0.9, 1.0, 20)])
def test_ea_restatement_bit_exact_vs_reference_build(dim, L, beta, J, ns):
    """+-J EA on the reference's own RegularRandomBond<BimodalPDF> lattice (bonds drawn and
    symmetrised there): its bonds are symmetric in the (+d, -d) slot order the restatement
    stores one bond per (site, +d) for, and the restatement run on those bonds equals the
    reference's Metropolis move bit for bit."""
    out = subprocess.run([ob.REF_METRO, "ea", str(dim), str(L), str(beta), str(J), str(ns), "12345"],
                         capture_output=True, text=True, check=True).stdout
    lines = out.strip().split("\n")
    N = L ** dim
    bonds = np.array([[int(x) for x in ln.split()] for ln in lines[:N]], np.int8)
    spins = np.array([int(x) for x in lines[N:]])
    nb = exact.neighbour_table(dim, L)
    for d in range(dim):
        assert np.array_equal(bonds[:, 2 * d + 1], bonds[nb[:, 2 * d + 1], 2 * d])
    assert 0.3 < (bonds == 1).mean() < 0.7
    mine, _, _ = ob.ref_run(ob.EA, dim, L, beta, J=J, nmetro=1, nwarm=ns, seed=12345,
                            bonds=np.ascontiguousarray(bonds[:, 0::2]).reshape(-1))
    assert np.array_equal(spins, mine)


@pytest.mark.skipif(not os.path.exists(ob.REF_METRO), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("dim,L,beta,lam,mass,ns", [(3, 6, 0.69, 4.5, 1.0, 20), (2, 8, 1.0, 1.0, -0.5, 20),
                                                    (3, 4, 0.5, 0.3, 0.2, 30)])
def test_phi4_restatement_bit_exact_vs_reference_build(dim, L, beta, lam, mass, ns):
    """Phi4 (config/Phi4.ini parameters and two others, one with a negative mass term):
    the reference's Metropolis move with its on-site energy assembly and the radial
    proposal, against the restatement, bit for bit; the chain moved."""
    out = subprocess.run([ob.REF_METRO, "phi4", str(dim), str(L), str(beta), str(lam), str(mass), str(ns), "12345"],
                         capture_output=True, text=True, check=True).stdout
    ref = np.array([[float(x) for x in ln.split()] for ln in out.strip().split("\n")])
    mine, _, _ = ob.ref_run(ob.PHI4, dim, L, beta, J=-1.0, lam=lam, mass=mass, nmetro=1, nwarm=ns, seed=12345)
    start, _, _ = ob.ref_run(ob.PHI4, dim, L, beta, J=-1.0, lam=lam, mass=mass, nmetro=1, nwarm=0, seed=12345)
    assert np.array_equal(ref, mine)
    assert (np.abs(mine - start).max(axis=1) > 0).mean() > 0.5


@pytest.mark.skipif(not os.path.exists(ob.REF_METRO), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("model,dim,L,beta,J,ns", [
    ("ising", 2, 16, 0.4406868, -1.0, 40), ("ising", 3, 8, 0.2216546, -1.0, 40),
    ("ising", 2, 10, 0.3, 1.0, 20), ("heis", 3, 6, 0.693, -1.0, 20), ("heis", 2, 8, 1.5, -1.0, 20)])
def test_restatement_bit_exact_vs_reference_build(model, dim, L, beta, J, ns):
    ref = _ref_metropolis(model, dim, L, beta, J, ns, 12345)
    m = ob.ISING if model == "ising" else ob.HEIS
    mine, _, _ = ob.ref_run(m, dim, L, beta, J=J, nmetro=1, nwarm=ns, seed=12345)
    if model == "ising":
        assert np.array_equal(ref[:, 0].astype(int), mine)
    else:
        assert np.array_equal(ref, mine)


def test_kaufman_matches_enumeration():
    for L in (2, 4):
        for b in (0.1, 0.4406868, 0.9):
            assert exact.kaufman_energy(b, L, L) == pytest.approx(exact.enumerate_ising(2, L, b)["e"], abs=1e-12)


def _mean_err(x, nblocks=20):
    x = np.asarray(x, np.float64)
    n = len(x) // nblocks * nblocks
    b = x[:n].reshape(nblocks, -1).mean(1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nblocks)


def _replay_chain(dim, L, beta, J, nsteps, seed, bonds=None, wolff=0):
    """CPU replay of the GPU schedule: returns per-step e and m."""
    N = L ** dim
    s = ob.init_ising(N, seed, 0)
    e, m = [], []
    nb = exact.neighbour_table(dim, L)
    B = exact.bond_table(dim, L, bonds)
    for t in range(nsteps):
        if wolff:
            s, _ = ob.wolff_ising(dim, L, s, beta, J, wolff, seed, t * wolff, 0)
        s = ob.ising_sweeps(dim, L, s, beta, J, 1, seed, t, 0, bonds=bonds)
        si = s.astype(np.int64)
        e.append(J * (si[:, None] * si[nb] * B).sum() / 2.0 / N)
        m.append(si.sum() / N)
    return np.array(e), np.array(m)


@pytest.mark.parametrize("dim,L,beta,wolff", [(2, 4, 0.4406868, 0), (2, 4, 0.3, 2), (3, 2, 0.5, 0)])
def test_replay_metropolis_matches_enumeration(dim, L, beta, wolff):
    ex = exact.enumerate_ising(dim, L, beta)
    e, m = _replay_chain(dim, L, beta, -1.0, 20000, seed=3, wolff=wolff)
    me, se = _mean_err(e[500:])
    mm, sm = _mean_err(m[500:] ** 2)
    assert abs(me - ex["e"]) < 4 * se + 1e-9, (me, se, ex["e"])
    assert abs(mm - ex["m2"]) < 4 * sm + 1e-9, (mm, sm, ex["m2"])


def test_replay_ea_matches_enumeration():
    dim, L, beta = 3, 2, 0.8
    bonds = ob.init_bonds(dim, L ** dim, 0.5, 44, 0)
    ex = exact.enumerate_ising(dim, L, beta, bonds=bonds)
    e, m = _replay_chain(dim, L, beta, -1.0, 20000, seed=5, bonds=bonds)
    me, se = _mean_err(e[500:])
    assert abs(me - ex["e"]) < 4 * se + 1e-9, (me, se, ex["e"])


def test_planted_sign_error_is_caught():
    """A chain run with the coupling sign flipped must fail the same check."""
    dim, L, beta = 2, 4, 0.4406868
    ex = exact.enumerate_ising(dim, L, beta, J=-1.0)
    e, _ = _replay_chain(dim, L, beta, +1.0, 4000, seed=3)
    e = -e  # energies measured with the ferromagnetic sign
    me, se = _mean_err(e[200:])
    assert abs(me - ex["e"]) > 10 * se


def _py_neighbour(dim, L, i, sl):
    """Neighbour over slot sl (2d: +e_d, 2d+1: -e_d) of site i = sum_j x_j L^j, periodic."""
    d = sl >> 1
    pw = L ** d
    x = (i // pw) % L
    return i + ((1 if x + 1 < L else 1 - L) if sl % 2 == 0 else (-1 if x > 0 else L - 1)) * pw


@pytest.mark.parametrize("dim,L,beta", [(2, 8, 0.4), (3, 4, 0.25)])
def test_wolff_cluster_is_reachable_set_of_passing_tests(dim, L, beta):
    """The oracle's Wolff cluster (grown depth-first from a stack) is the set of sites reachable
    from the seed along passing tests, found here breadth-first from nothing but raw Philox words:
    the test from site i to its neighbour over slot k uses word e & 3 of counter
    (e >> 2, cluster, replica, 4), e = i*2*dim + k, and passes when the word is below the add
    threshold and the neighbour holds the seed's value (one uniform per test, as the reference
    draws rng.real() per candidate, Ising.h:146-158).  Keying each test by its undirected bond
    instead (a plausible mistake) changes the clusters."""
    N = L ** dim
    seed, rep = 77, 3
    key = (seed & 0xFFFFFFFF, seed >> 32)
    tw = ob.lib().orc_wolff_threshold(-np.expm1(2.0 * beta * -1.0))

    def word(e, ncl):
        return ob.philox(((e >> 2) & 0xFFFFFFFF, ncl, rep, 4), key)[e & 3]

    def reach(s, s0, uniform):
        val, seen, front = s[s0], {s0}, [s0]
        while front:
            nxt = []
            for i in front:
                for sl in range(2 * dim):
                    j = _py_neighbour(dim, L, i, sl)
                    if j not in seen and s[j] == val and uniform(i, sl, j) < tw:
                        seen.add(j)
                        nxt.append(j)
            front = nxt
        return seen

    s = ob.init_ising(N, 5, rep)
    differs, sizes_seen = 0, []
    for ncl in range(8):
        s0 = (ob.philox((0, ncl, rep, 3), key)[0] * N) >> 32
        per_test = reach(s, s0, lambda i, sl, j: word(i * 2 * dim + sl, ncl))
        out, sizes = ob.wolff_ising(dim, L, s, beta, -1.0, 1, seed, ncl, rep)
        assert set(np.nonzero(out != s)[0].tolist()) == per_test and sizes[0] == len(per_test), ncl
        per_bond = reach(s, s0, lambda i, sl, j: word((i if sl % 2 == 0 else j) * dim + (sl >> 1), ncl))
        differs += per_bond != per_test
        sizes_seen.append(len(per_test))
        s = out
    assert max(sizes_seen) > 4 and differs >= 1, (sizes_seen, differs)


def test_replay_potts_matches_enumeration():
    dim, L, Q, beta = 3, 2, 3, 0.7
    N = L ** dim
    ex = exact.enumerate_potts(dim, L, Q, beta)
    nb = exact.neighbour_table(dim, L)
    s = ob.init_potts(N, Q, 2, 0)
    es = []
    for t in range(20000):
        if t % 3 == 0:
            s, _ = ob.wolff_potts(dim, L, s, Q, beta, -1.0, 1, 2, t, 0)
        s = ob.potts_sweeps(dim, L, s, Q, beta, -1.0, 1, 2, t, 0)
        es.append(-1.0 * (s[:, None] == s[nb]).sum() / 2.0 / N)
    me, se = _mean_err(np.array(es[500:]))
    assert abs(me - ex["e"]) < 4 * se, (me, se, ex["e"])


@pytest.mark.parametrize("model", [ob.ISING, ob.POTTS])
def test_restatement_matches_enumeration(model):
    dim, L, beta = (2, 4, 0.4406868) if model == ob.ISING else (3, 2, 0.7)
    ex = exact.enumerate_ising(dim, L, beta) if model == ob.ISING else exact.enumerate_potts(dim, L, 3, beta)
    _, obs, _ = ob.ref_run(model, dim, L, beta, Q=3, ncluster=1, nmetro=1, nwarm=200, nmeas=40000, seed=9)
    me, se = _mean_err(obs[:, 0])
    assert abs(me - ex["e"]) < 4 * se, (me, se, ex["e"])


def test_restatement_c1_energy_vs_kaufman():
    """2D Ising L=64 at beta_c, 10 Wolff + 1 sweep per EMCS, vs the exact torus energy."""
    L, beta = 64, 0.4406868
    _, obs, cl = ob.ref_run(ob.ISING, 2, L, beta, ncluster=10, nmetro=1, nwarm=300, nmeas=3000, seed=1)
    me, se = _mean_err(obs[:, 0])
    ka = exact.kaufman_energy(beta, L, L)
    assert abs(me - ka) < 3 * se, (me, se, ka)
    assert 1000 < cl.mean() < 2200


@pytest.mark.parametrize("kind", ["int8", "msc"])
@pytest.mark.parametrize("J,beta", [(-1.0, 0.3), (1.0, 0.25)])
def test_replay_acceptance_per_class(kind, J, beta):
    """One sweep from a random 3D configuration: colour-0 sites see their initial
    neighbours, so each one's flip is one Bernoulli draw with p = min(1, exp(-beta dE)),
    dE = -2 J s_i h_i.  Per class, the flip fraction must match p (binomial, 4 sigma):
    this pins the 16-bit/tie rule and the bit-serial rule, bit order included."""
    L = 64
    N = L ** 3
    s0 = ob.init_ising(N, 17, 0)
    if kind == "int8":
        s1 = ob.ising_sweeps(3, L, s0, beta, J, 1, 99, 0, 0)
    else:
        s1 = ob.ising_msc_sweeps(L, s0, beta, J, 1, 99, 0, 0)
    nb = exact.neighbour_table(3, L)
    idx = np.arange(N)
    x = np.stack([(idx // L ** d) % L for d in range(3)], 1)
    c0 = (x.sum(1) % 2) == 0
    q = s0.astype(np.int64) * s0[nb].astype(np.int64).sum(1)
    flipped = s1 != s0
    for qv in range(-6, 7, 2):
        sel = c0 & (q == qv)
        n = int(sel.sum())
        if n < 500:
            continue
        p = min(1.0, np.exp(-beta * (-2.0 * J * qv)))
        f = flipped[sel].mean()
        assert abs(f - p) <= 4 * np.sqrt(max(p * (1 - p), 1e-12) / n) + 1e-12, (kind, qv, f, p, n)


def test_ea_bits_replay_acceptance_per_class():
    """The 32-replica word rule: per (replica temperature, class) the flip fraction of
    colour-0 sites in one sweep matches min(1, exp(-beta dE))."""
    dim, L = 3, 16
    N = L ** dim
    betas = 0.5 * (3.3 / 0.5) ** (np.arange(32) / 31)
    bonds = ob.init_bonds(dim, N, 0.5, 4, 0)
    s0 = np.stack([ob.init_ising(N, 3, j) for j in range(32)])
    s1 = ob.ea_bits_sweeps(dim, L, s0, bonds, betas, -1.0, 1, 77, 0, 0)
    nb = exact.neighbour_table(dim, L)
    B = exact.bond_table(dim, L, bonds)
    idx = np.arange(N)
    x = np.stack([(idx // L ** d) % L for d in range(3)], 1)
    c0 = (x.sum(1) % 2) == 0
    tot = {}
    for j in (0, 10, 31):
        q = s0[j].astype(np.int64) * (B * s0[j][nb]).sum(1)
        fl = s1[j] != s0[j]
        for qv in (2, 4, 6):
            sel = c0 & (q == qv)
            n = int(sel.sum())
            if n < 100:
                continue
            p = np.exp(-betas[j] * 2.0 * qv)
            f = fl[sel].mean()
            assert abs(f - p) <= 4 * np.sqrt(max(p * (1 - p), 1e-12) / n) + 1e-12, (j, qv, f, p, n)
            tot[(j, qv)] = n
        sel = c0 & (q <= 0)
        assert fl[sel].all()
    assert len(tot) >= 5


# ---------------------------------------------------------------- graph replays
_SMALL_EDGES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 4), (2, 5), (3, 4), (3, 6), (4, 5), (4, 7), (5, 8), (6, 7),
                (7, 8), (8, 9), (9, 0), (2, 9), (6, 9)]


def _small_graph():
    """10 vertices, degrees 2..5, odd cycles: a small irregular graph as symmetric CSR
    with the greedy colouring in vertex order."""
    n = 10
    adj = [[] for _ in range(n)]
    for a, b in _SMALL_EDGES:
        adj[a].append(b)
        adj[b].append(a)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(x) for x in adj])
    ci = np.array([v for x in adj for v in sorted(x)], np.int32)
    col = np.full(n, -1, np.int32)
    for u in range(n):
        used = {col[v] for v in adj[u]}
        c = 0
        while c in used:
            c += 1
        col[u] = c
    return rp, ci, col


def _graph_chain(rp, ci, col, beta, J, nsteps, seed, rep, potts, bonds=None, wolff=0):
    n = len(rp) - 1
    src = np.repeat(np.arange(n), np.diff(rp))
    keep = src < ci
    a, b = src[keep], ci[keep]
    wb = np.ones(len(a)) if bonds is None else bonds[keep].astype(np.float64)
    s = ob.init_potts(n, 3, seed, rep) if potts else ob.init_ising(n, seed, rep)
    es, ms = [], []
    for t in range(nsteps):
        if wolff:
            s, _ = ob.csr_wolff(rp, ci, s, beta, J, wolff, seed, t * wolff, rep, potts=potts)
        s = ob.csr_sweeps(rp, ci, col, s, beta, J, 1, seed, t, rep, potts=potts, bonds=bonds)
        si = s.astype(np.int64)
        if potts:
            es.append(J * (si[a] == si[b]).sum() / n)
            cnt = np.bincount(si, minlength=3)
            ms.append((3 * cnt.max() - n) / (n * 2))
        else:
            es.append(J * (si[a] * si[b] * wb).sum() / n)
            ms.append((si.sum() / n) ** 2)
    return np.array(es), np.array(ms)


@pytest.mark.parametrize("potts,beta,rep,wolff,ea", [(False, 0.5, 5, 0, False), (False, 0.4, 13, 1, False),
                                                     (False, 0.8, 3, 0, True), (True, 1.0, 6, 0, False),
                                                     (True, 0.8, 7, 1, False)])
def test_graph_replay_matches_enumeration(potts, beta, rep, wolff, ea):
    """Colour-class sweeps (and Wolff clusters keyed by the tested edges' slots) on an
    irregular 10-vertex graph sample the exact Boltzmann distribution: <E/N> and
    <m^2> (Ising) or <m> (Potts) within 4 sigma of exhaustive enumeration.  The
    replica numbers exercise every word / half-word of the shared Philox blocks."""
    rp, ci, col = _small_graph()
    bonds = None
    if ea:
        src = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
        lo, hi = np.minimum(src, ci), np.maximum(src, ci)
        bonds = np.where((lo * 7 + hi * 3) % 4 == 0, -1, 1).astype(np.int8)
    ex = exact.enumerate_graph(rp, ci, beta, -1.0, Q=3 if potts else 0, bonds=bonds)
    e, m = _graph_chain(rp, ci, col, beta, -1.0, 20000, 11, rep, potts, bonds=bonds, wolff=wolff)
    me, se = _mean_err(e[500:])
    mm, sm = _mean_err(m[500:])
    assert abs(me - ex["e"]) < 4 * se + 1e-9, (me, se, ex["e"])
    key = "m" if potts else "m2"
    assert abs(mm - ex[key]) < 4 * sm + 1e-9, (mm, sm, ex[key])


def test_graph_replay_planted_error_is_caught():
    """A Potts chain with the coupling sign flipped fails the same check by > 10 sigma."""
    rp, ci, col = _small_graph()
    ex = exact.enumerate_graph(rp, ci, 1.0, -1.0, Q=3)
    e, _ = _graph_chain(rp, ci, col, 1.0, +1.0, 4000, 11, 6, True)
    me, se = _mean_err(-e[200:])  # energies of the sign-flipped chain, measured with J = -1
    assert abs(me - ex["e"]) > 10 * se, (me, se, ex["e"])


def test_magft_restatement_vs_fft_and_stripe():
    """oracle/observables.py's MagFTComp against numpy's FFT of the axis profile (an
    independent algorithm) and the half-and-half stripe's closed form."""
    import observables
    rng = np.random.default_rng(3)
    for dim, L in [(2, 6), (3, 4), (2, 16)]:
        N = L ** dim
        shape = (L,) * dim  # numpy order: axis 0 slowest, so natural axis d is numpy axis dim-1-d
        for ncomp in (1, 3):
            s = rng.normal(size=(N, ncomp)) if ncomp == 3 else np.where(rng.random((N, 1)) < 0.5, 1.0, -1.0)
            for d in range(dim):
                prof = s.reshape(shape + (ncomp,)).sum(axis=tuple(a for a in range(dim) if a != dim - 1 - d))
                f = np.fft.fft(prof, axis=0)[1] / N
                assert observables.magft(s, dim, L, d) == pytest.approx(float(np.sum(np.abs(f) ** 2)), rel=1e-12, abs=1e-15)
    for dim, L in [(2, 8), (3, 16)]:
        N = L ** dim
        x0 = np.arange(N) % L
        stripe = np.where(x0 < L // 2, 1.0, -1.0)
        assert observables.magft(stripe, dim, L, 0) == pytest.approx((2.0 / (L * np.sin(np.pi / L))) ** 2, rel=1e-12)
        assert observables.magft(stripe, dim, L, 1) == pytest.approx(0.0, abs=1e-28)
        assert observables.magft(np.ones(N), dim, L, dim - 1) == pytest.approx(0.0, abs=1e-28)


# ---------------------------------------------------------------------------
# Wolff updates of continuous spins pinned by the Boltzmann distribution.
# No closed form exists for Heisenberg or Phi4, and the reference's generic Wolff
# move and Embedder<Phi4> (wolff.h:48-125, Phi4.h:182-244) cannot be built here
# (they need core.h, which needs HDF5).  What fixes them is that a Wolff move
# must leave the Boltzmann distribution invariant: chains that add Wolff clusters
# to Metropolis sweeps must reproduce the averages of Metropolis-only chains,
# whose move equals the reference build's bit for bit
# (test_phi4_restatement_bit_exact_vs_reference_build).

PHI4_INI = dict(beta=0.69, lam=4.5, mass=1.0)  # config/Phi4.ini (beta from its list)


def _phi4_chains(ncluster, seed0, nmeas=3000):
    return ob.ref_chains(ob.PHI4, 3, 6, PHI4_INI["beta"], 16, seed0=seed0, ncluster=ncluster, nmetro=5, nwarm=200,
                         nmeas=nmeas, J=-1.0, lam=PHI4_INI["lam"], mass=PHI4_INI["mass"])


def _zscores(a, b, cols):
    import stat_util as su
    out = {}
    for k, nm in cols:
        g, sg = su.mean_err(su.per_chain(a[:, :, k]))
        c, sc = su.mean_err(su.per_chain(b[:, :, k]))
        out[nm] = (g - c) / np.hypot(sg, sc)
    return out


def test_phi4_wolff_restatement_vs_metropolis_only():
    """3D L = 6 Phi4 at the config/Phi4.ini parameters: 16 chains of EMCS = 3 random-direction
    Wolff clusters + 5 sweeps against 16 Metropolis-only chains; <E/N>, <|m|> and <|phi|^2>
    agree within 4 standard errors.  With the planted mistake (phi renormalised to unit length
    after each reflection, which Phi4.h:241-242 does not do) <E/N> and <|phi|^2> are off by
    more than 8."""
    cols = [(0, "E/N"), (1, "|m|"), (4, "|phi|^2")]
    metro = _phi4_chains(0, 300)
    ob.lib().ref_set_plant(0)
    z = _zscores(_phi4_chains(3, 200), metro, cols)
    print("[phi4 wolff vs metropolis-only]", {k: round(v, 2) for k, v in z.items()}, flush=True)
    assert all(abs(v) < 4.0 for v in z.values()), z
    ob.lib().ref_set_plant(1)
    try:
        zp = _zscores(_phi4_chains(3, 200), metro, cols)
    finally:
        ob.lib().ref_set_plant(0)
    print("[phi4 wolff vs metropolis-only, planted renormalisation]", {k: round(v, 2) for k, v in zp.items()},
          flush=True)
    assert abs(zp["E/N"]) > 8.0 and abs(zp["|phi|^2"]) > 8.0, zp


@pytest.mark.parametrize("model,small,delta,rdir", [("heisenberg", True, 0.6, True),
                                                    ("heisenberg", False, 0.5, False),
                                                    ("phi4", False, 0.5, True)])
def test_csr_o3_oracle_chains_vs_restatement(model, small, delta, rdir):
    """The graph path's FP64 restatements (orc_csr_o3_colour, orc_csr_o3_wolff), driven as
    Markov chains on the 3D L = 4 torus given as a CSR graph, against Metropolis-only chains of
    the reference-faithful restatement on the regular lattice (bit-exact to the reference build):
    <E/N>, <|m|> (and <|phi|^2> for Phi4) within 4 standard errors.  Heisenberg at beta = 0.693
    with small-angle or fresh-direction proposals and random-direction or Cartesian clusters;
    Phi4 at the config/Phi4.ini parameters."""
    import o3_chains
    from marqov_amd import graphs
    phi4 = model == "phi4"
    beta = PHI4_INI["beta"] if phi4 else 0.693
    rp, ci = graphs.regular_csr(3, 4)
    orc = o3_chains.oracle_chains(rp, ci, phi4, beta, 16, 17, 100, 1200, 2, 3, small=small, delta=delta, rdir=rdir)
    ref = ob.ref_chains(ob.PHI4 if phi4 else ob.HEIS, 3, 4, beta, 16, seed0=400, ncluster=0, nmetro=3, nwarm=300,
                        nmeas=4000, J=-1.0, lam=PHI4_INI["lam"], mass=PHI4_INI["mass"])
    cols = [(0, "E/N"), (1, "|m|")] + ([(4, "|phi|^2")] if phi4 else [])
    z = _zscores(orc, ref, cols)
    print(f"[csr o3 oracle chains] {model} small={small} rdir={rdir}:", {k: round(v, 2) for k, v in z.items()},
          flush=True)
    assert all(abs(v) < 4.0 for v in z.values()), z
