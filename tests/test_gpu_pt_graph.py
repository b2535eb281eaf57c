"""Parallel tempering and HIP-graph replay, bit-exact against a CPU replay."""
import numpy as np
import pytest

import oracle_bind as ob
import exact_path  # noqa: F401  (puts oracle/ on sys.path)
import exact

pytestmark = pytest.mark.gpu

DIM, L = 3, 4
N = L ** DIM
NT = 8          # temperatures per sample
NS = 3          # samples
BETAS = 0.5 * (3.3 / 0.5) ** (np.arange(NT) / (NT - 1))


def _ea_state(ctx):
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, DIM, L)
    st = gpu.State(ctx, "ea", [lat] * (NS * NT), J=-1.0)
    st._lat = lat
    st.set_beta(np.tile(BETAS, NS))
    st.random_bonds(rep_sample=np.repeat(np.arange(NS), NT), p_plus=0.5, seed=4)
    st.init(seed=1)
    return st


class Replay:
    """CPU replay of sweeps + even/odd PT swaps (marqovscheduler.h:44-57,389-397)."""

    def __init__(self):
        self.s = [ob.init_ising(N, 1, r) for r in range(NS * NT)]
        self.bonds = [ob.init_bonds(DIM, N, 0.5, 4, k) for k in range(NS)]
        self.rot = np.arange(NS * NT)          # slot -> replica
        self.tor = np.tile(np.arange(NT), NS)  # replica -> temperature index
        self.nb = exact.neighbour_table(DIM, L)
        self.B = [exact.bond_table(DIM, L, b) for b in self.bonds]

    def sweeps(self, n, seed, sweep0):
        for r in range(NS * NT):
            self.s[r] = ob.ising_sweeps(DIM, L, self.s[r], BETAS[self.tor[r]], -1.0, n, seed, sweep0, r,
                                        bonds=self.bonds[r // NT])

    def energy(self, r):
        si = self.s[r].astype(np.int64)
        return -1.0 * (si[:, None] * si[self.nb] * self.B[r // NT]).sum() / 2.0

    def swap(self, parity, seed, step):
        acc = np.zeros(NS, np.int32)
        for g in range(NS):
            for k in range(parity, NT - 1, 2):
                sa, sb = g * NT + k, g * NT + k + 1
                a, b = self.rot[sa], self.rot[sb]
                en = (BETAS[k] - BETAS[k + 1]) * (self.energy(a) - self.energy(b))
                ratio = 0.0 if en <= np.log(np.finfo(float).tiny) else (
                    np.finfo(float).max if en > np.log(np.finfo(float).max) else np.exp(en))
                w = ob.philox([k, step & 0xffffffff, g, 5], [seed & 0xffffffff, seed >> 32])
                u = (((w[0] << 32) | w[1]) >> 11) * 2.0 ** -53
                if u < min(1.0, ratio):
                    self.rot[sa], self.rot[sb] = b, a
                    self.tor[a], self.tor[b] = k + 1, k
                    acc[g] += 1
        return acc

    def swap_pairs(self, pairs, seed, step):
        """One step of a PT plan (createPTplan/ptstep, marqovscheduler.h:202-211,374-413):
        the disjoint slot pairs, lower slot first, offered in every group."""
        acc = np.zeros(NS, np.int32)
        for g in range(NS):
            for ka, kb in pairs:
                if ka == kb:
                    continue
                ka, kb = min(ka, kb), max(ka, kb)
                sa, sb = g * NT + ka, g * NT + kb
                a, b = self.rot[sa], self.rot[sb]
                en = (BETAS[ka] - BETAS[kb]) * (self.energy(a) - self.energy(b))
                ratio = 0.0 if en <= np.log(np.finfo(float).tiny) else (
                    np.finfo(float).max if en > np.log(np.finfo(float).max) else np.exp(en))
                w = ob.philox([ka, step & 0xffffffff, g, 5], [seed & 0xffffffff, seed >> 32])
                u = (((w[0] << 32) | w[1]) >> 11) * 2.0 ** -53
                if u < min(1.0, ratio):
                    self.rot[sa], self.rot[sb] = b, a
                    self.tor[a], self.tor[b] = kb, ka
                    acc[g] += 1
        return acc


def test_pt_pair_plan_bit_exact(ctx):
    """Random plans of disjoint pairs (neighbours and far temperatures, either order, a
    self pair) swap exactly as the CPU replay decides."""
    st = _ea_state(ctx)
    rp = Replay()
    rng = np.random.default_rng(12)
    for step in range(8):
        st.metropolis(2, seed=7, sweep0=2 * step)
        rp.sweeps(2, 7, 2 * step)
        perm = rng.permutation(NT)
        pairs = [(int(perm[2 * i]), int(perm[2 * i + 1])) for i in range(int(rng.integers(1, NT // 2 + 1)))]
        if step == 3:
            pairs.append((int(perm[-1]), int(perm[-1])))  # a self pair offers nothing
        acc = st.pt_swap_pairs(NT, pairs, seed=9, step=step)
        eacc = rp.swap_pairs(pairs, 9, step)
        assert np.array_equal(acc, eacc), (step, pairs, acc, eacc)
        assert np.array_equal(st.pt_temps(), rp.tor), step
    for r in range(NS * NT):
        assert np.array_equal(st.download(r), rp.s[r]), r


def test_pt_pairs_neighbours_equal_even_odd(ctx):
    """Neighbour pairs (k, k+1) of one parity as a plan reproduce mq_pt_swap exactly."""
    a, b = _ea_state(ctx), _ea_state(ctx)
    for step in range(6):
        for st in (a, b):
            st.metropolis(2, seed=7, sweep0=2 * step)
        par = step & 1
        acc_a = a.pt_swap(NT, par, seed=9, step=step)
        acc_b = b.pt_swap_pairs(NT, [(k + 1, k) for k in range(par, NT - 1, 2)], seed=9, step=step)
        assert np.array_equal(acc_a, acc_b) and np.array_equal(a.pt_temps(), b.pt_temps()), step


def test_pt_pair_plan_rejects_bad_plans(ctx):
    from marqov_amd.gpu import MarqovError
    st = _ea_state(ctx)
    with pytest.raises(MarqovError, match="disjoint"):
        st.pt_swap_pairs(NT, [(0, 1), (1, 2)], seed=1, step=0)
    with pytest.raises(MarqovError, match="outside"):
        st.pt_swap_pairs(NT, [(0, NT)], seed=1, step=0)


def test_pt_swaps_bit_exact(ctx):
    st = _ea_state(ctx)
    rp = Replay()
    for step in range(6):
        st.metropolis(2, seed=7, sweep0=2 * step)
        rp.sweeps(2, 7, 2 * step)
        acc = st.pt_swap(NT, step & 1, seed=9, step=step)
        eacc = rp.swap(step & 1, 9, step)
        assert np.array_equal(acc, eacc), (step, acc, eacc)
        assert np.array_equal(st.pt_temps(), rp.tor), step
    for r in range(NS * NT):
        assert np.array_equal(st.download(r), rp.s[r]), r
    b = np.zeros(NS * NT)
    ctx.check(ctx.lib.mq_state_get_beta(st.h, b.ctypes.data), "get_beta")
    assert np.array_equal(b, BETAS[rp.tor])


@pytest.mark.parametrize("mode", [1, 2], ids=["hip-graph", "stream"])
def test_graph_replays_match_eager(ctx, mode):
    """Each mq_run_graph call (with eager work and synchronisation in between) equals the
    eager run, with the steps captured in a HIP graph and replayed or launched on the stream."""
    g = _ea_state(ctx)
    g.set_tuning(graph=mode)
    e = _ea_state(ctx)
    nsw, steps = 2, 3
    done = 0
    for call, repeats in enumerate((1, 1, 2)):
        g.run_graph(nsw, NT, seed=11, sweep0=done * nsw, step0=done, steps=steps, repeats=repeats)
        for t in range(done, done + steps * repeats):
            e.metropolis(nsw, seed=11, sweep0=t * nsw)
            e.pt_swap(NT, t & 1, seed=11, step=t)
        done += steps * repeats
        ctx.sync()
        assert np.array_equal(g.pt_temps(), e.pt_temps()), call
        for r in range(NS * NT):
            assert np.array_equal(g.download(r), e.download(r)), (call, r)
        # eager work on the graph's state between replays
        g.measure()


@pytest.mark.parametrize("mode", [1, 2], ids=["hip-graph", "stream"])
def test_graph_sweeps_only_regular_ising(ctx, mode):
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, 3, 32)
    a = gpu.State(ctx, "ising", [lat, lat], J=-1.0)
    b = gpu.State(ctx, "ising", [lat, lat], J=-1.0)
    a.set_tuning(graph=mode)
    for st in (a, b):
        st.set_beta([0.22, 0.3])
        st.init(seed=3)
    a.run_graph(3, 0, seed=5, sweep0=0, step0=0, steps=2, repeats=3)
    b.metropolis(18, seed=5, sweep0=0)
    for r in range(2):
        assert np.array_equal(a.download(r), b.download(r))
