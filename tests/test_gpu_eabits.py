"""Bit-sliced EA batches (32 replicas of one sample per uint32 word): random start,
sweeps, observables, parallel tempering and graph replay, bit-exact against the
CPU replay."""
import numpy as np
import pytest

import oracle_bind as ob
import exact_path  # noqa: F401
import exact

pytestmark = pytest.mark.gpu

DIM = 3
NT = 32
BETAS = 0.5 * (3.3 / 0.5) ** (np.arange(NT) / (NT - 1))


def _state(ctx, Ls, words_per_L=1, J=-1.0):
    from marqov_amd import gpu
    lats = {L: gpu.RegularLattice(ctx, DIM, L) for L in Ls}
    reps = [lats[L] for L in Ls for _ in range(32 * words_per_L)]
    st = gpu.State(ctx, "ea", reps, J=J, spin_bits=1)
    st._lats = lats
    nw = len(reps) // 32
    st.set_beta(np.tile(BETAS, nw))
    st.random_bonds(rep_sample=np.repeat(np.arange(nw), 32), p_plus=0.5, seed=4)
    st.init(seed=9)
    return st, reps


def _replay_words(reps, nsweeps, sweep0, temps=None, J=-1.0, states=None):
    nw = len(reps) // 32
    out = []
    for w in range(nw):
        L = reps[32 * w].L
        N = L ** DIM
        bonds = ob.init_bonds(DIM, N, 0.5, 4, w)
        s0 = states[w] if states is not None else np.stack([ob.init_ising(N, 9, 32 * w + j) for j in range(32)])
        b = BETAS if temps is None else BETAS[temps[32 * w:32 * w + 32]]
        out.append(ob.ea_bits_sweeps(DIM, L, s0, bonds, b, J, nsweeps, 21, sweep0, w))
    return out


@pytest.mark.parametrize("J", [-1.0, 1.0])
def test_eabits_init_and_sweeps_bit_exact(ctx, J):
    st, reps = _state(ctx, [4, 6], words_per_L=2, J=J)
    for r in (0, 31, 64, 127):
        assert np.array_equal(st.download(r), ob.init_ising(reps[r].N, 9, r)), r
    st.metropolis(3, seed=21, sweep0=0)
    exp = _replay_words(reps, 3, 0, J=J)
    for r in range(len(reps)):
        assert np.array_equal(st.download(r), exp[r // 32][r % 32]), r


@pytest.mark.parametrize("Ls,words", [([6], 1), ([4, 8, 12, 24], 2)])
def test_eabits_measure(ctx, Ls, words):
    """One fused launch measures every lattice size; at L=24 a word spans 14
    wavefronts whose per-replica totals are combined with atomics."""
    st, reps = _state(ctx, Ls, words_per_L=words)
    st.metropolis(2, seed=21, sweep0=0)
    obs = st.measure()
    for w in range(len(reps) // 32):
        L = reps[32 * w].L
        N = L ** 3
        nb = exact.neighbour_table(DIM, L)
        B = exact.bond_table(DIM, L, ob.init_bonds(DIM, N, 0.5, 4, w))
        for j in (0, 5, 31):
            r = 32 * w + j
            s = st.download(r).astype(np.int64)
            e = -1.0 * (s[:, None] * s[nb] * B).sum() / 2.0 / N
            assert obs[r, 0] == pytest.approx(e, abs=1e-12), r
            assert obs[r, 1] == pytest.approx(abs(s.sum()) / N, abs=1e-12), r


@pytest.mark.parametrize("mode", [1, 2], ids=["hip-graph", "stream"])
def test_eabits_pt_and_graph(ctx, mode):
    """PT swaps move temperatures between bits (the per-bit threshold words are
    rebuilt), and mq_run_graph's {sweeps, PT} steps equal the eager sequence, captured
    in a HIP graph or launched on the stream."""
    L = 4
    g, reps = _state(ctx, [L], words_per_L=2)
    g.set_tuning(graph=mode)
    e, _ = _state(ctx, [L], words_per_L=2)
    nsw, steps = 2, 3
    g.run_graph(nsw, NT, 21, 0, 0, steps, 2)
    for t in range(steps * 2):
        e.metropolis(nsw, seed=21, sweep0=t * nsw)
        e.pt_swap(NT, t & 1, seed=21, step=t)
    assert np.array_equal(g.pt_temps(), e.pt_temps())
    for r in range(len(reps)):
        assert np.array_equal(g.download(r), e.download(r)), r
    # after swaps the sweeps use the new temperatures: replay one more sweep
    temps = e.pt_temps()
    before = [np.stack([e.download(32 * w + j) for j in range(32)]) for w in range(2)]
    e.metropolis(1, seed=21, sweep0=100)
    exp = _replay_words(reps, 1, 100, temps=temps, states=before)
    for r in range(len(reps)):
        assert np.array_equal(e.download(r), exp[r // 32][r % 32]), r
    assert (temps != np.tile(np.arange(NT), 2)).any()
