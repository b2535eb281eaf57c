"""1-bit Ising on graphs (spin_bits = 1 on a CSR lattice, 32 replicas per word):
the chain equals the int8 graph chain bit for bit (same uniforms, same exact
acceptance rule), the CPU replay of the graph sweep reproduces it, and the
measurements, parallel tempering and graph replays agree with int8 storage."""
import numpy as np
import pytest

import oracle_bind as ob
from marqov_amd import graphs

pytestmark = pytest.mark.gpu


def _colours(ctx, lat):
    c = np.zeros(lat.N, np.int32)
    ctx.check(ctx.lib.mq_lattice_color_of(lat.h, c.ctypes.data), "mq_lattice_color_of")
    return c


def _hub_graph(n, seed):
    """Random sparse graph with 20 isolated vertices and one hub of degree 127."""
    rng = np.random.default_rng(seed)
    edges = set()
    while len(edges) < 2 * n:
        u, v = rng.integers(1, n - 20, 2)  # vertex 0 is the hub, the last 20 vertices stay isolated
        if u != v:
            edges.add((min(u, v), max(u, v)))
    for v in rng.choice(np.arange(1, n - 20), 127, replace=False):
        edges.add((0, int(v)))
    adj = [[] for _ in range(n)]
    for u, v in edges:
        adj[u].append(v)
        adj[v].append(u)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(x) for x in adj])
    ci = np.array([v for x in adj for v in sorted(x)], np.int32)
    return rp, ci


def _graph(kind):
    if kind == "rgg":
        rp, ci, _ = graphs.rgg2d(3000, graphs.rgg_radius(3000), 6)
    elif kind == "torus":
        rp, ci = graphs.regular_csr(2, 16)
    elif kind == "bigtorus":  # two colour classes of 32,768 vertices: the one-lane-per-word kernel
        rp, ci = graphs.regular_csr(2, 256)
    else:
        rp, ci = _hub_graph(1200, 3)
    return np.asarray(rp, np.int32), np.asarray(ci, np.int32)


def _pair(ctx, lats, J):
    from marqov_amd import gpu
    bits = gpu.State(ctx, "ising", lats, J=J, spin_bits=1)
    i8 = gpu.State(ctx, "ising", lats, J=J, spin_bits=8)
    return bits, i8


def _same(a, b):
    return all(np.array_equal(a.download(r), b.download(r)) for r in range(a.nrep))


@pytest.mark.parametrize("kind,nrep,J", [("rgg", 32, -1.0), ("rgg", 64, -1.0), ("torus", 96, -1.0),
                                         ("hub", 64, -1.0), ("rgg", 64, 1.0), ("bigtorus", 64, -1.0)])
def test_chain_equals_int8_and_replay(ctx, kind, nrep, J):
    from marqov_amd import gpu
    rp, ci = _graph(kind)
    if kind == "hub":
        assert np.diff(rp).max() == 127 and (np.diff(rp) == 0).sum() >= 20
    lat = gpu.CsrLattice(ctx, rp, ci)
    betas = np.linspace(0.15, 1.2, nrep)
    bits, i8 = _pair(ctx, [lat] * nrep, J)
    for st in (bits, i8):
        st.set_beta(betas)
        st.init(seed=5)
    assert _same(bits, i8)
    for st in (bits, i8):
        st.metropolis(6, seed=9, sweep0=3)
    assert _same(bits, i8)
    np.testing.assert_array_equal(bits.measure(), i8.measure())
    col = _colours(ctx, lat)
    N = len(rp) - 1
    for r in (0, 31, nrep - 1):
        exp = ob.csr_sweeps(rp, ci, col, ob.init_ising(N, 5, r), betas[r], J, 6, 9, 3, r)
        assert np.array_equal(bits.download(r), exp), r


def test_two_graphs_in_one_state(ctx):
    """Groups of 32 and 64 replicas on two graphs: each group has its own word count."""
    from marqov_amd import gpu
    a = gpu.CsrLattice(ctx, *_graph("rgg"))
    b = gpu.CsrLattice(ctx, *_graph("torus"))
    lats = [a] * 32 + [b] * 64
    bits, i8 = _pair(ctx, lats, -1.0)
    for st in (bits, i8):
        st.set_beta(np.linspace(0.3, 0.9, 96))
        st.init(seed=11)
        st.metropolis(4, seed=2, sweep0=0)
    assert _same(bits, i8)
    np.testing.assert_array_equal(bits.measure(), i8.measure())


def test_upload_download_round_trip(ctx):
    from marqov_amd import gpu
    rp, ci = _graph("rgg")
    lat = gpu.CsrLattice(ctx, rp, ci)
    bits = gpu.State(ctx, "ising", [lat] * 64, J=-1.0, spin_bits=1)
    bits.set_beta(0.4)
    bits.init(seed=3)
    before = [bits.download(r) for r in range(64)]
    s = np.where(np.random.default_rng(1).random(lat.N) < 0.5, 1, -1).astype(np.int8)
    bits.upload(37, s)
    assert np.array_equal(bits.download(37), s)
    assert all(np.array_equal(bits.download(r), before[r]) for r in range(64) if r != 37)


@pytest.mark.parametrize("mode", [1, 2], ids=["hip-graph", "stream"])
def test_pt_and_graph_replay_equal_int8(ctx, mode):
    """Eager sweeps + energies + PT swaps, then mq_run_graph steps (captured in a HIP graph
    and replayed, or launched on the stream): 1-bit and int8 storage make the same swap
    decisions and end in the same states."""
    from marqov_amd import gpu
    lat = gpu.CsrLattice(ctx, *_graph("rgg"))
    nrep, ntemp = 64, 32
    ladder = np.geomspace(0.2, 1.1, ntemp)
    bits, i8 = _pair(ctx, [lat] * nrep, -1.0)
    accs = []
    for st in (bits, i8):
        st.set_tuning(graph=mode)
        st.set_beta(np.tile(ladder, nrep // ntemp))
        st.init(seed=7)
        acc = []
        for step in range(4):
            st.metropolis(3, seed=8, sweep0=3 * step)
            acc.append(st.pt_swap(ntemp, step & 1, seed=9, step=step))
        st.run_graph(3, ntemp, seed=8, sweep0=12, step0=4, steps=2, repeats=3)
        accs.append((np.array(acc), st.pt_temps()))
    assert np.array_equal(accs[0][0], accs[1][0]) and np.array_equal(accs[0][1], accs[1][1])
    assert _same(bits, i8)
    np.testing.assert_array_equal(bits.energy(), i8.energy())


def test_refusals(ctx):
    from marqov_amd import gpu
    from marqov_amd.gpu import MarqovError
    lat = gpu.CsrLattice(ctx, *_graph("rgg"))
    with pytest.raises(MarqovError, match="multiple of 32"):
        gpu.State(ctx, "ising", [lat] * 48, J=-1.0, spin_bits=1)
    st = gpu.State(ctx, "ising", [lat] * 32, J=-1.0, spin_bits=1)
    st.set_beta(0.4)
    st.init(seed=1)
    with pytest.raises(MarqovError, match="int8 storage"):
        st.wolff(1, seed=1)
    with pytest.raises(MarqovError, match="no couplings"):
        st.set_bonds(0, np.ones(len(_graph("rgg")[1]), np.int8))
    with pytest.raises(MarqovError, match="int8 or float4"):
        st.measure_ft()
    rows = [list(range(1, 129))] + [[0]] * 128  # vertex 0 joined to 1..128: degree 128
    rp = np.zeros(len(rows) + 1, np.int32)
    rp[1:] = np.cumsum([len(x) for x in rows])
    ci = np.array([v for x in rows for v in x], np.int32)
    star = gpu.CsrLattice(ctx, rp, ci)
    with pytest.raises(MarqovError, match="degrees <= 127"):
        gpu.State(ctx, "ising", [star] * 32, J=-1.0, spin_bits=1)
