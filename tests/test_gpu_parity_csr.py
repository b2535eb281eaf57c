"""Graph (CSR) path: greedy colouring is proper, and colour-class sweeps and Wolff
clusters replay bit-exactly on the CPU given the library's colouring."""
import numpy as np
import pytest

import oracle_bind as ob
from marqov_amd import graphs

pytestmark = pytest.mark.gpu


def _graph(n, seed):
    rp, ci, _ = graphs.rgg2d(n, graphs.rgg_radius(n), seed)
    return rp, ci


def _lat(ctx, rp, ci):
    from marqov_amd import gpu
    return gpu.CsrLattice(ctx, rp, ci)


def _colours(ctx, lat):
    c = np.zeros(lat.N, np.int32)
    ctx.check(ctx.lib.mq_lattice_color_of(lat.h, c.ctypes.data), "mq_lattice_color_of")
    return c


def test_colouring_is_proper(ctx):
    rp, ci = _graph(3000, 2)
    lat = _lat(ctx, rp, ci)
    col = _colours(ctx, lat)
    src = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    assert (col[src] != col[ci]).all()
    sizes = lat.colors()
    assert sizes.sum() == len(rp) - 1 and len(sizes) == col.max() + 1


@pytest.mark.parametrize("model,beta", [("ising", 0.6), ("ising", 2.0), ("potts", 1.2)])
def test_csr_sweeps_bit_exact(ctx, model, beta):
    from marqov_amd import gpu
    rp, ci = _graph(2500, 3)
    lat = _lat(ctx, rp, ci)
    col = _colours(ctx, lat)
    st = gpu.State(ctx, model, [lat, lat], J=-1.0, q=3)
    st.set_beta([beta, beta / 2])
    st.init(seed=8)
    st.metropolis(5, seed=21, sweep0=4)
    N = len(rp) - 1
    for r in range(2):
        s0 = ob.init_potts(N, 3, 8, r) if model == "potts" else ob.init_ising(N, 8, r)
        exp = ob.csr_sweeps(rp, ci, col, s0, [beta, beta / 2][r], -1.0, 5, 21, 4, r, potts=(model == "potts"))
        assert np.array_equal(st.download(r), exp), r


def test_csr_ea_bonds_bit_exact(ctx):
    from marqov_amd import gpu
    rp, ci = _graph(1500, 4)
    N = len(rp) - 1
    lat = _lat(ctx, rp, ci)
    col = _colours(ctx, lat)
    # symmetric random couplings per edge
    src = np.repeat(np.arange(N), np.diff(rp))
    lo, hi = np.minimum(src, ci), np.maximum(src, ci)
    h = (lo.astype(np.int64) * 2654435761 + hi * 40503) % 1000003
    bonds = np.where(h % 2 == 0, 1, -1).astype(np.int8)
    st = gpu.State(ctx, "ea", [lat], J=-1.0)
    st.set_beta(1.3)
    st.set_bonds(0, bonds)
    st.init(seed=5)
    st.metropolis(4, seed=6, sweep0=0)
    exp = ob.csr_sweeps(rp, ci, col, ob.init_ising(N, 5, 0), 1.3, -1.0, 4, 6, 0, 0, bonds=bonds)
    assert np.array_equal(st.download(0), exp)


@pytest.mark.parametrize("block", ["256", "1024"])
@pytest.mark.parametrize("model,beta", [("ising", 0.9), ("potts", 1.6)])
def test_csr_wolff_bit_exact(ctx, model, beta, block, monkeypatch):
    from marqov_amd import gpu
    rp, ci = _graph(2000, 7)
    N = len(rp) - 1
    lat = _lat(ctx, rp, ci)
    st = gpu.State(ctx, model, [lat, lat], J=-1.0, q=3)
    st.set_tuning(wolff_block=int(block))
    st.set_beta(beta)
    st.init(seed=2)
    sizes = st.wolff(6, seed=13, cluster0=2)
    for r in range(2):
        s0 = ob.init_potts(N, 3, 2, r) if model == "potts" else ob.init_ising(N, 2, r)
        exp, esz = ob.csr_wolff(rp, ci, s0, beta, -1.0, 6, 13, 2, r, potts=(model == "potts"))
        assert np.array_equal(sizes[r], esz), (sizes[r], esz)
        assert np.array_equal(st.download(r), exp)


def test_csr_regular_graph_observables(ctx):
    """The periodic square lattice given as a graph measures like the regular path."""
    from marqov_amd import gpu
    rp, ci = graphs.regular_csr(2, 8)
    lat = _lat(ctx, rp, ci)
    st = gpu.State(ctx, "ising", [lat], J=-1.0)
    st.set_beta(0.4)
    st.init(seed=1)
    st.metropolis(3, seed=2)
    s = st.download(0)
    e = -1.0 * ob.lib().orc_bond_sum_ising(2, 8, ob.p(s), None) / 64
    obs = st.measure()
    assert obs[0, 0] == pytest.approx(e, abs=1e-12)
    assert obs[0, 1] == pytest.approx(abs(int(s.astype(int).sum())) / 64, abs=1e-12)


@pytest.mark.parametrize("model", ["ising", "ea", "potts"])
def test_csr_octet_kernel_bit_exact(ctx, model):
    """16 replicas on one graph select the 8-replicas-per-lane kernels."""
    from marqov_amd import gpu
    rp, ci = _graph(1200, 9)
    N = len(rp) - 1
    lat = _lat(ctx, rp, ci)
    col = _colours(ctx, lat)
    R = 16
    st = gpu.State(ctx, model, [lat] * R, J=-1.0, q=3)
    betas = np.linspace(0.3, 1.5, R)
    st.set_beta(betas)
    bonds = None
    if model == "ea":
        src = np.repeat(np.arange(N), np.diff(rp))
        lo, hi = np.minimum(src, ci), np.maximum(src, ci)
        bonds = np.where(((lo.astype(np.int64) * 7919 + hi * 104729) % 3) == 0, -1, 1).astype(np.int8)
        st.set_samples(np.zeros(R, np.int32))  # one sample shared by all replicas
        st.set_bonds(0, bonds)
    st.init(seed=14)
    st.metropolis(4, seed=15, sweep0=2)
    for r in range(R):
        if model == "potts":
            exp = ob.csr_sweeps(rp, ci, col, ob.init_potts(N, 3, 14, r), betas[r], -1.0, 4, 15, 2, r, potts=True)
        else:
            exp = ob.csr_sweeps(rp, ci, col, ob.init_ising(N, 14, r), betas[r], -1.0, 4, 15, 2, r, bonds=bonds)
        assert np.array_equal(st.download(r), exp), r


def _random_graph(n, mean_degree, seed):
    """Erdos-Renyi-like simple undirected graph as sorted CSR (no geometry, so BFS
    fronts grow geometrically)."""
    rng = np.random.default_rng(seed)
    m = int(n * mean_degree / 2)
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    keep = a != b
    u = np.minimum(a[keep], b[keep]).astype(np.int64)
    v = np.maximum(a[keep], b[keep]).astype(np.int64)
    pairs = np.unique(u * n + v)
    u, v = pairs // n, pairs % n
    src = np.concatenate([u, v])
    dst = np.concatenate([v, u])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, src + 1, 1)
    return np.cumsum(rp).astype(np.int32), dst.astype(np.int32)


@pytest.mark.parametrize("block", ["256", "1024"])
@pytest.mark.parametrize("model", ["ising", "potts"])
def test_csr_wolff_frontier_spill_bit_exact(ctx, model, block, monkeypatch):
    """A random graph of mean degree 16 in an ordered start with bond probability
    1 - e^-2: BFS fronts of thousands of vertices overflow the 4096-entry LDS ring and
    spill to the global queue (with 256- and 1024-thread workgroups); the clusters must
    still equal the CPU replay."""
    from marqov_amd import gpu
    n = 24000
    rp, ci = _random_graph(n, 16, 5)
    lat = _lat(ctx, rp, ci)
    st = gpu.State(ctx, model, [lat, lat], J=-1.0, q=3)
    st.set_tuning(wolff_block=int(block))
    beta = 1.0 if model == "ising" else 2.0
    st.set_beta(beta)
    starts = []
    for r in range(2):
        s0 = np.ones(n, np.int8) if model == "ising" else np.zeros(n, np.int8)
        s0[::11] = -1 if model == "ising" else 1
        st.upload(r, s0)
        starts.append(s0)
    sizes = st.wolff(2, seed=3, cluster0=0)
    for r in range(2):
        exp, esz = ob.csr_wolff(rp, ci, starts[r], beta, -1.0, 2, 3, 0, r, potts=(model == "potts"))
        assert esz.max() > n // 2
        assert np.array_equal(sizes[r], esz), (sizes[r], esz)
        assert np.array_equal(st.download(r), exp)


@pytest.mark.parametrize("model,R", [("ising", 40), ("potts", 64)])
def test_csr_octet_batches_sweeps_wolff_sweeps_bit_exact(ctx, model, R):
    """Graph states of 5 and 8 replica octets (the 8-replicas-per-lane kernels) on a graph of
    at least three colour classes: sweeps, Wolff clusters, sweeps again and the measurement
    equal the CPU replay for every replica."""
    from marqov_amd import gpu
    rp, ci = _graph(2000, 13)
    N = len(rp) - 1
    lat = _lat(ctx, rp, ci)
    col = _colours(ctx, lat)
    assert col.max() + 1 >= 3
    st = gpu.State(ctx, model, [lat] * R, J=-1.0, q=3)
    betas = np.linspace(0.4, 1.4, R)
    st.set_beta(betas)
    st.init(seed=24)
    st.metropolis(3, seed=25, sweep0=0)
    sizes = st.wolff(2, seed=26, cluster0=0)
    st.metropolis(2, seed=25, sweep0=3)
    obs = st.measure()
    potts = model == "potts"
    for r in range(R):
        s = ob.init_potts(N, 3, 24, r) if potts else ob.init_ising(N, 24, r)
        s = ob.csr_sweeps(rp, ci, col, s, betas[r], -1.0, 3, 25, 0, r, potts=potts)
        s, esz = ob.csr_wolff(rp, ci, s, betas[r], -1.0, 2, 26, 0, r, potts=potts)
        s = ob.csr_sweeps(rp, ci, col, s, betas[r], -1.0, 2, 25, 3, r, potts=potts)
        assert np.array_equal(sizes[r], esz), r
        got = st.download(r)
        assert np.array_equal(got, s), r
        if not potts:
            assert obs[r, 1] == pytest.approx(abs(int(got.astype(np.int64).sum())) / N, abs=1e-12), r
