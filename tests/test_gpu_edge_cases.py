"""Edge cases of the update path on the GPU against the oracle.

* The smallest lattice, L = 2: both x0 neighbours of a site are the same site, so
  its bond is listed twice, as the reference's neighbour arithmetic lists it
  (regular_hypercubic.h:31-66).  Sweeps on the LDS and the streaming kernels, Wolff
  on the small-lattice and the workgroup kernels, and the observables.
* Graphs with isolated vertices (degree 0), a hub of degree 127 (the largest the
  8-replicas-per-lane kernels take) and of degree 128 (the per-replica kernels then
  run), a single edge, and the complete graphs K_64, K_65 and K_130 (as many colour
  classes of one vertex; above 64 classes the sweep's counter phase uses a wider
  stride).
* Potts with q = 8, the largest Potts<Q> the reference instantiates, on lattices
  and graphs (sweeps, Wolff, observables).
* Zero counts are no-ops, negative counts are refused, and one mq_emcs equals
  mq_wolff followed by mq_metropolis with the counters it documents
  (emcs.h:69-109; mean cluster size NaN when ncluster = 0, emcs.h:108).
"""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from marqov_amd import gpu
    return gpu.Context(0)


@pytest.mark.parametrize("kernels", ["0", "1"])
@pytest.mark.parametrize("dim", [2, 3])
def test_L2_sweeps_bit_exact(ctx, dim, kernels, monkeypatch):
    from marqov_amd import gpu
    monkeypatch.setenv("MARQOV_AMD_KERNELS", kernels)
    L, N, R = 2, 2 ** dim, 4
    lat = gpu.RegularLattice(ctx, dim, L)
    betas = [0.2, 0.5, 1.0, 2.0]
    for model in ("ising", "ea", "potts"):
        st = gpu.State(ctx, model, [lat] * R, J=-1.0, q=3)
        st.set_beta(betas)
        if model == "ea":
            st.random_bonds(rep_sample=np.arange(R, dtype=np.int32), p_plus=0.5, seed=4)
        st.init(seed=7)
        st.metropolis(6, seed=8, sweep0=1)
        for r in range(R):
            if model == "potts":
                exp = ob.potts_sweeps(dim, L, ob.init_potts(N, 3, 7, r), 3, betas[r], -1.0, 6, 8, 1, r)
            else:
                b = ob.init_bonds(dim, N, 0.5, 4, r) if model == "ea" else None
                exp = ob.ising_sweeps(dim, L, ob.init_ising(N, 7, r), betas[r], -1.0, 6, 8, 1, r, bonds=b)
            assert np.array_equal(st.download(r), exp), (model, kernels, r)


@pytest.mark.parametrize("kernels", ["0", "1"])
@pytest.mark.parametrize("dim", [2, 3])
def test_L2_wolff_and_measure(ctx, dim, kernels, monkeypatch):
    from marqov_amd import gpu
    monkeypatch.setenv("MARQOV_AMD_KERNELS", kernels)
    L, N = 2, 2 ** dim
    lat = gpu.RegularLattice(ctx, dim, L)
    for model in ("ising", "potts"):
        st = gpu.State(ctx, model, [lat] * 3, J=-1.0, q=3)
        st.set_beta(0.7)
        st.init(seed=9)
        sizes = st.wolff(8, seed=10, cluster0=2)
        obs = st.measure()
        for r in range(3):
            if model == "ising":
                exp, esz = ob.wolff_ising(dim, L, ob.init_ising(N, 9, r), 0.7, -1.0, 8, 10, 2, r)
                e = -1.0 * ob.lib().orc_bond_sum_ising(dim, L, ob.p(exp), None) / N
                m = abs(int(exp.astype(np.int64).sum())) / N
            else:
                exp, esz = ob.wolff_potts(dim, L, ob.init_potts(N, 3, 9, r), 3, 0.7, -1.0, 8, 10, 2, r)
                e = -1.0 * ob.lib().orc_bond_sum_potts(dim, L, ob.p(exp)) / N
                cnt = np.bincount(exp.astype(np.int64), minlength=3)
                m = (3 * cnt.max() - N) / (N * 2)
            assert np.array_equal(sizes[r], esz), (model, r, sizes[r], esz)
            assert np.array_equal(st.download(r), exp), (model, r)
            assert obs[r, 0] == pytest.approx(e, abs=1e-12)
            assert obs[r, 1] == pytest.approx(m, abs=1e-12)


def _csr(n, edges):
    """Sorted symmetric CSR of an undirected simple graph on n vertices."""
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    src = np.concatenate([e[:, 0], e[:, 1]])
    dst = np.concatenate([e[:, 1], e[:, 0]])
    order = np.lexsort((dst, src))
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, src + 1, 1)
    return np.cumsum(rp).astype(np.int32), dst[order].astype(np.int32)


def _odd_graph(hub_deg):
    """A hub with hub_deg leaves, a 20-cycle, one lone edge and 5 isolated vertices,
    with the vertex ids shuffled."""
    edges = [(0, 1 + k) for k in range(hub_deg)]
    c0 = hub_deg + 1
    edges += [(c0 + k, c0 + (k + 1) % 20) for k in range(20)]
    e0 = c0 + 20
    edges.append((e0, e0 + 1))
    n = e0 + 2 + 5
    perm = np.random.default_rng(hub_deg).permutation(n)
    return _csr(n, [(perm[a], perm[b]) for a, b in edges])


def _colours(ctx, lat):
    c = np.zeros(lat.N, np.int32)
    ctx.check(ctx.lib.mq_lattice_color_of(lat.h, c.ctypes.data), "mq_lattice_color_of")
    return c


@pytest.mark.parametrize("hub_deg", [127, 128])
@pytest.mark.parametrize("nrep", [1, 16])
def test_odd_graph_sweeps_and_wolff_bit_exact(ctx, hub_deg, nrep):
    """nrep = 16 takes the 8-replicas-per-lane sweeps when the hub has 127 leaves and
    the per-replica sweeps at 128; nrep = 1 always takes the per-replica sweeps."""
    from marqov_amd import gpu
    rp, ci = _odd_graph(hub_deg)
    N = len(rp) - 1
    lat = gpu.CsrLattice(ctx, rp, ci)
    col = _colours(ctx, lat)
    betas = np.linspace(0.2, 1.6, nrep)
    for model in ("ising", "potts"):
        potts = model == "potts"
        st = gpu.State(ctx, model, [lat] * nrep, J=-1.0, q=3)
        st.set_beta(betas)
        st.init(seed=3)
        st.metropolis(5, seed=4, sweep0=0)
        starts = [st.download(r) for r in range(nrep)]
        for r in range(nrep):
            s0 = ob.init_potts(N, 3, 3, r) if potts else ob.init_ising(N, 3, r)
            exp = ob.csr_sweeps(rp, ci, col, s0, betas[r], -1.0, 5, 4, 0, r, potts=potts)
            assert np.array_equal(starts[r], exp), (model, hub_deg, nrep, r)
        sizes = st.wolff(12, seed=5, cluster0=0)
        for r in range(nrep):
            exp, esz = ob.csr_wolff(rp, ci, starts[r], betas[r], -1.0, 12, 5, 0, r, potts=potts)
            assert np.array_equal(sizes[r], esz), (model, hub_deg, nrep, r)
            assert np.array_equal(st.download(r), exp), (model, hub_deg, nrep, r)
        assert sizes.min() >= 1


@pytest.mark.parametrize("n", [64, 65, 130])
def test_complete_graphs_many_colour_classes(ctx, n):
    """The complete graph K_n needs n colour classes of one vertex each.  Up to 64 classes
    a sweep's phase is 64*sweep + colour; K_65 and K_130 (130 classes, degree 129: the
    per-replica kernels run) use the stride 128 and 256, and sweeps, Wolff clusters and the
    sweep-number limit 2^32/stride follow it, equal to the CPU replay."""
    from marqov_amd import gpu
    rp, ci = _csr(n, [(a, b) for a in range(n) for b in range(a + 1, n)])
    lat = gpu.CsrLattice(ctx, rp, ci)
    assert len(lat.colors()) == n
    col = _colours(ctx, lat)
    R = 16
    betas = np.linspace(0.005, 0.05, R)
    st = gpu.State(ctx, "ising", [lat] * R, J=-1.0)
    st.set_beta(betas)
    st.init(seed=1)
    st.metropolis(3, seed=2, sweep0=5)
    sizes = st.wolff(3, seed=4, cluster0=0)
    st.metropolis(1, seed=2, sweep0=8)
    for r in range(R):
        exp = ob.csr_sweeps(rp, ci, col, ob.init_ising(n, 1, r), betas[r], -1.0, 3, 2, 5, r)
        exp, esz = ob.csr_wolff(rp, ci, exp, betas[r], -1.0, 3, 4, 0, r)
        exp = ob.csr_sweeps(rp, ci, col, exp, betas[r], -1.0, 1, 2, 8, r)
        assert np.array_equal(sizes[r], esz), r
        assert np.array_equal(st.download(r), exp), r
    stride = 64
    while stride < n:
        stride *= 2
    lim = (1 << 32) // stride
    st.metropolis(1, seed=2, sweep0=lim - 1)
    with pytest.raises(gpu.MarqovError, match="sweep counter exceeds"):
        st.metropolis(1, seed=2, sweep0=lim)


def test_zero_and_negative_counts(ctx):
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, 2, 16)
    st = gpu.State(ctx, "ising", [lat] * 2, J=-1.0)
    st.set_beta(0.44)
    st.init(seed=3)
    s0 = [st.download(r) for r in range(2)]
    st.metropolis(0, seed=1)
    assert st.wolff(0, seed=1).shape == (2, 0)
    mc = st.emcs(0, 0, seed=1, step=0)
    assert np.all(np.isnan(mc))
    for r in range(2):
        assert np.array_equal(st.download(r), s0[r])
    for bad in (lambda: st.metropolis(-1, seed=1), lambda: st.wolff(-1, seed=1, sizes=False),
                lambda: st.emcs(-1, 1, seed=1, step=0), lambda: st.emcs(1, -1, seed=1, step=0)):
        with pytest.raises(gpu.MarqovError):
            bad()
    for r in range(2):
        assert np.array_equal(st.download(r), s0[r])


@pytest.mark.parametrize("model", ["ising", "heisenberg"])
def test_emcs_is_wolff_then_metropolis(ctx, model):
    """Two states from the same start: mq_emcs(nc, nm, step) on one, mq_wolff(nc,
    cluster0 = step*nc) then mq_metropolis(nm, sweep0 = step*nm) on the other: the
    states are identical and the mean cluster size is the mean of the sizes."""
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, 3, 8)
    kw = dict(J=-1.0) if model == "ising" else dict(J=-1.0, proposal="small", delta=1.0)
    a = gpu.State(ctx, model, [lat] * 3, **kw)
    b = gpu.State(ctx, model, [lat] * 3, **kw)
    for st in (a, b):
        st.set_beta([0.2, 0.22, 0.7] if model == "ising" else [0.5, 0.69, 0.9])
        st.init(seed=6)
    nc, nm = 3, 2
    means = []
    sizes = []
    for step in range(4):
        means.append(a.emcs(nc, nm, seed=7, step=step))
        sizes.append(b.wolff(nc, seed=7, cluster0=step * nc))
        b.metropolis(nm, seed=7, sweep0=step * nm)
    for step in range(4):
        assert np.array_equal(means[step], sizes[step].mean(axis=1)), step
    for r in range(3):
        assert np.array_equal(a.download(r), b.download(r)), r


def _potts_obs(s, q, bond_equal, N):
    cnt = np.bincount(s.astype(np.int64), minlength=q)
    return -1.0 * bond_equal / N, (q * cnt.max() - N) / (N * (q - 1))


@pytest.mark.parametrize("dim,L", [(2, 8), (3, 4)])
def test_potts8_regular_bit_exact(ctx, dim, L):
    """q = 8, the largest Potts<Q> the reference instantiates (demo/Potts.cpp:88): the
    kernels count colours 0..6 and the eighth count is what they leave."""
    from marqov_amd import gpu
    N, q, betas = L ** dim, 8, [1.2, 2.5]
    lat = gpu.RegularLattice(ctx, dim, L)
    st = gpu.State(ctx, "potts", [lat] * 2, J=-1.0, q=q)
    st.set_beta(betas)
    st.init(seed=5)
    st.metropolis(4, seed=6, sweep0=0)
    sizes = st.wolff(5, seed=7, cluster0=0)
    obs = st.measure()
    for r in range(2):
        s1 = ob.potts_sweeps(dim, L, ob.init_potts(N, q, 5, r), q, betas[r], -1.0, 4, 6, 0, r)
        exp, esz = ob.wolff_potts(dim, L, s1, q, betas[r], -1.0, 5, 7, 0, r)
        assert np.array_equal(sizes[r], esz), r
        got = st.download(r)
        assert np.array_equal(got, exp), r
        assert got.min() >= 0 and got.max() <= 7
        e, m = _potts_obs(exp, q, ob.lib().orc_bond_sum_potts(dim, L, ob.p(exp)), N)
        assert obs[r, 0] == pytest.approx(e, abs=1e-12)
        assert obs[r, 1] == pytest.approx(m, abs=1e-12)


@pytest.mark.parametrize("nrep", [2, 16])
def test_potts8_graph_bit_exact(ctx, nrep):
    """q = 8 on a graph with the per-replica (2 replicas) and the 8-replicas-per-lane
    (16 replicas) sweeps, Wolff, and the observables."""
    from marqov_amd import gpu, graphs
    rp, ci, _ = graphs.rgg2d(1500, graphs.rgg_radius(1500), 12)
    N, q = len(rp) - 1, 8
    lat = gpu.CsrLattice(ctx, rp, ci)
    col = _colours(ctx, lat)
    betas = np.linspace(0.8, 2.4, nrep)
    st = gpu.State(ctx, "potts", [lat] * nrep, J=-1.0, q=q)
    st.set_beta(betas)
    st.init(seed=2)
    st.metropolis(3, seed=3, sweep0=1)
    sizes = st.wolff(4, seed=4, cluster0=0)
    obs = st.measure()
    src = np.repeat(np.arange(N), np.diff(rp))
    for r in range(nrep):
        s1 = ob.csr_sweeps(rp, ci, col, ob.init_potts(N, q, 2, r), betas[r], -1.0, 3, 3, 1, r, potts=True, Q=q)
        exp, esz = ob.csr_wolff(rp, ci, s1, betas[r], -1.0, 4, 4, 0, r, potts=True, Q=q)
        assert np.array_equal(sizes[r], esz), r
        assert np.array_equal(st.download(r), exp), r
        eq = int((exp[src] == exp[ci]).sum()) // 2
        e, m = _potts_obs(exp, q, eq, N)
        assert obs[r, 0] == pytest.approx(e, abs=1e-12)
        assert obs[r, 1] == pytest.approx(m, abs=1e-12)
