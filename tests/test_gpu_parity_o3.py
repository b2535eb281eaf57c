"""FP32 O(3) / Phi4 Metropolis sweeps against an FP64 restatement of the same
schedule (oracle/replay.c, orc_o3_halfsweep: same Philox counters, proposals,
energy difference and acceptance rule).  One sweep from a downloaded state is
replayed colour by colour, the second colour from the GPU's own first-colour
result, so FP32 rounding cannot cascade.  Decisions must agree wherever the
acceptance margin |p - u| exceeds 1e-4 (FP32 energy and exp errors are orders of
magnitude smaller); accepted spins must equal the FP64 proposal to 2e-5 (FP32
rounding of the direction and normalisation) and rejected spins must not move."""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu

TOL_MARGIN = 1e-4
# bond tests: FP32 coupling products and expm1f err by ~1e-7 relative; uniforms are
# multiples of 2^-24
TOL_WOLFF = 1e-5


@pytest.fixture(scope="module")
def ctx():
    from marqov_amd import gpu
    return gpu.Context(0)


def _colour(dim, L):
    idx = np.arange(L ** dim)
    return sum((idx // L ** d) % L for d in range(dim)) & 1


@pytest.mark.parametrize("kernels", ["0", "1"])
@pytest.mark.parametrize("model,proposal,dim,L,beta", [
    ("heisenberg", "fresh", 3, 6, 0.693), ("heisenberg", "small", 3, 8, 0.693), ("heisenberg", "small", 2, 8, 1.2),
    ("phi4", "fresh", 3, 6, 0.69), ("phi4", "fresh", 2, 10, 0.9)])
def test_o3_sweep_decisions_match_fp64_restatement(ctx, model, proposal, dim, L, beta, kernels, monkeypatch):
    from marqov_amd import gpu
    monkeypatch.setenv("MARQOV_AMD_KERNELS", kernels)
    lat = gpu.RegularLattice(ctx, dim, L)
    nrep, seed, sweep0 = 3, 9, 3
    st = gpu.State(ctx, model, [lat] * nrep, J=-1.0, proposal=proposal, delta=0.6, lam=4.5, mass=1.0)
    st.set_beta(beta)
    st.init(seed=5)
    st.metropolis(sweep0, seed=seed, sweep0=0)
    s0 = [st.download(r).astype(np.float64) for r in range(nrep)]
    st.metropolis(1, seed=seed, sweep0=sweep0)
    s1 = [st.download(r).astype(np.float64) for r in range(nrep)]
    col = _colour(dim, L)
    kw = dict(phi4=(model == "phi4"), small_angle=(proposal == "small"), delta=0.6, lam=4.5, mass=1.0)
    n_checked = n_acc = n_near = 0
    for r in range(nrep):
        old = s0[r]
        for c in (0, 1):
            cpu, m = ob.o3_halfsweep(dim, L, old, c, 2 * sweep0 + c, r, seed, beta, J=-1.0, **kw)
            sites = col == c
            moved = np.any(s1[r] != old, axis=1)
            sure = sites & (np.abs(m) > TOL_MARGIN)
            n_near += int((sites & ~sure).sum())
            assert np.array_equal(moved[sure], (m > 0)[sure]), (r, c)
            acc = sure & (m > 0)
            assert np.abs(s1[r][acc] - cpu[acc]).max(initial=0.0) < 2e-5, (r, c)
            n_checked += int(sure.sum())
            n_acc += int(acc.sum())
            # the next colour sees the GPU's result for this one
            old = np.where(sites[:, None], s1[r], old)
    assert n_near <= 3, n_near
    assert 0.1 * n_checked < n_acc < 0.95 * n_checked, (n_acc, n_checked)


@pytest.mark.parametrize("model,embedding,dim,L,beta", [
    ("heisenberg", "cartesian", 3, 8, 0.8), ("heisenberg", "random", 3, 8, 0.8), ("heisenberg", "cartesian", 2, 16, 1.3),
    ("phi4", "random", 3, 8, 0.69), ("phi4", "random", 2, 16, 1.0)])
def test_o3_wolff_clusters_match_fp64_restatement(ctx, model, embedding, dim, L, beta):
    """Each O(3)/Phi4 cluster, grown by the GPU from a downloaded FP32 state, equals the
    FP64 restatement's cluster, the sites reachable from the seed along passing tests (same
    seed site, embedding and test uniforms): equal size, the same sites flipped, flipped
    spins to FP32 rounding.  A cluster is compared when no test of it lies within 1e-5 of
    its threshold."""
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, dim, L)
    nrep, seed = 2, 13
    st = gpu.State(ctx, model, [lat] * nrep, J=-1.0, proposal="small", delta=0.6, embedding=embedding,
                   lam=4.5, mass=1.0)
    st.set_beta(beta)
    st.init(seed=7)
    st.metropolis(5, seed=seed, sweep0=0)
    compared = skipped = 0
    sizes_seen = []
    # at least 8 clusters, more (up to 32) until one of more than 10 sites has been compared
    for k in range(32):
        if k >= 8 and compared >= 12 and max(sizes_seen) > 10:
            break
        pre = [st.download(r).astype(np.float64) for r in range(nrep)]
        sizes = st.wolff(1, seed=seed, cluster0=k)
        post = [st.download(r).astype(np.float64) for r in range(nrep)]
        for r in range(nrep):
            cpu, size, margin = ob.o3_wolff(dim, L, pre[r], embedding == "random" or model == "phi4", -1.0, beta,
                                            k, r, seed)
            if margin <= TOL_WOLFF:
                skipped += 1
                continue
            compared += 1
            sizes_seen.append(size)
            assert int(sizes[r, 0]) == size, (k, r)
            moved_gpu = np.any(post[r] != pre[r], axis=1)
            moved_cpu = np.any(cpu != pre[r], axis=1)
            assert np.array_equal(moved_gpu, moved_cpu), (k, r)
            assert np.abs(post[r] - cpu).max() < 2e-5, (k, r)
    assert compared >= 12 and skipped <= max(4, compared // 4), (compared, skipped)
    assert max(sizes_seen) > 10, sizes_seen


@pytest.mark.parametrize("model,embedding,beta", [("heisenberg", "cartesian", 1.2), ("heisenberg", "random", 1.2),
                                                  ("phi4", "random", 1.0)])
def test_csr_o3_wolff_clusters_match_fp64_restatement(ctx, model, embedding, beta):
    """The graph kernel's O(3)/Phi4 clusters (test uniforms keyed by the tested edge's
    CSR slot) against the FP64 restatement on a 2D random geometric graph of mean degree 8."""
    from marqov_amd import graphs, gpu
    n = 3000
    rp, ci, _ = graphs.rgg2d(n, graphs.rgg_radius(n, 8.0), 17)
    lat = gpu.CsrLattice(ctx, rp, ci)
    nrep, seed = 2, 19
    st = gpu.State(ctx, model, [lat] * nrep, J=-1.0, proposal="small", delta=0.6, embedding=embedding,
                   lam=4.5, mass=1.0)
    st.set_beta(beta)
    st.init(seed=8)
    st.metropolis(5, seed=seed, sweep0=0)
    compared = skipped = 0
    sizes_seen = []
    # at least 8 clusters, more (up to 32) until one of more than 10 sites has been compared
    for k in range(32):
        if k >= 8 and compared >= 12 and max(sizes_seen) > 10:
            break
        pre = [st.download(r).astype(np.float64) for r in range(nrep)]
        sizes = st.wolff(1, seed=seed, cluster0=k)
        post = [st.download(r).astype(np.float64) for r in range(nrep)]
        for r in range(nrep):
            cpu, size, margin = ob.csr_o3_wolff(rp, ci, pre[r], embedding == "random" or model == "phi4", -1.0, beta,
                                                k, r, seed)
            if margin <= TOL_WOLFF:
                skipped += 1
                continue
            compared += 1
            sizes_seen.append(size)
            assert int(sizes[r, 0]) == size, (k, r)
            assert np.array_equal(np.any(post[r] != pre[r], axis=1), np.any(cpu != pre[r], axis=1)), (k, r)
            assert np.abs(post[r] - cpu).max() < 2e-5, (k, r)
    assert compared >= 12 and skipped <= max(4, compared // 4), (compared, skipped)
    assert max(sizes_seen) > 10, sizes_seen


@pytest.mark.parametrize("model,proposal,beta", [("heisenberg", "fresh", 1.2), ("heisenberg", "small", 1.2),
                                                 ("phi4", "fresh", 1.0)])
def test_csr_o3_sweep_decisions_match_fp64_restatement(ctx, model, proposal, beta):
    """The graph kernel's FP32 O(3)/Phi4 sweep, colour class by colour class (counter
    phase 64 * sweep + colour), against the FP64 restatement on a 2D random geometric
    graph; each class is replayed from the GPU's result for the classes before it."""
    import ctypes
    from marqov_amd import graphs, gpu
    n = 2500
    rp, ci, _ = graphs.rgg2d(n, graphs.rgg_radius(n, 6.0), 23)
    lat = gpu.CsrLattice(ctx, rp, ci)
    colour = np.zeros(n, np.int32)
    ctx.check(ctx.lib.mq_lattice_color_of(lat.h, colour.ctypes.data_as(ctypes.c_void_p)), "mq_lattice_color_of")
    ncol = int(colour.max()) + 1
    nrep, seed, sweep0 = 3, 29, 2
    st = gpu.State(ctx, model, [lat] * nrep, J=-1.0, proposal=proposal, delta=0.6, lam=4.5, mass=1.0)
    st.set_beta(beta)
    st.init(seed=4)
    st.metropolis(sweep0, seed=seed, sweep0=0)
    s0 = [st.download(r).astype(np.float64) for r in range(nrep)]
    st.metropolis(1, seed=seed, sweep0=sweep0)
    s1 = [st.download(r).astype(np.float64) for r in range(nrep)]
    kw = dict(phi4=(model == "phi4"), small_angle=(proposal == "small"), delta=0.6, lam=4.5, mass=1.0)
    n_checked = n_acc = n_near = 0
    for r in range(nrep):
        old = s0[r]
        for c in range(ncol):
            cpu, m = ob.csr_o3_colour(rp, ci, colour, old, c, 64 * sweep0 + c, r, seed, beta, J=-1.0, **kw)
            sites = colour == c
            moved = np.any(s1[r] != old, axis=1)
            sure = sites & (np.abs(m) > TOL_MARGIN)
            n_near += int((sites & ~sure).sum())
            assert np.array_equal(moved[sure], (m > 0)[sure]), (r, c)
            acc = sure & (m > 0)
            assert np.abs(s1[r][acc] - cpu[acc]).max(initial=0.0) < 2e-5, (r, c)
            n_checked += int(sure.sum())
            n_acc += int(acc.sum())
            old = np.where(sites[:, None], s1[r], old)
    assert n_near <= 3, n_near
    assert 0.1 * n_checked < n_acc < 0.95 * n_checked, (n_acc, n_checked)

