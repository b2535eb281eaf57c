"""Statistical parity of the HIP path (FP32 local arithmetic, FP64 observables)
against the reference-faithful FP64 CPU restatement and exact results.

Independent replicas give clean error bars: per-replica time averages, then
mean +- standard error over replicas (GPU) or seeds (CPU).  The tolerance is
4 combined standard errors."""
import numpy as np
import pytest

import oracle_bind as ob
import exact_path  # noqa: F401
import exact

pytestmark = pytest.mark.gpu


def gpu_series(ctx, model, dim, L, beta, nrep, nwarm, nmeas, ncluster=1, nmetro=1, seed=3, **kw):
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, dim, L)
    st = gpu.State(ctx, model, [lat] * nrep, **kw)
    st.set_beta(beta)
    st.init(seed=seed)
    for t in range(nwarm):
        st.emcs(ncluster, nmetro, seed, t)
    obs = np.zeros((nmeas, nrep, 4))
    for t in range(nmeas):
        st.emcs(ncluster, nmetro, seed, nwarm + t)
        obs[t] = st.measure()
    return obs, st


def cpu_series(model, dim, L, beta, nseeds, nwarm, nmeas, ncluster=1, nmetro=1, **kw):
    res = []
    for s in range(nseeds):
        _, o, _ = ob.ref_run(model, dim, L, beta, ncluster=ncluster, nmetro=nmetro, nwarm=nwarm, nmeas=nmeas,
                             seed=100 + s, **kw)
        res.append(o)
    return np.array(res)  # seeds x nmeas x 2


def agree(gpu_per_rep, cpu_per_seed, k=4.0, label=""):
    g, c = np.asarray(gpu_per_rep), np.asarray(cpu_per_seed)
    sg = g.std(ddof=1) / np.sqrt(len(g))
    sc = c.std(ddof=1) / np.sqrt(len(c))
    d = abs(g.mean() - c.mean())
    print(f"[fp32-vs-fp64] {label}: gpu {g.mean():.6f} +- {sg:.6f}  cpu {c.mean():.6f} +- {sc:.6f}  "
          f"diff {d:.6f} = {d / np.hypot(sg, sc):.2f} sigma", flush=True)
    return d < k * np.hypot(sg, sc), (g.mean(), sg, c.mean(), sc)


@pytest.mark.parametrize("proposal,embedding", [("fresh", "cartesian"), ("small", "random")])
def test_heisenberg_vs_fp64_restatement(ctx, proposal, embedding):
    beta = 0.693
    obs, _ = gpu_series(ctx, "heisenberg", 3, 6, beta, 64, 100, 400, J=-1.0, proposal=proposal, delta=0.6,
                        embedding=embedding)
    cpu = cpu_series(ob.HEIS, 3, 6, beta, 12, 100, 1500, J=-1.0)
    ok_e, ie = agree(obs[:, :, 0].mean(0), cpu[:, :, 0].mean(1), label=f"heisenberg {proposal}/{embedding} E/N")
    ok_m, im = agree(obs[:, :, 1].mean(0), cpu[:, :, 1].mean(1), label=f"heisenberg {proposal}/{embedding} |m|")
    assert ok_e, ie
    assert ok_m, im


def test_heisenberg_c3_schedule_L12_vs_fp64_restatement(ctx):
    """The C3 configuration's update (small-angle step delta = 4, about 50% acceptance, and
    Cartesian-embedding Wolff) at beta = 0.693 on 3D L = 12: FP32 GPU replicas against the
    FP64 restatement of the reference's fresh-direction Metropolis and Cartesian Wolff."""
    beta = 0.693
    obs, _ = gpu_series(ctx, "heisenberg", 3, 12, beta, 64, 100, 1500, J=-1.0, proposal="small", delta=4.0,
                        embedding="cartesian")
    cpu = cpu_series(ob.HEIS, 3, 12, beta, 16, 100, 1500, J=-1.0)
    ok_e, ie = agree(obs[:, :, 0].mean(0), cpu[:, :, 0].mean(1), label="heisenberg C3 schedule L=12 E/N")
    ok_m, im = agree(obs[:, :, 1].mean(0), cpu[:, :, 1].mean(1), label="heisenberg C3 schedule L=12 |m|")
    assert ok_e, ie
    assert ok_m, im


def test_phi4_vs_fp64_restatement(ctx):
    beta, lam, mass = 0.69, 4.5, 1.0  # config/Phi4.ini
    obs, _ = gpu_series(ctx, "phi4", 3, 6, beta, 64, 100, 400, J=-1.0, lam=lam, mass=mass)
    cpu = cpu_series(ob.PHI4, 3, 6, beta, 12, 100, 1500, J=-1.0, lam=lam, mass=mass)
    ok_e, ie = agree(obs[:, :, 0].mean(0), cpu[:, :, 0].mean(1), label="phi4 E/N")
    ok_m, im = agree(obs[:, :, 1].mean(0), cpu[:, :, 1].mean(1), label="phi4 |m|")
    assert ok_e, ie
    assert ok_m, im


def test_potts3_2d_vs_restatement(ctx):
    beta = 0.9
    obs, _ = gpu_series(ctx, "potts", 2, 8, beta, 64, 100, 400, J=-1.0, q=3)
    cpu = cpu_series(ob.POTTS, 2, 8, beta, 12, 100, 1500, J=-1.0, Q=3)
    ok_e, ie = agree(obs[:, :, 0].mean(0), cpu[:, :, 0].mean(1), label="potts3 E/N")
    ok_m, im = agree(obs[:, :, 1].mean(0), cpu[:, :, 1].mean(1), label="potts3 m")
    assert ok_e, ie
    assert ok_m, im


def test_ising_c1_energy_vs_kaufman(ctx):
    """2D Ising EMCS schedule (10 Wolff + 1 sweep per EMCS) at L=64, beta_c, many replicas."""
    L, beta = 64, 0.4406868
    obs, _ = gpu_series(ctx, "ising", 2, L, beta, 256, 200, 300, ncluster=10, nmetro=1, J=-1.0)
    e = obs[:, :, 0].mean(0)
    ka = exact.kaufman_energy(beta, L, L)
    se = e.std(ddof=1) / np.sqrt(len(e))
    assert abs(e.mean() - ka) < 4 * se, (e.mean(), se, ka)
    m2, m4 = obs[:, :, 2].mean(), obs[:, :, 3].mean()
    binder = 1 - m4 / (3 * m2 * m2)
    assert 0.55 < binder < 0.66, binder  # U* = 0.6107 (Salas & Sokal 2000) at the critical point


def test_potts3_critical_energy_extrapolates_to_exact(ctx):
    """2D Potts-3 at its exact critical point beta_c = ln(1 + sqrt 3): the energy per site of
    L = 32, 64, 128 tori, fitted as E(L) = E_inf + a L^((alpha - 1)/nu) with the q = 3
    exponents alpha = 1/3, nu = 5/6 (so L^-0.8), must extrapolate to the exact critical energy
    -(1 + 1/sqrt 3) (Baxter; two bonds per site, E = J sum delta(s_i, s_j), J = -1) within 4
    standard errors, and the three sizes must follow the fitted form."""
    bc = np.log(1.0 + np.sqrt(3.0))
    sizes = np.array([32.0, 64.0, 128.0])
    means, errs = [], []
    for L in sizes.astype(int):
        obs, _ = gpu_series(ctx, "potts", 2, int(L), bc, 256, 300, 1000, ncluster=4, nmetro=1, seed=4, J=-1.0, q=3)
        e = obs[:, :, 0].mean(0)
        means.append(e.mean())
        errs.append(e.std(ddof=1) / np.sqrt(len(e)))
    E, sd = np.array(means), np.array(errs)
    A = np.vstack([np.ones(3), sizes ** -0.8]).T
    W = np.diag(1.0 / sd ** 2)
    cov = np.linalg.inv(A.T @ W @ A)
    p = cov @ A.T @ W @ E
    chi2 = float((((E - A @ p) / sd) ** 2).sum())
    exact_e = -(1.0 + 1.0 / np.sqrt(3.0))
    z = (p[0] - exact_e) / np.sqrt(cov[0, 0])
    print(f"[potts3 critical] E(L) = {np.round(E, 5).tolist()}  E_inf = {p[0]:.5f} +- {np.sqrt(cov[0, 0]):.5f}  "
          f"exact {exact_e:.5f}  ({z:.2f} sigma)  chi2 = {chi2:.2f}", flush=True)
    assert abs(z) < 4.0, (p, cov, E, sd)
    assert chi2 < 10.0, (chi2, E, sd)


def test_ising_graph_path_torus_vs_kaufman(ctx):
    """The graph path (greedy-coloured CSR sweeps and CSR Wolff) on the 2D 32 x 32 torus given
    as a graph, at beta_c: the energy equals Kaufman's exact finite-torus value within 4
    standard errors."""
    from marqov_amd import gpu, graphs
    L, beta = 32, 0.4406868
    rp, ci = graphs.regular_csr(2, L)
    lat = gpu.CsrLattice(ctx, rp, ci)
    R = 256
    st = gpu.State(ctx, "ising", [lat] * R, J=-1.0)
    st.set_beta(beta)
    st.init(seed=6)
    for t in range(200):
        st.emcs(4, 1, 6, t)
    e = np.zeros(R)
    n = 800
    for t in range(n):
        st.emcs(4, 1, 6, 200 + t)
        e += st.measure()[:, 0]
    e /= n
    se = e.std(ddof=1) / np.sqrt(R)
    ka = exact.kaufman_energy(beta, L, L)
    print(f"[graph torus] E/N = {e.mean():.5f} +- {se:.5f}  Kaufman {ka:.5f}  ({(e.mean() - ka) / se:.2f} sigma)", flush=True)
    assert abs(e.mean() - ka) < 4 * se, (e.mean(), se, ka)


def test_ising3d_binder_at_critical_point(ctx):
    """3D Ising at beta_c = 0.2216546 (Ferrenberg-Xu-Landau 2018), L = 32 and 64 (the row-vectorised
    headline kernel and Wolff): the Binder cumulant U = 1 - <m^4>/(3 <m^2>^2) lies at the universal
    value U* = 0.46548 (Hasenbusch 2010) within 4 jackknife errors plus 0.01 for corrections to
    scaling at these sizes (a sanity band: there is no closed form in 3D)."""
    from marqov_amd import gpu
    ustar = 0.46548
    for L in (32, 64):
        lat = gpu.RegularLattice(ctx, 3, L)
        R, n = 128, 1000
        st = gpu.State(ctx, "ising", [lat] * R, J=-1.0)
        st.set_beta(0.2216546)
        st.init(seed=8)
        for t in range(200):
            st.emcs(2, 1, 8, t)
        m2 = np.zeros(R)
        m4 = np.zeros(R)
        for t in range(n):
            st.emcs(2, 1, 8, 200 + t)
            o = st.measure()
            m2 += o[:, 2]
            m4 += o[:, 3]
        m2 /= n
        m4 /= n
        u = 1 - m4.mean() / (3 * m2.mean() ** 2)
        jk = np.array([1 - np.delete(m4, i).mean() / (3 * np.delete(m2, i).mean() ** 2) for i in range(R)])
        err = np.sqrt((R - 1) / R * ((jk - jk.mean()) ** 2).sum())
        print(f"[3d ising binder] L={L}: U = {u:.5f} +- {err:.5f} (U* = {ustar})", flush=True)
        assert abs(u - ustar) < 4 * err + 0.01, (L, u, err)
        del st


def test_heisenberg_renormalize_keeps_unit_length(ctx):
    _, st = gpu_series(ctx, "heisenberg", 3, 8, 0.693, 2, 0, 50, J=-1.0, proposal="small", delta=0.5)
    s = st.download(0).astype(np.float64)
    drift = np.abs(np.linalg.norm(s, axis=1) - 1).max()
    st.renormalize()
    s2 = st.download(0).astype(np.float64)
    assert np.abs(np.linalg.norm(s2, axis=1) - 1).max() <= 1e-6
    assert drift < 1e-4


def test_full_size_c2_properties(ctx):
    """3D Ising at L=256: measured energy equals the energy of the downloaded
    state; a sweep at beta = 0 accepts every move, so it flips every spin exactly once."""
    from marqov_amd import gpu
    L = 256
    lat = gpu.RegularLattice(ctx, 3, L)
    st = gpu.State(ctx, "ising", [lat], J=-1.0)
    st.set_beta(0.2216546)
    st.init(seed=2)
    st.metropolis(20, seed=2)
    s = st.download(0)
    e_host = -1.0 * ob.lib().orc_bond_sum_ising(3, L, ob.p(s), None) / L ** 3
    obs = st.measure()
    assert obs[0, 0] == pytest.approx(e_host, abs=1e-12)
    assert -1.2 < obs[0, 0] < -0.9
    st.set_beta(0.0)
    st.metropolis(1, seed=3, sweep0=100)
    assert np.array_equal(st.download(0), -s)  # every site is offered one flip per sweep


@pytest.mark.parametrize("model", ["heisenberg", "phi4"])
def test_fp64_observable_accumulation(ctx, model):
    """GPU FP64 sums over FP32 spins equal an FP64 CPU recomputation from the
    downloaded spins to round-off (relative error stated: < 1e-12)."""
    from marqov_amd import gpu
    L, beta = 16, 0.693
    lat = gpu.RegularLattice(ctx, 3, L)
    st = gpu.State(ctx, model, [lat, lat], J=-1.0, lam=4.5, mass=1.0)
    st.set_beta(beta)
    st.init(seed=5)
    st.metropolis(10, seed=5)
    obs = st.measure()
    nb = exact.neighbour_table(3, L)
    for r in range(2):
        s = st.download(r).astype(np.float64)
        N = L ** 3
        e = -1.0 * (s[:, None, :] * s[nb]).sum() / (2 * N)
        if model == "phi4":
            n2 = (s * s).sum(1)
            e += (1.0 / beta) * n2.sum() / N + (4.5 / beta) * ((n2 - 1) ** 2).sum() / N
        m = np.linalg.norm(s.sum(0)) / N
        assert abs(obs[r, 0] - e) <= 1e-12 * max(1.0, abs(e)), (obs[r, 0], e)
        assert abs(obs[r, 1] - m) <= 1e-12, (obs[r, 1], m)


@pytest.mark.parametrize("model,L", [("heisenberg", 16), ("phi4", 16), ("heisenberg", 64)])
def test_batched_equals_plain(ctx, model, L):
    """Batching changes the launch, not the arithmetic: replica 0 of a batch and a
    one-replica state run the same FP32 sweeps and Wolff clusters to bit-identical
    spins, and their FP64 observables agree to summation round-off.  So the batched
    path's observable error is the plain path's."""
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, 3, L)
    kw = dict(J=-1.0, proposal="small", delta=0.6, lam=4.5, mass=1.0)
    nb = 8 if L == 64 else 64
    states = [gpu.State(ctx, model, [lat] * nb, **kw), gpu.State(ctx, model, [lat], **kw)]
    for st in states:
        st.set_beta(0.693)
        st.init(seed=5)
        st.metropolis(6, seed=9)
        st.wolff(3, seed=9, cluster0=0)
        st.metropolis(2, seed=9, sweep0=6)
    assert np.array_equal(states[0].download(0), states[1].download(0))
    ob_b, ob_p = states[0].measure()[0], states[1].measure()[0]
    assert np.allclose(ob_b, ob_p, rtol=1e-12, atol=1e-14), (ob_b, ob_p)


@pytest.mark.parametrize("model,embedding,lattice", [("heisenberg", "cartesian", "regular"),
                                                     ("phi4", "random", "regular"),
                                                     ("heisenberg", "random", "graph")])
def test_vec_wolff_block_invariant(ctx, model, embedding, lattice, monkeypatch):
    """The O(3)/Phi4 cluster is the connected component of activated bonds, whose
    activation depends only on the two end spins before the flip and on the bond's
    Philox uniform; so the FP32 spins after Wolff moves are bit-identical whether the
    workgroup growing a cluster has 256 or 1024 threads."""
    from marqov_amd import gpu
    if lattice == "regular":
        lat = gpu.RegularLattice(ctx, 3, 16)
    else:
        from marqov_amd import graphs
        n = 4096
        rp, ci, _ = graphs.rgg2d(n, graphs.rgg_radius(n, 8.0), 11)
        lat = gpu.CsrLattice(ctx, rp, ci)
    out = {}
    for block in ("256", "1024"):
        st = gpu.State(ctx, model, [lat] * 3, J=-1.0, proposal="small", delta=0.6, embedding=embedding)
        st.set_tuning(wolff_block=int(block))
        st.set_beta(0.9 if lattice == "regular" else 1.2)
        st.init(seed=21)
        sizes = []
        for t in range(4):
            st.metropolis(3, seed=22, sweep0=3 * t)
            sizes.append(st.wolff(5, seed=23, cluster0=5 * t))
        out[block] = (np.concatenate(sizes, axis=1), [st.download(r) for r in range(3)])
    assert out["256"][0].max() > 20, out["256"][0]
    assert np.array_equal(out["256"][0], out["1024"][0])
    for r in range(3):
        assert np.array_equal(out["256"][1][r], out["1024"][1][r]), r


@pytest.mark.parametrize("model,embedding,lattice", [("heisenberg", "cartesian", "regular"),
                                                     ("phi4", "random", "regular"),
                                                     ("heisenberg", "random", "graph"),
                                                     ("phi4", "cartesian", "graph")])
def test_vec_wolff_visited_set_invariant(ctx, model, embedding, lattice, monkeypatch):
    """O(3)/Phi4 clusters keep their visited set in LDS while they cannot outgrow
    the tuning's wolff_vset sites and move it to the global bitmap when a level could
    pass that limit.  Membership does not depend on where it is recorded, so the spins
    and cluster sizes are bit-identical with the set off (0), with a tiny limit (16: nearly
    every cluster moves mid-growth) and with the default (2048)."""
    from marqov_amd import gpu
    if lattice == "regular":
        lat = gpu.RegularLattice(ctx, 3, 16)
    else:
        from marqov_amd import graphs
        n = 4096
        rp, ci, _ = graphs.rgg2d(n, graphs.rgg_radius(n, 8.0), 11)
        lat = gpu.CsrLattice(ctx, rp, ci)
    out = {}
    for vset in ("0", "16", "2048"):
        st = gpu.State(ctx, model, [lat] * 3, J=-1.0, proposal="small", delta=0.6, embedding=embedding)
        st.set_tuning(wolff_vset=int(vset))
        st.set_beta(0.9 if lattice == "regular" else 1.2)
        st.init(seed=31)
        sizes = []
        for t in range(4):
            st.metropolis(3, seed=32, sweep0=3 * t)
            sizes.append(st.wolff(6, seed=33, cluster0=6 * t))
        out[vset] = (np.concatenate(sizes, axis=1), [st.download(r) for r in range(3)])
    s0 = out["0"][0]
    print(f"[vset] {model}/{embedding}/{lattice}: clusters > 16: {(s0 > 16).sum()} of {s0.size}, "
          f"largest {s0.max()}", flush=True)
    assert s0.max() > 20, s0  # the limit-16 run moved at least one cluster mid-growth
    for vset in ("16", "2048"):
        assert np.array_equal(out[vset][0], s0), vset
        for r in range(3):
            assert np.array_equal(out[vset][1][r], out["0"][1][r]), (vset, r)


def test_heisenberg_renormalisation_cadence_bounds_drift(ctx):
    """FP32 reflections about a random direction drift |s| by rounding.  3D L = 8 Heisenberg,
    16 replicas, 10^5 EMCS of 4 random-direction Wolff clusters each: without renormalisation
    the largest ||s| - 1| grows well past FP32 rounding; with the FP64 renormalisation every
    1000 EMCS (GpuEMCS's default cadence) it stays at the drift of at most 1000 EMCS."""
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, 3, 8)
    drift = {}
    for renorm in (False, True):
        st = gpu.State(ctx, "heisenberg", [lat] * 16, J=-1.0, embedding="random")
        st.set_beta(0.693)
        st.init(seed=12)
        worst = 0.0
        for call in range(100):  # 1000 EMCS x 4 clusters per call
            st.wolff(4000, seed=13, cluster0=4000 * call, sizes=False)
            if call == 99:  # 1000 EMCS after the last renormalisation
                worst = max(float(np.abs(np.linalg.norm(st.download(r).astype(np.float64), axis=1) - 1).max())
                            for r in range(16))
            if renorm:
                st.renormalize()
        drift[renorm] = worst
    print(f"[renormalisation] max ||s|-1| after 10^5 EMCS: without {drift[False]:.2e}, "
          f"with every 1000 EMCS {drift[True]:.2e}", flush=True)
    assert drift[True] < 1e-5 and drift[True] < drift[False] / 3, drift
