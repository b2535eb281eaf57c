"""Statistical parity of the HIP path against the reference CPU build, in the form
the criterion states: energy, magnetisation and the Binder cumulant of the GPU
chains agree with the chains of the reference-faithful CPU restatement within 2
standard errors (asserted as a pre-registered chi^2 over all compared
observables, with every single |z| < 4; see stat_util).

Where the GPU's update differs from the reference's (random-direction Wolff for
Phi4, the graph path for continuous spins) the CPU side runs Metropolis only:
that move equals the reference build's bit for bit, and any valid Wolff or
colour-class update must reproduce its Boltzmann averages.  So these tests pin
the GPU's continuous-spin Wolff and graph kernels without relying on a CPU
restatement of the same algorithm.
"""
import numpy as np
import pytest

import oracle_bind as ob
import stat_util as su

pytestmark = pytest.mark.gpu

PHI4_INI = dict(beta=0.69, lam=4.5, mass=1.0)  # config/Phi4.ini (beta from its list)


def gpu_chains(ctx, model, lats, beta, nwarm, nmeas, ncluster, nmetro, seed, phi2_every=0, **kw):
    """obs[replica, t, k] with k = E/N, |m|, m^2, m^4 (mq_measure) and, when phi2_every > 0,
    the mean |phi|^2 of the downloaded spins every phi2_every EMCS as phi2[replica, t']."""
    from marqov_amd import gpu
    st = gpu.State(ctx, model, lats, **kw)
    st.set_beta(beta)
    st.init(seed=seed)
    for t in range(nwarm):
        st.emcs(ncluster, nmetro, seed, t)
    obs = np.zeros((len(lats), nmeas, 4))
    phi2 = []
    for t in range(nmeas):
        st.emcs(ncluster, nmetro, seed, nwarm + t)
        obs[:, t, :] = st.measure()
        if phi2_every and t % phi2_every == 0:
            phi2.append([float((st.download(r).astype(np.float64) ** 2).sum(1).mean()) for r in range(len(lats))])
    return obs, (np.array(phi2).T if phi2 else None)


def test_phi4_gpu_wolff_vs_cpu_metropolis_only(ctx):
    """3D L = 6 Phi4 at the config/Phi4.ini parameters.  GPU: 64 replicas, EMCS = 3
    random-direction Wolff clusters (Phi4.h:182-244) + 5 radial Metropolis sweeps, FP32.
    CPU: 16 Metropolis-only chains (5 sweeps per EMCS) of the restatement that equals the
    reference build bit for bit.  E/N, |m|, Binder U and <|phi|^2> agree."""
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, 3, 6)
    g, gphi2 = gpu_chains(ctx, "phi4", [lat] * 64, PHI4_INI["beta"], 200, 2000, 3, 5, 41, phi2_every=10, J=-1.0,
                          lam=PHI4_INI["lam"], mass=PHI4_INI["mass"], embedding="random")
    c = ob.ref_chains(ob.PHI4, 3, 6, PHI4_INI["beta"], 16, seed0=500, ncluster=0, nmetro=5, nwarm=200, nmeas=3000,
                      J=-1.0, lam=PHI4_INI["lam"], mass=PHI4_INI["mass"])
    rows = su.compare_chains(g, c, "phi4 wolff+metro (gpu) vs metro-only (cpu)")
    gm, gs = su.mean_err(gphi2.mean(1))
    cm, cs = su.mean_err(su.per_chain(c[:, :, 4]))
    rows.append(("phi4 wolff+metro (gpu) vs metro-only (cpu)", "|phi|^2", gm, gs, cm, cs, (gm - cm) / np.hypot(gs, cs)))
    su.assert_within_2sigma_combined(rows, "phi4 wolff pin")


@pytest.mark.parametrize("model,proposal,embedding", [("heisenberg", "small", "random"),
                                                      ("heisenberg", "fresh", "cartesian"),
                                                      ("phi4", "fresh", "random")])
def test_o3_graph_torus_vs_cpu_metropolis_only(ctx, model, proposal, embedding):
    """The graph path for continuous spins (greedy-coloured CSR sweeps k_csr_sweep_o3 and CSR
    Wolff k_csr_wolff) on the 3D L = 6 torus given as a CSR graph, 64 GPU replicas, against
    16 Metropolis-only chains of the restatement on the regular lattice: Heisenberg at
    beta = 0.693, Phi4 at the config/Phi4.ini parameters.  E/N, |m| and Binder U agree."""
    from marqov_amd import gpu, graphs
    phi4 = model == "phi4"
    beta = PHI4_INI["beta"] if phi4 else 0.693
    rp, ci = graphs.regular_csr(3, 6)
    lat = gpu.CsrLattice(ctx, rp, ci)
    g, _ = gpu_chains(ctx, model, [lat] * 64, beta, 200, 2000, 2, 3, 43, J=-1.0, lam=PHI4_INI["lam"],
                      mass=PHI4_INI["mass"], proposal=proposal, delta=0.6, embedding=embedding)
    c = ob.ref_chains(ob.PHI4 if phi4 else ob.HEIS, 3, 6, beta, 16, seed0=600, ncluster=0, nmetro=3, nwarm=300,
                      nmeas=4000, J=-1.0, lam=PHI4_INI["lam"], mass=PHI4_INI["mass"])
    rows = su.compare_chains(g, c, f"{model} graph torus {proposal}/{embedding} (gpu) vs metro-only (cpu)")
    su.assert_within_2sigma_combined(rows, "o3 graph pin")


def test_binder_energy_magnetisation_gpu_vs_cpu_build(ctx):
    """E/N, |m| and the Binder cumulant U = 1 - <m^4>/(3<m^2>^2) of GPU chains against chains of
    the reference-faithful CPU restatement, for the four models of the path:
      Ising 2D L = 16 at beta_c (1 Wolff cluster + 1 sweep per EMCS, both sides);
      Potts-3 2D L = 8 at beta = 1.0 (1 cluster + 1 sweep);
      Heisenberg 3D L = 8 at beta = 0.693 with the C3 update on the GPU (small-angle step 4,
        Cartesian Wolff) and the reference's fresh-direction Metropolis + Cartesian Wolff on the CPU;
      Phi4 3D L = 6 at the config/Phi4.ini parameters (3 Wolff + 5 sweeps; the CPU's Wolff is
        pinned against its Metropolis-only chains on the CPU).
    12 z-scores, chi^2 below its 95% quantile, every |z| < 4."""
    from marqov_amd import gpu
    rows = []
    cases = [
        ("ising 2D L=16 beta_c", "ising", ob.ISING, 2, 16, 0.4406868, 1, 1, {}, {}),
        ("potts-3 2D L=8 beta=1.0", "potts", ob.POTTS, 2, 8, 1.0, 1, 1, dict(q=3), dict(Q=3)),
        ("heisenberg 3D L=8 C3 update", "heisenberg", ob.HEIS, 3, 8, 0.693, 1, 1,
         dict(proposal="small", delta=4.0, embedding="cartesian"), {}),
        ("phi4 3D L=6 Phi4.ini", "phi4", ob.PHI4, 3, 6, PHI4_INI["beta"], 3, 5,
         dict(lam=PHI4_INI["lam"], mass=PHI4_INI["mass"], embedding="random"),
         dict(lam=PHI4_INI["lam"], mass=PHI4_INI["mass"])),
    ]
    for k, (label, gm, cm, dim, L, beta, ncl, nme, gkw, ckw) in enumerate(cases):
        lat = gpu.RegularLattice(ctx, dim, L)
        g, _ = gpu_chains(ctx, gm, [lat] * 128, beta, 300, 2000, ncl, nme, 51 + k, J=-1.0, **gkw)
        c = ob.ref_chains(cm, dim, L, beta, 32, seed0=700 + 100 * k, ncluster=ncl, nmetro=nme, nwarm=300,
                          nmeas=3000, J=-1.0, **ckw)
        rows += su.compare_chains(g, c, label)
    su.assert_within_2sigma_combined(rows, "binder")
