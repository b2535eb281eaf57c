#!/usr/bin/env python3
"""Throughput of the MARQOV update path on one MI355X.

Headline (value): 3D simple-cubic Ising, L=256 (16.8M sites), periodic, J=-1,
beta=0.2216546, int8 spins in the colour-split checkerboard layout, Philox4x32-10,
R=32 replicas of that lattice swept together (working set 512 MiB, above the
256 MB Infinity Cache).  A step is one full checkerboard Metropolis sweep of
every replica; value = spin-flip attempts/ns over the timed steps.

Beside it on the same JSON line (skip with --quick):
  plain        the same sweep, one replica per launch
  wolff        Wolff clusters on the thermalised 32-replica state vs one replica, and the
               next 100 clusters per replica in one persistent launch (no per-call lockstep)
  emcs         20 EMCS (10 clusters + 1 sweep) of that state: separate launches for the
               batch vs one persistent launch with each replica at its own pace
  c1_2d_L64    2D Ising L=64 EMCS schedule (10 Wolff + 1 sweep per EMCS), batched vs plain
  c3_heis_L128 Heisenberg FP32 sweeps + Wolff (separate calls and one persistent launch),
               batched vs plain
  c4_ea_batch  512 +-J samples x 32 temperatures, L in {8,12,16,24}, HIP-graph
               captured {sweep, PT swap}, batched vs plain
  c5_rgg       2D random geometric graph N=2^20, Ising and Potts-3, 64 replicas vs 1
  roofline     the headline sweep kernel's algorithmic HBM rate (3 B/attempt)
  cpu_baseline the reference's own Metropolis move compiled from its headers
               (oracle/_ref; the reference-faithful CPU restatement when that is not
               built) and the restatement's stack Wolff on 2D L=64, one simulation
               per thread on this host's cores
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per Wolff cluster site: (z + 1) neighbour tests of s_r bytes + the
# flip write + a 4 B push and a 4 B pop, plus 4 index bytes per neighbour on a graph:
# 2D Ising int8 13 B, 3D Ising int8 15 B, 3D Heisenberg with the Cartesian component
# (4 B of FP32) 36 B, the RGG of mean degree 6 39 B.
WOLFF_BYTES = {"2d_int8": 13.0, "3d_int8": 15.0, "3d_cartesian": 36.0, "rgg_int8": 39.0}

# thermalised states handed from the GPU legs to the CPU baseline (the same workload)
CPU_WOLFF_INPUTS = {}


def wolff_frac(sites_per_ns, kind):
    return round(sites_per_ns * WOLFF_BYTES[kind] / HBM_PEAK_GBS, 6)


def tune_wolff(ctx, st, nclusters):
    """The project's Wolff autotuner on this state (same clusters for every candidate, state restored)."""
    from marqov_amd import autotune
    shape, _ = autotune.tune_wolff(ctx, st, nclusters=nclusters, force=True,
                                   cache_path=os.path.join("/tmp", f"marqov_amd_tuning_{os.getuid()}.json"))
    return shape

# HBM bytes per headline launch (3D L=256, 32 replicas) by launch shape (block, XCD remap), from
# rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (profiles/k_sweep_ising3_row16_3d_L256_R32.md)
HEADLINE_TRAFFIC = {(64, 0): 872.6e6, (64, 1): 807.6e6, (128, 0): 839.1e6, (128, 1): 807.6e6,
                    (256, 0): 822.3e6, (256, 1): 807.5e6}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--replicas", type=int, default=32)
    ap.add_argument("--L", type=int, default=256)
    ap.add_argument("--quick", action="store_true", help="headline only (no configuration legs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="headline with the default launch shape (no autotuner)")
    ap.add_argument("--shape", default="", help="headline launch shape BLOCK,XCD_REMAP instead of the autotuner's")
    ap.add_argument("--legs", default="", help="comma-separated subset of the extra legs to run")
    ap.add_argument("--legs-only", action="store_true",
                    help="with --legs: skip the headline state's Wolff and plain comparisons (for profiling legs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (counter calibration for HBM traffic)")
    ap.add_argument("--locate-c5-beta", action="store_true",
                    help="only locate the pseudo-critical betas of the 2D RGG legs by Binder-cumulant crossings")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed(ctx, fn):
    ctx.sync()
    t0 = time.perf_counter()
    ctx.timer_begin()
    fn()
    ms = ctx.timer_end()
    return ms, time.perf_counter() - t0


def timed_wolff(ctx, st, nclusters, seed, cluster0):
    """One timed mq_wolff call; returns (ms, sizes) of exactly the clusters timed.  The copy
    of the sizes (nrep x nclusters int64) is inside the timed region."""
    box = {}
    ms, _ = timed(ctx, lambda: box.setdefault("sizes", st.wolff(nclusters, seed, cluster0=cluster0, sizes=True)))
    return ms, box["sizes"]


def wolff_rate(ctx, st, rounds, nsw, ncl, seed, ncalls, sweep0, cluster0=0, tune_ncl=0):
    """Wolff throughput on a thermalised state: `rounds` x (nsw sweeps + ncl clusters)
    untimed, then (with tune_ncl > 0) the Wolff autotuner on tune_ncl clusters, then
    `ncalls` timed calls of ncl clusters per replica.  A call lasts as long as its slowest
    replica's clusters, so this is the batch's real rate.  Returns (sites/ns, mean cluster
    size of the timed clusters, the Wolff launch shape used, the next cluster counter)."""
    c, sw = cluster0, sweep0
    for _ in range(rounds):
        st.metropolis(nsw, seed, sweep0=sw)
        sw += nsw
        st.wolff(ncl, seed, cluster0=c, sizes=False)
        c += ncl
    shape = tune_wolff(ctx, st, tune_ncl) if tune_ncl > 0 else None
    tn = st.get_tuning()
    used = {k: tn[k] for k in ("wolff_block", "wolff_ring", "wolff_vset")}
    ms_tot, sites, n = 0.0, 0, 0
    for _ in range(ncalls):
        ms, sizes = timed_wolff(ctx, st, ncl, seed, c)
        c += ncl
        ms_tot += ms
        sites += int(sizes.sum())
        n += sizes.size
    return sites / (ms_tot * 1e6), sites / n, {"tuned": shape is not None, **used}, c


def persistent_wolff_rate(ctx, st, nsteps, ncl, seed, cluster0):
    """The same clusters as nsteps mq_wolff calls of ncl clusters (cluster0 a multiple of ncl),
    grown in one persistent mq_emcs_run launch with no sweeps: each replica runs its
    nsteps * ncl clusters without waiting for the other replicas at every call.  Returns
    (sites/ns, mean cluster size, max over replicas / mean of the sites grown)."""
    assert cluster0 % ncl == 0
    box = {}
    ms, _ = timed(ctx, lambda: box.setdefault("mc", st.emcs_run(nsteps, ncl, 0, seed, cluster0 // ncl)))
    per_rep = box["mc"].sum(axis=0) * ncl
    return float(per_rep.sum()) / (ms * 1e6), float(box["mc"].mean()), float(per_rep.max() / per_rep.mean())


def emcs_compare(ctx, st, nsteps, ncluster, nmetro, seed, step0):
    """nsteps EMCS (ncluster Wolff clusters + nmetro sweeps) of every replica, timed twice
    from the same state: as separate mq_wolff / mq_metropolis launches for the whole batch,
    and as one persistent mq_emcs_run launch in which each replica runs at its own pace.
    Both give the same states (replicas 0 and R-1 compared)."""
    saved = [st.download(r) for r in range(st.nrep)]

    def separate():
        for k in range(nsteps):
            st.wolff(ncluster, seed, cluster0=(step0 + k) * ncluster, sizes=False)
            st.metropolis(nmetro, seed, sweep0=(step0 + k) * nmetro)

    ms_sep, _ = timed(ctx, separate)
    ends = [st.download(r) for r in (0, st.nrep - 1)]
    for r, s in enumerate(saved):
        st.upload(r, s)
    box = {}
    ms_per, _ = timed(ctx, lambda: box.setdefault("mc", st.emcs_run(nsteps, ncluster, nmetro, seed, step0)))
    same = all(np.array_equal(a, st.download(r)) for a, r in zip(ends, (0, st.nrep - 1)))
    sites = float(box["mc"].sum()) * ncluster
    return {"emcs": f"{nsteps} x ({ncluster} Wolff clusters + {nmetro} sweep) per replica", "replicas": st.nrep,
            "separate_emcs_per_s": round(nsteps * 1e3 / ms_sep, 2),
            "persistent_emcs_per_s": round(nsteps * 1e3 / ms_per, 2),
            "persistent_over_separate": round(ms_sep / ms_per, 2),
            "persistent_wolff_sites_per_ns_incl_sweeps": round(sites / (ms_per * 1e6), 4),
            "mean_cluster": round(float(box["mc"].mean()), 1), "states_equal": bool(same)}


def leg_c2_msc(gpu, ctx, L=256, R=32, steps=100):
    """3D Ising L=256, multi-spin-coded variant: 1 bit per spin, R replicas, exact bit-serial acceptance."""
    lat = gpu.RegularLattice(ctx, 3, L)
    N = L ** 3
    res = {}
    for name, nrep in (("batched", R), ("plain", 1)):
        st = gpu.State(ctx, "ising", [lat] * nrep, J=-1.0, spin_bits=1)
        st.set_beta(0.2216546)
        st.init(seed=2)
        st.metropolis(10, seed=2, sweep0=0)
        ms, _ = timed(ctx, lambda: st.metropolis(steps, seed=2, sweep0=10))
        rate = nrep * N * steps / (ms * 1e6)
        res[name] = {"replicas": nrep, "attempts_per_ns": round(rate, 3),
                     "roofline_frac": round(0.375 * rate / HBM_PEAK_GBS, 4)}
        if name == "batched":
            res[name]["e"] = float(st.measure()[:, 0].mean())
        del st
    res["batched_over_plain"] = round(res["batched"]["attempts_per_ns"] / res["plain"]["attempts_per_ns"], 2)
    return res


C1_REPLICAS = int(os.environ.get("MARQOV_BENCH_C1_REPLICAS", "16384"))


def leg_c1(gpu, ctx):
    """2D Ising L=64 at beta_c, EMCS = 10 Wolff + 1 sweep; C1_REPLICAS independent replicas
    (Wolff throughput grows with the clusters in flight: one wavefront grows one replica's
    clusters, and 16,384 bit-packed replicas keep every SIMD at its wave limit)."""
    lat = gpu.RegularLattice(ctx, 2, 64)
    res = {}
    for name, nrep, nstep in (("batched", C1_REPLICAS, 20), ("plain", 1, 20)):
        st = gpu.State(ctx, "ising", [lat] * nrep, J=-1.0)
        st.set_beta(0.4406868)
        st.init(seed=1)
        for t in range(10):
            st.emcs(10, 1, 1, t)
        st.wolff(10, 1, cluster0=10 ** 6, sizes=False)
        ms_w, sizes = timed_wolff(ctx, st, 10 * nstep, 1, 2 * 10 ** 6)
        mean_cl = float(sizes.mean())
        ms_m, _ = timed(ctx, lambda: st.metropolis(nstep, 1, sweep0=10 ** 5))
        w = float(sizes.sum()) / (ms_w * 1e6)
        res[name] = {"replicas": nrep, "wolff_sites_per_ns": round(w, 4),
                     "wolff_roofline_frac": wolff_frac(w, "2d_int8"),
                     "attempts_per_ns": round(nrep * 4096 * nstep / (ms_m * 1e6), 3),
                     "mean_cluster": round(mean_cl, 1),
                     "emcs_per_s": round(nrep * nstep / ((ms_w + ms_m) * 1e-3), 1)}
        del st
    res["batched_over_plain_attempts"] = round(res["batched"]["attempts_per_ns"] / res["plain"]["attempts_per_ns"], 2)
    res["batched_over_plain_wolff"] = round(res["batched"]["wolff_sites_per_ns"] / res["plain"]["wolff_sites_per_ns"], 2)
    return res


C3_REPLICAS = int(os.environ.get("MARQOV_BENCH_C3_REPLICAS", "256"))
# small-angle step s' = (s + delta r) / |s + delta r|: at beta = 0.693 the FP64 restatement
# accepts 0.80 / 0.59 / 0.52 / 0.49 of the moves at delta = 0.6 / 1.5 / 3 / 5 (3D L=16, after
# 30 sweeps); delta = 4 gives the ~50% the configuration asks for
C3_DELTA = 4.0


def leg_c3(gpu, ctx):
    """3D Heisenberg L=128 beta=0.693, FP32 float4 spins, small-angle proposal; C3_REPLICAS
    replicas (8 GiB of spins at 256: far beyond the 256 MB last-level cache, so the sweep
    rate is an HBM rate, and one workgroup per replica grows 256 clusters at a time) vs 1."""
    L, beta = 128, 0.693
    lat = gpu.RegularLattice(ctx, 3, L)
    N = L ** 3
    res = {}
    for name, nrep, nstep in (("batched", C3_REPLICAS, 20), ("plain", 1, 20)):
        st = gpu.State(ctx, "heisenberg", [lat] * nrep, J=-1.0, proposal="small", delta=C3_DELTA,
                       embedding="cartesian")
        st.set_beta(beta)
        st.init(seed=3)
        from marqov_amd import autotune
        shape, _ = autotune.tune(ctx, st, cache_path=os.path.join("/tmp", f"marqov_amd_tuning_{os.getuid()}.json"),
                                 force=True, sweeps=3)
        st.metropolis(5, 3, 0)
        ms, _ = timed(ctx, lambda: st.metropolis(nstep, 3, sweep0=5))
        # (the Wolff autotuner would save and restore the 8 GiB batch per candidate: the
        # automatic shape is used here and reported)
        w_rate, w_mean, w_shape, w_next = wolff_rate(ctx, st, 30, 10, 10, 3, 10, sweep0=5 + nstep)
        w_pers = persistent_wolff_rate(ctx, st, 10, 10, 3, w_next) if name == "batched" else None
        if name == "batched":
            CPU_WOLFF_INPUTS["heis_3d_L128"] = {"model": "heisenberg", "beta": beta, "dim": 3, "L": L,
                                               "spins": st.download(0)}
        # acceptance of one sweep of replica 0 on the thermalised state
        before = st.download(0)
        st.metropolis(1, 3, sweep0=10 ** 6)
        acceptance = float(np.any(st.download(0) != before, axis=1).mean())
        res[name] = {"replicas": nrep, "attempts_per_ns": round(nrep * N * nstep / (ms * 1e6), 3),
                     "delta": C3_DELTA, "acceptance": round(acceptance, 3), "launch_shape": shape,
                     "wolff_sites_per_ns": round(w_rate, 4),
                     "wolff_roofline_frac": wolff_frac(w_rate, "3d_cartesian"), "wolff_launch_shape": w_shape,
                     "mean_cluster": round(w_mean, 1),
                     "roofline_frac": round(48.0 * nrep * N * nstep / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        if w_pers is not None:
            res[name]["wolff_persistent"] = {
                "sites_per_ns": round(w_pers[0], 4), "mean_cluster": round(w_pers[1], 1),
                "roofline_frac": wolff_frac(w_pers[0], "3d_cartesian"),
                "max_over_mean_replica_sites": round(w_pers[2], 2),
                "schedule": "the next 100 clusters per replica (10 x 10) in one mq_emcs_run launch without sweeps"}
        if name == "batched":
            obs = st.measure()
            res[name]["e"] = float(obs[:, 0].mean())
        del st
    res["batched_over_plain"] = round(res["batched"]["attempts_per_ns"] / res["plain"]["attempts_per_ns"], 2)
    res["wolff"] = "after 30 x (10 sweeps + 10 clusters); 10 timed calls of 10 clusters per replica"
    # finite-size-scaling shape: as many FP32 replicas as fill 2 GiB (8x the 256 MB last-level
    # cache) in one launch vs one per launch, L = 16, 32, 64 (a single L=128 replica already
    # saturates HBM)
    for Ls, nrep in ((16, 32768), (32, 4096), (64, 512)):
        latS = gpu.RegularLattice(ctx, 3, Ls)
        small = {}
        for name, nr, nstep in (("batched", nrep, 20), ("plain", 1, 20)):
            st = gpu.State(ctx, "heisenberg", [latS] * nr, J=-1.0, proposal="small", delta=C3_DELTA)
            st.set_beta(beta)
            st.init(seed=3)
            st.metropolis(5, 3, 0)
            ms, _ = timed(ctx, lambda: st.metropolis(nstep, 3, sweep0=5))
            small[name] = round(nr * Ls ** 3 * nstep / (ms * 1e6), 3)
            del st
        res[f"L{Ls}_batch"] = {"replicas": nrep, "batched_attempts_per_ns": small["batched"],
                               "plain_attempts_per_ns": small["plain"],
                               "batched_over_plain": round(small["batched"] / small["plain"], 1),
                               "roofline_frac": round(48.0 * small["batched"] / HBM_PEAK_GBS, 4)}
    return res


def leg_c4(gpu, ctx):
    """EA +-J batch: 512 samples (128 each of L=8,12,16,24) x 32 temperatures, HIP-graph
    captured {10 sweeps, energy, PT swap}; bit-sliced words (32 temperatures of a sample per
    uint32) and int8 spins, against one L=24 replica per launch."""
    Ls = [8, 12, 16, 24]
    nt, per = 32, 128
    betas = 0.5 * (3.3 / 0.5) ** (np.arange(nt) / (nt - 1))
    lats = {L: gpu.RegularLattice(ctx, 3, L) for L in Ls}
    reps = [lats[L] for L in Ls for _ in range(per) for _ in range(nt)]
    nrep = len(reps)
    sites = sum(lat.N for lat in reps)
    res = {"replicas": nrep, "spin_sites": sites}
    nsw, steps = 10, 5
    for name, bits in (("bits", 1), ("int8", 8)):
        st = gpu.State(ctx, "ea", reps, J=-1.0, spin_bits=bits)
        st.set_beta(np.tile(betas, nrep // nt))
        st.random_bonds(rep_sample=np.repeat(np.arange(nrep // nt), nt), p_plus=0.5, seed=4)
        st.init(seed=4)
        st.run_graph(nsw, nt, 4, 0, 0, steps, 1)
        ms, _ = timed(ctx, lambda: st.run_graph(nsw, nt, 4, nsw * steps, steps, steps, 4))
        # the same steps captured in a HIP graph and replayed (mq_tuning.graph = 1)
        st.set_tuning(graph=1)
        st.run_graph(nsw, nt, 4, 2 * nsw * steps * 5, 5 * steps, steps, 1)
        ms_g, _ = timed(ctx, lambda: st.run_graph(nsw, nt, 4, 2 * nsw * steps * 6, 6 * steps, steps, 4))
        # sweeps only (no energy or swap), replayed from a HIP graph: against the eager sweeps
        # below, the cost of a replayed kernel node over a stream launch
        st.run_graph(nsw, 1, 4, 2 * nsw * steps * 7, 0, steps, 1)
        ms_gs, _ = timed(ctx, lambda: st.run_graph(nsw, 1, 4, 2 * nsw * steps * 8, 0, steps, 4))
        st.set_tuning(graph=0)
        acc = st.pt_swap(nt, 0, 4, 10 ** 6)
        res[f"{name}_attempts_per_ns"] = round(sites * nsw * steps * 4 / (ms * 1e6), 3)
        res[f"{name}_hip_graph_attempts_per_ns"] = round(sites * nsw * steps * 4 / (ms_g * 1e6), 3)
        res[f"{name}_pt_accept_even_pairs"] = round(float(acc.sum()) / (len(acc) * 16), 3)
        # the same sweeps launched eagerly (no graph, no energy or swap): the sweep kernels alone
        ms_e, _ = timed(ctx, lambda: st.metropolis(nsw * steps, 4, sweep0=10 ** 6))
        res[f"{name}_eager_sweeps_attempts_per_ns"] = round(sites * nsw * steps / (ms_e * 1e6), 3)
        if bits == 1:  # 2 launches per sweep, 1 advance per repetition
            nodes = 2 * nsw * steps * 4 + 4
            res["hip_graph_node_overhead_us"] = round((ms_gs - ms_e * 4) * 1e3 / nodes, 2)
            res["bits_hip_graph_sweeps_only_attempts_per_ns"] = round(sites * nsw * steps * 4 / (ms_gs * 1e6), 3)
        del st
    res["graph"] = (f"mq_run_graph: {steps} x ({nsw} sweeps + energy + PT swap) per repetition, 4 repetitions, "
                    "launched on the stream (the default); hip_graph: the same steps captured in one HIP graph "
                    "and replayed")
    os.environ["MARQOV_AMD_KERNELS"] = "1"
    sq = gpu.State(ctx, "ea", [lats[24]], J=-1.0)
    del os.environ["MARQOV_AMD_KERNELS"]
    sq.set_beta(1.0)
    sq.random_bonds(p_plus=0.5, seed=4)
    sq.init(seed=4)
    ms_p, _ = timed(ctx, lambda: sq.metropolis(200, 4, 0))
    plain = 24 ** 3 * 200 / (ms_p * 1e6)
    del sq
    res["attempts_per_ns"] = res["bits_attempts_per_ns"]
    res["plain_attempts_per_ns"] = round(plain, 4)
    res["batched_over_plain"] = round(res["bits_attempts_per_ns"] / plain, 1)
    return res


def leg_c5(gpu, ctx):
    """2D random geometric graph N=2^20, mean degree 6, Ising and Potts-3, 64 replicas vs 1."""
    from marqov_amd import graphs
    n = 1 << 20
    t0 = time.perf_counter()
    rp, ci = ctx.rgg(2, n, graphs.rgg_radius(n), 5)  # built on the device (mq_rgg_build)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    lat = gpu.CsrLattice(ctx, rp, ci)  # host greedy colouring, renumbering, upload
    t_lat = time.perf_counter() - t0
    res = {"N": n, "edges": int(rp[-1] // 2), "colours": int(len(lat.colors())),
           "graph_build_ms": round(1e3 * t_build, 1), "colouring_and_upload_ms": round(1e3 * t_lat, 1)}
    # pseudo-critical points of these graphs: Binder-cumulant crossings of N = 16,384 and 65,536
    # (`bench.py --locate-c5-beta`): Ising 0.494, Potts-3 1.088
    for model, beta in (("ising", 0.494), ("potts", 1.088)):
        for name, nrep, nstep in (("batched", 64, 10), ("plain", 1, 10)):
            st = gpu.State(ctx, model, [lat] * nrep, J=-1.0, q=3)
            st.set_beta(beta)
            st.init(seed=5)
            st.metropolis(2, 5, 0)
            ms, _ = timed(ctx, lambda: st.metropolis(nstep, 5, sweep0=2))
            res[f"{model}_{name}_attempts_per_ns"] = round(nrep * n * nstep / (ms * 1e6), 3)
            if name == "batched":
                # thermalise with rounds of (4 sweeps + 4 clusters) until the mean energy moves
                # by less than 2e-4 per site over 100 rounds (at most 4000 rounds): the state
                # coarsens for hundreds of rounds from the random start
                sw, cl, rounds, e_prev = 2 + nstep, 0, 0, None
                while rounds < 4000:
                    for _ in range(100):
                        st.metropolis(4, 5, sweep0=sw)
                        st.wolff(4, 5, cluster0=cl, sizes=False)
                        sw, cl = sw + 4, cl + 4
                    rounds += 100
                    e = float(st.measure()[:, 0].mean())
                    if e_prev is not None and abs(e - e_prev) < 2e-4:
                        break
                    e_prev = e
                w_rate, w_mean, w_shape, w_next = wolff_rate(ctx, st, 0, 4, 4, 5, 8, sweep0=sw, cluster0=cl,
                                                             tune_ncl=4)
                res[f"{model}_wolff_sites_per_ns"] = round(w_rate, 4)
                res[f"{model}_wolff_roofline_frac"] = wolff_frac(w_rate, "rgg_int8")
                res[f"{model}_wolff_mean_cluster"] = round(w_mean, 1)
                res[f"{model}_wolff_thermalisation_rounds"] = rounds
                res[f"{model}_wolff_launch_shape"] = w_shape
                CPU_WOLFF_INPUTS[f"rgg_{model}"] = {"model": model, "beta": beta, "rp": rp, "ci": ci,
                                                    "spins": st.download(0)}
                # the next 32 clusters per replica in one launch: no lockstep every 4 clusters
                ms1, sz1 = timed_wolff(ctx, st, 32, 5, w_next)
                res[f"{model}_wolff_one_launch_sites_per_ns"] = round(float(sz1.sum()) / (ms1 * 1e6), 4)
                res[f"{model}_wolff_one_launch_max_over_mean_replica_sites"] = round(
                    float(sz1.sum(axis=1).max() / sz1.sum(axis=1).mean()), 2)
            del st
        res[f"{model}_batched_over_plain"] = round(res[f"{model}_batched_attempts_per_ns"] /
                                                   res[f"{model}_plain_attempts_per_ns"], 2)
    res["wolff"] = ("64 replicas thermalised by rounds of (4 sweeps + 4 clusters) until E/N moves by < 2e-4 "
                    "over 100 rounds; Wolff autotuner; 8 timed calls of 4 clusters per replica, then one "
                    "call of the next 32 clusters per replica (one_launch)")
    # the Ising sweep on 1-bit storage (32 replicas per uint32 word): the same chain as int8 storage,
    # (d+2)*4/32 B of spins + 4(d+1)/64 B of CSR indices per attempt instead of (d+2) + 4(d+1)/64
    st = gpu.State(ctx, "ising", [lat] * 64, J=-1.0, spin_bits=1)
    st.set_beta(0.494)
    st.init(seed=5)
    st.metropolis(2, 5, 0)
    ms, _ = timed(ctx, lambda: st.metropolis(10, 5, sweep0=2))
    rate = 64 * n * 10 / (ms * 1e6)
    dbar = res["edges"] * 2 / n
    bpa = (dbar + 2) * 4 / 32 + 4 * (dbar + 1) / 64
    res["ising_bits_batched_attempts_per_ns"] = round(rate, 3)
    res["ising_bits_bytes_per_attempt"] = round(bpa, 3)
    res["ising_bits_over_int8"] = round(rate / res["ising_batched_attempts_per_ns"], 2)
    del st
    return res


def locate_c5_beta(gpu, ctx, sizes=(4096, 16384, 65536), R=32, nwarm=300, nmeas=1500):
    """Pseudo-critical betas of Ising and Potts-3 on 2D random geometric graphs of mean degree 6
    (the RGG legs' graphs at smaller N, built on the device with the legs' seed): for each size,
    R replicas at each of 13 betas, EMCS = 5 Wolff clusters + 1 sweep, nwarm + nmeas EMCS with a
    measurement after each; U = 1 - <m^4> / (3 <m^2>^2) per beta, and the beta where the U(beta)
    curves of consecutive sizes cross (linear interpolation between grid points)."""
    from marqov_amd import graphs
    grids = {"ising": np.linspace(0.46, 0.54, 9), "potts": np.linspace(1.04, 1.20, 9)}
    out = {}
    for model, betas in grids.items():
        curves = {}
        for n in sizes:
            rp, ci = ctx.rgg(2, n, graphs.rgg_radius(n), 5)
            lat = gpu.CsrLattice(ctx, rp, ci)
            nb = len(betas)
            st = gpu.State(ctx, model, [lat] * (nb * R), J=-1.0, q=3)
            st.set_beta(np.repeat(betas, R))
            st.init(seed=11)
            for t in range(nwarm):
                st.emcs(5, 1, 11, t)
            m2 = np.zeros(nb * R)
            m4 = np.zeros(nb * R)
            for t in range(nmeas):
                st.emcs(5, 1, 11, nwarm + t)
                o = st.measure()
                m2 += o[:, 2]
                m4 += o[:, 3]
            m2 = (m2 / nmeas).reshape(nb, R).mean(1)
            m4 = (m4 / nmeas).reshape(nb, R).mean(1)
            curves[n] = 1.0 - m4 / (3.0 * m2 * m2)
            log(f"{model} N={n}: U = {np.round(curves[n], 3).tolist()}")
            del st, lat
        cross = {}
        for n1, n2 in zip(sizes[:-1], sizes[1:]):
            d = curves[n2] - curves[n1]
            idx = [i for i in range(len(betas) - 1) if d[i] * d[i + 1] < 0]
            cross[f"{n1}-{n2}"] = [round(float(betas[i] - d[i] * (betas[i + 1] - betas[i]) / (d[i + 1] - d[i])), 4)
                                   for i in idx]
        out[model] = {"betas": [round(float(b), 4) for b in betas],
                      "binder": {str(n): [round(float(u), 4) for u in curves[n]] for n in sizes},
                      "crossings": cross}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "marqov_amd runs on one GPU; sharding across GPUs is designed but not implemented",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.calibrate_copy:
        import torch
        src = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": src.numel() * 4, "copies": 5}))
        return
    from marqov_amd import gpu

    ctx = gpu.Context(0)
    if a.locate_c5_beta:
        print(json.dumps(locate_c5_beta(gpu, ctx)), flush=True)
        return
    L, R, beta = a.L, a.replicas, 0.2216546
    N = L ** 3
    lat = gpu.RegularLattice(ctx, 3, L)

    # ---- headline: batched Metropolis on 3D Ising L=256
    log("headline: 3D Ising batched sweeps")
    st = gpu.State(ctx, "ising", [lat] * R, J=-1.0)
    st.set_beta(beta)
    st.init(seed=2)
    tuning = None
    if a.shape:
        blk, xcd = (int(x) for x in a.shape.split(","))
        tuning = {"kernels": 0, "block": blk, "xcd_remap": xcd, "lds_block": 256}
        st.set_tuning(**tuning)
    elif not a.no_tune:
        # launch shape from the project's autotuner (its timing sweeps use counters far from
        # the run's, so they only add warm-up to this chain); retuned on every run, the choice
        # also written to a scratch file
        from marqov_amd import autotune
        shape, _ = autotune.tune(ctx, st, cache_path=os.path.join("/tmp", f"marqov_amd_tuning_{os.getuid()}.json"),
                                 force=True, sweeps=6)
        tuning = shape
    st.metropolis(a.warmup, seed=2, sweep0=0)
    ms, wall = timed(ctx, lambda: st.metropolis(a.steps, seed=2, sweep0=a.warmup))
    value = a.steps * R * N / (ms * 1e6)
    tn = st.get_tuning()
    shape_key = (tn["block"], tn["xcd_remap"])
    launch_ms = ms / (2 * a.steps)
    bytes_per_launch = 3.0 * R * (N // 2)
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    obs = st.measure()
    wolff_b = emcs_b = None
    if not a.quick and not a.legs_only:
        log("headline state: Wolff clusters")
        wolff_b = wolff_rate(ctx, st, 20, 4, 10, 2, 8, sweep0=a.warmup + a.steps, tune_ncl=10)
        log("headline state: Wolff clusters, one persistent launch")
        wolff_p = persistent_wolff_rate(ctx, st, 10, 10, 2, wolff_b[3])
        CPU_WOLFF_INPUTS["3d_L256"] = {"model": "ising", "beta": beta, "dim": 3, "L": L, "spins": st.download(0)}
        log("headline state: EMCS, separate launches and persistent")
        emcs_b = emcs_compare(ctx, st, 20, 10, 1, seed=4, step0=1 << 20)
    del st

    line = {
        "metric": "spin-flip attempts/ns and Wolff sites/ns, batched vs plain, 1 GPU",
        "value": round(value, 3),
        "unit": "attempts/ns",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms / a.steps,
        "wall_ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic (Philox-seeded random start, seed 2)",
        "config": {"workload": f"3D Ising L={L} periodic J=-1 beta={beta} int8 checkerboard Metropolis, "
                               f"{R} replicas per launch", "replicas": R, "L": L, "sites": N,
                   "launch_shape": tuning if tuning is not None else "default (block 256, XCD remap)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     # HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE of this configuration
                     # and launch shape, scaled by a 1 GiB copy (factors 2.0 / 1.0)
                     "traffic": HEADLINE_TRAFFIC.get(shape_key) if (L, R) == (256, 32) else None,
                     "traffic_source": "profiles/k_sweep_ising3_row16_3d_L256_R32.md (by launch shape)",
                     "kernel": "k_sweep_ising3_row16", "launch_ms": launch_ms,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "observables": {"e_mean": float(obs[:, 0].mean()), "m_abs_mean": float(obs[:, 1].mean())},
    }

    if not a.quick and not a.legs_only:
        log("plain 3D Ising")
        sp = gpu.State(ctx, "ising", [lat], J=-1.0)
        sp.set_beta(beta)
        sp.init(seed=2)
        sp.metropolis(4, seed=2, sweep0=0)
        k = max(a.steps // 4, 10)
        ms_p, _ = timed(ctx, lambda: sp.metropolis(k, seed=2, sweep0=4))
        line["plain"] = {"attempts_per_ns": round(k * N / (ms_p * 1e6), 3), "replicas": 1, "steps": k}
        line["batched_over_plain"] = round(value / line["plain"]["attempts_per_ns"], 3)
        w_rate, w_mean, _, _ = wolff_rate(ctx, sp, 20, 4, 10, 2, 8, sweep0=4 + k)
        line["plain"]["wolff_sites_per_ns"] = round(w_rate, 4)
        line["plain"]["wolff_mean_cluster"] = round(w_mean, 1)
        line["wolff"] = {"sites_per_ns": round(wolff_b[0], 4), "mean_cluster": round(wolff_b[1], 1),
                         "replicas": R, "plain_sites_per_ns": round(w_rate, 4),
                         "batched_over_plain": round(wolff_b[0] / w_rate, 2),
                         "roofline_frac": wolff_frac(wolff_b[0], "3d_int8"), "bytes_per_site": WOLFF_BYTES["3d_int8"],
                         "launch_shape": wolff_b[2],
                         "persistent": {"sites_per_ns": round(wolff_p[0], 4), "mean_cluster": round(wolff_p[1], 1),
                                        "roofline_frac": wolff_frac(wolff_p[0], "3d_int8"),
                                        "max_over_mean_replica_sites": round(wolff_p[2], 2),
                                        "schedule": "the next 100 clusters per replica (10 x 10) in one "
                                                    "mq_emcs_run launch without sweeps"},
                         "schedule": "after 20 x (4 sweeps + 10 clusters); 8 timed calls of 10 clusters per replica"}
        line["emcs"] = emcs_b
        del sp
    if not a.quick:
        for name, fn in (("c2_msc", leg_c2_msc), ("c1_2d_L64", leg_c1), ("c3_heis_L128", leg_c3),
                         ("c4_ea_batch", leg_c4), ("c5_rgg", leg_c5)):
            if a.legs and name not in a.legs.split(","):
                continue
            log(name)
            try:
                line[name] = fn(gpu, ctx)
            except Exception as ex:  # a failing leg is reported, the headline still prints
                line[name] = {"error": repr(ex)}

    if not a.no_cpu:
        log("cpu baseline")
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_bind as ob
        ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        nth = max(1, min(16, ncpu))
        ref_bin = os.path.join(ROOT, "oracle", "_ref", "ref_metropolis")
        ref = None
        if os.path.exists(ref_bin):
            # the reference's own Metropolis<H, RegularLattice>::move, compiled from its headers
            import subprocess
            try:
                r = subprocess.run([ref_bin, "bench", "ising", "3", str(L), repr(beta), "-1.0", "1", "3", "2", str(nth)],
                                   capture_output=True, text=True, timeout=600)
                if r.returncode == 0:
                    ref = json.loads(r.stdout.strip().splitlines()[-1])
            except (OSError, ValueError, subprocess.TimeoutExpired):
                ref = None
        if ref is not None:
            line["cpu_baseline"] = {"value": round(float(ref["attempts_per_ns"]), 5), "unit": "attempts/ns",
                                    "cores": nth, "kind": "reference",
                                    "per_core": round(float(ref["per_thread"]), 5),
                                    "sample": f"{nth} threads x (1 warm-up + 3 timed) sweeps of the reference's "
                                              f"Metropolis<Ising, RegularLattice>::move (random site, mt19937_64 "
                                              f"RNGCache) on 3D Ising L={L}, one simulation per thread, "
                                              f"{ref['seconds']:.1f} s timed"}
        else:
            out = np.zeros(3)
            ob.lib().ref_bench_regular(1, 3, L, -1.0, 2, beta, 1, 3, nth, 2, 0, ob.p(out))
            line["cpu_baseline"] = {"value": round(float(out[0]), 5), "unit": "attempts/ns", "cores": nth,
                                    "kind": "port", "per_core": round(float(out[1]), 5),
                                    "sample": f"{nth} threads x (1 warm-up + 3 timed) random-site Metropolis sweeps "
                                              f"of 3D Ising L={L}, one replica per thread, {out[2]:.1f} s timed"}
        # Wolff on the 2D L=64 critical Ising lattice, the configuration the CPU reference runs
        ow = np.zeros(3)
        ob.lib().ref_bench_regular(1, 2, 64, -1.0, 2, 0.4406868, 500, 20000, nth, 1, 1, ob.p(ow))
        line["cpu_baseline"]["wolff_2d_L64"] = {
            "value": round(float(ow[0]), 5), "unit": "sites/ns", "cores": nth, "per_core": round(float(ow[1]), 5),
            "sample": f"{nth} threads x (500 warm-up + 20000 timed) stack Wolff clusters of 2D Ising L=64 at "
                      f"beta_c, one replica per thread, {ow[2]:.1f} s timed"}
        if os.path.exists(ref_bin):
            # the reference's move for the FP64 Heisenberg model, fresh-direction proposal, 3D L=128
            import subprocess
            try:
                r = subprocess.run([ref_bin, "bench", "heis", "3", "128", "0.693", "-1.0", "1", "2", "3", str(nth)],
                                   capture_output=True, text=True, timeout=300)
                if r.returncode == 0:
                    hz = json.loads(r.stdout.strip().splitlines()[-1])
                    line["cpu_baseline"]["heisenberg_3d_L128"] = {
                        "value": round(float(hz["attempts_per_ns"]), 5), "unit": "attempts/ns", "cores": nth,
                        "kind": "reference", "per_core": round(float(hz["per_thread"]), 5),
                        "sample": f"{nth} threads x (1 warm-up + 2 timed) sweeps of the reference's "
                                  f"Metropolis<Heisenberg, RegularLattice>::move (FP64, fresh direction) on 3D L=128, "
                                  f"one simulation per thread, {hz['seconds']:.1f} s timed"}
            except (OSError, ValueError, subprocess.TimeoutExpired):
                pass
        # Wolff on the thermalised GPU states of the C2, C3 and C5 legs (the same workload):
        # the restatement's stack Wolff, one copy of the state per thread, a fixed time each
        models = {"ising": ob.ISING, "potts": ob.POTTS, "heisenberg": ob.HEIS}
        for key, inp in CPU_WOLFF_INPUTS.items():
            sp = inp["spins"]
            spins = sp.astype(np.float64) if sp.ndim == 2 else sp.astype(np.int32)
            r = ob.bench_wolff(models[inp["model"]], inp["beta"], spins, 2.0, nth, dim=inp.get("dim", 0),
                               L=inp.get("L", 0), rp=inp.get("rp"), ci=inp.get("ci"))
            line["cpu_baseline"][f"wolff_{key}"] = {
                "value": round(float(r[0]), 5), "unit": "sites/ns", "cores": nth, "kind": "port",
                "per_core": round(float(r[1]), 5), "mean_cluster": round(float(r[4]), 1),
                "sample": f"{nth} threads x stack Wolff clusters ({int(r[3])} in all) for {r[2]:.1f} s on copies of "
                          f"replica 0 of the GPU's thermalised {key} state, beta {inp['beta']}"}
        cpu_model = ""
        try:
            with open("/proc/cpuinfo") as f:
                cpu_model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
        except OSError:
            pass
        line["cpu_baseline"]["host"] = {"cpu_model": cpu_model, "cpus_usable": ncpu,
                                        "cpus_online": os.cpu_count(), "threads_used": nth}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
