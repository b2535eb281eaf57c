#!/usr/bin/env python3
"""Throughput of the MARQOV update path on one MI355X.

Headline workload: 3D simple-cubic Ising, L=256 (16.8M sites), periodic, J=-1,
beta=0.2216546, int8 spins in the colour-split checkerboard layout, Philox4x32-10,
R replicas of that lattice swept together (batched; R=32 puts the working set
above the 256 MB Infinity Cache).  A step is one full checkerboard Metropolis
sweep of every replica.  value = spin-flip attempts/ns over the timed steps.

Beside it on the same JSON line:
  plain     - the same sweep, one replica per launch (attempts/ns)
  wolff     - Wolff sites/ns: 2D Ising L=64 at beta_c, 10 clusters per replica
              per step, batched (many replicas, one workgroup each) vs plain (1)
  roofline  - the sweep kernel's algorithmic HBM rate (3 B per attempt: two reads
              and one write of the int8 lattice per sweep) vs 8 TB/s
  cpu_baseline - the reference-faithful CPU restatement (random-site Metropolis,
              mt19937_64) timed on this host's cores, one replica per thread.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--replicas", type=int, default=32)
    ap.add_argument("--L", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip plain and Wolff legs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (counter calibration for HBM traffic)")
    return ap.parse_args()


def time_sweeps(ctx, st, steps, warmup, seed):
    st.metropolis(warmup, seed=seed, sweep0=0)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.timer_begin()
    st.metropolis(steps, seed=seed, sweep0=warmup)
    ms = ctx.timer_end()
    wall = time.perf_counter() - t0
    return ms, wall


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is design-only",
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
    L, R, beta = a.L, a.replicas, 0.2216546
    N = L ** 3
    lat = gpu.RegularLattice(ctx, 3, L)

    # ---- batched Metropolis (headline)
    st = gpu.State(ctx, "ising", [lat] * R, J=-1.0)
    st.set_beta(beta)
    st.init(seed=2)
    ms, wall = time_sweeps(ctx, st, a.steps, a.warmup, seed=2)
    attempts = a.steps * R * N
    value = attempts / (ms * 1e6)
    launches = 2 * a.steps
    launch_ms = ms / launches
    bytes_per_launch = 3.0 * R * (N // 2)
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    obs = st.measure()
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
                               f"{R} replicas per launch", "replicas": R, "L": L, "sites": N},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_sweep_ising<3,false>", "launch_ms": launch_ms,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "observables": {"e_mean": float(obs[:, 0].mean()), "m_abs_mean": float(obs[:, 1].mean())},
    }

    if not a.no_extra:
        # ---- plain Metropolis: one replica per launch
        sp = gpu.State(ctx, "ising", [lat], J=-1.0)
        sp.set_beta(beta)
        sp.init(seed=2)
        k = max(a.steps // 4, 10)
        ms_p, _ = time_sweeps(ctx, sp, k, max(a.warmup // 4, 2), seed=2)
        line["plain"] = {"attempts_per_ns": round(k * N / (ms_p * 1e6), 3), "replicas": 1, "steps": k}
        line["batched_over_plain"] = round(value / line["plain"]["attempts_per_ns"], 3)
        del sp
        # ---- Wolff sites/ns on 2D L=64 at beta_c
        l2 = gpu.RegularLattice(ctx, 2, 64)
        wres = {}
        for name, nrep in (("batched", 2048), ("plain", 1)):
            sw = gpu.State(ctx, "ising", [l2] * nrep, J=-1.0)
            sw.set_beta(0.4406868)
            sw.init(seed=1)
            sw.metropolis(20, seed=1)
            sw.wolff(20, seed=1, cluster0=0, sizes=False)
            ctx.sync()
            ctx.timer_begin()
            sizes = sw.wolff(10 * 5, seed=1, cluster0=20, sizes=False)
            wms = ctx.timer_end()
            sizes = sw.wolff(10, seed=1, cluster0=70, sizes=True)
            mean = float(sizes.mean())
            wres[name] = {"sites_per_ns": round(nrep * 50 * mean / (wms * 1e6), 4), "replicas": nrep,
                          "mean_cluster": round(mean, 1)}
            del sw
        line["wolff"] = {"workload": "2D Ising L=64 beta=0.4406868, 50 clusters per replica", **wres}

    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import numpy as np

        import oracle_bind as ob
        ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        nth = max(1, min(16, ncpu))
        out = np.zeros(3)
        ob.lib().ref_bench_regular(1, 3, L, -1.0, 2, beta, 1, 3, nth, 2, 0, ob.p(out))
        line["cpu_baseline"] = {"value": round(float(out[0]), 5), "unit": "attempts/ns", "cores": nth,
                                "kind": "port", "per_core": round(float(out[1]), 5),
                                "sample": f"{nth} threads x (1 warm-up + 3 timed) random-site Metropolis sweeps "
                                          f"of 3D Ising L={L}, one replica per thread, {out[2]:.1f} s timed"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
