"""Markov chains built from the oracle's one-step FP64 restatements of the graph
path for O(3)/Phi4 spins (orc_csr_o3_colour: one colour class of a sweep,
orc_csr_o3_wolff: one cluster), on the GPU kernels' Philox schedule.  A chain of
such steps must sample the Boltzmann distribution, so its averages can be set
against the reference-faithful restatement's chains, whose Metropolis move
equals the reference build's bit for bit.  Test infrastructure only."""
import numpy as np

import oracle_bind as ob


def greedy_colouring(rp, ci):
    """Smallest colour not taken by an earlier neighbour, in vertex order (any proper
    colouring makes a valid colour-class sweep)."""
    n = len(rp) - 1
    col = np.full(n, -1, np.int32)
    for v in range(n):
        used = {int(col[u]) for u in ci[rp[v]:rp[v + 1]] if col[u] >= 0}
        c = 0
        while c in used:
            c += 1
        col[v] = c
    return col


def o3_obs(s, rp, ci, phi4, beta, lam, mass, J=-1.0):
    """E/N (observables.h:173-352 with the Phi4 on-site terms), |m|, m^2, m^4, mean |phi|^2."""
    N = len(s)
    src = np.repeat(np.arange(N), np.diff(rp))
    e = J * float((s[src] * s[ci]).sum()) / (2 * N)
    n2 = (s * s).sum(1)
    if phi4:
        e += float(((mass / beta) * n2 + (lam / beta) * (n2 - 1) ** 2).mean())
    m = float(np.linalg.norm(s.sum(0))) / N
    return e, m, m * m, m ** 4, float(n2.mean())


def oracle_chain(args):
    """One chain of EMCS = ncluster oracle clusters then nmetro oracle colour sweeps, with
    the counters the GPU would use for replica `rep`; returns obs[nmeas, 5]."""
    (rp, ci, col, phi4, beta, rep, seed, nwarm, nmeas, ncluster, nmetro, small, delta, rdir, lam, mass) = args
    rng = np.random.default_rng(1000 + rep)
    N = len(rp) - 1
    v = rng.normal(size=(N, 3))
    s = v / np.linalg.norm(v, axis=1)[:, None]
    nc = int(col.max()) + 1
    out = []
    sweep = ncl = 0
    for t in range(nwarm + nmeas):
        for _ in range(ncluster):
            s, _, _ = ob.csr_o3_wolff(rp, ci, s, rdir, -1.0, beta, ncl, rep, seed)
            ncl += 1
        for _ in range(nmetro):
            for c in range(nc):
                s, _ = ob.csr_o3_colour(rp, ci, col, s, c, 64 * sweep + c, rep, seed, beta, phi4=phi4,
                                        small_angle=small, delta=delta, lam=lam, mass=mass)
            sweep += 1
        if t >= nwarm:
            out.append(o3_obs(s, rp, ci, phi4, beta, lam, mass))
    return np.array(out)


def oracle_chains(rp, ci, phi4, beta, nchains, seed, nwarm, nmeas, ncluster, nmetro, small=False, delta=0.5,
                  rdir=True, lam=4.5, mass=1.0, workers=8):
    """nchains oracle chains (replicas 0..nchains-1) in worker processes; obs[chain, t, 5]."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    col = greedy_colouring(rp, ci)
    jobs = [(rp, ci, col, phi4, beta, r, seed, nwarm, nmeas, ncluster, nmetro, small, delta, int(rdir), lam, mass)
            for r in range(nchains)]
    with ProcessPoolExecutor(max_workers=max(1, min(workers, nchains)), mp_context=mp.get_context("fork")) as ex:
        return np.array(list(ex.map(oracle_chain, jobs)))
