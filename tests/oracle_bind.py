"""ctypes binding of the oracle (oracle/liboracle.so): used to check results, never by the product."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_METRO = os.path.join(ORACLE_DIR, "_ref", "ref_metropolis")

_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_D = ctypes.c_double

ISING, EA, POTTS, HEIS, PHI4 = 1, 2, 3, 4, 5

_SIG = {
    "orc_neighbour": (_I64, [_I, _I, _I64, _I]),
    "orc_thresholds_ising": (None, [_D, _D, _I, _P, _P]),
    "orc_thresholds_potts": (None, [_D, _D, _I, _P, _P]),
    "orc_wolff_threshold": (_U32, [_D]),
    "orc_init_ising": (None, [_I64, _P, _U64, _U32]),
    "orc_init_potts": (None, [_I64, _P, _I, _U64, _U32]),
    "orc_init_bonds": (None, [_I, _I64, _P, _D, _U64, _U32]),
    "orc_rgg_points": (None, [_I, _I64, _U64, _P]),
    "orc_o3_halfsweep": (None, [_I, _I, _P, _P, _P, _I, _I, _D, _D, _D, _D, _D, _I, _U32, _U32, _U64]),
    "orc_o3_wolff": (_I64, [_I, _I, _P, _P, _I, _D, _D, _U32, _U32, _U64, _P]),
    "orc_csr_o3_colour": (None, [_I64, _P, _P, _P, _P, _P, _P, _I, _I, _D, _D, _D, _D, _D, _I, _U32, _U32, _U64]),
    "orc_csr_o3_wolff": (_I64, [_I64, _P, _P, _P, _P, _I, _D, _D, _U32, _U32, _U64, _P]),
    "orc_ising_sweeps": (None, [_I, _I, _P, _P, _P, _P, _I, _U64, _U32, _U32]),
    "orc_potts_sweeps": (None, [_I, _I, _P, _I, _P, _P, _I, _U64, _U32, _U32]),
    "orc_wolff_ising": (_I, [_I, _I, _P, _U32, _I, _U64, _U32, _U32, _P]),
    "orc_wolff_potts": (_I, [_I, _I, _P, _I, _U32, _I, _U64, _U32, _U32, _P]),
    "orc_bond_sum_ising": (_I64, [_I, _I, _P, _P]),
    "orc_mag_sum_ising": (_I64, [_I64, _P]),
    "orc_bond_sum_potts": (_I64, [_I, _I, _P]),
    "orc_csr_sweeps": (None, [_I64, _P, _P, _P, _I, _P, _P, _P, _P, _I, _I, _I, _I, _U64, _U32, _U32]),
    "orc_csr_wolff": (_I, [_I64, _P, _P, _P, _I, _I, _U32, _I, _U64, _U32, _U32, _P]),
    "orc_ising_msc_sweeps": (None, [_I, _P, _P, _P, _I, _U64, _U32, _U32]),
    "orc_ea_bits_sweeps": (None, [_I, _I, _P, _P, _P, _P, _I, _U64, _U32, _U32]),
    "ref_run_regular": (_I, [_I, _I, _I, _D, _I, _D, _D, _P, _D, _I, _I, _I, _I, _U64, _I, _P, _P, _P, _P, _P]),
    "ref_run_csr": (_I, [_I, _I64, _P, _P, _D, _I, _D, _D, _P, _D, _I, _I, _I, _I, _U64, _I, _P, _P, _P, _P, _P]),
    "ref_mt19937_64_nth": (_U64, [_U64, _I64]),
    "ref_set_plant": (None, [_I]),
    "ref_rngcache_draws": (None, [_U64, _I, _U64, _I64, _P]),
    "ref_bench_regular": (_I, [_I, _I, _I, _D, _I, _D, _I, _I, _I, _U64, _I, _P]),
    "ref_bench_wolff": (_I, [_I, _I, _I, _I64, _P, _P, _D, _I, _D, _P, _P, _D, _I, _U64, _P]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True, capture_output=True)
        l = ctypes.CDLL(LIB)
        for k, (r, a) in _SIG.items():
            f = getattr(l, k)
            f.restype = r
            f.argtypes = a
        _lib = l
    return _lib


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def philox(ctr, key):
    """Philox4x32-10 through the oracle's C code (via a tiny numpy-free helper)."""
    from ctypes import c_uint32
    c = (c_uint32 * 4)(*ctr)
    k = (c_uint32 * 2)(*key)
    o = (c_uint32 * 4)()
    _philox_fn()(c, k, o)
    return list(o)


_ph = None


def _philox_fn():
    global _ph
    if _ph is None:
        f = lib().orc_philox4x32_10_export
        f.restype = None
        f.argtypes = [_P, _P, _P]
        _ph = f
    return _ph


def thresholds_ising(beta, J, qmax):
    hi = np.zeros(2 * qmax + 1, np.uint32)
    lo = np.zeros(2 * qmax + 1, np.uint32)
    lib().orc_thresholds_ising(beta, J, qmax, p(hi), p(lo))
    return hi, lo


def thresholds_potts(beta, k, zmax):
    hi = np.zeros(2 * zmax + 1, np.uint32)
    lo = np.zeros(2 * zmax + 1, np.uint32)
    lib().orc_thresholds_potts(beta, k, zmax, p(hi), p(lo))
    return hi, lo


def init_ising(N, seed, rep):
    s = np.zeros(N, np.int8)
    lib().orc_init_ising(N, p(s), seed, rep)
    return s


def init_potts(N, Q, seed, rep):
    s = np.zeros(N, np.int8)
    lib().orc_init_potts(N, p(s), Q, seed, rep)
    return s


def o3_halfsweep(dim, L, s, colour, phase, rep, seed, beta, J=-1.0, phi4=False, small_angle=False, delta=0.5,
                 lam=4.5, mass=1.0):
    """FP64 restatement of one colour half-sweep of the FP32 O(3)/Phi4 kernel:
    returns (new spins float64[N,3], margin p - u per site, 0 off the colour)."""
    s_in = np.ascontiguousarray(s, np.float64).reshape(-1, 3)
    out = np.zeros_like(s_in)
    margin = np.zeros(len(s_in), np.float64)
    lib().orc_o3_halfsweep(dim, L, p(s_in), p(out), p(margin), int(phi4), int(small_angle), J, delta, lam, mass,
                           beta, colour, phase, rep, seed)
    return out, margin


def o3_wolff(dim, L, s, random_dir, J, beta, ncl, rep, seed):
    """FP64 restatement of one O(3)/Phi4 Wolff cluster on the GPU's random numbers:
    returns (spins after the flip, cluster size, smallest decision margin)."""
    s_in = np.ascontiguousarray(s, np.float64).reshape(-1, 3)
    out = np.zeros_like(s_in)
    mm = np.zeros(1, np.float64)
    size = lib().orc_o3_wolff(dim, L, p(s_in), p(out), int(random_dir), J, beta, ncl, rep, seed, p(mm))
    return out, int(size), float(mm[0])


def csr_o3_colour(rp, ci, colour_of, s, colour, phase, rep, seed, beta, J=-1.0, phi4=False, small_angle=False,
                  delta=0.5, lam=4.5, mass=1.0):
    """orc_csr_o3_colour: one colour class of the FP32 O(3)/Phi4 graph sweep in FP64."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    col = np.ascontiguousarray(colour_of, np.int32)
    s_in = np.ascontiguousarray(s, np.float64).reshape(-1, 3)
    out = np.zeros_like(s_in)
    margin = np.zeros(len(s_in), np.float64)
    lib().orc_csr_o3_colour(len(rp) - 1, p(rp), p(ci), p(col), p(s_in), p(out), p(margin), int(phi4),
                            int(small_angle), J, delta, lam, mass, beta, colour, phase, rep, seed)
    return out, margin


def csr_o3_wolff(rp, ci, s, random_dir, J, beta, ncl, rep, seed):
    """orc_csr_o3_wolff: one O(3)/Phi4 cluster on a CSR graph, FP64."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    s_in = np.ascontiguousarray(s, np.float64).reshape(-1, 3)
    out = np.zeros_like(s_in)
    mm = np.zeros(1, np.float64)
    size = lib().orc_csr_o3_wolff(len(rp) - 1, p(rp), p(ci), p(s_in), p(out), int(random_dir), J, beta, ncl, rep,
                                  seed, p(mm))
    return out, int(size), float(mm[0])


def rgg_points(dim, n, seed):
    out = np.zeros((n, dim), np.float64)
    lib().orc_rgg_points(dim, n, seed, p(out))
    return out


def rgg_brute(pts, radius):
    """All-pairs minimum-image neighbour lists (the definition of the graph), as CSR:
    dx = |x_i - x_j|, 1 - dx above 1/2, s = dx_0^2 + dx_1^2 (+ dx_2^2) rounded in that
    order, edge when s < r^2."""
    n, dim = pts.shape
    r2 = radius * radius
    rows = []
    for i in range(n):
        dx = np.abs(pts[i] - pts)
        dx = np.where(dx > 0.5, 1.0 - dx, dx)
        s = dx[:, 0] * dx[:, 0]
        for d in range(1, dim):
            s = s + dx[:, d] * dx[:, d]
        nb = np.nonzero(s < r2)[0]
        rows.append(nb[nb != i])
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    return rp.astype(np.int32), np.concatenate(rows).astype(np.int32)


def init_bonds(dim, N, p_plus, seed, sample):
    b = np.zeros(N * dim, np.int8)
    lib().orc_init_bonds(dim, N, p(b), p_plus, seed, sample)
    return b


def ising_sweeps(dim, L, s, beta, J, nsweeps, seed, sweep0, rep, bonds=None):
    hi, lo = thresholds_ising(beta, J, 2 * dim)
    s = np.ascontiguousarray(s, np.int8).copy()
    lib().orc_ising_sweeps(dim, L, p(s), None if bonds is None else p(np.ascontiguousarray(bonds, np.int8)),
                           p(hi), p(lo), nsweeps, seed, sweep0, rep)
    return s


def potts_sweeps(dim, L, s, Q, beta, k, nsweeps, seed, sweep0, rep):
    hi, lo = thresholds_potts(beta, k, 2 * dim)
    s = np.ascontiguousarray(s, np.int8).copy()
    lib().orc_potts_sweeps(dim, L, p(s), Q, p(hi), p(lo), nsweeps, seed, sweep0, rep)
    return s


def wolff_ising(dim, L, s, beta, J, nclusters, seed, cl0, rep):
    s = np.ascontiguousarray(s, np.int8).copy()
    tw = lib().orc_wolff_threshold(-np.expm1(2.0 * beta * J))
    sizes = np.zeros(nclusters, np.int64)
    lib().orc_wolff_ising(dim, L, p(s), tw, nclusters, seed, cl0, rep, p(sizes))
    return s, sizes


def wolff_potts(dim, L, s, Q, beta, k, nclusters, seed, cl0, rep):
    s = np.ascontiguousarray(s, np.int8).copy()
    tw = lib().orc_wolff_threshold(-np.expm1(beta * k))
    sizes = np.zeros(nclusters, np.int64)
    lib().orc_wolff_potts(dim, L, p(s), Q, tw, nclusters, seed, cl0, rep, p(sizes))
    return s, sizes


def ref_run(model, dim, L, beta, J=-1.0, Q=3, lam=4.5, mass=1.0, ncluster=0, nmetro=1, nwarm=0, nmeas=0,
            seed=1, init=True, spins=None, bonds=None):
    """Reference-faithful CPU restatement: returns (spins, obs[nmeas, 5], mean_cluster[nmeas]);
    obs columns are E/N, |m| (Potts: its order parameter), m^2, m^4 and the mean |phi|^2."""
    N = L ** dim
    vec = model in (HEIS, PHI4)
    is_ = np.zeros(N, np.int32)
    ds = np.zeros(3 * N, np.float64)
    if spins is not None:
        if vec:
            ds[:] = np.asarray(spins, np.float64).reshape(-1)
        else:
            is_[:] = np.asarray(spins).reshape(-1)
    obs = np.zeros((max(nmeas, 1), 5))
    cl = np.zeros(max(nmeas, 1))
    b = None if bonds is None else np.ascontiguousarray(bonds, np.int8)
    lib().ref_run_regular(model, dim, L, J, Q, lam, mass, None if b is None else p(b), beta, ncluster, nmetro, nwarm,
                          nmeas, seed, int(init), p(is_), p(ds), p(obs), p(cl), None)
    return (ds.reshape(N, 3) if vec else is_), obs[:nmeas], cl[:nmeas]


def ref_chains(model, dim, L, beta, nseeds, seed0=100, threads=8, **kw):
    """nseeds independent chains of ref_run (seeds seed0, seed0+1, ...) on a thread pool (the
    library call releases the GIL); returns obs[nseeds, nmeas, 5]."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=max(1, min(threads, nseeds))) as ex:
        res = list(ex.map(lambda s: ref_run(model, dim, L, beta, seed=seed0 + s, **kw)[1], range(nseeds)))
    return np.array(res)


def bench_wolff(model, beta, spins, seconds, nthreads, dim=0, L=0, rp=None, ci=None, J=-1.0, Q=3, seed=1):
    """Timed CPU Wolff on copies of `spins` (int spins or (N, 3) vectors) on a regular lattice
    (dim, L) or a CSR graph (rp, ci): (aggregate sites/ns, per-thread sites/ns, wall s,
    clusters, mean cluster size)."""
    vec = model in (HEIS, PHI4)
    N = len(rp) - 1 if rp is not None else L ** dim
    is0 = None if vec else np.ascontiguousarray(spins, np.int32).reshape(-1)
    ds0 = np.ascontiguousarray(spins, np.float64).reshape(-1) if vec else None
    out = np.zeros(5)
    rpa = None if rp is None else np.ascontiguousarray(rp, np.int32)
    cia = None if ci is None else np.ascontiguousarray(ci, np.int32)
    lib().ref_bench_wolff(model, dim, L, N, None if rpa is None else p(rpa), None if cia is None else p(cia), J, Q,
                          beta, None if is0 is None else p(is0), None if ds0 is None else p(ds0), seconds, nthreads,
                          seed, p(out))
    return tuple(out)


def csr_sweeps(rp, ci, colour, s, beta, J, nsweeps, seed, sweep0, rep, potts=False, Q=3, bonds=None):
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    colour = np.ascontiguousarray(colour, np.int32)
    N = len(rp) - 1
    qmax = int(np.diff(rp).max())
    hi, lo = thresholds_potts(beta, J, qmax) if potts else thresholds_ising(beta, J, qmax)
    s = np.ascontiguousarray(s, np.int8).copy()
    b = None if bonds is None else np.ascontiguousarray(bonds, np.int8)
    lib().orc_csr_sweeps(N, p(rp), p(ci), p(colour), int(colour.max()) + 1, p(s), None if b is None else p(b),
                         p(hi), p(lo), qmax, int(potts), Q, nsweeps, seed, sweep0, rep)
    return s


def csr_wolff(rp, ci, s, beta, J, nclusters, seed, cl0, rep, potts=False, Q=3):
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    N = len(rp) - 1
    pr = -np.expm1(beta * J) if potts else -np.expm1(2.0 * beta * J)
    tw = lib().orc_wolff_threshold(pr)
    s = np.ascontiguousarray(s, np.int8).copy()
    sizes = np.zeros(nclusters, np.int64)
    lib().orc_csr_wolff(N, p(rp), p(ci), p(s), int(potts), Q, tw, nclusters, seed, cl0, rep, p(sizes))
    return s, sizes


def ising_msc_sweeps(L, s, beta, J, nsweeps, seed, sweep0, rep):
    hi, lo = thresholds_ising(beta, J, 6)
    s = np.ascontiguousarray(s, np.int8).copy()
    lib().orc_ising_msc_sweeps(L, p(s), p(hi), p(lo), nsweeps, seed, sweep0, rep)
    return s


def ea_bits_sweeps(dim, L, spins32, bonds, betas32, J, nsweeps, seed, sweep0, gw):
    """spins32: (32, N) int8; betas32: 32 inverse temperatures (one per bit)."""
    z = 2 * dim
    his, los = [], []
    for b in betas32:
        hi, lo = thresholds_ising(b, J, z)
        his.append(hi)
        los.append(lo)
    thi = np.ascontiguousarray(np.concatenate(his), np.uint32)
    tlo = np.ascontiguousarray(np.concatenate(los), np.uint32)
    s = np.ascontiguousarray(spins32, np.int8).copy()
    lib().orc_ea_bits_sweeps(dim, L, p(s), p(np.ascontiguousarray(bonds, np.int8)), p(thi), p(tlo), nsweeps, seed,
                             sweep0, gw)
    return s
