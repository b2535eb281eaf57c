"""Python host API over the C ABI (include/marqov_amd.h).

Mirrors the reference's objects for the update path: a lattice
(FixedDimRegularHypercubic / GraphFromCSV), a model with its couplings
(Ising / EdwardsAndersonIsing / Potts / Heisenberg / Phi4), a state space per
replica (Core::statespace) and the EMCS / Metropolis / Wolff moves.  Errors
raise MarqovError with the library's message, as the reference raises
std::runtime_error (src/libmarqov/core.h:338-339).
"""
import ctypes

import numpy as np

from . import _lib as L

MODELS = {
    "ising": L.MQ_ISING,
    "ea": L.MQ_EA_ISING,
    "potts": L.MQ_POTTS,
    "heisenberg": L.MQ_HEISENBERG,
    "phi4": L.MQ_PHI4,
}


class MarqovError(RuntimeError):
    pass


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Context:
    def __init__(self, device=0):
        self.lib = L.load()
        h = ctypes.c_void_p()
        rc = self.lib.mq_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise MarqovError(f"mq_ctx_create failed ({rc}): no usable HIP device {device}")
        self.h = h

    def check(self, rc, what):
        if rc != 0:
            msg = self.lib.mq_last_error(self.h)
            raise MarqovError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def sync(self):
        self.check(self.lib.mq_sync(self.h), "mq_sync")

    def timer_begin(self):
        self.check(self.lib.mq_timer_begin(self.h), "mq_timer_begin")

    def timer_end(self):
        ms = ctypes.c_double()
        self.check(self.lib.mq_timer_end(self.h, ctypes.byref(ms)), "mq_timer_end")
        return ms.value

    def rgg(self, dim, n, radius, seed, coords=False):
        """Random geometric graph built on the device (mq_rgg_build): returns
        (row_ptr int32[n+1], col_idx int32[nnz]) and, with coords=True, the points
        float64[n, dim]."""
        nnz = ctypes.c_int64()
        rp = np.zeros(n + 1, np.int32)
        self.check(self.lib.mq_rgg_build(self.h, dim, n, float(radius), seed, ctypes.byref(nnz), _ptr(rp), None, 0,
                                         None), "mq_rgg_build")
        ci = np.zeros(max(nnz.value, 1), np.int32)
        xyz = np.zeros((n, dim), np.float64) if coords else None
        self.check(self.lib.mq_rgg_build(self.h, dim, n, float(radius), seed, ctypes.byref(nnz), _ptr(rp), _ptr(ci),
                                         ci.size, _ptr(xyz) if coords else None), "mq_rgg_build")
        ci = ci[:nnz.value]
        return (rp, ci, xyz) if coords else (rp, ci)

    def close(self):
        if self.h:
            self.lib.mq_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RegularLattice:
    """Periodic hypercube (src/lattice/regular_hypercubic.h:15-100)."""

    def __init__(self, ctx, dim, L_):
        self.ctx, self.dim, self.L = ctx, dim, L_
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.mq_lattice_regular(ctx.h, dim, L_, ctypes.byref(h)), "mq_lattice_regular")
        self.h = h
        self.N = L_ ** dim

    def colors(self):
        n = ctypes.c_int()
        self.ctx.check(self.ctx.lib.mq_lattice_colors(self.h, ctypes.byref(n), None), "mq_lattice_colors")
        sizes = np.zeros(max(n.value, 1), np.int64)
        self.ctx.check(self.ctx.lib.mq_lattice_colors(self.h, ctypes.byref(n), _ptr(sizes)), "mq_lattice_colors")
        return sizes[: n.value].copy()

    def __del__(self):
        # a closed context has released the device; its lattices must not be touched
        try:
            if self.ctx.h:
                self.ctx.lib.mq_lattice_destroy(self.h)
        except Exception:
            pass


class CsrLattice(RegularLattice):
    """General graph from CSR adjacency (src/lattice/graph_from_csv.h:32-57)."""

    def __init__(self, ctx, row_ptr, col_idx):  # noqa: D401 - deliberate no super().__init__
        self.ctx = ctx
        self.row_ptr = np.ascontiguousarray(row_ptr, np.int32)
        self.col_idx = np.ascontiguousarray(col_idx, np.int32)
        self.N = len(self.row_ptr) - 1
        self.dim = 0
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.mq_lattice_csr(ctx.h, self.N, _ptr(self.row_ptr), _ptr(self.col_idx), ctypes.byref(h)),
                  "mq_lattice_csr")
        self.h = h


class State:
    """A batch of replicas of one model (one Core::statespace per replica)."""

    def __init__(self, ctx, model, lattices, J=-1.0, q=3, lam=4.5, mass=1.0, proposal="fresh", delta=0.5,
                 embedding="cartesian", spin_bits=8):
        self.ctx = ctx
        self.model = model
        self.lattices = list(lattices)
        self.nrep = len(self.lattices)
        self.J = J
        self.q = q
        p = L.ModelParams()
        p.model = MODELS[model]
        p.J = J
        p.q = q
        p.lambda_ = lam
        p.mass = mass
        p.proposal = {"fresh": L.MQ_PROPOSE_FRESH, "small": L.MQ_PROPOSE_SMALL_ANGLE}[proposal]
        p.delta = delta
        p.embedding = {"cartesian": L.MQ_EMBED_CARTESIAN, "random": L.MQ_EMBED_RANDOM_DIR}[embedding]
        p.spin_bits = spin_bits
        self.params = p
        arr = (ctypes.c_void_p * self.nrep)(*[lat.h.value for lat in self.lattices])
        self._arr = arr
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.mq_state_create(ctx.h, ctypes.byref(p), self.nrep, ctypes.cast(arr, ctypes.c_void_p),
                                          ctypes.byref(h)), "mq_state_create")
        self.h = h
        self.vec = model in ("heisenberg", "phi4")

    # -- setup
    def set_beta(self, beta):
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(beta, np.float64), (self.nrep,)))
        self.ctx.check(self.ctx.lib.mq_state_set_beta(self.h, _ptr(b)), "mq_state_set_beta")

    def init(self, seed):
        self.ctx.check(self.ctx.lib.mq_state_init(self.h, seed), "mq_state_init")

    def random_bonds(self, rep_sample=None, p_plus=0.5, seed=4):
        rs = None if rep_sample is None else np.ascontiguousarray(rep_sample, np.int32)
        self.ctx.check(self.ctx.lib.mq_state_random_bonds(self.h, None if rs is None else _ptr(rs), p_plus, seed),
                       "mq_state_random_bonds")

    def set_samples(self, rep_sample):
        rs = np.ascontiguousarray(rep_sample, np.int32)
        self.ctx.check(self.ctx.lib.mq_state_set_samples(self.h, _ptr(rs)), "mq_state_set_samples")

    def set_bonds(self, sample, bonds):
        b = np.ascontiguousarray(bonds, np.int8)
        self.ctx.check(self.ctx.lib.mq_state_set_bonds(self.h, sample, _ptr(b)), "mq_state_set_bonds")

    def upload(self, rep, spins):
        a = np.ascontiguousarray(spins, np.float32 if self.vec else np.int8)
        self.ctx.check(self.ctx.lib.mq_state_upload(self.h, rep, _ptr(a), a.nbytes), "mq_state_upload")

    def download(self, rep):
        N = self.lattices[rep].N
        a = np.zeros((N, 3), np.float32) if self.vec else np.zeros(N, np.int8)
        self.ctx.check(self.ctx.lib.mq_state_download(self.h, rep, _ptr(a), a.nbytes), "mq_state_download")
        return a

    # -- moves
    def metropolis(self, nsweeps, seed, sweep0=0):
        self.ctx.check(self.ctx.lib.mq_metropolis(self.h, nsweeps, seed, sweep0), "mq_metropolis")

    def wolff(self, nclusters, seed, cluster0=0, sizes=True):
        out = np.zeros((self.nrep, nclusters), np.int64) if sizes else None
        self.ctx.check(self.ctx.lib.mq_wolff(self.h, nclusters, seed, cluster0, None if out is None else _ptr(out)),
                       "mq_wolff")
        return out

    def emcs(self, ncluster, nmetro, seed, step):
        mc = np.zeros(self.nrep, np.float64)
        self.ctx.check(self.ctx.lib.mq_emcs(self.h, ncluster, nmetro, seed, step, _ptr(mc)), "mq_emcs")
        return mc

    def emcs_run(self, nsteps, ncluster, nmetro, seed, step0, means=True):
        """nsteps EMCS per replica in one persistent launch (mq_emcs_run); returns the
        (nsteps, nrep) mean cluster sizes or None."""
        mc = np.zeros((nsteps, self.nrep), np.float64) if means else None
        self.ctx.check(self.ctx.lib.mq_emcs_run(self.h, nsteps, ncluster, nmetro, seed, step0,
                                                None if mc is None else _ptr(mc)), "mq_emcs_run")
        return mc

    def measure(self):
        out = np.zeros((self.nrep, L.MQ_NOBS), np.float64)
        self.ctx.check(self.ctx.lib.mq_measure(self.h, _ptr(out)), "mq_measure")
        return out

    def measure_ft(self):
        """MagFTComp per replica and axis (observables.h:104-170): (nrep, 3), NaN for axes >= dim."""
        out = np.zeros((self.nrep, 3), np.float64)
        self.ctx.check(self.ctx.lib.mq_measure_ft(self.h, _ptr(out)), "mq_measure_ft")
        return out

    def energy(self):
        out = np.zeros(self.nrep, np.float64)
        self.ctx.check(self.ctx.lib.mq_energy(self.h, _ptr(out)), "mq_energy")
        return out

    def renormalize(self):
        self.ctx.check(self.ctx.lib.mq_renormalize(self.h), "mq_renormalize")

    def pt_swap(self, ntemp, parity, seed, step):
        acc = np.zeros(self.nrep // ntemp, np.int32)
        self.ctx.check(self.ctx.lib.mq_pt_swap(self.h, ntemp, parity, seed, step, _ptr(acc)), "mq_pt_swap")
        return acc

    def pt_swap_pairs(self, ntemp, pairs, seed, step):
        """Offer the disjoint temperature-slot pairs of one step of a PT plan in every group."""
        pr = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1, 2))
        acc = np.zeros(self.nrep // ntemp, np.int32)
        self.ctx.check(self.ctx.lib.mq_pt_swap_pairs(self.h, ntemp, len(pr), _ptr(pr), seed, step, _ptr(acc)),
                       "mq_pt_swap_pairs")
        return acc

    def pt_temps(self):
        t = np.zeros(self.nrep, np.int32)
        self.ctx.check(self.ctx.lib.mq_pt_get_temps(self.h, _ptr(t)), "mq_pt_get_temps")
        return t

    def set_tuning(self, **kw):
        """Launch shapes (mq_tuning); fields not given keep their current values."""
        cur = self.get_tuning()
        unknown = set(kw) - set(cur)
        if unknown:
            raise TypeError(f"unknown tuning fields {sorted(unknown)}")
        cur.update(kw)
        t = L.Tuning(*[int(cur[f]) for f, _ in L.Tuning._fields_])
        self.ctx.check(self.ctx.lib.mq_state_set_tuning(self.h, ctypes.byref(t)), "mq_state_set_tuning")

    def get_tuning(self):
        t = L.Tuning()
        self.ctx.check(self.ctx.lib.mq_state_get_tuning(self.h, ctypes.byref(t)), "mq_state_get_tuning")
        return {f: getattr(t, f) for f, _ in L.Tuning._fields_}

    def run_graph(self, nsweeps, ntemp, seed, sweep0, step0, steps, repeats):
        """`repeats` x `steps` steps of {nsweeps sweeps; PT swap over ntemp temperatures if
        ntemp > 1} with no host synchronisation in between (mq_run_graph): launched on the
        stream, or captured in a HIP graph and replayed with set_tuning(graph=1)."""
        self.ctx.check(self.ctx.lib.mq_run_graph(self.h, nsweeps, ntemp, seed, sweep0, step0, steps, repeats),
                       "mq_run_graph")

    def __del__(self):
        try:
            if self.ctx.h:
                self.ctx.lib.mq_state_destroy(self.h)
        except Exception:
            pass
