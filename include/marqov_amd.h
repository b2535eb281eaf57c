/*
 * marqov_amd.h -- C ABI of the MI355X update path (Metropolis sweeps and
 * Wolff single-cluster updates for Ising, +-J Edwards-Anderson, Potts,
 * Heisenberg and Phi4 on regular lattices and CSR graphs).
 *
 * Plain C: opaque handles, POD arguments, int status (0 = ok).  A failing call
 * returns a negative status and leaves a message in mq_last_error(ctx).
 * One context owns one HIP stream on one device; every call on a context is
 * ordered on that stream.  Handles are not thread-safe; distinct contexts may
 * be used from distinct threads (the reference runs one Core per worker thread,
 * src/libmarqov/marqovqueue.h:169-226).
 *
 * Sign convention follows the reference: E = J * sum_<ij> s_i.s_j, so J = -1 is
 * ferromagnetic (src/hamiltonian/util/observables.h:209-216, config/Ising.ini).
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef MARQOV_AMD_H
#define MARQOV_AMD_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__)
#define MQ_API __attribute__((visibility("default")))
#else
#define MQ_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define MQ_OK 0
#define MQ_ERR_ARG (-1)
#define MQ_ERR_HIP (-2)
#define MQ_ERR_STATE (-3)
#define MQ_ERR_UNSUPPORTED (-4)

/* models (src/hamiltonian/{Ising,EdwardsAndersonIsing,Potts,Heisenberg,Phi4}.h) */
#define MQ_ISING 1
#define MQ_EA_ISING 2
#define MQ_POTTS 3
#define MQ_HEISENBERG 4
#define MQ_PHI4 5

/* O(3) Metropolis proposals: the reference's fresh uniform direction
 * (src/hamiltonian/util/initializers.h:61-81) or s' = normalize(s + delta*r). */
#define MQ_PROPOSE_FRESH 0
#define MQ_PROPOSE_SMALL_ANGLE 1

/* Heisenberg Wolff embeddings: the reference's random Cartesian component
 * (src/hamiltonian/Heisenberg.h:191-243) or a random unit vector (Phi4.h:182-244). */
#define MQ_EMBED_CARTESIAN 0
#define MQ_EMBED_RANDOM_DIR 1

/* observables per replica written by mq_measure */
#define MQ_NOBS 4 /* e = E/N, |m|, m^2, m^4 (m = |sum s|/N; Potts: reference order parameter) */

typedef struct mq_ctx mq_ctx;
typedef struct mq_lattice mq_lattice;
typedef struct mq_state mq_state;

typedef struct mq_model_params {
    int model;      /* MQ_ISING ... MQ_PHI4 */
    double J;       /* coupling (Ising/EA/Heisenberg/Phi4) or Potts k */
    int q;          /* Potts number of states (2..8) */
    double lambda;  /* Phi4 quartic coefficient (Phi4.h:85-124) */
    double mass;    /* Phi4 quadratic coefficient */
    int proposal;   /* MQ_PROPOSE_* (Heisenberg) */
    double delta;   /* small-angle step (Heisenberg) */
    int embedding;  /* MQ_EMBED_* (Heisenberg Wolff) */
    int spin_bits;  /* storage: 0 or 8 = int8 per spin; 1 = one bit per spin (no Wolff):
                       Ising on 3D lattices with L a power of two >= 64 packs 32 sites
                       per word; EA (regular lattices) and Ising on graphs (degrees
                       <= 127) pack 32 replicas per word, in runs of multiples of 32
                       replicas per lattice.  Ising on graphs then makes the same moves
                       as int8 storage, bit for bit. */
} mq_model_params;

/* ---- context ---------------------------------------------------------- */
/* Replaces the per-simulation setup of Core (src/libmarqov/core.h:468-494). */
MQ_API int mq_ctx_create(int device, mq_ctx** out);
MQ_API void mq_ctx_destroy(mq_ctx* ctx);
MQ_API const char* mq_last_error(const mq_ctx* ctx);
MQ_API int mq_sync(mq_ctx* ctx);
/* device timing on the context's stream (hipEvent pair) */
MQ_API int mq_timer_begin(mq_ctx* ctx);
MQ_API int mq_timer_end(mq_ctx* ctx, double* ms);

/* ---- lattices --------------------------------------------------------- */
/* Periodic hypercube, dim 2 or 3, L even: FixedDimRegularHypercubic<dim>
 * (src/lattice/regular_hypercubic.h:15-100); site index sum_j x_j L^j. */
MQ_API int mq_lattice_regular(mq_ctx* ctx, int dim, int L, mq_lattice** out);
/* General graph as CSR adjacency (symmetric, no self loops): replaces the
 * nbrs() of GraphFromCSV / RandomGeometricGraph (src/lattice/graph_from_csv.h:32-57,
 * random_geometric_graph.h:258-290).  The library colours it greedily. */
MQ_API int mq_lattice_csr(mq_ctx* ctx, int64_t nsites, const int32_t* row_ptr, const int32_t* col_idx,
                   mq_lattice** out);
/* Random geometric graph built on the device, as CSR for mq_lattice_csr: n
 * points uniform in the periodic unit square (dim 2) or cube (dim 3), joined
 * below minimum-image distance radius; replaces RandomGeometricGraph's PCL
 * kd-tree construction (src/lattice/random_geometric_graph.h:29-168), in 2D as
 * well as 3D.  Point i's coordinate d is the 53-bit Philox uniform of counter
 * (i, d >> 1, 0, 8), key = seed; rows are sorted.  Writes *nnz and, when
 * non-NULL, row_ptr[n+1], col_idx[col_cap >= nnz] and coords[n*dim]; call
 * with col_idx = NULL to learn nnz first.  2 <= n <= 2^30, 0 < radius < 1/2. */
MQ_API int mq_rgg_build(mq_ctx* ctx, int dim, int64_t n, double radius, uint64_t seed, int64_t* nnz,
                        int32_t* row_ptr, int32_t* col_idx, int64_t col_cap, double* coords);
/* number of colour classes and, when sizes is non-NULL, their sizes
 * (sizes[ncolors]: call with sizes = NULL to learn the count first).  A graph
 * of more than 64 classes sweeps with phase stride*sweep + colour, stride the
 * next power of two of the class count (at most 65536 classes). */
MQ_API int mq_lattice_colors(const mq_lattice* lat, int* ncolors, int64_t* sizes);
MQ_API int64_t mq_lattice_nsites(const mq_lattice* lat);
/* colour class of every site in the reference's site order (regular lattices:
 * (x0+x1+x2)&1; graphs: greedy colouring in vertex order) */
MQ_API int mq_lattice_color_of(const mq_lattice* lat, int32_t* colour);
MQ_API void mq_lattice_destroy(mq_lattice* lat);

/* ---- states (replica batches) ------------------------------------------ */
/* nrep replicas of one model; replica r lives on lattices[r].  Replaces the
 * Space of Core (core.h:249-342) for a whole batch of simulations. */
MQ_API int mq_state_create(mq_ctx* ctx, const mq_model_params* params, int nrep, mq_lattice* const* lattices,
                    mq_state** out);
MQ_API void mq_state_destroy(mq_state* st);
/* inverse temperature per replica (Core::beta, core.h:372-494) */
MQ_API int mq_state_set_beta(mq_state* st, const double* beta);
MQ_API int mq_state_get_beta(const mq_state* st, double* beta);
/* EA couplings: rep_sample[r] names the disorder sample of replica r;
 * bonds of sample k are drawn +1 with probability p_plus from (seed, k)
 * (BimodalPDF, src/lattice/util/prob_density.h:25-35; symmetric by construction
 * as RegularRandomBond makes them, src/lattice/regular_random_bond.h:38-54). */
MQ_API int mq_state_random_bonds(mq_state* st, const int32_t* rep_sample, double p_plus, uint64_t seed);
/* replica -> disorder sample map without drawing couplings (then set each
 * sample with mq_state_set_bonds); replicas of one sample must share a lattice */
MQ_API int mq_state_set_samples(mq_state* st, const int32_t* rep_sample);
/* EA couplings from the host for one sample: regular lattices bonds[i*dim+d]
 * couples i and i+e_d; CSR lattices one value per CSR slot (must be symmetric). */
MQ_API int mq_state_set_bonds(mq_state* st, int sample, const int8_t* bonds);
/* random start (Ising.h:99-107, Potts.h:149-156, core.h:774-782) on the device */
MQ_API int mq_state_init(mq_state* st, uint64_t seed);
/* host <-> device in the reference's site order (AoS Space, core.h:270-300):
 * int8 per site for Ising/EA/Potts, float[3] per site for Heisenberg/Phi4. */
MQ_API int mq_state_upload(mq_state* st, int rep, const void* host, int64_t bytes);
MQ_API int mq_state_download(mq_state* st, int rep, void* host, int64_t bytes);

/* ---- updates ---------------------------------------------------------- */
/* nsweeps checkerboard / colour-class Metropolis sweeps of every replica;
 * replaces EMCS's Metropolis loop (emcs.h:93-106) with Metropolis::move
 * (metropolis.h:86-126).  Sweep t uses the 32-bit Philox counter phase
 * 2*(sweep0+t) + colour on regular lattices and 64*(sweep0+t) + colour on
 * graphs (stride*(sweep0+t) + colour for graphs of more than 64 colour classes,
 * see mq_lattice_colors), so every sweep number must stay below 2^31 (regular),
 * 2^26 (graphs of up to 64 classes) or 2^32/stride; calls that would pass it
 * are refused. */
MQ_API int mq_metropolis(mq_state* st, int nsweeps, uint64_t seed, uint64_t sweep0);
/* nclusters single-cluster Wolff moves of every replica; replaces EMCS's
 * cluster loop (emcs.h:83-89) with Wolff::move (wolff.h:48-125, Ising.h:122-169,
 * Heisenberg.h:191-243, Potts.h:262-322).  Cluster c is cluster number
 * cluster0 + c (refused from 2^32 on): its seed site and embedding come from
 * Philox counter (0, number, replica, 3), and every test from a member to a
 * neighbour has its own uniform, as the reference draws rng.real() per test, so
 * the cluster (the sites reachable from the seed along passing tests) depends
 * on the state and the number only, not on launch shapes.  sizes (host, may be
 * NULL) receives nrep*nclusters cluster sizes, replica-major. */
MQ_API int mq_wolff(mq_state* st, int nclusters, uint64_t seed, uint64_t cluster0, int64_t* sizes);
/* one EMCS (emcs.h:69-109): ncluster Wolff moves then nmetro sweeps; writes
 * the per-replica mean cluster size (NaN when ncluster = 0, as emcs.h:108).
 * Equal to mq_wolff(ncluster, cluster0 = step*ncluster) followed by
 * mq_metropolis(nmetro, sweep0 = step*nmetro); negative counts are refused. */
MQ_API int mq_emcs(mq_state* st, int ncluster, int nmetro, uint64_t seed, uint64_t step, double* mean_cluster);
/* nsteps EMCS of every replica in one persistent launch; replaces the EMCS
 * loops of Core::wrmploop / gameloop (core.h:1034-1050, 1001-1020) between
 * measurements.  Equal, bit for bit, to nsteps calls of
 * mq_emcs(ncluster, nmetro, seed, step0 + k), but each replica runs its steps
 * without waiting for the others, so a batch is not held to its slowest
 * replica at every step.  Regular lattices without couplings: Ising and Potts
 * (int8 storage), Heisenberg and Phi4.  No renormalisation happens inside a
 * run.  mean_cluster (host, may be NULL): [nsteps][nrep], NaN when
 * ncluster = 0. */
MQ_API int mq_emcs_run(mq_state* st, int nsteps, int ncluster, int nmetro, uint64_t seed, uint64_t step0,
                       double* mean_cluster);
/* observables, FP64 (observables.h:32-66,173-352; Potts.h:40-65): out[r*MQ_NOBS+k] */
MQ_API int mq_measure(mq_state* st, double* out);
/* No counterpart in the reference: mq_measure for callers that cannot allocate
 * (code under HIP graph capture, tight serving or sampling loops, harnesses that
 * forbid allocation).  Writes the observables of mq_measure to DEVICE memory,
 * d_out[r*MQ_NOBS+k] (nrep*MQ_NOBS doubles), asynchronously on the context's
 * stream (mq_sync before reading them on the host); integer models give the
 * same bits as mq_measure.  It allocates nothing, copies nothing to the host and
 * does not synchronise: its partial sums live in one static device array per
 * GPU (128 B per replica, at most MQ_NOALLOC_MAX_REP replicas, 4 MiB, held
 * until the process ends), which each call clears and refills.  Calls that
 * overlap on one GPU (from contexts on different streams) share that array and
 * race; order them, e.g. by using one context per GPU for these calls.  States
 * with more than MQ_NOALLOC_MAX_REP replicas are refused. */
#define MQ_NOALLOC_MAX_REP 32768
MQ_API int mq_measure_noalloc(mq_state* st, double* d_out);
/* Fourier component of the magnetisation at the smallest wavevector along each
 * axis, MagFTComp (observables.h:104-170; the reference measures it for Ising,
 * axes 0 and 1, Ising.h:82-94, and for Phi4, axes 0-2, Phi4.h:116-140):
 * out[r*3 + d] = sum_j |(1/N) sum_i s_i[j] exp(2 pi i x_d)|^2 with the site
 * coordinate x_d = (index_d + 1/2)/L (regular_hypercubic.h:73-85), NaN for
 * d >= dim.  Ising / EA / Heisenberg / Phi4 on regular lattices, FP64. */
MQ_API int mq_measure_ft(mq_state* st, double* out);
/* total energy per replica (sum over bonds, FP64), as Core::calcAction/beta (core.h:1145-1151) */
MQ_API int mq_energy(mq_state* st, double* out);
/* FP64 renormalisation of Heisenberg (O(3)) spins to unit length: FP32
 * reflections about a random direction drift |s| by rounding, where the
 * reference's doubles keep it (Heisenberg.h:191-243).  The C++ GpuEMCS calls it
 * every 1000 EMCS by default; Phi4 fields have free length and are refused. */
MQ_API int mq_renormalize(mq_state* st);

/* ---- parallel tempering ---------------------------------------------- */
/* Replica-exchange between neighbouring temperatures inside each group of
 * ntemp replicas (replicas g*ntemp .. g*ntemp+ntemp-1 share a sample and hold
 * the temperatures beta[perm]).  Pairs (k, k+1) with k = parity mod 2 are
 * offered a swap accepted with min(1, exp((b_a-b_b)(E_a-E_b))) computed in FP64
 * (calcMHratio / calcprob, src/libmarqov/marqovscheduler.h:44-57,445-465);
 * an accepted swap exchanges the temperatures of the two replicas (the
 * reference exchanges the state spaces, marqovscheduler.h:466-470; the Markov
 * chain is the same).  accepted (host, may be NULL) gets the number of
 * accepted swaps per group. */
MQ_API int mq_pt_swap(mq_state* st, int ntemp, int parity, uint64_t seed, uint64_t step, int32_t* accepted);
/* Replica exchange along an explicit plan: the reference's scheduler offers
 * the pairs of a user-made plan (createPTplan, ptstep,
 * src/libmarqov/marqovscheduler.h:202-211,374-413).  pairs[2p], pairs[2p+1]
 * (p < npairs) are temperature slots 0..ntemp-1 offered in every group with the
 * acceptance of mq_pt_swap; the pairs of one call must be disjoint, a pair of a
 * slot with itself offers nothing.  The uniform of pair (a < b) is keyed by a,
 * so the neighbour pairs (k, k+1) of one parity reproduce mq_pt_swap exactly. */
MQ_API int mq_pt_swap_pairs(mq_state* st, int ntemp, int npairs, const int32_t* pairs, uint64_t seed, uint64_t step,
                            int32_t* accepted);
/* temperature index currently held by each replica (after swaps) */
MQ_API int mq_pt_get_temps(mq_state* st, int32_t* temp_of_rep);

/* ---- launch shapes (set by the autotuner, marqov_amd/autotune.py) ------- */
/* No counterpart in the reference.  kernels: 0 = automatic (LDS multi-sweep
 * kernel for lattices that fit in LDS, row-vectorised 3D kernel where it
 * applies), 1 = streaming kernels only, 2 = generic streaming kernel only.
 * block: threads per block of the streaming sweep kernels (64..512, power of 2);
 * xcd_remap: give each XCD a contiguous slab of rows; lds_block: threads per
 * block of the LDS multi-sweep kernel (64..512).
 * Wolff: wolff_block: threads of the workgroup growing one replica's clusters on
 * large lattices and graphs (0 = automatic: 1024 up to half as many replicas as
 * compute units, else 256; or 256, 512, 1024); wolff_ring: LDS frontier entries
 * per replica of the small-lattice kernel (power of two, 64..2048, default 512);
 * wolff_vset: clusters up to this many sites keep their membership in an LDS
 * hash set, larger ones in global memory (0..4096, default 4096).
 * graph: how mq_run_graph issues its steps: 0 = automatic (the stream),
 * 1 = captured into one HIP graph and replayed, 2 = launched on the stream.
 * Results do not depend on any of these. */
typedef struct mq_tuning {
    int kernels;
    int block;
    int xcd_remap;
    int lds_block;
    int wolff_block;
    int wolff_ring;
    int wolff_vset;
    int graph;
} mq_tuning;
MQ_API int mq_state_set_tuning(mq_state* st, const mq_tuning* t);
MQ_API int mq_state_get_tuning(const mq_state* st, mq_tuning* t);

/* ---- graphs ------------------------------------------------------------ */
/* `steps` repetitions of {nsweeps Metropolis sweeps; optional PT swap with
 * ntemp temperatures}, run `repeats` times with no host synchronisation in
 * between; sweep and step counters advance per repetition from device memory.
 * mq_tuning.graph = 1 captures the repetitions into one HIP graph and replays
 * it; the default launches the same kernels, in the same order and with the
 * same arguments, on the state's stream (on MI355X with ROCm 7 a replayed
 * kernel node costs about 2.7 us more than a stream launch).  Both give the
 * same states bit for bit. */
MQ_API int mq_run_graph(mq_state* st, int nsweeps, int ntemp, uint64_t seed, uint64_t sweep0, uint64_t step0,
                 int steps, int repeats);

#ifdef __cplusplus
}
#endif
#endif
