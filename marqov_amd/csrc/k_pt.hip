// Parallel tempering, O(3) renormalisation and HIP-graph replay.
//
// PT follows the reference's acceptance (src/libmarqov/marqovscheduler.h:44-57,
// 389-397, 445-465): endiff = S_aa - S_ab + S_bb - S_ba with S_xy = beta_x E(y),
// i.e. (beta_a - beta_b)(E_a - E_b); ratio = calcMHratio(endiff) (exp with the
// double range clamped); accept iff u < min(1, ratio).  The reference swaps the
// state spaces of the two simulations; here the two replicas exchange their
// temperatures (acceptance tables, beta) instead, which moves 2 table rows
// rather than two lattices and leaves the Markov chain unchanged.  mq_pt_swap
// offers neighbouring temperatures (k, k+1) with k = parity mod 2 inside each
// group of ntemp replicas; mq_pt_swap_pairs offers the disjoint slot pairs of one
// step of a caller's plan, such as one drawn by the reference's createPTplan
// (marqovscheduler.h:202-211).  For models whose
// Hamiltonian depends on beta only through on-site terms scaled by 1/beta
// (Phi4, Phi4.h:110-124), beta_x E_x(y) = beta_x J I(y) + O(y), so only the
// interaction sum I enters endiff.
#include <cfloat>

#include "mq_internal.hpp"

int mq_metropolis_async(mq_state* st, int nsweeps, uint64_t seed, uint64_t sweep0);

namespace {

__global__ void k_zero64(unsigned long long* p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// The interaction energy J*I of replica r, read by the swap from the measurement
// sums (integer sums for int8/bit storage, FP32-accumulated doubles for O(3)/Phi4).
struct EnergySrc {
    const long long* isum;
    const double* fsum;
    int vec;
    double J;
};

__device__ __forceinline__ double replica_energy(const EnergySrc& es, int r)
{
    return es.vec ? es.J * es.fsum[(int64_t)r * 8] : es.J * (double)es.isum[(int64_t)r * 8];
}

__device__ __forceinline__ double mh_ratio(double en)
{
    // calcMHratio (marqovscheduler.h:44-57)
    if (en > log(DBL_MIN)) {
        if (en > log(DBL_MAX)) return DBL_MAX;
        return exp(en);
    }
    return 0.0;
}

// table words exchanged per round trip (2*tbl = 26 for 3D lattices)
constexpr int PT_CHUNK = 32;

// One offered exchange between temperature slots ka < kb of group g: the
// replicas a, b holding them swap temperatures with probability
// min(1, calcMHratio((beta_ka - beta_kb)(E_a - E_b))); the uniform is the
// 53-bit Philox draw keyed by (ka, step, g).

__device__ __forceinline__ void pt_offer(int g, int ntemp, int ka, int kb, uint32_t step, Key2 key,
                                         int32_t* __restrict__ rep_of_temp, int32_t* __restrict__ temp_of_rep,
                                         const EnergySrc es, const double* __restrict__ ladder,
                                         uint32_t* __restrict__ thr, int tbl, uint32_t* __restrict__ wthr,
                                         float* __restrict__ betaf, double* __restrict__ betad, int32_t* __restrict__ acc)
{
    const int64_t sa = (int64_t)g * ntemp + ka, sb = (int64_t)g * ntemp + kb;
    const int a = rep_of_temp[sa], b = rep_of_temp[sb];
    const double en = (ladder[sa] - ladder[sb]) * (replica_energy(es, a) - replica_energy(es, b));
    const double ratio = mh_ratio(en);
    const uint4 w = philox4x32_10((uint32_t)ka, step, (uint32_t)g, ST_PT, key);
    const double u = (double)((((uint64_t)w.x << 32) | w.y) >> 11) * 0x1.0p-53;
    if (!(u < fmin(1.0, ratio))) return;
    // Exchange the two replicas' temperature data.  Every value is loaded before any
    // is stored, PT_CHUNK table words at a time: the rows share one array, so loads
    // placed after stores would wait for them, one memory round trip per word.
    const uint32_t wa = wthr[a], wb = wthr[b];
    const float fa = betaf[a], fb = betaf[b];
    const double da = betad[a], db = betad[b];
    uint32_t* ra = thr + (int64_t)a * 2 * tbl;
    uint32_t* rb = thr + (int64_t)b * 2 * tbl;
    const int ne = 2 * tbl;
    for (int e0 = 0; e0 < ne; e0 += PT_CHUNK) {
        // loads at a clamped (in-row) index, unconditionally, so they issue together
        uint32_t xa[PT_CHUNK], xb[PT_CHUNK];
#pragma unroll
        for (int j = 0; j < PT_CHUNK; ++j) {
            const int e = min(e0 + j, ne - 1);
            xa[j] = ra[e];
            xb[j] = rb[e];
        }
#pragma unroll
        for (int j = 0; j < PT_CHUNK; ++j) {
            if (e0 + j < ne) {
                ra[e0 + j] = xb[j];
                rb[e0 + j] = xa[j];
            }
        }
    }
    rep_of_temp[sa] = b;
    rep_of_temp[sb] = a;
    temp_of_rep[a] = kb;
    temp_of_rep[b] = ka;
    wthr[a] = wb;
    wthr[b] = wa;
    betaf[a] = fb;
    betaf[b] = fa;
    betad[a] = db;
    betad[b] = da;
    atomicAdd(acc + g, 1);
}

// neighbouring temperatures (k, k+1), k = parity mod 2, in every group
__global__ void k_pt_swap(int ngroups, int ntemp, int parity, uint32_t step, const uint32_t* __restrict__ dyn, Key2 key,
                          int32_t* __restrict__ rep_of_temp, int32_t* __restrict__ temp_of_rep,
                          const EnergySrc es, const double* __restrict__ ladder,
                          uint32_t* __restrict__ thr, int tbl, uint32_t* __restrict__ wthr, float* __restrict__ betaf,
                          double* __restrict__ betad, int32_t* __restrict__ acc)
{
    if (dyn) step += dyn[1];
    if (parity < 0) parity = (int)(step & 1u);
    const int npairs = (ntemp - parity) / 2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ngroups * npairs) return;
    const int g = (int)(t / npairs);
    const int k = parity + 2 * (int)(t - (int64_t)g * npairs);
    if (k + 1 >= ntemp) return;
    pt_offer(g, ntemp, k, k + 1, step, key, rep_of_temp, temp_of_rep, es, ladder, thr, tbl, wthr, betaf, betad, acc);
}

// a plan of npairs disjoint slot pairs (lower slot first), the same in every group
__global__ void k_pt_swap_pairs(int ngroups, int ntemp, int npairs, const int32_t* __restrict__ pairs, uint32_t step,
                                Key2 key, int32_t* __restrict__ rep_of_temp, int32_t* __restrict__ temp_of_rep,
                                const EnergySrc es, const double* __restrict__ ladder,
                                uint32_t* __restrict__ thr, int tbl, uint32_t* __restrict__ wthr,
                                float* __restrict__ betaf, double* __restrict__ betad, int32_t* __restrict__ acc)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ngroups * npairs) return;
    const int g = (int)(t / npairs);
    const int p = (int)(t - (int64_t)g * npairs);
    pt_offer(g, ntemp, pairs[2 * p], pairs[2 * p + 1], step, key, rep_of_temp, temp_of_rep, es, ladder, thr, tbl,
             wthr, betaf, betad, acc);
}

__global__ void k_advance(uint32_t* dyn, uint32_t dphase, uint32_t dstep)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        dyn[0] += dphase;
        dyn[1] += dstep;
    }
}

// FP64 renormalisation of float4 spins: s <- s / |s|
__global__ void k_renorm(float4* __restrict__ s, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 v = s[i];
    const double nn = sqrt((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
    if (nn > 0.0) {
        v.x = (float)(v.x / nn);
        v.y = (float)(v.y / nn);
        v.z = (float)(v.z / nn);
    }
    s[i] = v;
}

hipError_t zero_fill(void* p, int64_t n64, hipStream_t s)
{
    if (n64 <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero64, dim3((unsigned)((n64 + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<unsigned long long*>(p), n64);
    return hipGetLastError();
}

bool vec_model(const mq_state* st) { return st->p.model == MQ_HEISENBERG || st->p.model == MQ_PHI4; }

// measurement sums on the device (no host synchronisation: usable inside a graph);
// the swap kernels read the energies from them (replica_energy)
hipError_t energy_async(mq_state* st)
{
    hipStream_t s = st->ctx->stream;
    hipError_t e;
    if ((e = zero_fill(st->d_isum, (int64_t)st->nrep * 8, s)) != hipSuccess) return e;
    // the FP32 sums exist for O(3)/Phi4 spins only; integer models neither write nor read them
    if (vec_model(st) && (e = zero_fill(st->d_fsum, (int64_t)st->nrep * 8, s)) != hipSuccess) return e;
    return launch_measure_all(st, st->d_isum, st->d_fsum);
}

EnergySrc energy_src(const mq_state* st) { return EnergySrc{st->d_isum, st->d_fsum, vec_model(st) ? 1 : 0, st->p.J}; }

int ensure_pt(mq_state* st, int ntemp)
{
    if (ntemp < 2 || st->nrep % ntemp != 0)
        return mq_fail(st->ctx, MQ_ERR_ARG, "ntemp must be >= 2 and divide the number of replicas");
    for (int g = 0; g < st->nrep / ntemp; ++g)
        for (int k = 1; k < ntemp; ++k)
            if (st->lat[g * ntemp + k] != st->lat[g * ntemp])
                return mq_fail(st->ctx, MQ_ERR_ARG, "replicas of one PT group must share a lattice");
    if (st->pt_ntemp == ntemp && st->d_rep_of_temp) return MQ_OK;
    if (!st->d_rep_of_temp) {
        MQ_HIP(st->ctx, hipMalloc(&st->d_rep_of_temp, sizeof(int32_t) * st->nrep));
        MQ_HIP(st->ctx, hipMalloc(&st->d_temp_of_rep, sizeof(int32_t) * st->nrep));
        MQ_HIP(st->ctx, hipMalloc(&st->d_ladder, sizeof(double) * st->nrep));
        MQ_HIP(st->ctx, hipMalloc(&st->d_pt_acc, sizeof(int32_t) * (st->nrep + 2)));
    }
    std::vector<int32_t> rot(st->nrep), tor(st->nrep);
    for (int r = 0; r < st->nrep; ++r) {
        rot[r] = r;
        tor[r] = r % ntemp;
    }
    MQ_HIP(st->ctx, hipMemcpy(st->d_rep_of_temp, rot.data(), sizeof(int32_t) * st->nrep, hipMemcpyHostToDevice));
    MQ_HIP(st->ctx, hipMemcpy(st->d_temp_of_rep, tor.data(), sizeof(int32_t) * st->nrep, hipMemcpyHostToDevice));
    MQ_HIP(st->ctx, hipMemcpy(st->d_ladder, st->beta.data(), sizeof(double) * st->nrep, hipMemcpyHostToDevice));
    st->pt_ntemp = ntemp;
    return MQ_OK;
}

hipError_t pt_swap_async(mq_state* st, int ntemp, int parity, uint32_t step, Key2 key)
{
    const int ngroups = st->nrep / ntemp;
    const int64_t work = (int64_t)ngroups * ((ntemp + 1) / 2);
    hipLaunchKernelGGL(k_pt_swap, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, st->ctx->stream, ngroups, ntemp,
                       parity, step, st->cur_dyn, key, st->d_rep_of_temp, st->d_temp_of_rep, energy_src(st), st->d_ladder,
                       st->d_thr, st->tbl, st->d_wthr, st->d_betaf, st->d_betad, st->d_pt_acc);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && st->eabits) e = launch_wtab_eabits(st);  // temperatures moved between bits
    return e;
}

}  // namespace

extern "C" int mq_pt_swap(mq_state* st, int ntemp, int parity, uint64_t seed, uint64_t step, int32_t* accepted)
{
    if (!st) return MQ_ERR_ARG;
    if (st->beta.empty()) return mq_fail(st->ctx, MQ_ERR_STATE, "beta not set");
    if (parity != 0 && parity != 1) return mq_fail(st->ctx, MQ_ERR_ARG, "parity must be 0 or 1");
    int rc = ensure_pt(st, ntemp);
    if (rc) return rc;
    const Key2 key{(uint32_t)seed, (uint32_t)(seed >> 32)};
    hipStream_t s = st->ctx->stream;
    MQ_HIP(st->ctx, zero_fill(st->d_pt_acc, (st->nrep + 1) / 2, s));
    MQ_HIP(st->ctx, energy_async(st));
    MQ_HIP(st->ctx, pt_swap_async(st, ntemp, parity, (uint32_t)step, key));
    if (accepted) {
        MQ_HIP(st->ctx, hipMemcpyAsync(accepted, st->d_pt_acc, sizeof(int32_t) * (st->nrep / ntemp),
                                       hipMemcpyDeviceToHost, s));
    }
    MQ_HIP(st->ctx, hipStreamSynchronize(s));
    return MQ_OK;
}

extern "C" int mq_pt_swap_pairs(mq_state* st, int ntemp, int npairs, const int32_t* pairs, uint64_t seed, uint64_t step,
                                int32_t* accepted)
{
    if (!st) return MQ_ERR_ARG;
    if (npairs < 0 || (npairs > 0 && !pairs)) return mq_fail(st->ctx, MQ_ERR_ARG, "pairs: npairs < 0 or pairs is NULL");
    if (st->beta.empty()) return mq_fail(st->ctx, MQ_ERR_STATE, "beta not set");
    int rc = ensure_pt(st, ntemp);
    if (rc) return rc;
    // validate and normalise: slots in range, lower slot first, every slot at most once;
    // a pair of a slot with itself offers nothing (as the reference's ptstep, which
    // skips self swaps) and is dropped
    std::vector<int32_t> plan;
    std::vector<char> used((size_t)ntemp, 0);
    for (int p = 0; p < npairs; ++p) {
        int a = pairs[2 * p], b = pairs[2 * p + 1];
        if (a < 0 || b < 0 || a >= ntemp || b >= ntemp)
            return mq_fail(st->ctx, MQ_ERR_ARG, "PT pair slot outside 0..ntemp-1");
        if (a == b) continue;
        if (a > b) std::swap(a, b);
        if (used[(size_t)a] || used[(size_t)b])
            return mq_fail(st->ctx, MQ_ERR_ARG, "PT pairs of one step must be disjoint");
        used[(size_t)a] = used[(size_t)b] = 1;
        plan.push_back(a);
        plan.push_back(b);
    }
    const int np = (int)plan.size() / 2;
    const Key2 key{(uint32_t)seed, (uint32_t)(seed >> 32)};
    hipStream_t s = st->ctx->stream;
    MQ_HIP(st->ctx, zero_fill(st->d_pt_acc, (st->nrep + 1) / 2, s));
    if (np > 0) {
        if ((int)st->pt_pairs_cap < np) {
            if (st->d_pt_pairs) MQ_HIP(st->ctx, hipFree(st->d_pt_pairs));
            st->d_pt_pairs = nullptr;
            MQ_HIP(st->ctx, hipMalloc(&st->d_pt_pairs, sizeof(int32_t) * 2 * (size_t)np));
            st->pt_pairs_cap = np;
        }
        MQ_HIP(st->ctx, hipMemcpyAsync(st->d_pt_pairs, plan.data(), sizeof(int32_t) * 2 * (size_t)np,
                                       hipMemcpyHostToDevice, s));
        MQ_HIP(st->ctx, energy_async(st));
        const int ngroups = st->nrep / ntemp;
        const int64_t work = (int64_t)ngroups * np;
        hipLaunchKernelGGL(k_pt_swap_pairs, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, ngroups, ntemp, np,
                           st->d_pt_pairs, (uint32_t)step, key, st->d_rep_of_temp, st->d_temp_of_rep, energy_src(st),
                           st->d_ladder, st->d_thr, st->tbl, st->d_wthr, st->d_betaf, st->d_betad, st->d_pt_acc);
        MQ_HIP(st->ctx, hipGetLastError());
        if (st->eabits) MQ_HIP(st->ctx, launch_wtab_eabits(st));
    }
    if (accepted) {
        MQ_HIP(st->ctx, hipMemcpyAsync(accepted, st->d_pt_acc, sizeof(int32_t) * (st->nrep / ntemp),
                                       hipMemcpyDeviceToHost, s));
    }
    MQ_HIP(st->ctx, hipStreamSynchronize(s));
    return MQ_OK;
}

extern "C" int mq_pt_get_temps(mq_state* st, int32_t* temp_of_rep)
{
    if (!st) return MQ_ERR_ARG;
    if (!temp_of_rep) return mq_fail(st->ctx, MQ_ERR_ARG, "temp_of_rep is NULL");
    if (!st->d_temp_of_rep) {
        for (int r = 0; r < st->nrep; ++r) temp_of_rep[r] = -1;
        return MQ_OK;
    }
    MQ_HIP(st->ctx, hipStreamSynchronize(st->ctx->stream));
    MQ_HIP(st->ctx, hipMemcpy(temp_of_rep, st->d_temp_of_rep, sizeof(int32_t) * st->nrep, hipMemcpyDeviceToHost));
    return MQ_OK;
}

extern "C" int mq_renormalize(mq_state* st)
{
    if (!st) return MQ_ERR_ARG;
    if (st->p.model != MQ_HEISENBERG)
        return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED,
                       "renormalisation applies to O(3) spins (Heisenberg; Phi4 fields have free length)");
    const int64_t n = st->spin_total / 16;
    hipLaunchKernelGGL(k_renorm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st->ctx->stream,
                       reinterpret_cast<float4*>(st->d_spins), n);
    MQ_HIP(st->ctx, hipGetLastError());
    return MQ_OK;
}

extern "C" int mq_run_graph(mq_state* st, int nsweeps, int ntemp, uint64_t seed, uint64_t sweep0, uint64_t step0,
                            int steps, int repeats)
{
    if (!st) return MQ_ERR_ARG;
    if (st->beta.empty()) return mq_fail(st->ctx, MQ_ERR_STATE, "beta not set");
    if (st->p.model == MQ_EA_ISING && !st->d_bonds) return mq_fail(st->ctx, MQ_ERR_STATE, "EA couplings not set");
    if (nsweeps < 0 || steps <= 0 || repeats <= 0) return mq_fail(st->ctx, MQ_ERR_ARG, "bad graph sizes");
    // the replay advances every sweep kernel's phase by one device offset
    const uint64_t per_sweep = max_color_stride(st);
    for (const auto* l : st->lat)
        if (l->color_stride != per_sweep)
            return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED,
                           "graph replay needs one colour stride for all lattices of the state (graphs of up to 64 "
                           "colour classes, or of one class count)");
    const uint64_t total = sweep0 + (uint64_t)nsweeps * steps * repeats;
    if (sweep0 >= sweep_limit(st) || total > sweep_limit(st)) return mq_fail(st->ctx, MQ_ERR_ARG, sweep_limit_msg(st));
    if (ntemp > 1) {
        int rc = ensure_pt(st, ntemp);
        if (rc) return rc;
    }
    hipStream_t s = st->ctx->stream;
    if (!st->d_dyn) MQ_HIP(st->ctx, hipMalloc(&st->d_dyn, 2 * sizeof(uint32_t)));
    MQ_HIP(st->ctx, zero_fill(st->d_dyn, 1, s));
    if (ntemp > 1) MQ_HIP(st->ctx, zero_fill(st->d_pt_acc, (st->nrep + 1) / 2, s));
    MQ_HIP(st->ctx, hipStreamSynchronize(s));
    const Key2 key{(uint32_t)seed, (uint32_t)(seed >> 32)};
    // One repetition of the steps: every launch reads its sweep and step offsets
    // from d_dyn, which the last launch advances, so issuing it again (or replaying
    // its capture) continues the counters.
    hipError_t e = hipSuccess;
    auto issue = [&]() -> int {
        st->cur_dyn = st->d_dyn;
        int r = MQ_OK;
        for (int t = 0; t < steps && r == MQ_OK && e == hipSuccess; ++t) {
            r = mq_metropolis_async(st, nsweeps, seed, sweep0 + (uint64_t)t * nsweeps);
            if (r == MQ_OK && ntemp > 1) {
                e = energy_async(st);
                if (e == hipSuccess) e = pt_swap_async(st, ntemp, -1, (uint32_t)(step0 + t), key);
            }
        }
        if (r == MQ_OK && e == hipSuccess) {
            hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s, st->d_dyn, (uint32_t)(per_sweep * nsweeps * steps),
                               (uint32_t)steps);
            e = hipGetLastError();
        }
        st->cur_dyn = nullptr;
        return r;
    };
    if (st->step_graph != 1) {  // on the stream
        int rc = MQ_OK;
        for (int i = 0; i < repeats && rc == MQ_OK && e == hipSuccess; ++i) rc = issue();
        if (rc != MQ_OK) return rc;
        hipError_t es = hipStreamSynchronize(s);
        if (e != hipSuccess || es != hipSuccess)
            return mq_fail(st->ctx, MQ_ERR_HIP, std::string("step launch: ") + hipGetErrorString(e != hipSuccess ? e : es));
        return MQ_OK;
    }
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    MQ_HIP(st->ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = issue();
    hipError_t ec = hipStreamEndCapture(s, &graph);
    if (rc != MQ_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return rc;
    }
    if (e != hipSuccess || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return mq_fail(st->ctx, MQ_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : ec));
    }
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return mq_fail(st->ctx, MQ_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    for (int i = 0; i < repeats && e == hipSuccess; ++i) e = hipGraphLaunch(exec, s);
    hipError_t es = hipStreamSynchronize(s);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess || es != hipSuccess)
        return mq_fail(st->ctx, MQ_ERR_HIP, std::string("graph launch: ") + hipGetErrorString(e != hipSuccess ? e : es));
    return MQ_OK;
}
