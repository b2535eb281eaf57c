// Host side of the C ABI: contexts, lattices, replica batches, acceptance
// tables and the update / measurement entry points declared in
// include/marqov_amd.h.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <numeric>

#include "mq_internal.hpp"

int mq_fail(mq_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

namespace {

template <class T>
hipError_t dalloc(T** p, size_t n)
{
    *p = nullptr;
    if (n == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
}

template <class T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

Key2 key_of(uint64_t seed) { return Key2{(uint32_t)seed, (uint32_t)(seed >> 32)}; }

bool is_vec(int model) { return model == MQ_HEISENBERG || model == MQ_PHI4; }
bool is_int(int model) { return model == MQ_ISING || model == MQ_EA_ISING || model == MQ_POTTS; }

void split_threshold(double p, uint32_t* hi, uint32_t* lo)
{
    if (!(p < 1.0)) {
        *hi = 65536u;
        *lo = 0u;
        return;
    }
    double t = std::floor(std::ldexp(p, 32));
    uint64_t T = (t <= 0.0) ? 0u : (uint64_t)t;
    if (T > 0xFFFFFFFFull) T = 0xFFFFFFFFull;
    *hi = (uint32_t)(T >> 16);
    *lo = (uint32_t)(T & 0xFFFFu);
}

uint32_t prob_threshold(double p)
{
    if (!(p > 0.0)) return 0u;
    double t = std::floor(std::ldexp(p, 32));
    if (t >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)t;
}

int qmax_of(const mq_state* st)
{
    int q = 0;
    for (auto* l : st->lat) q = std::max(q, l->kind == LAT_REGULAR ? 2 * l->dim : l->maxdeg);
    return q;
}

// host natural order <-> device layout position of site i
int64_t dev_pos(const mq_lattice* lat, int64_t i)
{
    if (lat->kind == LAT_REGULAR) return (int64_t)regular_colour(lat->dim, lat->L, i) * (lat->N / 2) + (i >> 1);
    return lat->pos_of[i];
}

__global__ void k_zero_u64(unsigned long long* p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// clear n 64-bit words with a fill kernel (graph-capture safe)
hipError_t zero64(void* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<unsigned long long*>(p), n);
    return hipGetLastError();
}

// measurement staging: per replica isum[8], fsum[8] (bit patterns) and the
// current beta, so one copy brings everything to the host
__global__ void k_pack_obs(const unsigned long long* __restrict__ isum, const double* __restrict__ fsum,
                           const double* __restrict__ betad, int nrep, uint64_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nrep * 17) return;
    const int64_t r = i / 17;
    const int k = (int)(i - r * 17);
    uint64_t v;
    if (k < 8) v = isum[r * 8 + k];
    else if (k < 16) v = (uint64_t)__double_as_longlong(fsum[r * 8 + (k - 8)]);
    else v = (uint64_t)__double_as_longlong(betad[r]);
    out[i] = v;
}

int check_state(mq_state* st)
{
    if (!st) return MQ_ERR_ARG;
    if (st->beta.empty()) return mq_fail(st->ctx, MQ_ERR_STATE, "beta not set (call mq_state_set_beta)");
    if (st->p.model == MQ_EA_ISING && !st->d_bonds)
        return mq_fail(st->ctx, MQ_ERR_STATE, "EA couplings not set (mq_state_random_bonds / mq_state_set_bonds)");
    return MQ_OK;
}

int ensure_bond_storage(mq_state* st)
{
    if (st->d_bonds) return MQ_OK;
    if (st->csrbits)
        return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED, "1-bit Ising on graphs has no couplings (use int8 storage for +-J)");
    st->sample_bond_off.assign(st->nsamples, -1);
    int64_t off = 0;
    for (int r = 0; r < st->nrep; ++r) {
        const int k = st->rep_sample[r];
        if (st->sample_bond_off[k] >= 0) continue;
        const mq_lattice* l = st->lat[r];
        st->sample_bond_off[k] = off;
        const int64_t n = (l->kind == LAT_REGULAR) ? l->N * l->dim : l->nnz;
        off += (n + 255) / 256 * 256;
    }
    MQ_HIP(st->ctx, dalloc(&st->d_bonds, (size_t)std::max<int64_t>(off, 256)));
    MQ_HIP(st->ctx, hipMemsetAsync(st->d_bonds, 0, (size_t)std::max<int64_t>(off, 256), st->ctx->stream));
    std::vector<int64_t> rb(st->nrep);
    for (int r = 0; r < st->nrep; ++r) rb[r] = st->sample_bond_off[st->rep_sample[r]];
    MQ_HIP(st->ctx, hipMemcpyAsync(st->d_rep_bond_off, rb.data(), sizeof(int64_t) * st->nrep, hipMemcpyHostToDevice,
                                   st->ctx->stream));
    MQ_HIP(st->ctx, hipStreamSynchronize(st->ctx->stream));
    return MQ_OK;
}

int ensure_wolff_scratch(mq_state* st, int nclusters)
{
    if (!st->d_queue) {
        st->rep_site_off.resize(st->nrep);
        int64_t off = 0;
        for (int r = 0; r < st->nrep; ++r) {
            st->rep_site_off[r] = off;
            off += (st->lat[r]->N + 63) / 64 * 64;
        }
        st->total_sites = off;
        MQ_HIP(st->ctx, dalloc(&st->d_queue, (size_t)off));
        MQ_HIP(st->ctx, dalloc(&st->d_visit, (size_t)(off / 32)));
        MQ_HIP(st->ctx, hipMemsetAsync(st->d_visit, 0, sizeof(uint32_t) * (size_t)(off / 32), st->ctx->stream));
        MQ_HIP(st->ctx, dalloc(&st->d_rep_site_off, (size_t)st->nrep));
        MQ_HIP(st->ctx, hipMemcpyAsync(st->d_rep_site_off, st->rep_site_off.data(), sizeof(int64_t) * st->nrep,
                                       hipMemcpyHostToDevice, st->ctx->stream));
        MQ_HIP(st->ctx, hipStreamSynchronize(st->ctx->stream));
    }
    const size_t need = (size_t)st->nrep * (size_t)std::max(nclusters, 1);
    if (need > st->csize_cap) {
        dfree(st->d_csize);
        MQ_HIP(st->ctx, dalloc(&st->d_csize, need));
        st->csize_cap = need;
    }
    return MQ_OK;
}

}  // namespace

// ============================================================== scratch
// Partial sums of mq_measure_noalloc: per replica 8 integer sums, then (at word
// MQ_NOALLOC_MAX_REP*8) 8 FP64 sums, 128 B per replica, 4 MiB in all.  Every
// device the library runs on holds its own copy until the process ends.  In host
// code `scratch` names a host-side stand-in, not device memory: host code only
// ever uses the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned long long scratch[(size_t)MQ_NOALLOC_MAX_REP * 16];

namespace {
constexpr int MQ_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_of_dev[MQ_MAX_DEVICES] = {};

// Device address of `scratch` on `device` (which must be current), looked up once
// per device.
hipError_t scratch_address(int device, void** out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= MQ_MAX_DEVICES) return hipErrorInvalidDevice;
    if (!g_scratch_of_dev[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_of_dev[device] = p;
    }
    *out = g_scratch_of_dev[device];
    return hipSuccess;
}
}  // namespace

// ============================================================== context
extern "C" int mq_ctx_create(int device, mq_ctx** out)
{
    if (!out) return MQ_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MQ_ERR_HIP;
    if (device < 0 || device >= n) return MQ_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return MQ_ERR_HIP;
    auto* c = new mq_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        if (c->ev0) (void)hipEventDestroy(c->ev0);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return MQ_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->num_cus = prop.multiProcessorCount;
        if (prop.maxGridSize[1] > 0) c->max_grid_y = prop.maxGridSize[1];
    }
    if (scratch_address(device, &c->scratch) != hipSuccess) {
        (void)hipEventDestroy(c->ev0);
        (void)hipEventDestroy(c->ev1);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return MQ_ERR_HIP;
    }
    *out = c;
    return MQ_OK;
}

extern "C" void mq_ctx_destroy(mq_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char* mq_last_error(const mq_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int mq_sync(mq_ctx* ctx)
{
    if (!ctx) return MQ_ERR_ARG;
    MQ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MQ_OK;
}

extern "C" int mq_timer_begin(mq_ctx* ctx)
{
    if (!ctx) return MQ_ERR_ARG;
    MQ_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    return MQ_OK;
}

extern "C" int mq_timer_end(mq_ctx* ctx, double* ms)
{
    if (!ctx || !ms) return MQ_ERR_ARG;
    MQ_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    MQ_HIP(ctx, hipEventSynchronize(ctx->ev1));
    float f = 0.f;
    MQ_HIP(ctx, hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
    *ms = f;
    return MQ_OK;
}

// ============================================================== lattices
extern "C" int mq_lattice_regular(mq_ctx* ctx, int dim, int L, mq_lattice** out)
{
    if (!ctx || !out) return MQ_ERR_ARG;
    *out = nullptr;
    if (dim != 2 && dim != 3) return mq_fail(ctx, MQ_ERR_UNSUPPORTED, "regular lattices: dim must be 2 or 3");
    if (L < 2 || (L & 1)) return mq_fail(ctx, MQ_ERR_ARG, "regular lattices: L must be even and >= 2 (checkerboard)");
    int64_t N = 1;
    for (int d = 0; d < dim; ++d) N *= L;
    if (N > (int64_t)1 << 31) return mq_fail(ctx, MQ_ERR_ARG, "regular lattice too large (N > 2^31)");
    auto* l = new mq_lattice();
    l->ctx = ctx;
    l->kind = LAT_REGULAR;
    l->dim = dim;
    l->L = L;
    l->N = N;
    l->ncolors = 2;
    l->color_off = {0, N / 2, N};
    l->maxdeg = 2 * dim;
    *out = l;
    return MQ_OK;
}

extern "C" int64_t mq_lattice_nsites(const mq_lattice* lat) { return lat ? lat->N : -1; }

extern "C" int mq_lattice_colors(const mq_lattice* lat, int* ncolors, int64_t* sizes)
{
    if (!lat) return MQ_ERR_ARG;
    if (!ncolors) return mq_fail(lat->ctx, MQ_ERR_ARG, "mq_lattice_colors: ncolors is NULL");
    *ncolors = lat->ncolors;
    if (sizes)
        for (int c = 0; c < lat->ncolors; ++c) sizes[c] = lat->color_off[c + 1] - lat->color_off[c];
    return MQ_OK;
}

extern "C" void mq_lattice_destroy(mq_lattice* lat)
{
    if (!lat) return;
    dfree(lat->d_rp);
    dfree(lat->d_ci);
    dfree(lat->d_edge);
    dfree(lat->d_pos_of);
    dfree(lat->d_vert_of);
    if (lat->d_ell) (void)hipFree(lat->d_ell);
    delete lat;
}

// =============================================================== states
extern "C" int mq_state_create(mq_ctx* ctx, const mq_model_params* params, int nrep, mq_lattice* const* lattices,
                               mq_state** out)
{
    if (!ctx || !params || !lattices || !out || nrep <= 0) return MQ_ERR_ARG;
    *out = nullptr;
    const int model = params->model;
    if (model < MQ_ISING || model > MQ_PHI4) return mq_fail(ctx, MQ_ERR_ARG, "unknown model");
    if (model == MQ_POTTS && (params->q < 2 || params->q > 8))
        return mq_fail(ctx, MQ_ERR_UNSUPPORTED, "Potts: q must be in [2,8]");
    const int kind0 = lattices[0] ? lattices[0]->kind : -1;
    const int dim0 = lattices[0] ? lattices[0]->dim : -1;
    for (int r = 0; r < nrep; ++r) {
        if (!lattices[r]) return mq_fail(ctx, MQ_ERR_ARG, "null lattice");
        if (lattices[r]->kind != kind0 || (kind0 == LAT_REGULAR && lattices[r]->dim != dim0))
            return mq_fail(ctx, MQ_ERR_UNSUPPORTED, "all replicas of a state must use lattices of one kind and dim");
    }
    const bool eabits = params->spin_bits == 1 && model == MQ_EA_ISING;
    const bool csrbits = params->spin_bits == 1 && model == MQ_ISING && kind0 == LAT_CSR;
    const bool msc = params->spin_bits == 1 && !eabits && !csrbits;
    if (params->spin_bits != 0 && params->spin_bits != 1 && params->spin_bits != 8)
        return mq_fail(ctx, MQ_ERR_ARG, "spin_bits must be 0, 1 or 8");
    if (eabits) {
        for (int r = 0; r < nrep; ++r)
            if (lattices[r]->kind != LAT_REGULAR)
                return mq_fail(ctx, MQ_ERR_UNSUPPORTED, "bit-sliced EA batches need regular lattices");
        for (int r = 0; r < nrep;) {
            int e = r + 1;
            while (e < nrep && lattices[e] == lattices[r]) ++e;
            if ((e - r) % 32 != 0)
                return mq_fail(ctx, MQ_ERR_ARG, "bit-sliced EA: each run of replicas on one lattice must be a multiple of 32");
            r = e;
        }
        if (params->J == 0.0) return mq_fail(ctx, MQ_ERR_UNSUPPORTED, "bit-sliced EA needs J != 0");
    }
    if (csrbits) {
        for (int r = 0; r < nrep;) {
            int e = r + 1;
            while (e < nrep && lattices[e] == lattices[r]) ++e;
            if ((e - r) % 32 != 0)
                return mq_fail(ctx, MQ_ERR_ARG, "1-bit Ising on graphs: each run of replicas on one graph must be a multiple of 32");
            if (!csrbits_supported(lattices[r]))
                return mq_fail(ctx, MQ_ERR_UNSUPPORTED, "1-bit Ising on graphs needs vertex degrees <= 127");
            if (lattices[r]->N * (int64_t)((e - r) / 8) >= (1ll << 31))
                return mq_fail(ctx, MQ_ERR_UNSUPPORTED, "1-bit Ising on graphs: vertices x octets of replicas must stay below 2^31");
            r = e;
        }
    }
    if (msc) {
        if (model != MQ_ISING) return mq_fail(ctx, MQ_ERR_UNSUPPORTED, "multi-spin coding is for the Ising model");
        for (int r = 0; r < nrep; ++r)
            if (!msc_supported(lattices[r]))
                return mq_fail(ctx, MQ_ERR_UNSUPPORTED, "multi-spin coding needs 3D lattices with L a power of two >= 64");
        if (params->J == 0.0) return mq_fail(ctx, MQ_ERR_UNSUPPORTED, "multi-spin coding needs J != 0");
    }
    auto* st = new mq_state();
    st->ctx = ctx;
    st->p = *params;
    st->nrep = nrep;
    st->msc = msc;
    st->eabits = eabits;
    st->csrbits = csrbits;
    st->lat.assign(lattices, lattices + nrep);
    st->site_bytes = is_vec(model) ? 16 : 1;
    auto rep_bytes = [&](const mq_lattice* l) {
        const int64_t b = msc ? l->N / 8 : l->N * st->site_bytes;
        return (b + 255) / 256 * 256;
    };
    // groups of consecutive replicas on one lattice; regular lattices store each
    // replica contiguously, graphs interleave the replicas of a group site by site
    st->rep_off.resize(nrep);
    st->rep_pos_stride.resize(nrep);
    int64_t off = 0;
    for (int r = 0; r < nrep;) {
        int e = r + 1;
        while (e < nrep && st->lat[e] == st->lat[r]) ++e;
        Group g;
        g.lat = st->lat[r];
        g.rep0 = r;
        g.nrep = e - r;
        g.off = off;
        if (eabits) {
            // one uint32 word per site for each group of 32 replicas
            g.stride = (g.lat->N * 4 + 255) / 256 * 256;
            g.pos_stride = 4;
            off += g.stride * (g.nrep / 32);
        } else if (csrbits) {
            // one uint32 word per vertex and word of 32 replicas, words of a vertex adjacent
            g.stride = 4;
            g.pos_stride = 4 * (g.nrep / 32);
            off += (g.lat->N * g.pos_stride + 255) / 256 * 256;
        } else if (g.lat->kind == LAT_CSR) {
            g.stride = st->site_bytes;
            g.pos_stride = (int64_t)st->site_bytes * g.nrep;
            off += (g.lat->N * g.pos_stride + 255) / 256 * 256;
        } else {
            g.stride = rep_bytes(g.lat);
            g.pos_stride = st->site_bytes;
            off += g.stride * g.nrep;
        }
        for (int k = r; k < e; ++k) {
            st->rep_off[k] = g.off + ((eabits || csrbits) ? (int64_t)((k - r) / 32) : (int64_t)(k - r)) * g.stride;
            st->rep_pos_stride[k] = g.pos_stride;
        }
        if (((eabits || csrbits) ? g.nrep / 32 : g.nrep) > ctx->max_grid_y) {  // replicas run along the grid's y dimension
            const std::string msg = "at most " + std::to_string(ctx->max_grid_y) +
                                    ((eabits || csrbits) ? " words of 32 replicas" : " replicas") + " per lattice in one state";
            delete st;
            return mq_fail(ctx, MQ_ERR_UNSUPPORTED, msg);
        }
        st->groups.push_back(g);
        r = e;
    }
    st->spin_total = off;
    st->tbl = 2 * qmax_of(st) + 1;
    if (const char* kc = std::getenv("MARQOV_AMD_KERNELS")) st->kernel_choice = std::atoi(kc);
    st->rep_sample.resize(nrep);
    std::iota(st->rep_sample.begin(), st->rep_sample.end(), 0);
    st->nsamples = nrep;
    if (eabits) {  // one disorder sample per word of 32 replicas by default
        for (int r = 0; r < nrep; ++r) st->rep_sample[r] = r / 32;
        st->nsamples = nrep / 32;
    }
    auto fail = [&](hipError_t e) {
        ctx->err = std::string("allocation failed: ") + hipGetErrorString(e);
        mq_state_destroy(st);
        return MQ_ERR_HIP;
    };
    hipError_t e;
    if ((e = dalloc(&st->d_spins, (size_t)off)) != hipSuccess) return fail(e);
    if ((e = dalloc(&st->d_thr, (size_t)nrep * st->tbl * 2)) != hipSuccess) return fail(e);
    if ((e = dalloc(&st->d_wthr, (size_t)nrep)) != hipSuccess) return fail(e);
    if ((e = dalloc(&st->d_betaf, (size_t)nrep)) != hipSuccess) return fail(e);
    if ((e = dalloc(&st->d_betad, (size_t)nrep)) != hipSuccess) return fail(e);
    if ((e = dalloc(&st->d_isum, (size_t)nrep * 8)) != hipSuccess) return fail(e);
    if ((e = dalloc(&st->d_fsum, (size_t)nrep * 8)) != hipSuccess) return fail(e);
    if ((e = dalloc(&st->d_rep_bond_off, (size_t)nrep)) != hipSuccess) return fail(e);
    if ((e = hipMemsetAsync(st->d_spins, 0, (size_t)off, ctx->stream)) != hipSuccess) return fail(e);
    if ((e = hipMemsetAsync(st->d_rep_bond_off, 0, sizeof(int64_t) * nrep, ctx->stream)) != hipSuccess) return fail(e);
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return fail(e);
    if ((e = prepare_lds(st)) != hipSuccess) return fail(e);
    if (eabits && (e = hipMalloc(&st->d_wtab, wtab_bytes() * (size_t)(nrep / 32))) != hipSuccess) return fail(e);
    if (eabits && (e = prepare_eabits(st)) != hipSuccess) return fail(e);
    *out = st;
    return MQ_OK;
}

extern "C" void mq_state_destroy(mq_state* st)
{
    if (!st) return;
    (void)hipStreamSynchronize(st->ctx->stream);
    dfree(st->d_spins);
    dfree(st->d_thr);
    dfree(st->d_wthr);
    dfree(st->d_betaf);
    dfree(st->d_betad);
    dfree(st->d_bonds);
    dfree(st->d_rep_bond_off);
    dfree(st->d_queue);
    dfree(st->d_visit);
    dfree(st->d_csize);
    dfree(st->d_rep_site_off);
    dfree(st->d_isum);
    dfree(st->d_fsum);
    dfree(st->d_ft);
    dfree(st->d_temp_of_rep);
    dfree(st->d_rep_of_temp);
    dfree(st->d_pt_acc);
    dfree(st->d_pt_pairs);
    dfree(st->d_pack);
    if (st->h_pack) (void)hipHostFree(st->h_pack);
    dfree(st->d_ladder);
    dfree(st->d_dyn);
    if (st->d_msc_cls) (void)hipFree(st->d_msc_cls);
    if (st->d_wtab) (void)hipFree(st->d_wtab);
    if (st->d_bits_groups) (void)hipFree(st->d_bits_groups);
    if (st->d_lds_groups) (void)hipFree(st->d_lds_groups);
    delete st;
}

// Acceptance tables (host, FP64): Ising/EA entry e = q + qmax with
// q = s_i sum_j b_ij s_j and dE = -2 J q; Potts entry e = dn + qmax with
// dE = k dn.  The Wolff add probability is -expm1(2 beta J) for Ising
// (Ising.h:139) and 1 - exp(beta k) for Potts (Potts.h:237-242, pref 2).
static int upload_tables(mq_state* st)
{
    const int nrep = st->nrep, tbl = st->tbl, qmax = (tbl - 1) / 2;
    std::vector<uint32_t> thr((size_t)nrep * tbl * 2);
    std::vector<uint32_t> wthr(nrep);
    std::vector<float> bf(nrep);
    for (int r = 0; r < nrep; ++r) {
        const double beta = st->beta[r];
        bf[r] = (float)beta;
        for (int e = 0; e < tbl; ++e) {
            double dE;
            if (st->p.model == MQ_POTTS) dE = st->p.J * (double)(e - qmax);
            else dE = -2.0 * st->p.J * (double)(e - qmax);
            const double p = (dE <= 0.0) ? 1.0 : std::exp(-beta * dE);
            split_threshold(p, &thr[((size_t)r * tbl + e) * 2], &thr[((size_t)r * tbl + e) * 2 + 1]);
        }
        double pw = 0.0;
        if (st->p.model == MQ_ISING || st->p.model == MQ_EA_ISING) pw = -std::expm1(2.0 * beta * st->p.J);
        else if (st->p.model == MQ_POTTS) pw = -std::expm1(beta * st->p.J);
        wthr[r] = prob_threshold(pw);
    }
    hipStream_t s = st->ctx->stream;
    MQ_HIP(st->ctx, hipMemcpyAsync(st->d_thr, thr.data(), thr.size() * 4, hipMemcpyHostToDevice, s));
    if (st->msc) MQ_HIP(st->ctx, msc_upload_classes(st, thr));
    if (st->eabits) MQ_HIP(st->ctx, launch_wtab_eabits(st));
    MQ_HIP(st->ctx, hipMemcpyAsync(st->d_wthr, wthr.data(), wthr.size() * 4, hipMemcpyHostToDevice, s));
    MQ_HIP(st->ctx, hipMemcpyAsync(st->d_betaf, bf.data(), bf.size() * 4, hipMemcpyHostToDevice, s));
    MQ_HIP(st->ctx, hipMemcpyAsync(st->d_betad, st->beta.data(), st->beta.size() * 8, hipMemcpyHostToDevice, s));
    MQ_HIP(st->ctx, hipStreamSynchronize(s));
    return MQ_OK;
}

extern "C" int mq_state_set_beta(mq_state* st, const double* beta)
{
    if (!st) return MQ_ERR_ARG;
    if (!beta) return mq_fail(st->ctx, MQ_ERR_ARG, "beta is NULL");
    for (int r = 0; r < st->nrep; ++r)
        if (!(beta[r] >= 0.0) || !std::isfinite(beta[r])) return mq_fail(st->ctx, MQ_ERR_ARG, "beta must be finite and >= 0");
    st->beta.assign(beta, beta + st->nrep);
    st->pt_ntemp = 0;  // temperatures are re-assigned: restart the PT bookkeeping
    return upload_tables(st);
}

extern "C" int mq_state_get_beta(const mq_state* st, double* beta)
{
    if (!st) return MQ_ERR_ARG;
    if (!beta) return mq_fail(st->ctx, MQ_ERR_ARG, "beta is NULL");
    if (st->beta.empty()) return mq_fail(st->ctx, MQ_ERR_STATE, "beta not set (call mq_state_set_beta)");
    // current temperature of every replica (parallel tempering moves them)
    MQ_HIP(st->ctx, hipStreamSynchronize(st->ctx->stream));
    MQ_HIP(st->ctx, hipMemcpy(beta, st->d_betad, sizeof(double) * st->nrep, hipMemcpyDeviceToHost));
    return MQ_OK;
}

static int set_samples(mq_state* st, const int32_t* rep_sample)
{
    int ns = 0;
    for (int r = 0; r < st->nrep; ++r) {
        if (rep_sample[r] < 0) return mq_fail(st->ctx, MQ_ERR_ARG, "negative sample id");
        ns = std::max(ns, rep_sample[r] + 1);
    }
    std::vector<mq_lattice*> lat_of(ns, nullptr);
    for (int r = 0; r < st->nrep; ++r) {
        auto*& l = lat_of[rep_sample[r]];
        if (l && l != st->lat[r]) return mq_fail(st->ctx, MQ_ERR_ARG, "replicas of one sample must share a lattice");
        l = st->lat[r];
    }
    if (st->eabits)
        for (int r = 0; r < st->nrep; ++r)
            if (rep_sample[r] != rep_sample[r - r % 32])
                return mq_fail(st->ctx, MQ_ERR_ARG, "bit-sliced EA: the 32 replicas of a word must share a sample");
    dfree(st->d_bonds);
    st->rep_sample.assign(rep_sample, rep_sample + st->nrep);
    st->nsamples = ns;
    return MQ_OK;
}

extern "C" int mq_state_set_samples(mq_state* st, const int32_t* rep_sample)
{
    if (!st) return MQ_ERR_ARG;
    if (!rep_sample) return mq_fail(st->ctx, MQ_ERR_ARG, "rep_sample is NULL");
    int rc = set_samples(st, rep_sample);
    if (rc) return rc;
    return ensure_bond_storage(st);
}

extern "C" int mq_state_random_bonds(mq_state* st, const int32_t* rep_sample, double p_plus, uint64_t seed)
{
    if (!st) return MQ_ERR_ARG;
    if (!(p_plus >= 0.0 && p_plus <= 1.0)) return mq_fail(st->ctx, MQ_ERR_ARG, "p_plus must be in [0,1]");
    if (rep_sample) {
        int rc0 = set_samples(st, rep_sample);
        if (rc0) return rc0;
    }
    int rc = ensure_bond_storage(st);
    if (rc) return rc;
    std::vector<int> first(st->nsamples, -1);
    for (int r = 0; r < st->nrep; ++r)
        if (first[st->rep_sample[r]] < 0) first[st->rep_sample[r]] = r;
    for (int k = 0; k < st->nsamples; ++k) {
        if (first[k] < 0) continue;
        const mq_lattice* l = st->lat[first[k]];
        Group g{const_cast<mq_lattice*>(l), first[k], 1, 0, 0};
        if (l->kind != LAT_REGULAR) return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED, "random bonds on graphs: use mq_state_set_bonds");
        MQ_HIP(st->ctx, launch_bonds_regular(st, g, k, p_plus, key_of(seed)));
    }
    MQ_HIP(st->ctx, hipStreamSynchronize(st->ctx->stream));
    return MQ_OK;
}

extern "C" int mq_state_set_bonds(mq_state* st, int sample, const int8_t* bonds)
{
    if (!st) return MQ_ERR_ARG;
    if (!bonds) return mq_fail(st->ctx, MQ_ERR_ARG, "bonds is NULL");
    if (sample < 0 || sample >= st->nsamples) return mq_fail(st->ctx, MQ_ERR_ARG, "sample out of range");
    int rc = ensure_bond_storage(st);
    if (rc) return rc;
    int r0 = -1;
    for (int r = 0; r < st->nrep; ++r)
        if (st->rep_sample[r] == sample) { r0 = r; break; }
    if (r0 < 0) return mq_fail(st->ctx, MQ_ERR_ARG, "no replica uses this sample");
    const mq_lattice* l = st->lat[r0];
    std::vector<int8_t> dev;
    if (l->kind == LAT_REGULAR) {
        const int64_t N = l->N, half = N / 2;
        const int dim = l->dim;
        dev.resize((size_t)(N * dim));
        for (int64_t i = 0; i < N; ++i) {
            const int c = regular_colour(dim, l->L, i);
            for (int d = 0; d < dim; ++d) {
                const int8_t b = bonds[i * dim + d];
                if (b != 1 && b != -1) return mq_fail(st->ctx, MQ_ERR_ARG, "bonds must be +1 or -1");
                dev[(size_t)(c * dim * half + d * half + (i >> 1))] = b;
            }
        }
    } else {
        const int64_t nnz = l->nnz;
        dev.resize((size_t)std::max<int64_t>(nnz, 1));
        for (int64_t k = 0; k < nnz; ++k) {
            const int8_t b = bonds[k];
            if (b != 1 && b != -1) return mq_fail(st->ctx, MQ_ERR_ARG, "bonds must be +1 or -1");
            if (bonds[l->h_mirror[k]] != b) return mq_fail(st->ctx, MQ_ERR_ARG, "graph couplings must be symmetric");
            dev[(size_t)l->h_slot_of_orig[k]] = b;
        }
        dev.resize((size_t)nnz);
    }
    MQ_HIP(st->ctx, hipMemcpy(st->d_bonds + st->sample_bond_off[sample], dev.data(), dev.size(), hipMemcpyHostToDevice));
    return MQ_OK;
}

extern "C" int mq_state_init(mq_state* st, uint64_t seed)
{
    if (!st) return MQ_ERR_ARG;
    for (const auto& g : st->groups) {
        if (st->eabits) MQ_HIP(st->ctx, launch_init_eabits(st, g, key_of(seed)));
        else if (st->msc) MQ_HIP(st->ctx, launch_init_msc(st, g, key_of(seed)));
        else if (st->csrbits) MQ_HIP(st->ctx, launch_init_csrbits(st, g, key_of(seed)));
        else if (g.lat->kind == LAT_REGULAR) MQ_HIP(st->ctx, launch_init_regular(st, g, key_of(seed)));
        else MQ_HIP(st->ctx, launch_init_csr(st, g, key_of(seed)));
    }
    MQ_HIP(st->ctx, hipStreamSynchronize(st->ctx->stream));
    return MQ_OK;
}


static const Group* group_of(const mq_state* st, int rep)
{
    for (const auto& g : st->groups)
        if (rep >= g.rep0 && rep < g.rep0 + g.nrep) return &g;
    return nullptr;
}

// Copy one replica's sites between a dense host array (elem bytes per site, in
// device position order) and the device: one contiguous copy when the replica
// is stored contiguously, else through a host image of the whole group block.
static int copy_replica(mq_state* st, int rep, void* host, size_t elem, bool to_device)
{
    const mq_lattice* l = st->lat[rep];
    const int64_t N = l->N;
    const int64_t ps = st->rep_pos_stride[rep];
    char* dev = st->d_spins + st->rep_off[rep];
    if (ps == (int64_t)elem) {
        if (to_device) MQ_HIP(st->ctx, hipMemcpy(dev, host, (size_t)N * elem, hipMemcpyHostToDevice));
        else MQ_HIP(st->ctx, hipMemcpy(host, dev, (size_t)N * elem, hipMemcpyDeviceToHost));
        return MQ_OK;
    }
    const Group* g = group_of(st, rep);
    const size_t bytes = (size_t)(N * ps);
    std::vector<char> img(bytes);
    char* gbase = st->d_spins + g->off;
    const int64_t within = (int64_t)(rep - g->rep0) * g->stride;
    MQ_HIP(st->ctx, hipMemcpy(img.data(), gbase, bytes, hipMemcpyDeviceToHost));
    char* h = static_cast<char*>(host);
    for (int64_t p = 0; p < N; ++p) {
        char* slot = img.data() + p * ps + within;
        if (to_device) std::memcpy(slot, h + p * elem, elem);
        else std::memcpy(h + p * elem, slot, elem);
    }
    if (to_device) MQ_HIP(st->ctx, hipMemcpy(gbase, img.data(), bytes, hipMemcpyHostToDevice));
    return MQ_OK;
}

extern "C" int mq_state_upload(mq_state* st, int rep, const void* host, int64_t bytes)
{
    if (!st) return MQ_ERR_ARG;
    if (!host) return mq_fail(st->ctx, MQ_ERR_ARG, "host buffer is NULL");
    if (rep < 0 || rep >= st->nrep) return mq_fail(st->ctx, MQ_ERR_ARG, "replica index out of range");
    const mq_lattice* l = st->lat[rep];
    const int64_t N = l->N;
    if (is_vec(st->p.model)) {
        if (bytes != N * 12) return mq_fail(st->ctx, MQ_ERR_ARG, "upload: expected 3 floats per site");
        const float* h = static_cast<const float*>(host);
        std::vector<float> dev((size_t)N * 4, 0.f);
        for (int64_t i = 0; i < N; ++i) {
            const int64_t p = dev_pos(l, i);
            dev[p * 4 + 0] = h[3 * i];
            dev[p * 4 + 1] = h[3 * i + 1];
            dev[p * 4 + 2] = h[3 * i + 2];
        }
        int rc = copy_replica(st, rep, dev.data(), 16, true);
        if (rc) return rc;
    } else if (st->eabits) {
        if (bytes != N) return mq_fail(st->ctx, MQ_ERR_ARG, "upload: expected one int8 per site");
        const int8_t* h = static_cast<const int8_t*>(host);
        std::vector<uint32_t> dev((size_t)N);
        MQ_HIP(st->ctx, hipStreamSynchronize(st->ctx->stream));
        MQ_HIP(st->ctx, hipMemcpy(dev.data(), st->d_spins + st->rep_off[rep], (size_t)N * 4, hipMemcpyDeviceToHost));
        const uint32_t bit = 1u << (rep % 32);
        for (int64_t i = 0; i < N; ++i) {
            if (h[i] != 1 && h[i] != -1) return mq_fail(st->ctx, MQ_ERR_ARG, "upload: spin value out of range");
            uint32_t& w = dev[(size_t)dev_pos(l, i)];
            w = (h[i] < 0) ? (w | bit) : (w & ~bit);
        }
        MQ_HIP(st->ctx, hipMemcpy(st->d_spins + st->rep_off[rep], dev.data(), (size_t)N * 4, hipMemcpyHostToDevice));
    } else if (st->csrbits) {
        if (bytes != N) return mq_fail(st->ctx, MQ_ERR_ARG, "upload: expected one int8 per site");
        const int8_t* h = static_cast<const int8_t*>(host);
        for (int64_t i = 0; i < N; ++i)
            if (h[i] != 1 && h[i] != -1) return mq_fail(st->ctx, MQ_ERR_ARG, "upload: spin value out of range");
        // read-modify-write of the group's words: bit (rep - rep0) % 32 of word group (rep - rep0) / 32
        const Group* g = group_of(st, rep);
        const int W = g->nrep / 32, w = (rep - g->rep0) / 32;
        const uint32_t bit = 1u << ((rep - g->rep0) % 32);
        std::vector<uint32_t> dev((size_t)(N * W));
        MQ_HIP(st->ctx, hipStreamSynchronize(st->ctx->stream));
        MQ_HIP(st->ctx, hipMemcpy(dev.data(), st->d_spins + g->off, dev.size() * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < N; ++i) {
            uint32_t& x = dev[(size_t)(dev_pos(l, i) * W + w)];
            x = (h[i] < 0) ? (x | bit) : (x & ~bit);
        }
        MQ_HIP(st->ctx, hipMemcpy(st->d_spins + g->off, dev.data(), dev.size() * 4, hipMemcpyHostToDevice));
    } else if (st->msc) {
        if (bytes != N) return mq_fail(st->ctx, MQ_ERR_ARG, "upload: expected one int8 per site");
        const int8_t* h = static_cast<const int8_t*>(host);
        std::vector<uint32_t> dev((size_t)(N / 32), 0u);
        for (int64_t i = 0; i < N; ++i) {
            if (h[i] != 1 && h[i] != -1) return mq_fail(st->ctx, MQ_ERR_ARG, "upload: spin value out of range");
            const int64_t p = dev_pos(l, i);
            if (h[i] < 0) dev[(size_t)(p >> 5)] |= 1u << (p & 31);
        }
        MQ_HIP(st->ctx, hipMemcpy(st->d_spins + st->rep_off[rep], dev.data(), dev.size() * 4, hipMemcpyHostToDevice));
    } else {
        if (bytes != N) return mq_fail(st->ctx, MQ_ERR_ARG, "upload: expected one int8 per site");
        const int8_t* h = static_cast<const int8_t*>(host);
        std::vector<int8_t> dev((size_t)N);
        for (int64_t i = 0; i < N; ++i) {
            const int8_t v = h[i];
            if (st->p.model == MQ_POTTS ? (v < 0 || v >= st->p.q) : (v != 1 && v != -1))
                return mq_fail(st->ctx, MQ_ERR_ARG, "upload: spin value out of range");
            dev[dev_pos(l, i)] = v;
        }
        int rc = copy_replica(st, rep, dev.data(), 1, true);
        if (rc) return rc;
    }
    return MQ_OK;
}

extern "C" int mq_state_download(mq_state* st, int rep, void* host, int64_t bytes)
{
    if (!st) return MQ_ERR_ARG;
    if (!host) return mq_fail(st->ctx, MQ_ERR_ARG, "host buffer is NULL");
    if (rep < 0 || rep >= st->nrep) return mq_fail(st->ctx, MQ_ERR_ARG, "replica index out of range");
    const mq_lattice* l = st->lat[rep];
    const int64_t N = l->N;
    MQ_HIP(st->ctx, hipStreamSynchronize(st->ctx->stream));
    if (is_vec(st->p.model)) {
        if (bytes != N * 12) return mq_fail(st->ctx, MQ_ERR_ARG, "download: expected 3 floats per site");
        std::vector<float> dev((size_t)N * 4);
        int rc = copy_replica(st, rep, dev.data(), 16, false);
        if (rc) return rc;
        float* h = static_cast<float*>(host);
        for (int64_t i = 0; i < N; ++i) {
            const int64_t p = dev_pos(l, i);
            h[3 * i] = dev[p * 4];
            h[3 * i + 1] = dev[p * 4 + 1];
            h[3 * i + 2] = dev[p * 4 + 2];
        }
    } else if (st->eabits) {
        if (bytes != N) return mq_fail(st->ctx, MQ_ERR_ARG, "download: expected one int8 per site");
        std::vector<uint32_t> dev((size_t)N);
        MQ_HIP(st->ctx, hipMemcpy(dev.data(), st->d_spins + st->rep_off[rep], (size_t)N * 4, hipMemcpyDeviceToHost));
        int8_t* h = static_cast<int8_t*>(host);
        const int b = rep % 32;
        for (int64_t i = 0; i < N; ++i) h[i] = ((dev[(size_t)dev_pos(l, i)] >> b) & 1u) ? (int8_t)-1 : (int8_t)1;
    } else if (st->csrbits) {
        if (bytes != N) return mq_fail(st->ctx, MQ_ERR_ARG, "download: expected one int8 per site");
        const Group* g = group_of(st, rep);
        const int W = g->nrep / 32, w = (rep - g->rep0) / 32, b = (rep - g->rep0) % 32;
        std::vector<uint32_t> dev((size_t)(N * W));
        MQ_HIP(st->ctx, hipMemcpy(dev.data(), st->d_spins + g->off, dev.size() * 4, hipMemcpyDeviceToHost));
        int8_t* h = static_cast<int8_t*>(host);
        for (int64_t i = 0; i < N; ++i)
            h[i] = ((dev[(size_t)(dev_pos(l, i) * W + w)] >> b) & 1u) ? (int8_t)-1 : (int8_t)1;
    } else if (st->msc) {
        if (bytes != N) return mq_fail(st->ctx, MQ_ERR_ARG, "download: expected one int8 per site");
        std::vector<uint32_t> dev((size_t)(N / 32));
        MQ_HIP(st->ctx, hipMemcpy(dev.data(), st->d_spins + st->rep_off[rep], dev.size() * 4, hipMemcpyDeviceToHost));
        int8_t* h = static_cast<int8_t*>(host);
        for (int64_t i = 0; i < N; ++i) {
            const int64_t p = dev_pos(l, i);
            h[i] = ((dev[(size_t)(p >> 5)] >> (p & 31)) & 1u) ? (int8_t)-1 : (int8_t)1;
        }
    } else {
        if (bytes != N) return mq_fail(st->ctx, MQ_ERR_ARG, "download: expected one int8 per site");
        std::vector<int8_t> dev((size_t)N);
        int rc = copy_replica(st, rep, dev.data(), 1, false);
        if (rc) return rc;
        int8_t* h = static_cast<int8_t*>(host);
        for (int64_t i = 0; i < N; ++i) h[i] = dev[dev_pos(l, i)];
    }
    return MQ_OK;
}

// ============================================================== updates
int mq_metropolis_async(mq_state* st, int nsweeps, uint64_t seed, uint64_t sweep0)
{
    const Key2 key = key_of(seed);
    if (nsweeps > 0 && (sweep0 >= sweep_limit(st) || (uint64_t)nsweeps > sweep_limit(st) - sweep0))
        return mq_fail(st->ctx, MQ_ERR_ARG, sweep_limit_msg(st));
    if (nsweeps > 0 && st->kernel_choice == 0 && !st->msc && !st->eabits && !st->csrbits) {
        hipError_t e = launch_sweeps_lds(st, nsweeps, (uint32_t)sweep0, key);
        if (e == hipSuccess) return MQ_OK;
        if (e != hipErrorNotSupported) return mq_fail(st->ctx, MQ_ERR_HIP, std::string("lds sweep: ") + hipGetErrorString(e));
    }
    for (int t = 0; t < nsweeps; ++t) {
        const uint64_t sw = sweep0 + (uint64_t)t;
        if (st->eabits) {  // one fused launch per colour over every lattice size
            for (int c = 0; c < 2; ++c) {
                hipError_t e = launch_sweep_eabits_all(st, c, (uint32_t)(2 * sw + c), key);
                if (e != hipSuccess) return mq_fail(st->ctx, MQ_ERR_HIP, std::string("sweep launch: ") + hipGetErrorString(e));
            }
            continue;
        }
        for (const auto& g : st->groups) {
            const int nc = g.lat->ncolors;
            for (int c = 0; c < nc; ++c) {
                const uint32_t phase = (uint32_t)(sw * (uint64_t)g.lat->color_stride + c);
                hipError_t e = st->msc       ? launch_sweep_msc(st, g, c, phase, key)
                               : st->csrbits ? launch_sweep_csrbits(st, g, c, phase, key)
                               : (g.lat->kind == LAT_REGULAR) ? launch_sweep_regular(st, g, c, phase, key)
                                                              : launch_sweep_csr(st, g, c, phase, key);
                if (e != hipSuccess) return mq_fail(st->ctx, MQ_ERR_HIP, std::string("sweep launch: ") + hipGetErrorString(e));
            }
        }
    }
    return MQ_OK;
}

extern "C" int mq_metropolis(mq_state* st, int nsweeps, uint64_t seed, uint64_t sweep0)
{
    int rc = check_state(st);
    if (rc) return rc;
    if (nsweeps < 0) return mq_fail(st->ctx, MQ_ERR_ARG, "nsweeps < 0");
    // colour-ordered sweeps per group; groups are independent replicas
    return mq_metropolis_async(st, nsweeps, seed, sweep0);
}

extern "C" int mq_wolff(mq_state* st, int nclusters, uint64_t seed, uint64_t cluster0, int64_t* sizes)
{
    int rc = check_state(st);
    if (rc) return rc;
    if (nclusters < 0) return mq_fail(st->ctx, MQ_ERR_ARG, "nclusters < 0");
    if (nclusters == 0) return MQ_OK;
    if (st->p.model == MQ_EA_ISING)
        return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED, "Wolff is not defined for the +-J spin glass (frustrated bonds)");
    if (st->msc || st->csrbits) return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED, "Wolff runs on int8 storage (spin_bits = 8)");
    if (cluster0 + (uint64_t)nclusters > (1ull << 32)) return mq_fail(st->ctx, MQ_ERR_ARG, "cluster counter exceeds 2^32");
    for (const auto* l : st->lat)
        if (l->kind == LAT_REGULAR && l->dim == 3 && l->L > 1024)
            return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED, "Wolff on 3D lattices needs L <= 1024 (10-bit site coordinates)");
    rc = ensure_wolff_scratch(st, nclusters);
    if (rc) return rc;
    for (const auto& g : st->groups) {
        hipError_t e = (g.lat->kind == LAT_REGULAR) ? launch_wolff_regular(st, g, nclusters, (uint32_t)cluster0, key_of(seed))
                                                    : launch_wolff_csr(st, g, nclusters, (uint32_t)cluster0, key_of(seed));
        if (e != hipSuccess) return mq_fail(st->ctx, MQ_ERR_HIP, std::string("wolff launch: ") + hipGetErrorString(e));
    }
    if (sizes) {
        MQ_HIP(st->ctx, hipMemcpyAsync(sizes, st->d_csize, sizeof(int64_t) * (size_t)st->nrep * nclusters,
                                       hipMemcpyDeviceToHost, st->ctx->stream));
        MQ_HIP(st->ctx, hipStreamSynchronize(st->ctx->stream));
    }
    return MQ_OK;
}

extern "C" int mq_emcs(mq_state* st, int ncluster, int nmetro, uint64_t seed, uint64_t step, double* mean_cluster)
{
    int rc = check_state(st);
    if (rc) return rc;
    if (ncluster < 0 || nmetro < 0) return mq_fail(st->ctx, MQ_ERR_ARG, "ncluster and nmetro must be >= 0");
    std::vector<int64_t> sizes;
    if (ncluster > 0) {
        sizes.resize((size_t)st->nrep * ncluster);
        rc = mq_wolff(st, ncluster, seed, step * (uint64_t)ncluster, mean_cluster ? sizes.data() : nullptr);
        if (rc) return rc;
    }
    rc = mq_metropolis(st, nmetro, seed, step * (uint64_t)nmetro);
    if (rc) return rc;
    if (mean_cluster) {
        for (int r = 0; r < st->nrep; ++r) {
            double a = 0;
            for (int c = 0; c < ncluster; ++c) a += (double)sizes[(size_t)r * ncluster + c];
            mean_cluster[r] = a / ncluster;  // NaN for ncluster = 0, as emcs.h:108
        }
    }
    return MQ_OK;
}

extern "C" int mq_emcs_run(mq_state* st, int nsteps, int ncluster, int nmetro, uint64_t seed, uint64_t step0,
                           double* mean_cluster)
{
    int rc = check_state(st);
    if (rc) return rc;
    if (nsteps < 0 || ncluster < 0 || nmetro < 0)
        return mq_fail(st->ctx, MQ_ERR_ARG, "nsteps, ncluster and nmetro must be >= 0");
    if (nsteps == 0) return MQ_OK;
    if (st->p.model == MQ_EA_ISING)
        return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED, "Wolff is not defined for the +-J spin glass (frustrated bonds)");
    if (st->msc || st->eabits || st->csrbits)
        return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED, "mq_emcs_run: int8 storage (spin_bits = 8)");
    if (st->d_bonds) return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED, "mq_emcs_run: no couplings");
    for (const auto* l : st->lat) {
        if (l->kind != LAT_REGULAR) return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED, "mq_emcs_run: regular lattices");
        if (l->dim == 3 && l->L > 1024)
            return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED, "Wolff on 3D lattices needs L <= 1024 (10-bit site coordinates)");
    }
    const uint64_t last = step0 + (uint64_t)nsteps;  // one past the last step
    if (ncluster > 0 && (step0 >= (1ull << 32) || last * (uint64_t)ncluster > (1ull << 32)))
        return mq_fail(st->ctx, MQ_ERR_ARG, "cluster counter exceeds 2^32");
    if (nmetro > 0 && (step0 >= sweep_limit(st) || last * (uint64_t)nmetro > sweep_limit(st)))
        return mq_fail(st->ctx, MQ_ERR_ARG, sweep_limit_msg(st));
    const int64_t ncl_all = (int64_t)ncluster * nsteps;
    if (ncl_all > (1 << 30)) return mq_fail(st->ctx, MQ_ERR_ARG, "nsteps * ncluster too large");
    rc = ensure_wolff_scratch(st, (int)std::max<int64_t>(ncl_all, 1));
    if (rc) return rc;
    for (const auto& g : st->groups) {
        hipError_t e = launch_emcs_regular(st, g, nsteps, ncluster, nmetro, step0, key_of(seed));
        if (e != hipSuccess) return mq_fail(st->ctx, MQ_ERR_HIP, std::string("emcs launch: ") + hipGetErrorString(e));
    }
    if (mean_cluster) {
        std::vector<int64_t> sizes((size_t)std::max<int64_t>(ncl_all, 1) * st->nrep);
        if (ncluster > 0)
            MQ_HIP(st->ctx, hipMemcpyAsync(sizes.data(), st->d_csize, sizeof(int64_t) * (size_t)(ncl_all * st->nrep),
                                           hipMemcpyDeviceToHost, st->ctx->stream));
        MQ_HIP(st->ctx, hipStreamSynchronize(st->ctx->stream));
        for (int k = 0; k < nsteps; ++k)
            for (int r = 0; r < st->nrep; ++r) {
                double a = 0;
                for (int c = 0; c < ncluster; ++c)
                    a += (double)sizes[((size_t)k * st->nrep + r) * ncluster + c];
                mean_cluster[(size_t)k * st->nrep + r] = a / ncluster;  // NaN for ncluster = 0, as emcs.h:108
            }
    }
    return MQ_OK;
}

hipError_t launch_measure_all(mq_state* st, long long* isum, double* fsum)
{
    if (st->eabits) return launch_measure_eabits_all(st, isum);
    for (const auto& g : st->groups) {
        hipError_t e = st->msc       ? launch_measure_msc(st, g, isum)
                       : st->csrbits ? launch_measure_csrbits(st, g, isum)
                       : (g.lat->kind == LAT_REGULAR) ? launch_measure_regular(st, g, isum, fsum)
                                                      : launch_measure_csr(st, g, isum, fsum);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Runs the measurement kernels and brings isum, fsum and the current betas to
// the pinned host buffer h_pack ([nrep][17] words) with one copy.
static int measure_sums(mq_state* st)
{
    hipStream_t s = st->ctx->stream;
    if (!st->d_pack) {
        const size_t n = (size_t)st->nrep * 17;
        MQ_HIP(st->ctx, dalloc(&st->d_pack, n));
        MQ_HIP(st->ctx, hipHostMalloc(reinterpret_cast<void**>(&st->h_pack), n * sizeof(uint64_t), hipHostMallocDefault));
    }
    MQ_HIP(st->ctx, zero64(st->d_isum, (int64_t)st->nrep * 8, s));
    MQ_HIP(st->ctx, zero64(st->d_fsum, (int64_t)st->nrep * 8, s));
    hipError_t me = launch_measure_all(st, st->d_isum, st->d_fsum);
    if (me != hipSuccess) return mq_fail(st->ctx, MQ_ERR_HIP, std::string("measure launch: ") + hipGetErrorString(me));
    const int64_t n = (int64_t)st->nrep * 17;
    hipLaunchKernelGGL(k_pack_obs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const unsigned long long*>(st->d_isum), st->d_fsum, st->d_betad, st->nrep,
                       st->d_pack);
    MQ_HIP(st->ctx, hipGetLastError());
    MQ_HIP(st->ctx, hipMemcpyAsync(st->h_pack, st->d_pack, sizeof(uint64_t) * (size_t)n, hipMemcpyDeviceToHost, s));
    MQ_HIP(st->ctx, hipStreamSynchronize(s));
    return MQ_OK;
}

extern "C" int mq_measure(mq_state* st, double* out)
{
    if (!st) return MQ_ERR_ARG;
    if (!out) return mq_fail(st->ctx, MQ_ERR_ARG, "out is NULL");
    if (st->beta.empty()) return mq_fail(st->ctx, MQ_ERR_STATE, "beta not set");
    int rc = measure_sums(st);
    if (rc) return rc;
    for (int r = 0; r < st->nrep; ++r) {
        const double N = (double)st->lat[r]->N;
        const uint64_t* P = st->h_pack + (size_t)r * 17;
        long long I[8];
        double F[8];
        std::memcpy(I, P, sizeof(I));
        std::memcpy(F, P + 8, sizeof(F));
        double bcur_r;
        std::memcpy(&bcur_r, P + 16, sizeof(double));
        double e = 0, m = 0;
        if (st->p.model == MQ_POTTS) {
            e = st->p.J * (double)I[0] / N;
            // the kernels count colours 0..6; colour 7 (q = 8) is what they leave
            long long mx = 0, counted = 0;
            for (int q = 0; q < std::min(st->p.q, 7); ++q) {
                mx = std::max(mx, I[1 + q]);
                counted += I[1 + q];
            }
            if (st->p.q == 8) mx = std::max(mx, (long long)st->lat[r]->N - counted);
            m = ((double)st->p.q * (double)mx - N) / (N * (st->p.q - 1));  // Potts.h:40-65
        } else if (is_int(st->p.model)) {
            e = st->p.J * (double)I[0] / N;
            m = std::fabs((double)I[1]) / N;
        } else {
            e = st->p.J * F[0] / N;
            if (st->p.model == MQ_PHI4) {
                const double b = bcur_r;
                e += (st->p.mass / b) * F[4] / N + (st->p.lambda / b) * F[5] / N;
            }
            m = std::sqrt(F[1] * F[1] + F[2] * F[2] + F[3] * F[3]) / N;
        }
        out[(size_t)r * MQ_NOBS + 0] = e;
        out[(size_t)r * MQ_NOBS + 1] = m;
        out[(size_t)r * MQ_NOBS + 2] = m * m;
        out[(size_t)r * MQ_NOBS + 3] = m * m * m * m;
    }
    return MQ_OK;
}

// Observables of replicas rep0 .. rep0+nrep-1 (one lattice of N sites) from the
// sums of the measurement kernels, with the arithmetic and the operation order
// of mq_measure's host loop (no contraction), so integer models give the same
// bits.
__global__ void k_finish_obs(const long long* __restrict__ isum, const double* __restrict__ fsum,
                             const double* __restrict__ betad, int rep0, int nrep, double N, int model, int q,
                             double J, double mass, double lambda, double* __restrict__ out)
{
#pragma clang fp contract(off)
    const int r = rep0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= rep0 + nrep) return;
    const long long* I = isum + (size_t)r * 8;
    const double* F = fsum + (size_t)r * 8;
    double e, m;
    if (model == MQ_POTTS) {
        e = J * (double)I[0] / N;
        long long mx = 0, counted = 0;
        const int qc = q < 7 ? q : 7;
        for (int k = 0; k < qc; ++k) {
            mx = I[1 + k] > mx ? I[1 + k] : mx;
            counted += I[1 + k];
        }
        if (q == 8) {
            const long long rest = (long long)N - counted;
            mx = rest > mx ? rest : mx;
        }
        m = ((double)q * (double)mx - N) / (N * (q - 1));
    } else if (model == MQ_ISING || model == MQ_EA_ISING) {
        e = J * (double)I[0] / N;
        m = fabs((double)I[1]) / N;
    } else {
        e = J * F[0] / N;
        if (model == MQ_PHI4) {
            const double b = betad[r];
            e += (mass / b) * F[4] / N + (lambda / b) * F[5] / N;
        }
        m = sqrt(F[1] * F[1] + F[2] * F[2] + F[3] * F[3]) / N;
    }
    out[(size_t)r * MQ_NOBS + 0] = e;
    out[(size_t)r * MQ_NOBS + 1] = m;
    out[(size_t)r * MQ_NOBS + 2] = m * m;
    out[(size_t)r * MQ_NOBS + 3] = m * m * m * m;
}

extern "C" int mq_measure_noalloc(mq_state* st, double* d_out)
{
    if (!st) return MQ_ERR_ARG;
    if (!d_out) return mq_fail(st->ctx, MQ_ERR_ARG, "out is NULL");
    if (st->beta.empty()) return mq_fail(st->ctx, MQ_ERR_STATE, "beta not set");
    if (st->nrep > MQ_NOALLOC_MAX_REP)
        return mq_fail(st->ctx, MQ_ERR_ARG,
                       "mq_measure_noalloc: more replicas than its static scratch holds (MQ_NOALLOC_MAX_REP = " +
                           std::to_string(MQ_NOALLOC_MAX_REP) + ")");
    if (!st->ctx->scratch) return mq_fail(st->ctx, MQ_ERR_STATE, "scratch address unknown");
    hipStream_t s = st->ctx->stream;
    long long* isum = static_cast<long long*>(st->ctx->scratch);
    double* fsum = reinterpret_cast<double*>(isum + (size_t)MQ_NOALLOC_MAX_REP * 8);
    // the array keeps the previous call's sums: clear this state's part with fill kernels
    MQ_HIP(st->ctx, zero64(isum, (int64_t)st->nrep * 8, s));
    MQ_HIP(st->ctx, zero64(fsum, (int64_t)st->nrep * 8, s));
    hipError_t me = launch_measure_all(st, isum, fsum);
    if (me != hipSuccess) return mq_fail(st->ctx, MQ_ERR_HIP, std::string("measure launch: ") + hipGetErrorString(me));
    for (const auto& g : st->groups) {
        hipLaunchKernelGGL(k_finish_obs, dim3((unsigned)((g.nrep + 255) / 256)), dim3(256), 0, s, isum, fsum,
                           st->d_betad, g.rep0, g.nrep, (double)g.lat->N, st->p.model, st->p.q, st->p.J, st->p.mass,
                           st->p.lambda, d_out);
        MQ_HIP(st->ctx, hipGetLastError());
    }
    return MQ_OK;
}

extern "C" int mq_measure_ft(mq_state* st, double* out)
{
    if (!st) return MQ_ERR_ARG;
    if (!out) return mq_fail(st->ctx, MQ_ERR_ARG, "out is NULL");
    if (st->p.model == MQ_POTTS)
        return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED,
                       "MagFTComp: Ising, EA, Heisenberg and Phi4 spins only (the reference measures none for Potts)");
    if (st->msc || st->eabits || st->csrbits)
        return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED, "MagFTComp: int8 or float4 storage only (spin_bits = 8)");
    for (const auto* l : st->lat)
        if (l->kind != LAT_REGULAR)
            return mq_fail(st->ctx, MQ_ERR_UNSUPPORTED, "MagFTComp needs site coordinates (regular lattices)");
    hipStream_t s = st->ctx->stream;
    const int64_t n = (int64_t)st->nrep * 18;
    if (!st->d_ft) MQ_HIP(st->ctx, dalloc(&st->d_ft, (size_t)n));
    MQ_HIP(st->ctx, zero64(st->d_ft, n, s));
    for (const auto& g : st->groups) {
        hipError_t e = launch_measure_ft_regular(st, g);
        if (e != hipSuccess) return mq_fail(st->ctx, MQ_ERR_HIP, std::string("measure launch: ") + hipGetErrorString(e));
    }
    std::vector<double> h((size_t)n);
    MQ_HIP(st->ctx, hipMemcpyAsync(h.data(), st->d_ft, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, s));
    MQ_HIP(st->ctx, hipStreamSynchronize(s));
    for (int r = 0; r < st->nrep; ++r) {
        const double N = (double)st->lat[r]->N;
        for (int d = 0; d < 3; ++d) {
            double v = std::numeric_limits<double>::quiet_NaN();
            if (d < st->lat[r]->dim) {
                // |(1/N) sum_i s_i[j] e^{2 pi i x_d}|^2 summed over components (observables.h:127-140)
                v = 0.0;
                for (int j = 0; j < 3; ++j) {
                    const double re = h[(size_t)r * 18 + (size_t)(d * 3 + j) * 2] / N;
                    const double im = h[(size_t)r * 18 + (size_t)(d * 3 + j) * 2 + 1] / N;
                    v += re * re + im * im;
                }
            }
            out[(size_t)r * 3 + d] = v;
        }
    }
    return MQ_OK;
}

extern "C" int mq_energy(mq_state* st, double* out)
{
    if (!st) return MQ_ERR_ARG;
    if (!out) return mq_fail(st->ctx, MQ_ERR_ARG, "out is NULL");
    std::vector<double> obs((size_t)st->nrep * MQ_NOBS);
    int rc = mq_measure(st, obs.data());
    if (rc) return rc;
    for (int r = 0; r < st->nrep; ++r) out[r] = obs[(size_t)r * MQ_NOBS] * (double)st->lat[r]->N;
    return MQ_OK;
}

extern "C" int mq_state_set_tuning(mq_state* st, const mq_tuning* t)
{
    if (!st) return MQ_ERR_ARG;
    if (!t) return mq_fail(st->ctx, MQ_ERR_ARG, "tuning is NULL");
    auto pow2 = [](int v) { return v >= 64 && v <= 512 && (v & (v - 1)) == 0; };
    if (t->kernels < 0 || t->kernels > 2 || !pow2(t->block) || !pow2(t->lds_block) || (t->xcd_remap != 0 && t->xcd_remap != 1))
        return mq_fail(st->ctx, MQ_ERR_ARG, "tuning: kernels in 0..2, block and lds_block powers of two in 64..512");
    const bool ring_ok = t->wolff_ring >= 64 && t->wolff_ring <= 2048 && (t->wolff_ring & (t->wolff_ring - 1)) == 0;
    if ((t->wolff_block != 0 && t->wolff_block != 256 && t->wolff_block != 512 && t->wolff_block != 1024) || !ring_ok ||
        t->wolff_vset < 0 || t->wolff_vset > MQ_VSET / 2)
        return mq_fail(st->ctx, MQ_ERR_ARG,
                       "tuning: wolff_block 0, 256, 512 or 1024; wolff_ring a power of two in 64..2048; wolff_vset in 0..4096");
    if (t->graph < 0 || t->graph > 2) return mq_fail(st->ctx, MQ_ERR_ARG, "tuning: graph in 0..2");
    st->kernel_choice = t->kernels;
    st->tune_block = t->block;
    st->tune_xcd = t->xcd_remap;
    st->lds_block = t->lds_block;
    st->wolff_block = t->wolff_block;
    st->wolff_ring = t->wolff_ring;
    st->wolff_vset = t->wolff_vset;
    st->step_graph = t->graph;
    return MQ_OK;
}

extern "C" int mq_state_get_tuning(const mq_state* st, mq_tuning* t)
{
    if (!st) return MQ_ERR_ARG;
    if (!t) return mq_fail(st->ctx, MQ_ERR_ARG, "tuning is NULL");
    t->kernels = st->kernel_choice;
    t->block = st->tune_block;
    t->xcd_remap = st->tune_xcd;
    t->lds_block = st->lds_block;
    t->wolff_block = st->wolff_block;
    t->wolff_ring = st->wolff_ring;
    t->wolff_vset = st->wolff_vset;
    t->graph = st->step_graph;
    return MQ_OK;
}
