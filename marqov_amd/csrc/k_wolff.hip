// Wolff single-cluster updates on regular lattices as a workgroup-cooperative
// level-synchronous frontier BFS, one workgroup per replica.
//
// The cluster is grown in natural site indices (x0 fastest); spins are read
// and written through the colour-split layout.  A site joins the cluster when
// the bond to an in-cluster neighbour is active; activation uses one Philox
// uniform per bond keyed by (bond = owner*dim + d, cluster number, replica),
// so the cluster is the connected component of active bonds and does not
// depend on visit order.  Membership is an atomicOr on a per-replica bitmap,
// cleared again from the cluster's own queue (O(cluster), not O(N)).
//
// Reference semantics: Ising.h:122-169 (aligned neighbours, p = -expm1(2 beta J)),
// Potts.h:189-322 (neighbours equal to the old colour, relabel to a different
// colour), Heisenberg.h:191-243 (random Cartesian component; add with
// 1 - exp(-2 beta cpl), cpl = J s_cur'[a] s_j[a]), Phi4.h:182-244 (random
// unit vector reflection s -= 2 (s.r) r, coupling (s_cur'.r)(s_j.r)).
#include "mq_internal.hpp"

namespace {

enum { WK_ISING = 0, WK_POTTS = 1, WK_CART = 2, WK_RDIR = 3 };

template <int DIM>
__device__ __forceinline__ int64_t nat_nbr(int L, int64_t i, int slot)
{
    const int d = slot >> 1;
    int64_t pw = 1;
    if (d >= 1) pw = L;
    if (d >= 2) pw = (int64_t)L * L;
    const int x = (int)((i / pw) % L);
    if ((slot & 1) == 0) return (x == L - 1) ? i - (int64_t)(L - 1) * pw : i + pw;
    return (x == 0) ? i + (int64_t)(L - 1) * pw : i - pw;
}

template <int DIM>
__device__ __forceinline__ int64_t split_addr(int L, int64_t half, int64_t i)
{
    const int64_t x0 = i % L;
    const int64_t r = i / L;
    int s = (int)x0 + (int)(r % L);
    if (DIM == 3) s += (int)(r / L);
    return (int64_t)(s & 1) * half + (i >> 1);
}

__device__ __forceinline__ float3 rnddir_w(uint32_t w0, uint32_t w1)
{
    const float a0 = 2.0f * u01f(w0);
    const float a1 = 2.0f * u01f(w1);
    const float rho = sqrtf(fmaxf(a0 * (2.0f - a0), 0.0f));
    float sn, cs;
    sincospif(a1, &sn, &cs);
    return make_float3(1.0f - a0, rho * sn, rho * cs);
}

__device__ __forceinline__ float comp(const float4& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

template <int DIM, int KIND>
__global__ void __launch_bounds__(256) k_wolff_regular(char* __restrict__ base, int64_t stride, int L, int nclusters,
                                                       uint32_t cl0, int rep0, int Q,
                                                       const uint32_t* __restrict__ wthr,
                                                       const float* __restrict__ betaf, float J,
                                                       int32_t* __restrict__ queue, uint32_t* __restrict__ visit,
                                                       const int64_t* __restrict__ rep_site_off,
                                                       int64_t* __restrict__ csize, Key2 key)
{
    constexpr bool VEC = (KIND == WK_CART || KIND == WK_RDIR);
    __shared__ int s_tail;
    const int r = blockIdx.x;
    const int grep = rep0 + r;
    int64_t N = (int64_t)L * L;
    if (DIM == 3) N *= L;
    const int64_t half = N >> 1;
    int8_t* sp8 = reinterpret_cast<int8_t*>(base + r * stride);
    float4* sp4 = reinterpret_cast<float4*>(base + r * stride);
    int32_t* q = queue + rep_site_off[grep];
    uint32_t* vis = visit + (rep_site_off[grep] >> 5);
    const uint32_t tw = VEC ? 0u : wthr[grep];
    const float beta = VEC ? betaf[grep] : 0.f;

    for (int c = 0; c < nclusters; ++c) {
        const uint32_t ncl = cl0 + (uint32_t)c;
        const uint4 sw = philox4x32_10(0u, ncl, (uint32_t)grep, ST_WSEED, key);
        const int64_t seed = (int64_t)(((uint64_t)sw.x * (uint64_t)N) >> 32);
        const int64_t seed_a = split_addr<DIM>(L, half, seed);
        int val = 0, newq = 0, axis = 0;
        float3 rd = make_float3(0.f, 0.f, 0.f);
        if (KIND == WK_ISING || KIND == WK_POTTS) {
            val = sp8[seed_a];
            if (KIND == WK_POTTS) newq = (val + 1 + (int)(((uint64_t)sw.y * (uint64_t)(Q - 1)) >> 32)) % Q;
        } else if (KIND == WK_CART) {
            axis = (int)(((uint64_t)sw.y * 3ull) >> 32);
        } else {
            rd = rnddir_w(sw.y, sw.z);
        }
        __syncthreads();  // every lane has read the seed's old value
        if (threadIdx.x == 0) {
            if (KIND == WK_ISING) sp8[seed_a] = (int8_t)(-val);
            else if (KIND == WK_POTTS) sp8[seed_a] = (int8_t)newq;
            else {
                float4 v = sp4[seed_a];
                if (KIND == WK_CART) {
                    if (axis == 0) v.x = -v.x;
                    else if (axis == 1) v.y = -v.y;
                    else v.z = -v.z;
                } else {
                    const float d = v.x * rd.x + v.y * rd.y + v.z * rd.z;
                    v.x -= 2.f * d * rd.x;
                    v.y -= 2.f * d * rd.y;
                    v.z -= 2.f * d * rd.z;
                }
                sp4[seed_a] = v;
            }
            q[0] = (int32_t)seed;
            atomicOr(&vis[seed >> 5], 1u << (seed & 31));
            s_tail = 1;
        }
        __syncthreads();
        int head = 0;
        while (true) {
            const int tail = s_tail;
            __syncthreads();
            if (head >= tail) break;
            for (int idx = head + (int)threadIdx.x; idx < tail; idx += blockDim.x) {
                const int64_t cur = q[idx];
                float pc = 0.f;
                if (VEC) {
                    const float4 vc = sp4[split_addr<DIM>(L, half, cur)];
                    pc = (KIND == WK_CART) ? comp(vc, axis) : (vc.x * rd.x + vc.y * rd.y + vc.z * rd.z);
                }
#pragma unroll
                for (int sl = 0; sl < 2 * DIM; ++sl) {
                    const int64_t j = nat_nbr<DIM>(L, cur, sl);
                    const int64_t ja = split_addr<DIM>(L, half, j);
                    const int64_t owner = (sl & 1) ? j : cur;
                    const int64_t b = owner * DIM + (sl >> 1);
                    bool add;
                    if (!VEC) {
                        if (sp8[ja] != val) continue;
                        const uint4 bw = philox4x32_10((uint32_t)(b >> 2), ncl, (uint32_t)grep, ST_WBOND, key);
                        add = u4_get(bw, (int)(b & 3)) < tw;
                    } else {
                        const float4 vj = sp4[ja];
                        const float pj = (KIND == WK_CART) ? comp(vj, axis) : (vj.x * rd.x + vj.y * rd.y + vj.z * rd.z);
                        const float cpl = J * pc * pj;
                        if (!(cpl > 0.f)) continue;
                        const uint4 bw = philox4x32_10((uint32_t)(b >> 2), ncl, (uint32_t)grep, ST_WBOND, key);
                        add = u01f(u4_get(bw, (int)(b & 3))) < -expm1f(-2.0f * beta * cpl);
                    }
                    if (!add) continue;
                    const uint32_t bit = 1u << (j & 31);
                    const uint32_t old = atomicOr(&vis[j >> 5], bit);
                    if (old & bit) continue;
                    if (KIND == WK_ISING) sp8[ja] = (int8_t)(-val);
                    else if (KIND == WK_POTTS) sp8[ja] = (int8_t)newq;
                    else {
                        float4 v = sp4[ja];
                        if (KIND == WK_CART) {
                            if (axis == 0) v.x = -v.x;
                            else if (axis == 1) v.y = -v.y;
                            else v.z = -v.z;
                        } else {
                            const float d = v.x * rd.x + v.y * rd.y + v.z * rd.z;
                            v.x -= 2.f * d * rd.x;
                            v.y -= 2.f * d * rd.y;
                            v.z -= 2.f * d * rd.z;
                        }
                        sp4[ja] = v;
                    }
                    const int pos = atomicAdd(&s_tail, 1);
                    q[pos] = (int32_t)j;
                }
            }
            __syncthreads();
            head = tail;
        }
        const int size = s_tail;
        for (int idx = threadIdx.x; idx < size; idx += blockDim.x) vis[q[idx] >> 5] = 0u;
        __syncthreads();
        if (threadIdx.x == 0 && csize) csize[(int64_t)grep * nclusters + c] = size;
    }
}


// LDS-resident variant for small int8 lattices (N <= 16384): the replica's
// spins, its visited bitmap and a 16-bit cluster queue live in LDS for all
// nclusters clusters; spins are written back once.  Same bond counters and
// membership rule as k_wolff_regular, so the clusters are identical.
template <int DIM, int KIND>
__global__ void __launch_bounds__(256) k_wolff_lds(char* __restrict__ base, int64_t stride, int L, int nclusters,
                                                   uint32_t cl0, int rep0, int Q, const uint32_t* __restrict__ wthr,
                                                   int64_t* __restrict__ csize, Key2 key)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_tail;
    const int r = blockIdx.x;
    const int grep = rep0 + r;
    int N = L * L;
    if (DIM == 3) N *= L;
    const int half = N >> 1;
    int8_t* S = reinterpret_cast<int8_t*>(smem);
    uint32_t* vis = reinterpret_cast<uint32_t*>(smem + ((N + 15) & ~15));
    uint16_t* q = reinterpret_cast<uint16_t*>(smem + ((N + 15) & ~15) + (((N + 31) / 32) * 4 + 15) / 16 * 16);
    int8_t* gs = reinterpret_cast<int8_t*>(base + r * stride);
    for (int i = threadIdx.x; i < N; i += blockDim.x) S[i] = gs[i];
    for (int i = threadIdx.x; i < (N + 31) / 32; i += blockDim.x) vis[i] = 0u;
    const uint32_t tw = wthr[grep];
    __syncthreads();
    for (int c = 0; c < nclusters; ++c) {
        const uint32_t ncl = cl0 + (uint32_t)c;
        const uint4 sw = philox4x32_10(0u, ncl, (uint32_t)grep, ST_WSEED, key);
        const int seed = (int)(((uint64_t)sw.x * (uint64_t)N) >> 32);
        const int seed_a = (int)split_addr<DIM>(L, half, seed);
        const int val = S[seed_a];
        int newq = 0;
        if (KIND == WK_POTTS) newq = (val + 1 + (int)(((uint64_t)sw.y * (uint64_t)(Q - 1)) >> 32)) % Q;
        __syncthreads();
        if (threadIdx.x == 0) {
            S[seed_a] = (KIND == WK_ISING) ? (int8_t)(-val) : (int8_t)newq;
            q[0] = (uint16_t)seed;
            vis[seed >> 5] |= 1u << (seed & 31);
            s_tail = 1;
        }
        __syncthreads();
        int head = 0;
        while (true) {
            const int tail = s_tail;
            __syncthreads();
            if (head >= tail) break;
            for (int idx = head + (int)threadIdx.x; idx < tail; idx += blockDim.x) {
                const int cur = q[idx];
#pragma unroll
                for (int sl = 0; sl < 2 * DIM; ++sl) {
                    const int j = (int)nat_nbr<DIM>(L, cur, sl);
                    const int ja = (int)split_addr<DIM>(L, half, j);
                    if (S[ja] != val) continue;
                    const int owner = (sl & 1) ? j : cur;
                    const uint32_t b = (uint32_t)owner * DIM + (uint32_t)(sl >> 1);
                    const uint4 bw = philox4x32_10(b >> 2, ncl, (uint32_t)grep, ST_WBOND, key);
                    if (!(u4_get(bw, (int)(b & 3u)) < tw)) continue;
                    const uint32_t bit = 1u << (j & 31);
                    if (atomicOr(&vis[j >> 5], bit) & bit) continue;
                    S[ja] = (KIND == WK_ISING) ? (int8_t)(-val) : (int8_t)newq;
                    q[atomicAdd(&s_tail, 1)] = (uint16_t)j;
                }
            }
            __syncthreads();
            head = tail;
        }
        const int size = s_tail;
        for (int idx = threadIdx.x; idx < size; idx += blockDim.x) vis[q[idx] >> 5] = 0u;
        __syncthreads();
        if (threadIdx.x == 0 && csize) csize[(int64_t)grep * nclusters + c] = size;
    }
    for (int i = threadIdx.x; i < N; i += blockDim.x) gs[i] = S[i];
}

}  // namespace

hipError_t launch_wolff_regular(mq_state* st, const Group& g, int nclusters, uint32_t cl0, Key2 key)
{
    const int L = g.lat->L, dim = g.lat->dim;
    const int model = st->p.model;
    int kind;
    if (model == MQ_ISING) kind = WK_ISING;
    else if (model == MQ_POTTS) kind = WK_POTTS;
    else if (model == MQ_HEISENBERG && st->p.embedding == MQ_EMBED_CARTESIAN) kind = WK_CART;
    else if (model == MQ_HEISENBERG || model == MQ_PHI4) kind = WK_RDIR;
    else return hipErrorInvalidValue;
    const dim3 grid(g.nrep), block(256);
    hipStream_t s = st->ctx->stream;
    char* base = st->d_spins + g.off;
    const float J = (float)st->p.J;
    const int64_t N = g.lat->N;
    if ((kind == WK_ISING || kind == WK_POTTS) && N <= 16384 && st->kernel_choice == 0) {
        const size_t lds = (size_t)((N + 15) & ~15ll) + (size_t)(((N + 31) / 32) * 4 + 15) / 16 * 16 + (size_t)N * 2;
#define MQ_WL(D, K)                                                                                                   \
    hipLaunchKernelGGL((k_wolff_lds<D, K>), grid, block, lds, s, base, g.stride, L, nclusters, cl0, g.rep0, st->p.q, \
                       st->d_wthr, st->d_csize, key)
        if (dim == 3) {
            if (kind == WK_ISING) MQ_WL(3, WK_ISING);
            else MQ_WL(3, WK_POTTS);
        } else {
            if (kind == WK_ISING) MQ_WL(2, WK_ISING);
            else MQ_WL(2, WK_POTTS);
        }
#undef MQ_WL
        return hipGetLastError();
    }
#define MQ_W(D, K)                                                                                               \
    hipLaunchKernelGGL((k_wolff_regular<D, K>), grid, block, 0, s, base, g.stride, L, nclusters, cl0, g.rep0,  \
                       st->p.q, st->d_wthr, st->d_betaf, J, st->d_queue, st->d_visit, st->d_rep_site_off,       \
                       st->d_csize, key)
    if (dim == 3) {
        if (kind == WK_ISING) MQ_W(3, WK_ISING);
        else if (kind == WK_POTTS) MQ_W(3, WK_POTTS);
        else if (kind == WK_CART) MQ_W(3, WK_CART);
        else MQ_W(3, WK_RDIR);
    } else {
        if (kind == WK_ISING) MQ_W(2, WK_ISING);
        else if (kind == WK_POTTS) MQ_W(2, WK_POTTS);
        else if (kind == WK_CART) MQ_W(2, WK_CART);
        else MQ_W(2, WK_RDIR);
    }
#undef MQ_W
    return hipGetLastError();
}
