// Wolff single-cluster updates on regular lattices as a workgroup-cooperative
// level-synchronous frontier BFS, one workgroup per replica.
//
// The cluster is grown in natural site indices (x0 fastest); spins are read
// and written through the colour-split layout.  A site joins the cluster when
// the test from an in-cluster neighbour passes; every test has its own Philox
// uniform keyed by (site * 2 * dim + neighbour slot, cluster number, replica),
// as the reference draws a fresh random number per test, so the cluster is the
// set of sites reachable from the seed along passing tests and does not depend
// on visit order.  Ising and Potts sites join with one atomic on their
// spin byte.  O(3)/Phi4 membership is an LDS hash set (an LDS compare-and-swap
// per claim, no global round trip) while the cluster cannot outgrow vlim
// sites, then an atomicOr on a per-replica bitmap, cleared again from the
// cluster's own queue (O(cluster), not O(N)).
//
// Reference semantics: Ising.h:122-169 (aligned neighbours, p = -expm1(2 beta J)),
// Potts.h:189-322 (neighbours equal to the old colour, relabel to a different
// colour), Heisenberg.h:191-243 (random Cartesian component; add with
// 1 - exp(-2 beta cpl), cpl = J s_cur'[a] s_j[a]), Phi4.h:182-244 (random
// unit vector reflection s -= 2 (s.r) r, coupling (s_cur'.r)(s_j.r)).
#include <algorithm>

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

template <int KIND>
__device__ __forceinline__ float proj(const float4& v, int axis, const float3& rd)
{
    return (KIND == WK_CART) ? comp(v, axis) : (v.x * rd.x + v.y * rd.y + v.z * rd.z);
}

// Cartesian embedding: negate component axis; random direction r: s -= 2 (s.r) r.
template <int KIND>
__device__ __forceinline__ float4 reflect(float4 v, int axis, const float3& rd)
{
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
    return v;
}

// Uniform of the Wolff test from site i (natural index) to its neighbour over
// slot sl (2d: +e_d, 2d+1: -e_d): word e & 3 of block e >> 2 of the test stream,
// with e = i * 2 * DIM + sl (for 3D written as t = 3i + d, block t >> 1, so that
// it stays in 32 bits for N <= 2^30).  The two tests along one axis share a block.
template <int DIM>
__device__ __forceinline__ uint32_t test_uniform(uint32_t i, int sl, uint32_t ncl, uint32_t grep, Key2 key)
{
    if (DIM == 2) return u4_get(philox4x32_10(i, ncl, grep, ST_WBOND, key), sl);
    const uint32_t t = 3u * i + (uint32_t)(sl >> 1);
    return u4_get(philox4x32_10(t >> 1, ncl, grep, ST_WBOND, key), (int)(((t & 1u) << 1) | (uint32_t)(sl & 1)));
}

// All 2 * DIM test uniforms of site i: one Philox block in 2D, two in 3D
// (e = 6i ... 6i + 5 spans blocks (3i) >> 1 and the next one).
template <int DIM>
__device__ __forceinline__ void site_test_uniforms(uint32_t i, uint32_t ncl, uint32_t grep, Key2 key,
                                                   uint32_t (&u)[2 * DIM])
{
    if constexpr (DIM == 2) {
        const uint4 w = philox4x32_10(i, ncl, grep, ST_WBOND, key);
        u[0] = w.x;
        u[1] = w.y;
        u[2] = w.z;
        u[3] = w.w;
    } else {
        const uint32_t b0 = (3u * i) >> 1;
        const uint4 w0 = philox4x32_10(b0, ncl, grep, ST_WBOND, key);
        const uint4 w1 = philox4x32_10(b0 + 1u, ncl, grep, ST_WBOND, key);
        const bool odd = (i & 1u) != 0u;  // 6i = 2 mod 4: slots 0, 1 are words 2, 3 of the first block
        u[0] = odd ? w0.z : w0.x;
        u[1] = odd ? w0.w : w0.y;
        u[2] = odd ? w1.x : w0.z;
        u[3] = odd ? w1.y : w0.w;
        u[4] = odd ? w1.z : w1.x;
        u[5] = odd ? w1.w : w1.y;
    }
}

// Frontier ring in LDS: entry pos written during the level whose first entry
// is `head` goes to ring slot pos mod WCAP when pos - head < WCAP, else to the
// global queue.  Entries of the level being read lie in [head, tail) and the
// ones being written in [tail, tail + WCAP - (tail - head)), so a slot is never
// overwritten before it is read.  The reader of that level knows `head` as its
// own `hprev`.
constexpr int WCAP = 4096;
// pair tasks (the two tests of a site along one axis) a lane takes at once
constexpr int WKP = 2;
// Levels of more than SITE_MIN * blockDim tests with claims in global memory run
// one lane per site (two Philox blocks for a 3D site's six tests); narrower levels,
// and levels whose members are in the LDS set, one lane per test.  Same clusters,
// sites/ns with one lane per test only / SITE_MIN 1 / 2 / 4 / 6: thermalised 3D Ising
// L=256 x 32 0.583 / 0.537 / 0.589 / 0.604 / 0.579; Heisenberg L=128 x 128 (Cartesian)
// 0.498 / 0.449 / 0.484 / 0.500; one lane per site on every level: 0.454 and 0.297.
constexpr int SITE_MIN = 4;

// LDS bytes of one replica's spins in the small-lattice kernel: a byte per
// site, or a bit per site when packed (16-byte multiples)
__host__ __device__ inline uint32_t wave_spin_bytes(uint32_t N, bool pack) { return pack ? (N / 8 + 15) & ~15u : (N + 15) & ~15u; }

// Sites are handled as packed coordinates: 10 bits per axis in 3D (L <= 1024),
// 16 bits in 2D.  Frontier entries and LDS set keys are packed sites, so a
// neighbour, its natural index and its colour-split address come from adds and
// multiplies, with no integer division per test.
template <int DIM>
struct SitePack {
    static constexpr int SB = DIM == 3 ? 10 : 16;
    static constexpr uint32_t CM = (1u << SB) - 1u;
    __device__ static uint32_t pack(uint32_t x0, uint32_t x1, uint32_t x2)
    {
        return DIM == 3 ? x0 | (x1 << SB) | (x2 << (2 * (DIM == 3 ? SB : 0))) : x0 | (x1 << SB);
    }
    __device__ static uint32_t z(uint32_t p) { return DIM == 3 ? p >> (2 * (DIM == 3 ? SB : 0)) : 0u; }
    // natural site index sum_j x_j L^j
    __device__ static uint32_t natural(uint32_t p, uint32_t uL, uint32_t LL)
    {
        return (p & CM) + uL * ((p >> SB) & CM) + LL * z(p);
    }
};

// Grows nclusters clusters of replica r (group-local index) with the calling
// workgroup; the body of k_wolff_regular and of the persistent EMCS kernel.
template <int DIM, int KIND>
__device__ __forceinline__ void wolff_regular_clusters(const int r, char* __restrict__ base, int64_t stride, int L,
                                                       int nclusters, uint32_t cl0, int rep0, int Q,
                                                       const uint32_t* __restrict__ wthr,
                                                       const float* __restrict__ betaf, float J,
                                                       int32_t* __restrict__ queue, uint32_t* __restrict__ visit,
                                                       const int64_t* __restrict__ rep_site_off,
                                                       int64_t* __restrict__ csize, int vlim, Key2 key)
{
    constexpr bool VEC = (KIND == WK_CART || KIND == WK_RDIR);
    constexpr int SB = SitePack<DIM>::SB;
    constexpr uint32_t CM = SitePack<DIM>::CM;
    constexpr int NS = 2 * DIM;
    __shared__ int s_tail;
    __shared__ uint32_t s_ring[WCAP];
    __shared__ float s_proj[VEC ? WCAP : 1];
    __shared__ uint32_t s_vset[MQ_VSET];
    const int grep = rep0 + r;
    const uint32_t uL = (uint32_t)L;
    const uint32_t LL = uL * uL;
    const uint32_t N = (DIM == 3) ? LL * uL : LL;  // <= 2^31 (checked at lattice creation)
    const uint32_t half = N >> 1;
    int8_t* sp8 = reinterpret_cast<int8_t*>(base + r * stride);
    float4* sp4 = reinterpret_cast<float4*>(base + r * stride);
    int32_t* q = queue + rep_site_off[grep];
    uint32_t* vis = visit + (rep_site_off[grep] >> 5);
    const uint32_t tw = VEC ? 0u : wthr[grep];
    const float beta = VEC ? betaf[grep] : 0.f;
    if (vlim > 0)
        for (int i = threadIdx.x; i < MQ_VSET; i += blockDim.x) s_vset[i] = 0xFFFFFFFFu;

    for (int c = 0; c < nclusters; ++c) {
        const uint32_t ncl = cl0 + (uint32_t)c;
        const uint4 sw = philox4x32_10(0u, ncl, (uint32_t)grep, ST_WSEED, key);
        const uint32_t seed = (uint32_t)(((uint64_t)sw.x * (uint64_t)N) >> 32);
        const uint32_t sx0 = seed % uL, srow = seed / uL;
        const uint32_t sx1 = (DIM == 3) ? srow % uL : srow, sx2 = (DIM == 3) ? srow / uL : 0u;
        const uint32_t seed_p = SitePack<DIM>::pack(sx0, sx1, sx2);
        const uint32_t seed_a = ((sx0 + sx1 + sx2) & 1u) * half + (seed >> 1);
        // Membership: while the cluster cannot outgrow vlim sites, an LDS hash set of
        // packed sites (one LDS compare-and-swap per claim, and for int8 spins the
        // alignment test is a plain load issued before the test's Philox draw); after
        // that, the spin byte itself (Ising, Potts: "still holds the old value" is "not
        // in the cluster", claimed with one global atomic) or the global bitmap
        // (O(3)/Phi4).  The choice is block-uniform.
        bool glob = vlim == 0;
        int val = 0, newq = 0, axis = 0;
        float3 rd = make_float3(0.f, 0.f, 0.f);
        if (KIND == WK_ISING || KIND == WK_POTTS) {
            // atomic load: the previous cluster changed bytes with stores and atomics
            uint32_t sh;
            val = (int8_t)(peek_word(sp8 + seed_a, sh) >> sh);
            newq = (KIND == WK_POTTS) ? (val + 1 + (int)(((uint64_t)sw.y * (uint64_t)(Q - 1)) >> 32)) % Q : -val;
        } else if (KIND == WK_CART) {
            axis = (int)(((uint64_t)sw.y * 3ull) >> 32);
        } else {
            rd = rnddir_w(sw.y, sw.z);
        }
        __syncthreads();  // every lane has read the seed's old value
        if (threadIdx.x == 0) {
            if (KIND == WK_ISING || KIND == WK_POTTS) {
                if (glob) {
                    if (KIND == WK_ISING) {
                        claim_ising(sp8 + seed_a, val);
                    } else {
                        uint32_t sh;
                        claim_potts(sp8 + seed_a, peek_word(sp8 + seed_a, sh), val, newq);
                    }
                } else {
                    sp8[seed_a] = (int8_t)newq;
                    vset_insert(s_vset, seed_p);
                }
            } else {
                const float4 v = reflect<KIND>(sp4[seed_a], axis, rd);
                sp4[seed_a] = v;
                s_proj[0] = proj<KIND>(v, axis, rd);
                if (glob) atomicOr(&vis[seed >> 5], 1u << (seed & 31));
                else vset_insert(s_vset, seed_p);
                q[0] = (int32_t)seed_p;
            }
            s_ring[0] = seed_p;
            s_tail = 1;
        }
        __syncthreads();
        int head = 0, hprev = 0;
        while (true) {
            const int tail = s_tail;
            __syncthreads();
            if (head >= tail) break;
            // One lane per (frontier site, neighbour slot): a level costs one Philox
            // evaluation and one memory round trip per lane; new entries are placed
            // with one LDS atomic per wavefront.
            const int ntask = (tail - head) * NS;
            if (!glob && tail + ntask > vlim) {
                // this level could take the cluster past vlim sites: from now on claim
                // in global memory.  Ising/Potts members already hold the new value;
                // O(3)/Phi4 members (all listed in q) move to the bitmap.
                if (VEC) {
                    for (int idx = threadIdx.x; idx < tail; idx += blockDim.x) {
                        const uint32_t m = (uint32_t)q[idx];
                        const uint32_t mn = SitePack<DIM>::natural(m, uL, LL);
                        atomicOr(&vis[mn >> 5], 1u << (mn & 31));
                    }
                }
                glob = true;
                __syncthreads();
            }
            if ((KIND == WK_ISING || KIND == WK_POTTS) && glob && ntask > SITE_MIN * (int)blockDim.x) {
                // A wide level with claims in global memory: one lane per frontier site.  The
                // site's coordinates are unpacked once, its 2*DIM test uniforms come from two
                // Philox blocks (one in 2D) instead of one block per test, and its 2*DIM
                // claims are in flight together.
                const int nsite = tail - head;
                for (int e = (int)threadIdx.x; e < nsite; e += (int)blockDim.x) {
                    const int idx = head + e;
                    const uint32_t cur = (idx - hprev < WCAP) ? s_ring[idx & (WCAP - 1)] : (uint32_t)q[idx];
                    const uint32_t x0 = cur & CM, x1 = (cur >> SB) & CM, x2 = SitePack<DIM>::z(cur);
                    const uint32_t curn = x0 + uL * x1 + LL * x2;
                    const uint32_t nh = ((x0 + x1 + x2) & 1u) ? 0u : half;  // the neighbours' colour half
                    uint32_t jp[NS], jn[NS], jan[NS], u[NS], seen[NS], prev[NS];
                    bool ok[NS];
#pragma unroll
                    for (int sl = 0; sl < NS; ++sl) {
                        const int d = sl >> 1;
                        const uint32_t xd = d == 0 ? x0 : (d == 1 ? x1 : x2);
                        uint32_t nx;
                        if ((sl & 1) == 0) nx = (xd == uL - 1) ? 0u : xd + 1u;
                        else nx = (xd == 0) ? uL - 1u : xd - 1u;
                        jp[sl] = (cur & ~(CM << (SB * d))) | (nx << (SB * d));
                        jn[sl] = curn + (nx - xd) * (d == 0 ? 1u : (d == 1 ? uL : LL));  // mod 2^32
                        jan[sl] = nh + (jn[sl] >> 1);
                        seen[sl] = 0u;
                        if (KIND == WK_POTTS) {
                            uint32_t sh;
                            seen[sl] = peek_word(sp8 + jan[sl], sh);  // in flight during the Philox rounds
                        }
                    }
                    site_test_uniforms<DIM>(curn, ncl, (uint32_t)grep, key, u);
                    if (KIND == WK_ISING) {
                        const bool up = __builtin_amdgcn_readfirstlane(val) > 0;
#pragma unroll
                        for (int sl = 0; sl < NS; ++sl) {
                            prev[sl] = 0u;
                            if (u[sl] < tw) {  // the claim itself tests alignment
                                uint32_t sh;
                                uint32_t* w = spin_word(sp8 + jan[sl], sh);
                                const uint32_t m = 0xFEu << sh;
                                prev[sl] = up ? atomicOr(w, m) : atomicAnd(w, ~m);
                            }
                        }
#pragma unroll
                        for (int sl = 0; sl < NS; ++sl)
                            ok[sl] = u[sl] < tw && (int8_t)(prev[sl] >> ((jan[sl] & 3u) * 8u)) == (int8_t)val;
                    } else {
                        bool tr[NS];
#pragma unroll
                        for (int sl = 0; sl < NS; ++sl) {
                            const uint32_t sh = (jan[sl] & 3u) * 8u;
                            tr[sl] = (int8_t)(seen[sl] >> sh) == (int8_t)val && u[sl] < tw;
                            prev[sl] = 0u;
                            if (tr[sl]) {
                                uint32_t sh2;
                                uint32_t* w = spin_word(sp8 + jan[sl], sh2);
                                prev[sl] = atomicCAS(w, seen[sl], (seen[sl] & ~(0xFFu << sh)) | ((uint32_t)(uint8_t)newq << sh));
                            }
                        }
#pragma unroll
                        for (int sl = 0; sl < NS; ++sl)  // a word changed under the first swap: retry from what it held
                            ok[sl] = tr[sl] && (prev[sl] == seen[sl] || claim_potts(sp8 + jan[sl], prev[sl], val, newq));
                    }
                    uint32_t nok = 0;
#pragma unroll
                    for (int sl = 0; sl < NS; ++sl) nok += ok[sl] ? 1u : 0u;
                    int pos = wave_reserve(&s_tail, nok);
#pragma unroll
                    for (int sl = 0; sl < NS; ++sl) {
                        if (!ok[sl]) continue;
                        if (pos - head < WCAP) s_ring[pos & (WCAP - 1)] = jp[sl];
                        else q[pos] = (int32_t)jp[sl];
                        ++pos;
                    }
                }
            } else if (VEC && glob && ntask > SITE_MIN * (int)blockDim.x) {
                // The same for O(3)/Phi4 with the visited set in the global bitmap: the site's
                // projection is read once, its 2*DIM neighbour spins are loaded together, and
                // its bitmap claims are in flight together.
                const int nsite = tail - head;
                for (int e = (int)threadIdx.x; e < nsite; e += (int)blockDim.x) {
                    const int idx = head + e;
                    const bool in_ring = idx - hprev < WCAP;
                    const uint32_t cur = in_ring ? s_ring[idx & (WCAP - 1)] : (uint32_t)q[idx];
                    const uint32_t x0 = cur & CM, x1 = (cur >> SB) & CM, x2 = SitePack<DIM>::z(cur);
                    const uint32_t curn = x0 + uL * x1 + LL * x2;
                    const uint32_t par = (x0 + x1 + x2) & 1u;
                    const uint32_t nh = par ? 0u : half;  // the neighbours' colour half
                    const float pc = in_ring ? s_proj[idx & (WCAP - 1)]
                                             : proj<KIND>(sp4[(par ? half : 0u) + (curn >> 1)], axis, rd);
                    uint32_t jp[NS], jn[NS], jan[NS], u[NS], old[NS];
                    float4 vj[NS];
                    bool ok[NS], tr[NS];
#pragma unroll
                    for (int sl = 0; sl < NS; ++sl) {
                        const int d = sl >> 1;
                        const uint32_t xd = d == 0 ? x0 : (d == 1 ? x1 : x2);
                        uint32_t nx;
                        if ((sl & 1) == 0) nx = (xd == uL - 1) ? 0u : xd + 1u;
                        else nx = (xd == 0) ? uL - 1u : xd - 1u;
                        jp[sl] = (cur & ~(CM << (SB * d))) | (nx << (SB * d));
                        jn[sl] = curn + (nx - xd) * (d == 0 ? 1u : (d == 1 ? uL : LL));  // mod 2^32
                        jan[sl] = nh + (jn[sl] >> 1);
                        vj[sl] = sp4[jan[sl]];  // in flight during the Philox rounds
                    }
                    site_test_uniforms<DIM>(curn, ncl, (uint32_t)grep, key, u);
#pragma unroll
                    for (int sl = 0; sl < NS; ++sl) {
                        const float cpl = J * pc * proj<KIND>(vj[sl], axis, rd);
                        tr[sl] = cpl > 0.f && u01f(u[sl]) < -expm1f(-2.0f * beta * cpl);
                        old[sl] = 0u;
                        if (tr[sl]) old[sl] = atomicOr(&vis[jn[sl] >> 5], 1u << (jn[sl] & 31));
                    }
                    uint32_t nok = 0;
#pragma unroll
                    for (int sl = 0; sl < NS; ++sl) {
                        ok[sl] = tr[sl] && !(old[sl] & (1u << (jn[sl] & 31)));
                        nok += ok[sl] ? 1u : 0u;
                    }
                    int pos = wave_reserve(&s_tail, nok);
#pragma unroll
                    for (int sl = 0; sl < NS; ++sl) {
                        if (!ok[sl]) continue;
                        // the claim succeeded, so nobody changed vj since it was read
                        const float4 v = reflect<KIND>(vj[sl], axis, rd);
                        sp4[jan[sl]] = v;
                        if (pos - head < WCAP) {
                            s_ring[pos & (WCAP - 1)] = jp[sl];
                            s_proj[pos & (WCAP - 1)] = proj<KIND>(v, axis, rd);
                        }
                        q[pos] = (int32_t)jp[sl];  // every member, for the bitmap
                        ++pos;
                    }
                }
            } else if (ntask > (int)blockDim.x) {
                // A wide level in pair tasks: a task is the two tests of one frontier site
                // along one axis (slots 2d and 2d + 1), whose uniforms are two words of one
                // Philox block.  A lane takes up to WKP tasks p, p + blockDim, ... at once,
                // issues all their loads, draws their uniforms and then issues all their
                // claims before using any result, so the level waits on one memory round
                // trip per lane.  Claims of one lane still act in order, and a site claimed
                // twice is taken once.
                const int npair = (tail - head) * DIM;
                constexpr int NT = 2 * WKP;  // tests per lane
                for (int p0 = (int)threadIdx.x; p0 < npair; p0 += WKP * (int)blockDim.x) {
                    uint32_t jp[NT], jn[NT], jan[NT], u[NT], old[NT], seen[NT];
                    float pc[WKP];
                    int sv[NT];
                    float4 vj[NT];
                    bool act[NT], ok[NT], try_[NT];
#pragma unroll
                    for (int k = 0; k < WKP; ++k) {
                        const int t = p0 + k * (int)blockDim.x;
                        pc[k] = 0.f;
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const int m = 2 * k + h;
                            act[m] = t < npair;
                            ok[m] = try_[m] = false;
                            jp[m] = jn[m] = jan[m] = 0u;
                            u[m] = 0xFFFFFFFFu;
                            old[m] = 0u;
                            seen[m] = 0u;
                            sv[m] = 0;
                            vj[m] = make_float4(0.f, 0.f, 0.f, 0.f);
                        }
                        if (t < npair) {
                            const int e = t / DIM;
                            const int d = t - e * DIM;
                            const int idx = head + e;
                            const bool in_ring = idx - hprev < WCAP;
                            const uint32_t cur = in_ring ? s_ring[idx & (WCAP - 1)] : (uint32_t)q[idx];
                            const uint32_t x0 = cur & CM, x1 = (cur >> SB) & CM, x2 = SitePack<DIM>::z(cur);
                            const uint32_t xd = d == 0 ? x0 : (d == 1 ? x1 : x2);
                            const uint32_t curn = x0 + uL * x1 + LL * x2;
                            const uint32_t pw = d == 0 ? 1u : (d == 1 ? uL : LL);
                            const uint32_t nh = ((x0 + x1 + x2) & 1u) ? 0u : half;  // the neighbours' colour half
#pragma unroll
                            for (int h = 0; h < 2; ++h) {
                                const int m = 2 * k + h;
                                uint32_t nx;
                                if (h == 0) nx = (xd == uL - 1) ? 0u : xd + 1u;
                                else nx = (xd == 0) ? uL - 1u : xd - 1u;
                                jp[m] = (cur & ~(CM << (SB * d))) | (nx << (SB * d));
                                jn[m] = curn + (nx - xd) * pw;  // mod 2^32
                                jan[m] = nh + (jn[m] >> 1);
                                // in flight during the Philox rounds: the spin byte while members
                                // are in the LDS set, the word a Potts claim swaps in global memory
                                if (!VEC && !glob) sv[m] = sp8[jan[m]];
                                if (KIND == WK_POTTS && glob) {
                                    uint32_t sh;
                                    seen[m] = peek_word(sp8 + jan[m], sh);
                                }
                                if (VEC) vj[m] = sp4[jan[m]];
                            }
                            if (VEC) {
                                const uint32_t cura = (((x0 + x1 + x2) & 1u) ? half : 0u) + (curn >> 1);
                                pc[k] = in_ring ? s_proj[idx & (WCAP - 1)] : proj<KIND>(sp4[cura], axis, rd);
                            }
                            // words e = curn*2*DIM + 2d and e + 1 of the test stream: one block
                            uint32_t blk, w0;
                            if (DIM == 2) {
                                blk = curn;
                                w0 = 2u * (uint32_t)d;
                            } else {
                                const uint32_t t3 = 3u * curn + (uint32_t)d;
                                blk = t3 >> 1;
                                w0 = (t3 & 1u) << 1;
                            }
                            const uint4 bw = philox4x32_10(blk, ncl, (uint32_t)grep, ST_WBOND, key);
                            u[2 * k] = w0 == 0u ? bw.x : bw.z;
                            u[2 * k + 1] = w0 == 0u ? bw.y : bw.w;
                        }
                    }
                    if (KIND == WK_ISING && glob) {
                        // all claims in flight together (the claim tests alignment)
                        const bool up = __builtin_amdgcn_readfirstlane(val) > 0;
#pragma unroll
                        for (int m = 0; m < NT; ++m) {
                            try_[m] = act[m] && u[m] < tw;
                            if (try_[m]) {
                                uint32_t sh;
                                uint32_t* w = spin_word(sp8 + jan[m], sh);
                                const uint32_t mk = 0xFEu << sh;
                                old[m] = up ? atomicOr(w, mk) : atomicAnd(w, ~mk);
                            }
                        }
#pragma unroll
                        for (int m = 0; m < NT; ++m)
                            ok[m] = try_[m] && (int8_t)(old[m] >> ((jan[m] & 3u) * 8u)) == (int8_t)val;
                    } else if (KIND == WK_POTTS && glob) {
#pragma unroll
                        for (int m = 0; m < NT; ++m) {
                            if (act[m] && u[m] < tw) {
                                const uint32_t sh = (jan[m] & 3u) * 8u;
                                if ((int8_t)(seen[m] >> sh) == (int8_t)val) {
                                    uint32_t sh2;
                                    uint32_t* w = spin_word(sp8 + jan[m], sh2);
                                    const uint32_t prev = atomicCAS(w, seen[m], (seen[m] & ~(0xFFu << sh)) | ((uint32_t)(uint8_t)newq << sh));
                                    // the word changed under the first swap: retry from what it held
                                    ok[m] = (prev == seen[m]) || claim_potts(sp8 + jan[m], prev, val, newq);
                                }
                            }
                        }
                    } else if (!VEC) {
#pragma unroll
                        for (int m = 0; m < NT; ++m) {
                            if (act[m] && sv[m] == val && u[m] < tw) {
                                ok[m] = vset_insert(s_vset, jp[m]);
                                if (ok[m]) sp8[jan[m]] = (int8_t)newq;
                            }
                        }
                    } else {
                        float cpl[NT];
#pragma unroll
                        for (int m = 0; m < NT; ++m) {
                            cpl[m] = J * pc[m >> 1] * proj<KIND>(vj[m], axis, rd);
                            try_[m] = act[m] && cpl[m] > 0.f && u01f(u[m]) < -expm1f(-2.0f * beta * cpl[m]);
                            if (try_[m] && glob) old[m] = atomicOr(&vis[jn[m] >> 5], 1u << (jn[m] & 31));
                        }
#pragma unroll
                        for (int m = 0; m < NT; ++m) {
                            if (!try_[m]) continue;
                            ok[m] = glob ? !(old[m] & (1u << (jn[m] & 31))) : vset_insert(s_vset, jp[m]);
                        }
                    }
                    uint32_t nok = 0;
#pragma unroll
                    for (int m = 0; m < NT; ++m) nok += ok[m] ? 1u : 0u;
                    int pos = wave_reserve(&s_tail, nok);
#pragma unroll
                    for (int m = 0; m < NT; ++m) {
                        if (!ok[m]) continue;
                        float pnew = 0.f;
                        if (VEC) {
                            // the claim succeeded, so nobody changed vj since it was read
                            const float4 v = reflect<KIND>(vj[m], axis, rd);
                            sp4[jan[m]] = v;
                            pnew = proj<KIND>(v, axis, rd);
                        }
                        if (pos - head < WCAP) {
                            s_ring[pos & (WCAP - 1)] = jp[m];
                            if (VEC) s_proj[pos & (WCAP - 1)] = pnew;
                        }
                        if (VEC || pos - head >= WCAP) q[pos] = (int32_t)jp[m];  // VEC: every member, for the bitmap
                        ++pos;
                    }
                }
            } else
            for (int t = (int)threadIdx.x; t < ntask; t += (int)blockDim.x) {
                const int e = t / NS;
                const int sl = t - e * NS;
                const int idx = head + e;
                const bool in_ring = idx - hprev < WCAP;
                const uint32_t cur = in_ring ? s_ring[idx & (WCAP - 1)] : (uint32_t)q[idx];
                const uint32_t x0 = cur & CM, x1 = (cur >> SB) & CM, x2 = SitePack<DIM>::z(cur);
                const int d = sl >> 1;
                const uint32_t xd = d == 0 ? x0 : (d == 1 ? x1 : x2);
                uint32_t nx;
                if ((sl & 1) == 0) nx = (xd == uL - 1) ? 0u : xd + 1u;
                else nx = (xd == 0) ? uL - 1u : xd - 1u;
                const uint32_t jp = (cur & ~(CM << (SB * d))) | (nx << (SB * d));
                const uint32_t curn = x0 + uL * x1 + LL * x2;
                const uint32_t jn = curn + (nx - xd) * (d == 0 ? 1u : (d == 1 ? uL : LL));  // mod 2^32
                // the neighbours of a site all have the other colour
                const uint32_t jan = (((x0 + x1 + x2) & 1u) ? 0u : half) + (jn >> 1);
                bool ok = false;
                float pnew = 0.f;
                if (KIND == WK_ISING || KIND == WK_POTTS) {
                    if (!glob) {
                        const int sv = sp8[jan];  // in flight during the Philox rounds
                        if (sv == val && test_uniform<DIM>(curn, sl, ncl, (uint32_t)grep, key) < tw) {
                            ok = vset_insert(s_vset, jp);
                            if (ok) sp8[jan] = (int8_t)newq;
                        }
                    } else if (KIND == WK_ISING) {
                        // the claim itself tests alignment
                        if (test_uniform<DIM>(curn, sl, ncl, (uint32_t)grep, key) < tw) ok = claim_ising(sp8 + jan, val);
                    } else {
                        uint32_t sh;
                        const uint32_t seen = peek_word(sp8 + jan, sh);  // in flight during the Philox rounds
                        if ((int8_t)(seen >> sh) == (int8_t)val && test_uniform<DIM>(curn, sl, ncl, (uint32_t)grep, key) < tw) {
                            uint32_t* w = spin_word(sp8 + jan, sh);
                            const uint32_t prev = atomicCAS(w, seen, (seen & ~(0xFFu << sh)) | ((uint32_t)(uint8_t)newq << sh));
                            // the word changed under the first swap: retry from what it held
                            ok = (prev == seen) || claim_potts(sp8 + jan, prev, val, newq);
                        }
                    }
                } else {
                    const uint32_t cura = (((x0 + x1 + x2) & 1u) ? half : 0u) + (curn >> 1);
                    const float pc = in_ring ? s_proj[idx & (WCAP - 1)] : proj<KIND>(sp4[cura], axis, rd);
                    const float4 vj = sp4[jan];  // in flight during the Philox rounds
                    const uint32_t ub = test_uniform<DIM>(curn, sl, ncl, (uint32_t)grep, key);
                    const float cpl = J * pc * proj<KIND>(vj, axis, rd);
                    if (cpl > 0.f && u01f(ub) < -expm1f(-2.0f * beta * cpl)) {
                        if (glob) {
                            const uint32_t ob = atomicOr(&vis[jn >> 5], 1u << (jn & 31));
                            ok = !(ob & (1u << (jn & 31)));
                        } else {
                            ok = vset_insert(s_vset, jp);
                        }
                        if (ok) {
                            // the claim succeeded, so nobody changed vj since it was read
                            const float4 v = reflect<KIND>(vj, axis, rd);
                            sp4[jan] = v;
                            pnew = proj<KIND>(v, axis, rd);
                        }
                    }
                }
                int pos = wave_reserve(&s_tail, ok ? 1u : 0u);
                if (ok) {
                    if (pos - head < WCAP) {
                        s_ring[pos & (WCAP - 1)] = jp;
                        if (VEC) s_proj[pos & (WCAP - 1)] = pnew;
                    }
                    if (VEC || pos - head >= WCAP) q[pos] = (int32_t)jp;  // VEC: every member, for the bitmap
                }
            }
            __syncthreads();
            hprev = head;
            head = tail;
        }
        const int size = s_tail;
        if (VEC && glob)
            for (int idx = threadIdx.x; idx < size; idx += blockDim.x) {
                const uint32_t m = (uint32_t)q[idx];
                vis[SitePack<DIM>::natural(m, uL, LL) >> 5] = 0u;
            }
        if (vlim > 0)
            for (int i = threadIdx.x; i < MQ_VSET; i += blockDim.x) s_vset[i] = 0xFFFFFFFFu;
        __syncthreads();
        if (threadIdx.x == 0 && csize) csize[(int64_t)grep * nclusters + c] = size;
    }
}

template <int DIM, int KIND>
__global__ void __launch_bounds__(1024) k_wolff_regular(char* __restrict__ base, int64_t stride, int L, int nclusters,
                                                       uint32_t cl0, int rep0, int Q,
                                                       const uint32_t* __restrict__ wthr,
                                                       const float* __restrict__ betaf, float J,
                                                       int32_t* __restrict__ queue, uint32_t* __restrict__ visit,
                                                       const int64_t* __restrict__ rep_site_off,
                                                       int64_t* __restrict__ csize, int vlim, Key2 key)
{
    wolff_regular_clusters<DIM, KIND>((int)blockIdx.x, base, stride, L, nclusters, cl0, rep0, Q, wthr, betaf, J, queue,
                                      visit, rep_site_off, csize, vlim, key);
}

// Workgroup sweeps of one replica for the persistent EMCS kernel: the decisions
// of the Ising and Potts sweep kernels (the same Philox counters and exact
// acceptance), Ising one group of 8 half-sites per thread and Philox block,
// Potts one half-site per thread and block.
template <int DIM>
__device__ __forceinline__ void wg_sweep_ising(int8_t* sp, const RegGeom& G, int colour, uint32_t phase, uint32_t grep,
                                               const uint32_t* s_thr, Key2 key)
{
    int8_t* own = sp + colour * G.half;
    const int8_t* oth = sp + (1 - colour) * G.half;
    const int64_t ngr = (G.half + 7) >> 3;
    constexpr int QMAX = 2 * DIM;
    for (int64_t g = threadIdx.x; g < ngr; g += blockDim.x) {
        const int64_t h0 = g << 3;
        const uint4 w = philox4x32_10((uint32_t)g, phase, grep, ST_METRO, key);
        uint4 tie = make_uint4(0, 0, 0, 0);
        bool have_tie = false;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int64_t h = h0 + m;
            if (h >= G.half) break;
            int64_t nb[2 * DIM];
            reg_nbrs<DIM>(G, colour, h, nb);
            int hs = 0;
#pragma unroll
            for (int sl = 0; sl < 2 * DIM; ++sl) hs += oth[nb[sl]];
            const int s = own[h];
            const int e = s * hs + QMAX;
            const uint32_t hi16 = (u4_get(w, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
            const uint32_t thi = s_thr[2 * e];
            bool acc = hi16 < thi;
            if (hi16 == thi) {
                if (!have_tie) {
                    tie = philox4x32_10((uint32_t)g, phase, grep, ST_TIE, key);
                    have_tie = true;
                }
                acc = ((u4_get(tie, m >> 1) >> (16 * (m & 1))) & 0xFFFFu) < s_thr[2 * e + 1];
            }
            if (acc) own[h] = (int8_t)(-s);
        }
    }
}

template <int DIM>
__device__ __forceinline__ void wg_sweep_potts(int8_t* sp, const RegGeom& G, int Q, int colour, uint32_t phase,
                                               uint32_t grep, const uint32_t* s_thr, Key2 key)
{
    int8_t* own = sp + colour * G.half;
    const int8_t* oth = sp + (1 - colour) * G.half;
    for (int64_t h = threadIdx.x; h < G.half; h += blockDim.x) {
        int64_t nb[2 * DIM];
        reg_nbrs<DIM>(G, colour, h, nb);
        const uint4 w = philox4x32_10((uint32_t)h, phase, grep, ST_METRO, key);
        const int snew = (int)(((uint64_t)w.x * (uint64_t)Q) >> 32);
        const int sold = own[h];
        int cn = 0, co = 0;
#pragma unroll
        for (int sl = 0; sl < 2 * DIM; ++sl) {
            const int v = oth[nb[sl]];
            cn += (v == snew);
            co += (v == sold);
        }
        const int e = cn - co + 2 * DIM;
        const uint64_t T = ((uint64_t)s_thr[2 * e] << 16) | s_thr[2 * e + 1];
        if ((uint64_t)w.y < T) own[h] = (int8_t)snew;
    }
}

// Persistent EMCS: nsteps x (ncluster Wolff clusters, then nmetro sweeps) of one
// replica per workgroup (int8 Ising / Potts, FP32 Heisenberg / Phi4), with no synchronisation between replicas, so a replica
// whose clusters are small goes on with its next steps while another still grows
// a large one.  Step k has the counters of mq_emcs(step0 + k): clusters
// (step0 + k) * ncluster + c, sweeps (step0 + k) * nmetro + t; the chain equals
// nsteps calls of mq_emcs bit for bit.  Replicas change only their own spins,
// and a workgroup's own writes are visible to its own later reads after its
// barriers.
template <int DIM, int KIND>
__global__ void __launch_bounds__(1024) k_emcs_regular(char* __restrict__ base, int64_t stride, int L, int nsteps,
                                                      int ncluster, int nmetro, uint64_t step0, int rep0, int nrep_all,
                                                      int Q, const uint32_t* __restrict__ wthr,
                                                      const uint32_t* __restrict__ thr, int tbl,
                                                      const float* __restrict__ betaf, float J, int model, int proposal,
                                                      float delta, float lambda, float mass,
                                                      int32_t* __restrict__ queue, uint32_t* __restrict__ visit,
                                                      const int64_t* __restrict__ rep_site_off,
                                                      int64_t* __restrict__ csize, int vlim, int lgL16, Key2 key)
{
    __shared__ uint32_t s_thr[2 * 16];
    __shared__ uint32_t s_thi[13];
    __shared__ uint32_t s_tlo[13];
    const int r = (int)blockIdx.x;
    const uint32_t grep = (uint32_t)(rep0 + r);
    if ((int)threadIdx.x < 2 * tbl) s_thr[threadIdx.x] = thr[(int64_t)grep * 2 * tbl + threadIdx.x];
    if (lgL16 && threadIdx.x < 13) {
        const uint32_t hi = thr[grep * 26u + 2u * threadIdx.x];
        const uint32_t lo = thr[grep * 26u + 2u * threadIdx.x + 1u];
        s_thi[threadIdx.x] = hi > 0xFFFFu ? 0xFFFFu : hi;
        s_tlo[threadIdx.x] = hi > 0xFFFFu ? 0x10000u : lo;
    }
    const RegGeom G = reg_geom(DIM, L);
    int8_t* sp = reinterpret_cast<int8_t*>(base + r * stride);
    __syncthreads();
    for (int k = 0; k < nsteps; ++k) {
        const uint64_t step = step0 + (uint64_t)k;
        if (ncluster > 0) {
            wolff_regular_clusters<DIM, KIND>(r, base, stride, L, ncluster, (uint32_t)(step * (uint64_t)ncluster), rep0,
                                              Q, wthr, betaf, J, queue, visit, rep_site_off,
                                              csize ? csize + (int64_t)k * nrep_all * ncluster : nullptr, vlim, key);
            __syncthreads();
        }
        for (int t = 0; t < nmetro; ++t) {
            const uint64_t sw = step * (uint64_t)nmetro + (uint64_t)t;
            for (int colour = 0; colour < 2; ++colour) {
                const uint32_t phase = (uint32_t)(2 * sw + (uint64_t)colour);
                if (KIND == WK_CART || KIND == WK_RDIR) {
                    float4* sp4 = reinterpret_cast<float4*>(base + r * stride);
                    float4* own = sp4 + colour * G.half;
                    const float4* oth = sp4 + (1 - colour) * G.half;
                    const float beta = betaf[grep];
                    for (int64_t h = threadIdx.x; h < G.half; h += blockDim.x)
                        o3_site_update<DIM, false>(own, oth, nullptr, G, colour, h, phase, grep, beta, model, proposal,
                                                   J, delta, lambda, mass, key);
                } else if (KIND == WK_ISING && DIM == 3 && lgL16) {
                    // 3D Ising, L a power of two >= 32: the headline sweep's 16-site chunks
                    const uint32_t nch = (uint32_t)(G.half >> 4);
                    for (uint32_t ch = threadIdx.x; ch < nch; ch += blockDim.x)
                        row16_chunk(sp, ch, lgL16, colour, phase, grep, s_thi, s_tlo, key);
                } else if (KIND == WK_ISING) {
                    wg_sweep_ising<DIM>(sp, G, colour, phase, grep, s_thr, key);
                } else {
                    wg_sweep_potts<DIM>(sp, G, Q, colour, phase, grep, s_thr, key);
                }
                __syncthreads();
            }
        }
    }
}


// One wavefront per replica for small int8 lattices (N <= 16384): the
// replica's spins and a 16-bit frontier ring (spilling to the global queue)
// live in the wave's slice of LDS for all nclusters clusters and are written
// back once.  A level is processed 64 frontier entries at a time, one per
// lane.  2D: the lane reads its four neighbours' spins, draws the entry's four
// test uniforms (one Philox block) and claims every neighbour that still holds
// the cluster's old value and whose test passes, the claims in flight
// together.  3D: (A) lanes list the candidate tests (neighbours that still
// hold the old value) in LDS; (B) lanes take candidates, one per lane, draw the
// test's uniform and claim.  Single-pass 3D (two Philox blocks per entry) and
// two-pass 2D measured 6% and 4% slower (3D L=16 x 4,096, 2D L=64 x 16,384, the
// same clusters).  New entries are placed by wave-wide prefix counts; nothing
// synchronises the workgroup.  Same test counters and membership rule as
// k_wolff_regular, so the clusters are identical.
// PACK (Ising, N a multiple of 32): the replica's spins sit in LDS as one bit
// per site (set = down), 1/8 of the bytes, so a CU holds up to twice as many
// replicas' waves and a batch larger than the CUs' share of the byte layout
// stays resident; a claim is then an atomic OR / AND on the site's bit.
template <int DIM, int KIND, bool PACK>
__global__ void __launch_bounds__(256) k_wolff_wave(char* __restrict__ base, int64_t stride, int L, int nrep,
                                                    int nclusters, uint32_t cl0, int rep0, int Q,
                                                    const uint32_t* __restrict__ wthr, int32_t* __restrict__ queue,
                                                    const int64_t* __restrict__ rep_site_off,
                                                    int64_t* __restrict__ csize, Key2 key, int cap, int rep_lds)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int r = (int)blockIdx.x * (int)(blockDim.x >> 6) + wave;
    if (r >= nrep) return;  // whole wave; nothing below synchronises the workgroup
    const int grep = rep0 + r;
    const uint32_t uL = (uint32_t)L;
    const uint32_t LL = uL * uL;
    const uint32_t N = (DIM == 3) ? LL * uL : LL;
    const uint32_t half = N >> 1;
    const float invL = 1.0f / (float)L;
    unsigned char* mine = smem + (size_t)wave * rep_lds;
    uint32_t* SW = reinterpret_cast<uint32_t*>(mine);  // PACK: bit h of word h>>5 = site address h is down
    int8_t* S = reinterpret_cast<int8_t*>(mine);
    // 3D only: the candidate list of the two-pass level (test word, natural index, address)
    constexpr int NCB = DIM == 3 ? 64 * 2 * DIM : 0;
    uint32_t* cb_b = reinterpret_cast<uint32_t*>(mine + wave_spin_bytes(N, PACK));
    uint16_t* cb_j = reinterpret_cast<uint16_t*>(cb_b + NCB);
    uint16_t* cb_a = cb_j + NCB;
    uint16_t* ring = cb_a + NCB;
    const uint32_t cmask = (uint32_t)cap - 1u;
    uint32_t* gw = reinterpret_cast<uint32_t*>(base + r * stride);  // N is a multiple of 4 (L even)
    int32_t* q = queue + rep_site_off[grep];
    if (PACK) {
        // 32 spin bytes -> one word: bit 1 of each byte (set iff -1) gathered by a multiply
        const uint4* g4 = reinterpret_cast<const uint4*>(base + r * stride);
        for (uint32_t w = lane; w < N / 32; w += 64) {
            const uint4 a = g4[2 * w], b = g4[2 * w + 1];
            const uint32_t d[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
            uint32_t word = 0u;
#pragma unroll
            for (int k = 0; k < 8; ++k) word |= ((((d[k] >> 1) & 0x01010101u) * 0x08102040u) >> 27) << (4 * k);
            SW[w] = word;
        }
    } else {
        for (uint32_t i = lane; i < N / 4; i += 64) SW[i] = gw[i];
    }
    const uint32_t tw = wthr[grep];
    const uint64_t lt = __lanemask_lt();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");

    for (int c = 0; c < nclusters; ++c) {
        const uint32_t ncl = cl0 + (uint32_t)c;
        const uint4 sw = philox4x32_10(0u, ncl, (uint32_t)grep, ST_WSEED, key);
        const uint32_t seed = (uint32_t)(((uint64_t)sw.x * (uint64_t)N) >> 32);
        const uint32_t seed_a = (uint32_t)split_addr<DIM>(L, half, seed);
        const int val = PACK ? (((SW[seed_a >> 5] >> (seed_a & 31u)) & 1u) ? -1 : 1) : (int)S[seed_a];
        const int newq = (KIND == WK_POTTS) ? (val + 1 + (int)(((uint64_t)sw.y * (uint64_t)(Q - 1)) >> 32)) % Q : -val;
        const int vbit = val < 0 ? 1 : 0;  // PACK: the old value's bit
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        if (lane == 0) {
            if (PACK) SW[seed_a >> 5] ^= 1u << (seed_a & 31u);
            else S[seed_a] = (int8_t)newq;
            ring[0] = (uint16_t)seed;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        int head = 0, hprev = 0, tail = 1;
        while (head < tail) {
            int next = tail;
            constexpr int NS = 2 * DIM;
            if constexpr (DIM == 2) {
                // one pass: a 2D entry's four test uniforms are one Philox block
                for (int chunk = head; chunk < tail; chunk += 64) {
                    // up to 64 frontier entries, one per lane
                    const int idx = chunk + lane;
                    const bool act = idx < tail;
                    // ring slots are always valid addresses: read unconditionally, fix up spills
                    uint32_t cur = (uint32_t)ring[(uint32_t)idx & cmask];
                    if (act && idx - hprev >= cap) cur = (uint32_t)q[idx];
                    if (!act) cur = 0u;
                    uint32_t x[3] = {0u, 0u, 0u};
                    {
                        uint32_t rem;
                        const uint32_t r1 = fdiv_u24(cur, uL, invL, rem);
                        x[0] = rem;
                        if (DIM == 3) x[2] = fdiv_u24(r1, uL, invL, x[1]);
                        else x[1] = r1;
                    }
                    const uint32_t nbase = ((x[0] + x[1] + x[2]) & 1u) ? 0u : half;
                    // neighbours of every lane (inactive lanes use site 0): the 2*DIM spin reads
                    // are issued together, unconditionally, before any is tested
                    uint32_t jn[NS], jav[NS], u[NS];
                    int sv[NS];
                    bool ok[NS];
    #pragma unroll
                    for (int sl = 0; sl < NS; ++sl) {
                        const int d = sl >> 1;
                        const uint32_t pw = d == 0 ? 1u : (d == 1 ? uL : LL);
                        uint32_t j;
                        if ((sl & 1) == 0) j = (x[d] == uL - 1) ? cur - (uL - 1) * pw : cur + pw;
                        else j = (x[d] == 0) ? cur + (uL - 1) * pw : cur - pw;
                        jn[sl] = j;
                        jav[sl] = nbase + (j >> 1);
                    }
    #pragma unroll
                    for (int sl = 0; sl < NS; ++sl) sv[sl] = PACK ? (int)((SW[jav[sl] >> 5] >> (jav[sl] & 31u)) & 1u) : (int)S[jav[sl]];
                    // the entry's 2*DIM test uniforms: one Philox block in 2D, two in 3D
                    site_test_uniforms<DIM>(cur, ncl, (uint32_t)grep, key, u);
                    // a passing test to a neighbour that still holds the old value claims it (the
                    // claim re-checks: another lane may have taken the site since it was read)
    #pragma unroll
                    for (int sl = 0; sl < NS; ++sl) {
                        ok[sl] = false;
                        if (act && sv[sl] == (PACK ? vbit : val) && u[sl] < tw) {
                            const uint32_t ja = jav[sl];
                            const uint32_t sh = (ja & 3u) * 8u;
                            if (PACK) {
                                const uint32_t mb = 1u << (ja & 31u);
                                const uint32_t old = vbit ? atomicAnd(&SW[ja >> 5], ~mb) : atomicOr(&SW[ja >> 5], mb);
                                ok[sl] = ((old & mb) != 0u) == (vbit != 0);
                            } else if (KIND == WK_ISING) {
                                const uint32_t mm = 0xFEu << sh;
                                const uint32_t old = val > 0 ? atomicOr(&SW[ja >> 2], mm) : atomicAnd(&SW[ja >> 2], ~mm);
                                ok[sl] = (int8_t)(old >> sh) == (int8_t)val;
                            } else {
                                uint32_t seen = SW[ja >> 2];
                                while ((int8_t)(seen >> sh) == (int8_t)val) {
                                    const uint32_t nw = (seen & ~(0xFFu << sh)) | ((uint32_t)(uint8_t)newq << sh);
                                    const uint32_t prev = atomicCAS(&SW[ja >> 2], seen, nw);
                                    if (prev == seen) {
                                        ok[sl] = true;
                                        break;
                                    }
                                    seen = prev;
                                }
                            }
                        }
                    }
                    // new entries placed by wave-wide prefix counts, slot by slot
    #pragma unroll
                    for (int sl = 0; sl < NS; ++sl) {
                        const uint64_t m = __ballot(ok[sl]);
                        if (ok[sl]) {
                            const int pos = next + __popcll(m & lt);
                            if (pos - head < cap) ring[(uint32_t)pos & cmask] = (uint16_t)jn[sl];
                            else q[pos] = (int32_t)jn[sl];
                        }
                        next += __popcll(m);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                }
            } else {
                // two passes: (A) lanes take entries and list the neighbours that still
                // hold the old value (candidate tests) in LDS; (B) lanes take candidates,
                // draw the test's Philox uniform and claim, one test per lane (a 3D entry's
                // six uniforms would cost two blocks per lane, more than the candidates need)
                for (int chunk = head; chunk < tail; chunk += 64) {
                    // pass A: candidate tests of up to 64 frontier entries
                    const int idx = chunk + lane;
                    const bool act = idx < tail;
                    // ring slots are always valid addresses: read unconditionally, fix up spills
                    uint32_t cur = (uint32_t)ring[(uint32_t)idx & cmask];
                    if (act && idx - hprev >= cap) cur = (uint32_t)q[idx];
                    if (!act) cur = 0u;
                    uint32_t x[3] = {0u, 0u, 0u};
                    {
                        uint32_t rem;
                        const uint32_t r1 = fdiv_u24(cur, uL, invL, rem);
                        x[0] = rem;
                        if (DIM == 3) x[2] = fdiv_u24(r1, uL, invL, x[1]);
                        else x[1] = r1;
                    }
                    const uint32_t nbase = ((x[0] + x[1] + x[2]) & 1u) ? 0u : half;
                    int ncand = 0;
                    // neighbours of every lane (inactive lanes use site 0): the 2*DIM byte
                    // reads are issued together, unconditionally, before any is tested
                    uint32_t jn[NS], jav[NS];
                    int sv[NS];
    #pragma unroll
                    for (int sl = 0; sl < NS; ++sl) {
                        const int d = sl >> 1;
                        const uint32_t pw = d == 0 ? 1u : (d == 1 ? uL : LL);
                        uint32_t j;
                        if ((sl & 1) == 0) j = (x[d] == uL - 1) ? cur - (uL - 1) * pw : cur + pw;
                        else j = (x[d] == 0) ? cur + (uL - 1) * pw : cur - pw;
                        jn[sl] = j;
                        jav[sl] = nbase + (j >> 1);
                    }
    #pragma unroll
                    for (int sl = 0; sl < NS; ++sl) sv[sl] = PACK ? (int)((SW[jav[sl] >> 5] >> (jav[sl] & 31u)) & 1u) : (int)S[jav[sl]];
    #pragma unroll
                    for (int sl = 0; sl < NS; ++sl) {
                        const bool cand = act && sv[sl] == (PACK ? vbit : val);
                        const uint64_t m = __ballot(cand);
                        if (cand) {
                            const int p = ncand + __popcll(m & lt);
                            cb_b[p] = cur * (uint32_t)NS + (uint32_t)sl;  // the test's word e (N <= 16384)
                            cb_j[p] = (uint16_t)jn[sl];
                            cb_a[p] = (uint16_t)jav[sl];
                        }
                        ncand += __popcll(m);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                    // pass B: one candidate test per lane
                    for (int cc = 0; cc < ncand; cc += 64) {
                        const int k = cc + lane;
                        bool ok = false;
                        uint32_t j = 0;
                        if (k < ncand) {
                            const uint32_t b = cb_b[k];
                            j = cb_j[k];
                            const uint32_t ja = cb_a[k];
                            const uint4 bw = philox4x32_10(b >> 2, ncl, (uint32_t)grep, ST_WBOND, key);
                            if (u4_get(bw, (int)(b & 3u)) < tw) {
                                const uint32_t sh = (ja & 3u) * 8u;
                                if (PACK) {
                                    const uint32_t mb = 1u << (ja & 31u);
                                    const uint32_t old = vbit ? atomicAnd(&SW[ja >> 5], ~mb) : atomicOr(&SW[ja >> 5], mb);
                                    ok = ((old & mb) != 0u) == (vbit != 0);
                                } else if (KIND == WK_ISING) {
                                    const uint32_t mm = 0xFEu << sh;
                                    const uint32_t old = val > 0 ? atomicOr(&SW[ja >> 2], mm) : atomicAnd(&SW[ja >> 2], ~mm);
                                    ok = (int8_t)(old >> sh) == (int8_t)val;
                                } else {
                                    uint32_t seen = SW[ja >> 2];
                                    while ((int8_t)(seen >> sh) == (int8_t)val) {
                                        const uint32_t nw = (seen & ~(0xFFu << sh)) | ((uint32_t)(uint8_t)newq << sh);
                                        const uint32_t prev = atomicCAS(&SW[ja >> 2], seen, nw);
                                        if (prev == seen) {
                                            ok = true;
                                            break;
                                        }
                                        seen = prev;
                                    }
                                }
                            }
                        }
                        const uint64_t m = __ballot(ok);
                        if (ok) {
                            const int pos = next + __popcll(m & lt);
                            if (pos - head < cap) ring[(uint32_t)pos & cmask] = (uint16_t)j;
                            else q[pos] = (int32_t)j;
                        }
                        next += __popcll(m);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                }
            }
            hprev = head;
            head = tail;
            tail = next;
        }
        if (lane == 0 && csize) csize[(int64_t)grep * nclusters + c] = tail;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    if (PACK) {
        // one word -> 32 spin bytes: +1 = 0x01, -1 = 0xFF
        uint4* g4 = reinterpret_cast<uint4*>(base + r * stride);
        for (uint32_t w = lane; w < N / 32; w += 64) {
            const uint32_t word = SW[w];
            uint32_t d[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) d[k] = (((((word >> (4 * k)) & 15u) * 0x00204081u) & 0x01010101u) * 0xFEu) | 0x01010101u;
            g4[2 * w] = make_uint4(d[0], d[1], d[2], d[3]);
            g4[2 * w + 1] = make_uint4(d[4], d[5], d[6], d[7]);
        }
    } else {
        for (uint32_t i = lane; i < N / 4; i += 64) gw[i] = SW[i];
    }
}

}  // namespace

hipError_t launch_emcs_regular(mq_state* st, const Group& g, int nsteps, int ncluster, int nmetro, uint64_t step0,
                               Key2 key)
{
    const int L = g.lat->L, dim = g.lat->dim;
    const dim3 grid(g.nrep), block(wolff_block(st, g.nrep));
    hipStream_t s = st->ctx->stream;
    char* base = st->d_spins + g.off;
    int lgL16 = 0;  // the row-vectorised sweep applies: 3D Ising, L a power of two in [32, 1024]
    if (st->p.model == MQ_ISING && dim == 3 && (L & (L - 1)) == 0 && L >= 32 && L <= 1024 && st->tbl == 13)
        while ((1 << lgL16) < L) ++lgL16;
#define MQ_E(D, K)                                                                                                 \
    hipLaunchKernelGGL((k_emcs_regular<D, K>), grid, block, 0, s, base, g.stride, L, nsteps, ncluster, nmetro,       \
                       step0, g.rep0, st->nrep, st->p.q, st->d_wthr, st->d_thr, st->tbl, st->d_betaf, J, model,     \
                       st->p.proposal, (float)st->p.delta, (float)st->p.lambda, (float)st->p.mass, st->d_queue,     \
                       st->d_visit, st->d_rep_site_off, st->d_csize, st->wolff_vset, lgL16, key)
    const int model = st->p.model;
    const float J = (float)st->p.J;
    int kind;
    if (model == MQ_ISING) kind = WK_ISING;
    else if (model == MQ_POTTS) kind = WK_POTTS;
    else if (model == MQ_HEISENBERG && st->p.embedding == MQ_EMBED_CARTESIAN) kind = WK_CART;
    else kind = WK_RDIR;
    if (dim == 3) {
        if (kind == WK_ISING) MQ_E(3, WK_ISING);
        else if (kind == WK_POTTS) MQ_E(3, WK_POTTS);
        else if (kind == WK_CART) MQ_E(3, WK_CART);
        else MQ_E(3, WK_RDIR);
    } else {
        if (kind == WK_ISING) MQ_E(2, WK_ISING);
        else if (kind == WK_POTTS) MQ_E(2, WK_POTTS);
        else if (kind == WK_CART) MQ_E(2, WK_CART);
        else MQ_E(2, WK_RDIR);
    }
#undef MQ_E
    return hipGetLastError();
}

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
    const dim3 grid(g.nrep), block(wolff_block(st, g.nrep));
    if (dim == 3 && L > 1024) return hipErrorNotSupported;  // packed sites: 10 bits per axis
    hipStream_t s = st->ctx->stream;
    char* base = st->d_spins + g.off;
    const float J = (float)st->p.J;
    const int64_t N = g.lat->N;
    if ((kind == WK_ISING || kind == WK_POTTS) && N <= 16384 && st->kernel_choice == 0) {
        // ring: the whole cluster when N <= wolff_ring (512 by default), else
        // wolff_ring entries with spill to the global queue
        int cap = 64;
        while (cap < N && cap < st->wolff_ring) cap <<= 1;
        const bool pack = kind == WK_ISING && N % 32 == 0;
        const int ncb = dim == 3 ? 64 * 2 * dim : 0;  // 3D candidate list: uint32 test word + two uint16
        const int rep_lds = (int)(wave_spin_bytes((uint32_t)N, pack) + 8 * ncb + (2 * cap + 15) / 16 * 16);
        const int wpb = std::max(1, std::min(4, 65536 / rep_lds));
        const dim3 wgrid((unsigned)((g.nrep + wpb - 1) / wpb)), wblock((unsigned)(64 * wpb));
        const size_t lds = (size_t)wpb * (size_t)rep_lds;
#define MQ_WL(D, K, P)                                                                                              \
    hipLaunchKernelGGL((k_wolff_wave<D, K, P>), wgrid, wblock, lds, s, base, g.stride, L, g.nrep, nclusters, cl0,  \
                       g.rep0, st->p.q, st->d_wthr, st->d_queue, st->d_rep_site_off, st->d_csize, key, cap, rep_lds)
        if (dim == 3) {
            if (kind == WK_ISING) {
                if (pack) MQ_WL(3, WK_ISING, true);
                else MQ_WL(3, WK_ISING, false);
            } else {
                MQ_WL(3, WK_POTTS, false);
            }
        } else {
            if (kind == WK_ISING) {
                if (pack) MQ_WL(2, WK_ISING, true);
                else MQ_WL(2, WK_ISING, false);
            } else {
                MQ_WL(2, WK_POTTS, false);
            }
        }
#undef MQ_WL
        return hipGetLastError();
    }
#define MQ_W(D, K)                                                                                               \
    hipLaunchKernelGGL((k_wolff_regular<D, K>), grid, block, 0, s, base, g.stride, L, nclusters, cl0, g.rep0,  \
                       st->p.q, st->d_wthr, st->d_betaf, J, st->d_queue, st->d_visit, st->d_rep_site_off,       \
                       st->d_csize, st->wolff_vset, key)
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
