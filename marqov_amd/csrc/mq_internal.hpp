// Internal declarations shared by the HIP translation units of libmarqov_amd.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/marqov_amd.h"

// Philox stream ids (counter word 3).  The CPU replay in oracle/replay.c uses
// the same layout.
enum : uint32_t {
    ST_INIT = 0,
    ST_METRO = 1,
    ST_TIE = 2,
    ST_WSEED = 3,
    ST_WBOND = 4,
    ST_PT = 5,
    ST_DISORDER = 6,
    ST_PROPOSE = 7,
};

enum { LAT_REGULAR = 0, LAT_CSR = 1 };

// Colour classes of a graph lattice: the Philox phase of a sweep is
// stride*sweep + colour, stride = 64 up to 64 classes, else the next power of two
// (at most MQ_MAX_COLOR_STRIDE, which leaves 2^16 sweep numbers).
#define MQ_MAX_COLOR_STRIDE (1 << 16)

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Key bumps are uniform per launch, so
// the compiler keeps them in SGPRs; each round is two 32x32->64 multiplies and
// two three-way xors per lane.
struct Key2 {
    uint32_t k0, k1;
};

// a ^ b ^ k in one VALU instruction: v_bitop3_b32 with truth table 0x96 (odd
// parity of the three inputs); k is wave-uniform (the round key) and stays in
// an SGPR.  The compiler otherwise emits two v_xor_b32 per XOR3 here.
__device__ __forceinline__ uint32_t xor3_k(uint32_t a, uint32_t b, uint32_t k)
{
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "s"(k));
    return r;
}

__device__ __forceinline__ uint4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, Key2 key)
{
    uint32_t k0 = key.k0, k1 = key.k1;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = xor3_k((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t n2 = xor3_k((uint32_t)(p0 >> 32), c3, k1);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return make_uint4(c0, c1, c2, c3);
}

__host__ __device__ __forceinline__ uint32_t u4_get(const uint4& v, int i)
{
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

// n / d for n < 2^24 through a float reciprocal and one exact correction step
// (the float quotient is within 1 of the true one there); rem gets n - q*d.
__device__ __forceinline__ uint32_t fdiv_u24(uint32_t n, uint32_t d, float inv, uint32_t& rem)
{
    uint32_t q = (uint32_t)((float)n * inv);
    int32_t r = (int32_t)(n - q * d);
    // branch-free correction by one either way
    const bool lo = r < 0, hi = r >= (int32_t)d;
    q = q - (lo ? 1u : 0u) + (hi ? 1u : 0u);
    r = r + (lo ? (int32_t)d : 0) - (hi ? (int32_t)d : 0);
    rem = (uint32_t)r;
    return q;
}

// [0,1) float from the top 24 bits
__device__ __forceinline__ float u01f(uint32_t w) { return (float)(w >> 8) * 0x1.0p-24f; }

// ---------------------------------------------------------------------------
// Regular-lattice geometry in the colour-split layout: colour c of site i is
// (x0+x1+x2)&1, half index h = i >> 1; rows of H = L/2 half-sites.
struct RegGeom {
    int L, H;
    int64_t plane, half;
};

__device__ __forceinline__ RegGeom reg_geom(int dim, int L)
{
    RegGeom g;
    g.L = L;
    g.H = L >> 1;
    g.plane = (int64_t)g.H * L;
    g.half = (dim == 3) ? g.plane * L : g.plane;
    return g;
}

// neighbour half-indices (other colour) of colour-c site h, slot order
// +x0,-x0,+x1,-x1,+x2,-x2 (regular_hypercubic.h:31-66)
template <int DIM>
__device__ __forceinline__ void reg_nbrs(const RegGeom& G, int colour, int64_t h, int64_t nb[2 * DIM])
{
    int64_t row;
    int k, x1, x2 = 0;
    if (G.half < (1 << 24)) {  // 32-bit indices, reciprocal division
        uint32_t kk, xx1;
        const uint32_t r32 = fdiv_u24((uint32_t)h, (uint32_t)G.H, 1.0f / (float)G.H, kk);
        if (DIM == 3) x2 = (int)fdiv_u24(r32, (uint32_t)G.L, 1.0f / (float)G.L, xx1);
        else xx1 = r32;
        row = r32;
        k = (int)kk;
        x1 = (int)xx1;
    } else {
        row = h / G.H;
        k = (int)(h - row * G.H);
        x1 = (int)(row % G.L);
        if (DIM == 3) x2 = (int)(row / G.L);
    }
    const int p = (colour + x1 + x2) & 1;
    const int64_t rb = h - k;
    nb[0] = rb + (p ? ((k + 1 == G.H) ? 0 : k + 1) : k);
    nb[1] = rb + (p ? k : ((k == 0) ? G.H - 1 : k - 1));
    nb[2] = (x1 == G.L - 1) ? h - (int64_t)(G.L - 1) * G.H : h + G.H;
    nb[3] = (x1 == 0) ? h + (int64_t)(G.L - 1) * G.H : h - G.H;
    if (DIM == 3) {
        nb[4] = (x2 == G.L - 1) ? h - (int64_t)(G.L - 1) * G.plane : h + G.plane;
        nb[5] = (x2 == 0) ? h + (int64_t)(G.L - 1) * G.plane : h - G.plane;
    }
}

// ---------------------------------------------------------------------------
// fresh uniform direction in the reference's component order (randomdir.h:45-65)
__device__ __forceinline__ float3 rnddir_f(uint32_t w0, uint32_t w1)
{
    const float a0 = 2.0f * u01f(w0);
    const float a1 = 2.0f * u01f(w1);
    const float rho = sqrtf(fmaxf(a0 * (2.0f - a0), 0.0f));
    float sn, cs;
    sincospif(a1, &sn, &cs);
    return make_float3(1.0f - a0, rho * sn, rho * cs);
}

// Metropolis update of site h of colour `colour` (FP32 O(3) / Phi4): proposal from
// the words of Philox(h, phase, replica, METRO) (fresh direction, small angle, or
// the Phi4 radial step), energy change against the 2*DIM neighbours of the other
// colour (times their +-1 couplings with BONDS; bs = the replica's coupling
// sample), acceptance u < exp(-beta dE) with the word 2 uniform.
template <int DIM, bool BONDS>
__device__ __forceinline__ void o3_site_update(float4* __restrict__ own, const float4* __restrict__ oth,
                                               const int8_t* __restrict__ bs, const RegGeom& G, int colour, int64_t h,
                                               uint32_t phase, uint32_t grep, float beta, int model, int proposal,
                                               float J, float delta, float lambda, float mass, Key2 key)
{
    int64_t nb[2 * DIM];
    reg_nbrs<DIM>(G, colour, h, nb);
    float nx = 0.f, ny = 0.f, nz = 0.f;
#pragma unroll
    for (int sl = 0; sl < 2 * DIM; ++sl) {
        const float4 v = oth[nb[sl]];
        float b = 1.0f;
        if (BONDS) {
            const int d = sl >> 1;
            b = (sl & 1) ? (float)bs[(int64_t)(1 - colour) * DIM * G.half + (int64_t)d * G.half + nb[sl]]
                         : (float)bs[(int64_t)colour * DIM * G.half + (int64_t)d * G.half + h];
        }
        nx += b * v.x;
        ny += b * v.y;
        nz += b * v.z;
    }
    const float4 so = own[h];
    const uint4 w = philox4x32_10((uint32_t)h, phase, grep, ST_METRO, key);
    const float3 rd = rnddir_f(w.x, w.y);
    float3 sn;
    if (model == MQ_PHI4) {
        sn = make_float3(so.x + 0.5f * rd.x, so.y + 0.5f * rd.y, so.z + 0.5f * rd.z);
    } else if (proposal == MQ_PROPOSE_SMALL_ANGLE) {
        float3 t = make_float3(so.x + delta * rd.x, so.y + delta * rd.y, so.z + delta * rd.z);
        const float inv = rsqrtf(t.x * t.x + t.y * t.y + t.z * t.z);
        sn = make_float3(t.x * inv, t.y * inv, t.z * inv);
    } else {
        sn = rd;
    }
    float dE = J * ((sn.x - so.x) * nx + (sn.y - so.y) * ny + (sn.z - so.z) * nz);
    if (model == MQ_PHI4) {
        const float n2 = sn.x * sn.x + sn.y * sn.y + sn.z * sn.z;
        const float o2 = so.x * so.x + so.y * so.y + so.z * so.z;
        dE += (mass / beta) * (n2 - o2) + (lambda / beta) * ((n2 - 1.0f) * (n2 - 1.0f) - (o2 - 1.0f) * (o2 - 1.0f));
    }
    bool acc = dE <= 0.0f;
    if (!acc) acc = u01f(w.z) < __expf(-beta * dE);
    if (acc) own[h] = make_float4(sn.x, sn.y, sn.z, 0.0f);
}

// ---------------------------------------------------------------------------
// One chunk of the row-vectorised 3D Ising sweep (L a power of two, 32 <= L <= 1024,
// no couplings): 16 consecutive half-sites of one row of colour `colour` of the
// replica at rbase.  A chunk's own 16 bytes, five 16-byte chunks of the other
// colour (same row, y+-1, z+-1) and one dword of the adjacent chunk for the
// x-shifted neighbour; spins are +-1 bytes (bit 1 set iff -1), so the doubled
// count of down neighbours is a byte-wise sum of (word & 0x02020202) and the
// table index e = q + 6 is own_down ? 2n : 12 - 2n per byte.  Random numbers
// and acceptance are those of the generic Ising sweep (group g = h >> 3):
// 16-bit uniforms against the threshold's high half, the tie stream on equality.
// s_thi / s_tlo: the replica's 13 threshold halves ("always" as 0xFFFF / 0x10000).
__device__ __forceinline__ void row16_chunk(int8_t* __restrict__ rbase, uint32_t chunk, int lgL, int colour,
                                            uint32_t phase, uint32_t grep, const uint32_t* s_thi,
                                            const uint32_t* s_tlo, Key2 key)
{
    const uint32_t L = 1u << lgL;
    const uint32_t H = L >> 1;
    const uint32_t lgcpr = (uint32_t)lgL - 5u;        // chunks of 16 per row: H/16
    const uint32_t lgplane = 2u * (uint32_t)lgL - 1u;  // H*L
    const uint32_t plane = 1u << lgplane;
    const uint32_t half = plane << lgL;
    const uint32_t row = chunk >> lgcpr;
    const uint32_t c = chunk & ((1u << lgcpr) - 1u);
    const uint32_t y = row & (L - 1u);
    const uint32_t z = row >> lgL;
    const uint32_t p = (colour + y + z) & 1u;
    const uint32_t rb = row * H;
    const uint32_t h0 = rb + (c << 4);
    int8_t* own = rbase + (int64_t)colour * half;
    const int8_t* oth = rbase + (int64_t)(1 - colour) * half;

    const uint4 S = *reinterpret_cast<const uint4*>(own + h0);
    const uint4 O = *reinterpret_cast<const uint4*>(oth + h0);
    const uint32_t cmax = (1u << lgcpr) - 1u;
    const uint32_t xo = p ? rb + ((c == cmax) ? 0u : ((c + 1u) << 4)) : rb + ((c == 0u) ? H : (c << 4)) - 4u;
    const uint32_t X = *reinterpret_cast<const uint32_t*>(oth + xo);
    const uint32_t hym = (y == 0u) ? h0 + (L - 1u) * H : h0 - H;
    const uint32_t hyp = (y == L - 1u) ? h0 - (L - 1u) * H : h0 + H;
    const uint32_t hzm = (z == 0u) ? h0 + (L - 1u) * plane : h0 - plane;
    const uint32_t hzp = (z == L - 1u) ? h0 - (L - 1u) * plane : h0 + plane;
    const uint4 Ym = *reinterpret_cast<const uint4*>(oth + hym);
    const uint4 Yp = *reinterpret_cast<const uint4*>(oth + hyp);
    const uint4 Zm = *reinterpret_cast<const uint4*>(oth + hzm);
    const uint4 Zp = *reinterpret_cast<const uint4*>(oth + hzp);

    // x-shifted neighbour words
    uint32_t xs0, xs1, xs2, xs3;
    if (p) {
        xs0 = __builtin_amdgcn_alignbyte(O.y, O.x, 1);
        xs1 = __builtin_amdgcn_alignbyte(O.z, O.y, 1);
        xs2 = __builtin_amdgcn_alignbyte(O.w, O.z, 1);
        xs3 = __builtin_amdgcn_alignbyte(X, O.w, 1);
    } else {
        xs0 = __builtin_amdgcn_alignbyte(O.x, X, 3);
        xs1 = __builtin_amdgcn_alignbyte(O.y, O.x, 3);
        xs2 = __builtin_amdgcn_alignbyte(O.z, O.y, 3);
        xs3 = __builtin_amdgcn_alignbyte(O.w, O.z, 3);
    }
    const uint32_t M2 = 0x02020202u;
    // per byte: 4*e with e = q + 6 (LDS byte address of the threshold entry)
    auto eword = [&](uint32_t s, uint32_t o, uint32_t xs, uint32_t ym, uint32_t yp, uint32_t zm, uint32_t zp) {
        const uint32_t n2 = (o & M2) + (xs & M2) + (ym & M2) + (yp & M2) + (zm & M2) + (zp & M2);
        const uint32_t mask = ((s & M2) >> 1) * 0xFFu;
        return ((n2 & mask) | ((0x0C0C0C0Cu - n2) & ~mask)) << 2;
    };
    uint32_t E[4];
    E[0] = eword(S.x, O.x, xs0, Ym.x, Yp.x, Zm.x, Zp.x);
    E[1] = eword(S.y, O.y, xs1, Ym.y, Yp.y, Zm.y, Zp.y);
    E[2] = eword(S.z, O.z, xs2, Ym.z, Yp.z, Zm.z, Zp.z);
    E[3] = eword(S.w, O.w, xs3, Ym.w, Yp.w, Zm.w, Zp.w);

    const uint32_t g0 = h0 >> 3;
    const char* thi_b = reinterpret_cast<const char*>(s_thi);
    uint32_t acc = 0u;  // bit 15-site: site accepted
    bool any_tie = false;
#pragma unroll
    for (int half8 = 0; half8 < 2; ++half8) {
        const uint4 w = philox4x32_10(g0 + half8, phase, grep, ST_METRO, key);
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int site = half8 * 8 + m;
            const uint32_t ea = (E[site >> 2] >> (8 * (site & 3))) & 0xFFu;
            const uint32_t u = (u4_get(w, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
            const uint32_t t = *reinterpret_cast<const uint32_t*>(thi_b + ea);
            acc = acc + acc + (u < t ? 1u : 0u);
            any_tie |= (u == t);
        }
    }
    if (__builtin_expect(any_tie, 0)) {
        // an equal high half: redo this thread's 16 decisions with the low halves
        acc = 0u;
#pragma unroll
        for (int half8 = 0; half8 < 2; ++half8) {
            const uint4 w = philox4x32_10(g0 + half8, phase, grep, ST_METRO, key);
            const uint4 tw = philox4x32_10(g0 + half8, phase, grep, ST_TIE, key);
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const int site = half8 * 8 + m;
                const uint32_t ea = (E[site >> 2] >> (8 * (site & 3))) & 0xFFu;
                const uint32_t u = (u4_get(w, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
                const uint32_t lo = (u4_get(tw, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
                const uint32_t t = *reinterpret_cast<const uint32_t*>(thi_b + ea);
                const uint32_t tl = *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(s_tlo) + ea);
                const bool a = (u < t) || (u == t && lo < tl);
                acc = acc + acc + (a ? 1u : 0u);
            }
        }
    }
    // acc holds site 0 in bit 15 ... site 15 in bit 0: expand 4 bits -> 4 bytes
    const uint32_t rev = __builtin_bitreverse32(acc) >> 16;  // site k in bit k
    auto expand = [](uint32_t nib) { return ((nib * 0x00204081u) & 0x01010101u) * 0xFEu; };
    const uint4 Snew = make_uint4(S.x ^ expand(rev & 15u), S.y ^ expand((rev >> 4) & 15u),
                                  S.z ^ expand((rev >> 8) & 15u), S.w ^ expand((rev >> 12) & 15u));
    *reinterpret_cast<uint4*>(own + h0) = Snew;
}

// ---------------------------------------------------------------------------
// Bit-sliced counters: plane p of a count holds bit p of 32 per-bit counts.
// Bit-sliced add of Y into X over NP planes (values that overflow NP planes are
// not produced by the callers).
template <int NP>
__device__ __forceinline__ void bs_add(uint32_t (&X)[NP], const uint32_t (&Y)[NP])
{
    uint32_t carry = 0u;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const uint32_t x = X[p], y = Y[p];
        X[p] = x ^ y ^ carry;
        carry = (x & y) | (x & carry) | (y & carry);
    }
}

// Reduce-scatter of bit-sliced per-bit counts across the 64 lanes of a
// wavefront: after five exchange steps (partners lane ^ 16, 8, 4, 2, 1, each
// lane keeping the half of the 32 bit positions its own lane bit selects) lane
// l holds the sum over its 32-lane half of bit position l & 31; the two halves
// are added as integers.  Returns that count.
template <int NP>
__device__ __forceinline__ int bs_reduce_scatter(uint32_t (&A)[NP])
{
    const int lane = (int)__lane_id();
#pragma unroll
    for (int m = 16; m >= 1; m >>= 1) {
        // bit positions j with (j & m) != 0
        const uint32_t hi = m == 16 ? 0xFFFF0000u : m == 8 ? 0xFF00FF00u : m == 4 ? 0xF0F0F0F0u : m == 2 ? 0xCCCCCCCCu : 0xAAAAAAAAu;
        const uint32_t keep = (lane & m) ? hi : ~hi;
        uint32_t Y[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            Y[p] = (uint32_t)__shfl_xor((int)A[p], m, 64) & keep;
            A[p] &= keep;
        }
        bs_add<NP>(A, Y);
    }
    const int j = lane & 31;
    int c = 0;
#pragma unroll
    for (int p = 0; p < NP; ++p) c += (int)((A[p] >> j) & 1u) << p;
    return c + __shfl_xor(c, 32, 64);
}

// ---------------------------------------------------------------------------
// Graph rows are visited in chunks of NBCH neighbours: the chunk's column
// indices are loaded together, then its spins, so a vertex costs two dependent
// memory round trips per chunk instead of two per neighbour.  Slots past the
// end of the row repeat the row's last slot, so every load of a chunk is
// unconditional (the compiler then issues them back to back); callers mask
// those slots out of their sums.
constexpr int NBCH = 8;
__device__ __forceinline__ void chunk_cols(const int32_t* __restrict__ ci, int k0, int b, int (&nb)[NBCH])
{
#pragma unroll
    for (int t = 0; t < NBCH; ++t) nb[t] = ci[min(k0 + t, b - 1)];
}

// ---------------------------------------------------------------------------
// Wolff cluster membership on int8 spins.  A cluster claims a site by changing
// its spin byte with one atomic on the containing 32-bit word (spin buffers are
// 256-byte aligned, so the word never leaves the buffer); the returned old
// byte says whether this lane made the change.  During one cluster a byte only
// ever goes from the cluster's old value to its new one, so "still equal to
// the old value" is exactly "not yet in the cluster" and no visited bitmap is
// needed.
__device__ __forceinline__ uint32_t* spin_word(int8_t* p, uint32_t& sh)
{
    // offset arithmetic on p itself (not an integer round trip), so the compiler
    // keeps p's global address space and emits global rather than flat accesses
    const uint32_t o = (uint32_t)(uintptr_t)p & 3u;
    sh = o * 8u;
    return reinterpret_cast<uint32_t*>(p - o);
}

// Ising bytes are +1 = 0x01 and -1 = 0xFF: OR 0xFE takes +1 to -1 and AND 0x01
// takes -1 to +1, each leaving the other value unchanged.  val (the cluster's
// old value) must be the same in every lane of the wavefront: the direction is
// then a scalar branch rather than two masked atomics in a row.
__device__ __forceinline__ bool claim_ising(int8_t* p, int val)
{
    uint32_t sh;
    uint32_t* w = spin_word(p, sh);
    const uint32_t m = 0xFEu << sh;
    const uint32_t old = __builtin_amdgcn_readfirstlane(val) > 0 ? atomicOr(w, m) : atomicAnd(w, ~m);
    return (int8_t)(old >> sh) == (int8_t)val;
}

// Current word holding the byte at p (relaxed atomic load).
__device__ __forceinline__ uint32_t peek_word(int8_t* p, uint32_t& sh)
{
    uint32_t* w = spin_word(p, sh);
    return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reserves n (< 16) consecutive entries for each active lane of the wavefront
// with a single LDS atomic on *tail, placing the lanes in lane order; returns
// this lane's first entry.  Every active lane of the wavefront must call it.
__device__ __forceinline__ int wave_reserve(int* tail, uint32_t n)
{
    const uint64_t lt = __lanemask_lt();
    int pre = 0, tot = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const uint64_t m = __ballot((int)((n >> b) & 1u));
        pre += __popcll(m & lt) << b;
        tot += __popcll(m) << b;
    }
    const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
    int base = 0;
    if ((int)__lane_id() == leader && tot > 0) base = atomicAdd(tail, tot);
    base = __shfl(base, leader);
    return base + pre;
}

// Visited set of a growing O(3)/Phi4 cluster in LDS: open addressing with
// linear probing over MQ_VSET site indices (empty = 0xFFFFFFFF), Fibonacci
// hashing of the index.  Callers keep at most MQ_VSET/2 members, so a free slot
// always exists and probe runs stay short.  Returns true when j was not a
// member and is one now; concurrent inserts of one j let exactly one lane win.
constexpr int MQ_VSET = 8192;
constexpr int MQ_VSET_LG = 13;
__device__ __forceinline__ bool vset_insert(uint32_t* set, uint32_t j)
{
    uint32_t h = (j * 2654435769u) >> (32 - MQ_VSET_LG);
    while (true) {
        const uint32_t old = atomicCAS(&set[h], 0xFFFFFFFFu, j);
        if (old == 0xFFFFFFFFu) return true;
        if (old == j) return false;
        h = (h + 1u) & (uint32_t)(MQ_VSET - 1);
    }
}

// Potts relabel val -> newq by compare-and-swap, starting from a word seen
// earlier; returns false once the byte no longer holds val.  Terminates: every
// failed swap means the word changed, and each byte changes at most once per
// cluster.
__device__ __forceinline__ bool claim_potts(int8_t* p, uint32_t seen, int val, int newq)
{
    uint32_t sh;
    uint32_t* w = spin_word(p, sh);
    while ((int8_t)(seen >> sh) == (int8_t)val) {
        const uint32_t nw = (seen & ~(0xFFu << sh)) | ((uint32_t)(uint8_t)newq << sh);
        const uint32_t prev = atomicCAS(w, seen, nw);
        if (prev == seen) return true;
        seen = prev;
    }
    return false;
}

// ---------------------------------------------------------------------------
struct mq_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    int num_cus = 256;
    int max_grid_y = 65535;  // workgroups along y, the replica dimension of most launches
    void* scratch = nullptr; // device address of this device's static scratch array (api.hip)
};

struct mq_lattice {
    mq_ctx* ctx = nullptr;
    int kind = LAT_REGULAR;
    int dim = 0, L = 0;
    int64_t N = 0;
    // CSR graphs: vertices renumbered so that each colour class is contiguous
    int ncolors = 2;
    uint32_t color_stride = 2;       // phase = color_stride*sweep + colour (2 on regular lattices)
    std::vector<int64_t> color_off;  // ncolors+1 (positions in the new numbering)
    std::vector<int32_t> pos_of;     // original vertex -> new position
    std::vector<int32_t> vert_of;    // new position -> original vertex
    std::vector<int32_t> h_rp, h_ci; // original numbering (host copy)
    int32_t* d_rp = nullptr;         // new numbering
    int32_t* d_ci = nullptr;
    int32_t* d_edge = nullptr;       // original slot id of each slot (keys the Wolff test uniforms)
    int32_t* d_pos_of = nullptr;     // original vertex -> position
    int32_t* d_vert_of = nullptr;    // position -> original vertex
    void* d_ell = nullptr;           // padded row table of the int8 graph Wolff (built on first use)
    std::vector<int32_t> colour;     // colour of each original vertex
    std::vector<int32_t> h_slot_of_orig;  // original CSR slot -> device slot
    std::vector<int32_t> h_mirror;        // original slot (u->v) -> original slot (v->u)
    int maxdeg = 0;
    int64_t nnz = 0;
};

// A run of consecutive replicas on the same lattice (uniform stride).
struct Group {
    mq_lattice* lat;
    int rep0, nrep;
    int64_t off;         // byte offset of replica rep0 in d_spins
    int64_t stride;      // bytes between consecutive replicas
    int64_t pos_stride;  // bytes between consecutive sites of one replica
};

struct mq_state {
    mq_ctx* ctx = nullptr;
    mq_model_params p{};
    int nrep = 0;
    std::vector<mq_lattice*> lat;
    std::vector<Group> groups;
    int site_bytes = 1;  // 1 (int8) or 16 (float4)
    std::vector<int64_t> rep_off;        // byte offset of site 0 of each replica
    std::vector<int64_t> rep_pos_stride; // bytes between consecutive sites of each replica
    char* d_spins = nullptr;
    int64_t spin_total = 0;
    // per-replica temperatures and acceptance tables
    std::vector<double> beta;
    int tbl = 0;                 // entries per replica
    uint32_t* d_thr = nullptr;   // [nrep][tbl] pairs (hi, lo)
    uint32_t* d_wthr = nullptr;  // [nrep] Wolff add thresholds (integer models)
    float* d_betaf = nullptr;    // [nrep]
    double* d_betad = nullptr;   // [nrep]
    // disorder (EA)
    std::vector<int32_t> rep_sample;
    int nsamples = 0;
    int8_t* d_bonds = nullptr;
    std::vector<int64_t> sample_bond_off;
    int64_t* d_rep_bond_off = nullptr;  // [nrep] byte offset of the replica's sample
    // Wolff scratch
    int32_t* d_queue = nullptr;
    uint32_t* d_visit = nullptr;
    int64_t* d_csize = nullptr;
    size_t csize_cap = 0;
    std::vector<int64_t> rep_site_off;  // per replica site offset into queue/visit scratch
    int64_t* d_rep_site_off = nullptr;
    int64_t total_sites = 0;
    // measurement
    long long* d_isum = nullptr;  // [nrep][8]
    double* d_fsum = nullptr;     // [nrep][8]
    double* d_ft = nullptr;       // [nrep][3 axes][3 components][re, im] MagFTComp sums
    // parallel tempering
    int32_t* d_temp_of_rep = nullptr;  // [nrep]
    int32_t* d_rep_of_temp = nullptr;  // [nrep] (grouped by ntemp)
    int32_t* d_pt_acc = nullptr;
    int32_t* d_pt_pairs = nullptr;  // slot pairs of the last mq_pt_swap_pairs plan
    int pt_pairs_cap = 0;
    uint64_t* d_pack = nullptr;        // [nrep][17] measurement staging: isum[8], fsum[8] and beta as raw words
    uint64_t* h_pack = nullptr;        // pinned host copy of d_pack
    std::vector<double> beta_ladder;   // per replica index at creation (temperature k)
    double* d_ladder = nullptr;
    int pt_ntemp = 0;
    bool generic_only = false;
    const uint32_t* cur_dyn = nullptr;  // device phase/step offsets while a graph is captured
    uint32_t* d_dyn = nullptr;          // [0] sweep-phase offset, [1] PT-step offset
    void* d_lds_groups = nullptr;       // descriptor table of the LDS multi-sweep kernel
    bool lds_ok = false;
    int lds_ngroups = 0, lds_blocks = 0;
    size_t lds_need[2] = {0, 0};        // LDS bytes per block without / with couplings
    bool msc = false;                   // 1-bit multi-spin-coded Ising storage
    bool eabits = false;                // EA: 32 replicas of one sample per uint32 word
    bool csrbits = false;               // Ising on graphs: 32 replicas of one graph per uint32 word
    void* d_wtab = nullptr;             // per word group: per-bit threshold words
    void* d_bits_groups = nullptr;      // lattice-group table of the fused bit-sliced sweep
    int bits_ngroups = 0, bits_blocks = 0;
    int bits_mwaves = 0;                // wavefronts of the fused bit-sliced measurement (table entries ngroups..2*ngroups-1)
    void* d_msc_cls = nullptr;          // per-replica Bernoulli classes of the MSC sweep
    int tune_block = 256;               // threads per block of the streaming sweep kernels
    int tune_xcd = 1;                   // XCD-aware block remap of the row-vectorised kernel
    int lds_block = 256;                // threads per block of the LDS multi-sweep kernel
    int kernel_choice = 0;              // 0 auto, 1 streaming kernels only, 2 generic streaming kernel only  // force the generic kernels (parity of the specialised ones)
    int wolff_ring = 512;               // LDS frontier ring entries of the small-lattice Wolff kernel (power of two, 64..2048)
    int wolff_block = 0;                // threads per replica of the global-memory Wolff kernels (0 = automatic)
    int wolff_vset = MQ_VSET / 2;       // Wolff on large lattices: clusters up to this size keep their membership in LDS (0..MQ_VSET/2)
    int step_graph = 0;                 // mq_run_graph: 0 automatic (stream), 1 HIP graph, 2 stream
};

// --------------------------------------------------------------------------
// error helpers
int mq_fail(mq_ctx* ctx, int code, const std::string& msg);
#define MQ_HIP(ctx, expr)                                                                          \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return mq_fail((ctx), MQ_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// --------------------------------------------------------------------------
// kernel launchers (each returns hipError_t of the launch)
hipError_t launch_init_regular(mq_state* st, const Group& g, Key2 key);
hipError_t launch_sweep_regular(mq_state* st, const Group& g, int colour, uint32_t phase, Key2 key);
hipError_t launch_measure_regular(mq_state* st, const Group& g, long long* isum, double* fsum);
hipError_t launch_measure_ft_regular(mq_state* st, const Group& g);
hipError_t launch_wolff_regular(mq_state* st, const Group& g, int nclusters, uint32_t cl0, Key2 key);
hipError_t launch_emcs_regular(mq_state* st, const Group& g, int nsteps, int ncluster, int nmetro, uint64_t step0,
                               Key2 key);
hipError_t launch_bonds_regular(mq_state* st, const Group& g, int sample, double p_plus, Key2 key);
hipError_t launch_sweeps_lds(mq_state* st, int nsweeps, uint32_t sweep0, Key2 key);
hipError_t prepare_lds(mq_state* st);
hipError_t launch_wtab_eabits(mq_state* st);
size_t wtab_bytes();
hipError_t launch_init_eabits(mq_state* st, const Group& g, Key2 key);
hipError_t prepare_eabits(mq_state* st);
hipError_t launch_sweep_eabits_all(mq_state* st, int colour, uint32_t phase, Key2 key);
hipError_t launch_measure_eabits_all(mq_state* st, long long* isum);
bool msc_supported(const mq_lattice* lat);
hipError_t msc_upload_classes(mq_state* st, const std::vector<uint32_t>& thr);
hipError_t launch_sweep_msc(mq_state* st, const Group& g, int colour, uint32_t phase, Key2 key);
hipError_t launch_init_msc(mq_state* st, const Group& g, Key2 key);
hipError_t launch_measure_msc(mq_state* st, const Group& g, long long* isum);

hipError_t launch_init_csr(mq_state* st, const Group& g, Key2 key);
hipError_t launch_sweep_csr(mq_state* st, const Group& g, int colour, uint32_t phase, Key2 key);
hipError_t launch_measure_csr(mq_state* st, const Group& g, long long* isum, double* fsum);
hipError_t launch_wolff_csr(mq_state* st, const Group& g, int nclusters, uint32_t cl0, Key2 key);
bool csrbits_supported(const mq_lattice* lat);
hipError_t launch_init_csrbits(mq_state* st, const Group& g, Key2 key);
hipError_t launch_sweep_csrbits(mq_state* st, const Group& g, int colour, uint32_t phase, Key2 key);
hipError_t launch_measure_csrbits(mq_state* st, const Group& g, long long* isum);
// measurement sums of every replica into isum / fsum ([nrep][8] each, zeroed by
// the caller), one launch per lattice group or one fused launch for bit-sliced EA
hipError_t launch_measure_all(mq_state* st, long long* isum, double* fsum);

// Threads of the workgroup that grows one replica's clusters in global memory.
// A wide frontier level is worked through blockDim entries at a time, each
// waiting on its claims' round trips, so the widest block keeps the most claims
// in flight while there is at most one replica per CU; beyond that, two
// 512-thread blocks per CU do better.  Measured on the same clusters (3D,
// sites/ns for 256 / 512 / 1024 threads): Heisenberg L=128 x 256 replicas
// 0.57 / 0.74 / 0.86, L=64 x 512 1.09 / 1.33 / 1.27, L=64 x 1024
// 1.73 / 2.12 / 1.96; Ising L=128 x 256 1.72 / 2.47 / 2.77, L=64 x 1024
// 6.18 / 7.01 / 6.39.
inline int wolff_block(const mq_state* st, int nwg)
{
    if (st->wolff_block > 0) return st->wolff_block;
    return nwg <= st->ctx->num_cus ? 1024 : 512;
}

// Sweep numbers a state accepts: the Philox phase of a sweep is one 32-bit
// counter word, color_stride*sweep + colour (2 on regular lattices, 64 or more
// on graphs), so sweeps must stay below 2^32 / stride: 2^31 on regular lattices,
// 2^26 on graphs of up to 64 colour classes.
inline uint32_t max_color_stride(const mq_state* st)
{
    uint32_t m = 2;
    for (const auto* l : st->lat) m = l->color_stride > m ? l->color_stride : m;
    return m;
}
inline uint64_t sweep_limit(const mq_state* st) { return (1ull << 32) / max_color_stride(st); }
inline std::string sweep_limit_msg(const mq_state* st)
{
    const uint32_t s = max_color_stride(st);
    int lg = 0;
    while ((1ull << lg) < sweep_limit(st)) ++lg;
    return "sweep counter exceeds 2^" + std::to_string(lg) + " (phase = " + std::to_string(s) + "*sweep + colour)";
}

hipError_t launch_energy_all(mq_state* st);
hipError_t launch_pt_swap(mq_state* st, int ntemp, int parity, uint32_t step, Key2 key);
hipError_t launch_renorm(mq_state* st);

// host-side layout helpers
inline int regular_colour(int dim, int L, int64_t i)
{
    int64_t s = 0;
    for (int d = 0; d < dim; ++d) {
        s += i % L;
        i /= L;
    }
    return (int)(s & 1);
}
