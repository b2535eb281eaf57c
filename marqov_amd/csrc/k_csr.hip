// General graphs in CSR form: greedy colouring on the host, colour-class
// Metropolis sweeps, frontier-BFS Wolff and FP64 observables on the device.
//
// Layout: vertices are renumbered so that every colour class is contiguous
// (position = class offset + breadth-first rank within the class); the device
// CSR holds positions.  The replicas of one graph are interleaved: the spin of replica r
// at position p sits at p * R + r (R replicas), so a wavefront sweeping one
// vertex for 64 replicas makes 64-byte coalesced neighbour reads while the CSR
// row and column indices are wave-uniform loads shared by all replicas.
// Kernels take (replica stride, position stride) in elements.  Random numbers are
// keyed by original vertex ids and original CSR slots, so a CPU replay needs
// only the colour of every vertex, not the renumbering.
//
//   Metropolis: Ising/EA draw one Philox block per vertex and 8 replicas:
//     counter (vertex, 64*sweep + colour, replica >> 3, METRO), 16-bit uniform
//     (replica & 7) against the high half of the 32-bit threshold of
//     dE = -2 J q (q = s_i sum b_ij s_j), the low half from the TIE stream on a
//     tie; Potts proposes (w * Q) >> 32 with word (replica & 3) of counter
//     (vertex, phase, replica >> 2, PROPOSE) and accepts with the Ising rule's
//     16-bit halves; O(3)/Phi4 use counter (vertex, phase, replica, METRO),
//     words 0,1 for the direction and word 2 for acceptance.
//   Wolff tests: the test from u to its neighbour v is keyed by the original CSR
//     slot of v in row u, counter (slot >> 2, cluster, replica, WBOND), word
//     slot & 3: every test has its own uniform, as the reference draws a fresh
//     random number per test.
#include <algorithm>
#include <numeric>

#include "mq_internal.hpp"

namespace {

__device__ __forceinline__ float3 rnddir_c(uint32_t w0, uint32_t w1)
{
    const float a0 = 2.0f * u01f(w0);
    const float a1 = 2.0f * u01f(w1);
    const float rho = sqrtf(fmaxf(a0 * (2.0f - a0), 0.0f));
    float sn, cs;
    sincospif(a1, &sn, &cs);
    return make_float3(1.0f - a0, rho * sn, rho * cs);
}

__global__ void __launch_bounds__(256) k_csr_init(char* __restrict__ base, int64_t rs, int64_t ps,
                                                  const int32_t* __restrict__ vert_of, int64_t N, int model, int Q,
                                                  int rep0, Key2 key)
{
    const int r = blockIdx.y;
    const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= N) return;
    const uint32_t v = (uint32_t)vert_of[pos];
    if (model == MQ_HEISENBERG || model == MQ_PHI4) {
        const uint4 w = philox4x32_10(v, 0u, (uint32_t)(rep0 + r), ST_INIT, key);
        const float3 d = rnddir_c(w.x, w.y);
        reinterpret_cast<float4*>(base)[r * rs + pos * ps] = make_float4(d.x, d.y, d.z, 0.f);
    } else {
        const uint4 w = philox4x32_10(v >> 2, 0u, (uint32_t)(rep0 + r), ST_INIT, key);
        const uint32_t x = u4_get(w, (int)(v & 3u));
        int8_t s;
        if (model == MQ_POTTS) s = (int8_t)(((uint64_t)x * (uint64_t)Q) >> 32);
        else s = (x >> 31) ? (int8_t)1 : (int8_t)-1;
        reinterpret_cast<int8_t*>(base)[r * rs + pos * ps] = s;
    }
}

__device__ __forceinline__ void gather_i8(const int8_t* s, int64_t ps, const int32_t* __restrict__ ci, int k0, int b,
                                          int (&v)[NBCH])
{
    int nb[NBCH];
    chunk_cols(ci, k0, b, nb);
#pragma unroll
    for (int t = 0; t < NBCH; ++t) v[t] = (int)s[(int64_t)nb[t] * ps];
}

// one thread per (vertex of the colour class, replica), replica fastest
template <bool BONDS>
__global__ void __launch_bounds__(512) k_csr_sweep_int8(int8_t* __restrict__ base, int64_t rs, int64_t ps, int R,
                                                        const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                        const int32_t* __restrict__ vert_of,
                                                        const int8_t* __restrict__ bonds,
                                                        const int64_t* __restrict__ rep_bond_off,
                                                        const uint32_t* __restrict__ thr, int tbl, int64_t c0,
                                                        int64_t nc, int model, int Q, uint32_t phase,
                                                        const uint32_t* __restrict__ dyn, int rep0, Key2 key)
{
    if (dyn) phase += dyn[0];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc * R) return;
    const int64_t vl = t / R;
    const int r = (int)(t - vl * R);
    const int grep = rep0 + r;
    const int64_t pos = c0 + vl;
    int8_t* s = base + r * rs;
    const int qmax = (tbl - 1) >> 1;
    const int a = rp[pos], b = rp[pos + 1];
    const uint32_t* T = thr + (int64_t)grep * 2 * tbl;
    const int si = s[pos * ps];
    if (model == MQ_POTTS) {
        const uint32_t v = (uint32_t)vert_of[pos];
        const uint4 pw = philox4x32_10(v, phase, (uint32_t)grep >> 2, ST_PROPOSE, key);
        const int snew = (int)(((uint64_t)u4_get(pw, grep & 3) * (uint64_t)Q) >> 32);
        int dn = 0;
        for (int k0 = a; k0 < b; k0 += NBCH) {
            int v[NBCH];
            gather_i8(s, ps, ci, k0, b, v);
#pragma unroll
            for (int t = 0; t < NBCH; ++t) dn += (k0 + t < b) ? (v[t] == snew) - (v[t] == si) : 0;
        }
        const int e = dn + qmax;
        const uint4 w8 = philox4x32_10(v, phase, (uint32_t)grep >> 3, ST_METRO, key);
        const int m = grep & 7;
        const uint32_t hi16 = (u4_get(w8, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
        bool acc = hi16 < T[2 * e];
        if (hi16 == T[2 * e]) {
            const uint4 tw = philox4x32_10(v, phase, (uint32_t)grep >> 3, ST_TIE, key);
            acc = ((u4_get(tw, m >> 1) >> (16 * (m & 1))) & 0xFFFFu) < T[2 * e + 1];
        }
        if (acc) s[pos * ps] = (int8_t)snew;
    } else {
        const uint32_t v = (uint32_t)vert_of[pos];
        const uint4 w8 = philox4x32_10(v, phase, (uint32_t)grep >> 3, ST_METRO, key);  // independent of the gathers
        int h = 0;
        const int8_t* bb = BONDS ? bonds + rep_bond_off[grep] : nullptr;
        for (int k0 = a; k0 < b; k0 += NBCH) {
            int v[NBCH];
            gather_i8(s, ps, ci, k0, b, v);
#pragma unroll
            for (int t = 0; t < NBCH; ++t) h += (k0 + t < b) ? (BONDS ? (int)bb[min(k0 + t, b - 1)] : 1) * v[t] : 0;
        }
        const int e = si * h + qmax;
        const int m = grep & 7;
        const uint32_t hi16 = (u4_get(w8, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
        bool acc = hi16 < T[2 * e];
        if (hi16 == T[2 * e]) {
            const uint4 tw = philox4x32_10(v, phase, (uint32_t)grep >> 3, ST_TIE, key);
            acc = ((u4_get(tw, m >> 1) >> (16 * (m & 1))) & 0xFFFFu) < T[2 * e + 1];
        }
        if (acc) s[pos * ps] = (int8_t)(-si);
    }
}

template <bool BONDS>
__global__ void __launch_bounds__(512) k_csr_sweep_o3(float4* __restrict__ base, int64_t rs, int64_t ps, int R,
                                                      const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                      const int32_t* __restrict__ vert_of,
                                                      const int8_t* __restrict__ bonds,
                                                      const int64_t* __restrict__ rep_bond_off,
                                                      const float* __restrict__ betaf, int64_t c0, int64_t nc,
                                                      int model, int proposal, float J, float delta, float lambda,
                                                      float mass, uint32_t phase, const uint32_t* __restrict__ dyn,
                                                      int rep0, Key2 key)
{
    if (dyn) phase += dyn[0];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc * R) return;
    const int64_t vl = t / R;
    const int r = (int)(t - vl * R);
    const int grep = rep0 + r;
    const int64_t pos = c0 + vl;
    float4* s = base + r * rs;
    const int a = rp[pos], b = rp[pos + 1];
    float nx = 0.f, ny = 0.f, nz = 0.f;
    const int8_t* bb = BONDS ? bonds + rep_bond_off[grep] : nullptr;
    for (int k0 = a; k0 < b; k0 += NBCH) {
        int nb[NBCH];
        float4 v[NBCH];
        chunk_cols(ci, k0, b, nb);
#pragma unroll
        for (int t = 0; t < NBCH; ++t) v[t] = s[(int64_t)nb[t] * ps];
#pragma unroll
        for (int t = 0; t < NBCH; ++t) {
            const float f = (k0 + t < b) ? (BONDS ? (float)bb[min(k0 + t, b - 1)] : 1.0f) : 0.0f;
            nx += f * v[t].x;
            ny += f * v[t].y;
            nz += f * v[t].z;
        }
    }
    const float4 so = s[pos * ps];
    const uint4 w = philox4x32_10((uint32_t)vert_of[pos], phase, (uint32_t)grep, ST_METRO, key);
    const float3 rd = rnddir_c(w.x, w.y);
    float3 sn;
    if (model == MQ_PHI4) {
        sn = make_float3(so.x + 0.5f * rd.x, so.y + 0.5f * rd.y, so.z + 0.5f * rd.z);
    } else if (proposal == MQ_PROPOSE_SMALL_ANGLE) {
        const float3 q = make_float3(so.x + delta * rd.x, so.y + delta * rd.y, so.z + delta * rd.z);
        const float inv = rsqrtf(q.x * q.x + q.y * q.y + q.z * q.z);
        sn = make_float3(q.x * inv, q.y * inv, q.z * inv);
    } else {
        sn = rd;
    }
    const float beta = betaf[grep];
    float dE = J * ((sn.x - so.x) * nx + (sn.y - so.y) * ny + (sn.z - so.z) * nz);
    if (model == MQ_PHI4) {
        const float n2 = sn.x * sn.x + sn.y * sn.y + sn.z * sn.z;
        const float o2 = so.x * so.x + so.y * so.y + so.z * so.z;
        dE += (mass / beta) * (n2 - o2) + (lambda / beta) * ((n2 - 1.f) * (n2 - 1.f) - (o2 - 1.f) * (o2 - 1.f));
    }
    bool acc = dE <= 0.f;
    if (!acc) acc = u01f(w.z) < __expf(-beta * dE);
    if (acc) s[pos * ps] = make_float4(sn.x, sn.y, sn.z, 0.f);
}


// Ising / EA on a graph, 8 replicas per lane (R % 8 == 0, rep0 % 8 == 0): the
// lane reads its 8 replicas' spins of each neighbour as one 8-byte word, counts
// down neighbours byte-wise ((word ^ bond sign) & 0x02 per byte, doubled count
// <= 2*maxdeg <= 254), forms the table index e = q + qmax per byte, and accepts
// with the 8 x 16-bit uniforms of one Philox block (tie -> TIE stream).  The
// thresholds are read through the caches (a block sweeps only 256/(R/8)
// vertices, too few to amortise staging all R tables in LDS).  Same decisions
// as the per-replica kernel above.
template <bool BONDS>
__global__ void __launch_bounds__(256) k_csr_sweep_ising8(int8_t* __restrict__ base, int R,
                                                          const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                          const int32_t* __restrict__ vert_of,
                                                          const int8_t* __restrict__ bonds,
                                                          const int64_t* __restrict__ rep_bond_off,
                                                          const uint32_t* __restrict__ thr, int tbl, int64_t c0,
                                                          int64_t nc, uint32_t phase, const uint32_t* __restrict__ dyn,
                                                          int rep0, Key2 key)
{
    if (dyn) phase += dyn[0];
    const int oct = R >> 3;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc * oct) return;
    const int64_t vl = t / oct;
    const int o = (int)(t - vl * oct);
    const int64_t pos = c0 + vl;
    const int qmax = (tbl - 1) >> 1;
    const int grep0 = rep0 + 8 * o;
    // own spins, vertex id and the Philox block do not depend on the gathers:
    // issue them first so their latency overlaps the neighbour loads
    uint2* ownp = reinterpret_cast<uint2*>(base + pos * R + 8 * o);
    const uint2 S = *ownp;
    const uint32_t vid = (uint32_t)vert_of[pos];
    const uint4 w = philox4x32_10(vid, phase, (uint32_t)grep0 >> 3, ST_METRO, key);
    const int a = rp[pos], b = rp[pos + 1];
    const int deg = b - a;
    const int8_t* bb = BONDS ? bonds + rep_bond_off[grep0] : nullptr;
    uint32_t n0 = 0u, n1 = 0u;
    for (int k0 = a; k0 < b; k0 += NBCH) {
        int nb[NBCH];
        uint2 v[NBCH];
        chunk_cols(ci, k0, b, nb);
#pragma unroll
        for (int t = 0; t < NBCH; ++t) v[t] = *reinterpret_cast<const uint2*>(base + (int64_t)nb[t] * R + 8 * o);
#pragma unroll
        for (int t = 0; t < NBCH; ++t) {
            const uint32_t in = (k0 + t < b) ? 0x02020202u : 0u;
            const uint32_t f = (BONDS && bb[min(k0 + t, b - 1)] < 0) ? 0xFEFEFEFEu : 0u;  // -x = x ^ 0xFE for x = +-1
            n0 += (v[t].x ^ f) & in;
            n1 += (v[t].y ^ f) & in;
        }
    }
    const uint32_t bias = (uint32_t)(qmax - deg) * 0x01010101u;
    const uint32_t two_deg = (uint32_t)(2 * deg) * 0x01010101u;
    auto eword = [&](uint32_t sw, uint32_t n2) {
        const uint32_t mask = ((sw & 0x02020202u) >> 1) * 0xFFu;  // own spin down
        return ((n2 & mask) | ((two_deg - n2) & ~mask)) + bias;     // 2*aligned - deg + qmax
    };
    const uint32_t E0 = eword(S.x, n0), E1 = eword(S.y, n1);
    const uint32_t* T = thr + (int64_t)grep0 * tbl * 2;  // the lane's 8 replicas' (hi, lo) rows
    uint32_t thv[8];  // all 8 table reads in flight together
#pragma unroll
    for (int m = 0; m < 8; ++m) thv[m] = T[((int64_t)m * tbl + (((m < 4 ? E0 : E1) >> (8 * (m & 3))) & 0xFFu)) * 2];
    uint32_t acc = 0u;
    bool tie = false;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const uint32_t u = (u4_get(w, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
        const uint32_t tt = thv[m] > 0xFFFFu ? 0xFFFFu : thv[m];  // "always" (65536) as 0xFFFF + lo 0x10000
        acc |= (u < tt ? 1u : 0u) << m;
        tie |= (u == tt);
    }
    if (tie) {
        const uint4 tw = philox4x32_10(vid, phase, (uint32_t)grep0 >> 3, ST_TIE, key);
        acc = 0u;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const uint32_t e = ((m < 4 ? E0 : E1) >> (8 * (m & 3))) & 0xFFu;
            const uint32_t u = (u4_get(w, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
            const uint32_t lo = (u4_get(tw, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
            const uint32_t th = T[((int64_t)m * tbl + e) * 2];
            const uint32_t tt = th > 0xFFFFu ? 0xFFFFu : th;
            const uint32_t tl = th > 0xFFFFu ? 0x10000u : T[((int64_t)m * tbl + e) * 2 + 1];
            acc |= ((u < tt) || (u == tt && lo < tl) ? 1u : 0u) << m;
        }
    }
    auto expand = [](uint32_t nib) { return ((nib * 0x00204081u) & 0x01010101u) * 0xFEu; };
    *ownp = make_uint2(S.x ^ expand(acc & 15u), S.y ^ expand(acc >> 4));
}

// q-state Potts on a graph, 8 replicas per lane (R % 8 == 0, rep0 % 8 == 0,
// maxdeg <= 127): the lane proposes 8 new colours from two PROPOSE blocks,
// counts per byte the neighbours equal to the new colour (cn) and to the old
// one (co) with a byte-wise zero test of (word ^ colours) (colours < 128, so
// x + 0x7F has its byte's top bit clear exactly when the byte is 0), forms
// e = cn + (qmax - co) per byte and accepts with the 16-bit halves of one
// METRO block (tie -> TIE stream).  Same decisions as k_csr_sweep_int8.
__global__ void __launch_bounds__(256) k_csr_sweep_potts8(int8_t* __restrict__ base, int R,
                                                          const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                          const int32_t* __restrict__ vert_of,
                                                          const uint32_t* __restrict__ thr, int tbl, int64_t c0,
                                                          int64_t nc, int Q, uint32_t phase,
                                                          const uint32_t* __restrict__ dyn, int rep0, Key2 key)
{
    if (dyn) phase += dyn[0];
    const int oct = R >> 3;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc * oct) return;
    const int64_t vl = t / oct;
    const int o = (int)(t - vl * oct);
    const int64_t pos = c0 + vl;
    const int qmax = (tbl - 1) >> 1;
    const int grep0 = rep0 + 8 * o;
    const uint32_t v = (uint32_t)vert_of[pos];
    // proposals of the lane's 8 replicas (bytes, replica grep0 + m in byte m)
    uint32_t SN[2];
#pragma unroll
    for (int hblk = 0; hblk < 2; ++hblk) {
        const uint4 pw = philox4x32_10(v, phase, (uint32_t)(grep0 >> 2) + (uint32_t)hblk, ST_PROPOSE, key);
        uint32_t word = 0u;
#pragma unroll
        for (int k = 0; k < 4; ++k) word |= (uint32_t)(((uint64_t)u4_get(pw, k) * (uint64_t)Q) >> 32) << (8 * k);
        SN[hblk] = word;
    }
    uint2* ownp = reinterpret_cast<uint2*>(base + pos * R + 8 * o);
    const uint2 S = *ownp;
    const int a = rp[pos], b = rp[pos + 1];
    const uint32_t K7 = 0x7F7F7F7Fu, H8 = 0x80808080u;
    auto eqb = [&](uint32_t x) { return (~(x + K7) & H8) >> 7; };  // 0x01 per zero byte (bytes < 0x80)
    uint32_t cn0 = 0u, cn1 = 0u, co0 = 0u, co1 = 0u;
    for (int k0 = a; k0 < b; k0 += NBCH) {
        int nb[NBCH];
        uint2 w[NBCH];
        chunk_cols(ci, k0, b, nb);
#pragma unroll
        for (int tt = 0; tt < NBCH; ++tt) w[tt] = *reinterpret_cast<const uint2*>(base + (int64_t)nb[tt] * R + 8 * o);
#pragma unroll
        for (int tt = 0; tt < NBCH; ++tt) {
            const uint32_t in = (k0 + tt < b) ? 0xFFFFFFFFu : 0u;
            cn0 += eqb(w[tt].x ^ SN[0]) & in;
            cn1 += eqb(w[tt].y ^ SN[1]) & in;
            co0 += eqb(w[tt].x ^ S.x) & in;
            co1 += eqb(w[tt].y ^ S.y) & in;
        }
    }
    const uint32_t qb = (uint32_t)qmax * 0x01010101u;  // co <= deg <= qmax per byte: no borrow
    const uint32_t E0 = cn0 + (qb - co0), E1 = cn1 + (qb - co1);
    const uint4 wm = philox4x32_10(v, phase, (uint32_t)grep0 >> 3, ST_METRO, key);
    const uint32_t* T = thr + (int64_t)grep0 * tbl * 2;
    uint32_t thv[8];  // all 8 table reads in flight together
#pragma unroll
    for (int m = 0; m < 8; ++m) thv[m] = T[((int64_t)m * tbl + (((m < 4 ? E0 : E1) >> (8 * (m & 3))) & 0xFFu)) * 2];
    uint32_t acc = 0u;
    bool tie = false;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const uint32_t u = (u4_get(wm, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
        const uint32_t tt = thv[m] > 0xFFFFu ? 0xFFFFu : thv[m];  // "always" (65536) as 0xFFFF + lo 0x10000
        acc |= (u < tt ? 1u : 0u) << m;
        tie |= (u == tt);
    }
    if (tie) {
        const uint4 tw = philox4x32_10(v, phase, (uint32_t)grep0 >> 3, ST_TIE, key);
        acc = 0u;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const uint32_t e = ((m < 4 ? E0 : E1) >> (8 * (m & 3))) & 0xFFu;
            const uint32_t u = (u4_get(wm, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
            const uint32_t lo = (u4_get(tw, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
            const uint32_t th = T[((int64_t)m * tbl + e) * 2];
            const uint32_t tt = th > 0xFFFFu ? 0xFFFFu : th;
            const uint32_t tl = th > 0xFFFFu ? 0x10000u : T[((int64_t)m * tbl + e) * 2 + 1];
            acc |= ((u < tt) || (u == tt && lo < tl) ? 1u : 0u) << m;
        }
    }
    auto bytemask = [](uint32_t nib) { return ((nib * 0x00204081u) & 0x01010101u) * 0xFFu; };
    const uint32_t M0 = bytemask(acc & 15u), M1 = bytemask(acc >> 4);
    *ownp = make_uint2((S.x & ~M0) | (SN[0] & M0), (S.y & ~M1) | (SN[1] & M1));
}

template <typename T>
__device__ __forceinline__ T wsum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// isum: [0] sum over edges b s_i s_j (Ising/EA) or #equal edges (Potts), [1] sum s,
// Potts [1+q] colour counts; fsum (O(3)): [0] sum over edges b s_i.s_j, [1..3] M,
// [4] sum |s|^2, [5] sum (|s|^2-1)^2.  Each edge counted once (from its lower end).
template <bool BONDS, bool VEC>
__global__ void __launch_bounds__(256) k_csr_measure(const char* __restrict__ base, int64_t rs, int64_t ps,
                                                     const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                     const int8_t* __restrict__ bonds,
                                                     const int64_t* __restrict__ rep_bond_off, int64_t N, int model,
                                                     int Q, int rep0, long long* __restrict__ isum,
                                                     double* __restrict__ fsum)
{
    __shared__ double red[8][4];
    const int r = blockIdx.y;
    const int grep = rep0 + r;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int8_t* bb = BONDS ? bonds + rep_bond_off[grep] : nullptr;
    // grid-stride over positions: a few blocks per replica keep the atomics few
    for (int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < N; pos += (int64_t)gridDim.x * blockDim.x) {
        const int a = rp[pos], b = rp[pos + 1];
        if (VEC) {
            const float4* s = reinterpret_cast<const float4*>(base) + r * rs;
            const float4 v = s[pos * ps];
            double e = 0;
            for (int k = a; k < b; ++k) {
                if (ci[k] <= pos) continue;
                const float4 u = s[(int64_t)ci[k] * ps];
                e += (BONDS ? (double)bb[k] : 1.0) * ((double)v.x * u.x + (double)v.y * u.y + (double)v.z * u.z);
            }
            const double n2 = (double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z;
            acc[0] += e;
            acc[1] += v.x;
            acc[2] += v.y;
            acc[3] += v.z;
            acc[4] += n2;
            acc[5] += (n2 - 1) * (n2 - 1);
        } else {
            const int8_t* s = reinterpret_cast<const int8_t*>(base) + r * rs;
            const int si = s[pos * ps];
            long long e = 0;
            for (int k = a; k < b; ++k) {
                if (ci[k] <= pos) continue;
                const int sj = s[(int64_t)ci[k] * ps];
                if (model == MQ_POTTS) e += (si == sj);
                else e += (long long)(BONDS ? bb[k] : 1) * si * sj;
            }
            acc[0] += (double)e;  // integer-valued doubles: exact below 2^53
            if (model == MQ_POTTS) {
#pragma unroll
                for (int q = 0; q < 7; ++q)  // static indices keep acc in registers
                    if (q < Q) acc[1 + q] += (si == q);
            } else {
                acc[1] += si;
            }
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const double x = wsum(acc[k]);
        if (lane == 0) red[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        double t = 0;
        for (int wv = 0; wv < (int)(blockDim.x >> 6); ++wv) t += red[threadIdx.x][wv];
        if (VEC) atomicAdd(fsum + (int64_t)grep * 8 + threadIdx.x, t);
        else atomicAdd(reinterpret_cast<unsigned long long*>(isum + (int64_t)grep * 8 + threadIdx.x),
                       (unsigned long long)(long long)llrint(t));
    }
}

enum { CK_ISING = 0, CK_POTTS = 1, CK_CART = 2, CK_RDIR = 3 };

template <int KIND>
__device__ __forceinline__ float cproj(const float4& v, int axis, const float3& rd)
{
    return (KIND == CK_CART) ? (axis == 0 ? v.x : (axis == 1 ? v.y : v.z)) : (v.x * rd.x + v.y * rd.y + v.z * rd.z);
}

template <int KIND>
__device__ __forceinline__ float4 creflect(float4 v, int axis, const float3& rd)
{
    if (KIND == CK_CART) {
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

// Frontier ring in LDS as in the regular-lattice kernel: entry pos written in
// the level that starts at `head` sits in ring slot pos mod CAP when
// pos - head < CAP, else in the global queue (CAP = CWCAP for int8 models,
// CWCAP/2 for O(3)/Phi4).
constexpr int CWCAP = 4096;

// O(3)/Phi4 Wolff on a CSR graph, one workgroup per replica, level-synchronous
// BFS (Ising and Potts: k_csr_wolff8 below).  The visited set is kept in LDS
// while the cluster cannot outgrow vlim sites (then in the global bitmap), and
// each frontier entry carries the flipped spin's projection.  A claimed site's
// row start and degree are loaded together with the claim and kept with its
// ring entry, so a level waits on two dependent global round trips (neighbour
// index, claim) instead of three.  Test randoms are keyed by the tested edge's
// original slot in the tested-from vertex's row.
template <int KIND>
__global__ void __launch_bounds__(1024) k_csr_wolff(char* __restrict__ base, int64_t rs, int64_t ps, int64_t N,
                                                   const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                   const int32_t* __restrict__ edge, const int32_t* __restrict__ pos_of,
                                                   int nclusters, uint32_t cl0, int rep0, int Q,
                                                   const uint32_t* __restrict__ wthr, const float* __restrict__ betaf,
                                                   float J, int32_t* __restrict__ queue, uint32_t* __restrict__ visit,
                                                   const int64_t* __restrict__ rep_site_off, int64_t* __restrict__ csize,
                                                   int maxdeg, int vlim, Key2 key)
{
    static_assert(KIND == CK_CART || KIND == CK_RDIR, "int8 models use k_csr_wolff8");
    constexpr int CAP = CWCAP / 2;  // ring entries (40 KB of LDS with the set)
    __shared__ int s_tail;
    // a ring entry is the site's row (start, degree) and, for O(3)/Phi4, its
    // flipped projection; the site index itself is needed only when spilled
    __shared__ int32_t s_kb[CAP];
    __shared__ int32_t s_deg[CAP];
    __shared__ float s_proj[CAP];
    __shared__ uint32_t s_vset[MQ_VSET];
    // the level chunk being worked: entry rows, degree scan, projections; wavefront sums
    __shared__ int32_t s_ckb[1024];
    __shared__ int32_t s_cpre[1024];
    __shared__ float s_cproj[1024];
    __shared__ int s_wsum[16];
    const int r = blockIdx.x;
    const int grep = rep0 + r;
    // element (position p) of this replica: sp4[p * ps]
    float4* sp4 = reinterpret_cast<float4*>(base) + r * rs;
    int32_t* q = queue + rep_site_off[grep];
    uint32_t* vis = visit + (rep_site_off[grep] >> 5);
    const float beta = betaf[grep];
    for (int i = threadIdx.x; i < MQ_VSET; i += blockDim.x) s_vset[i] = 0xFFFFFFFFu;
    for (int c = 0; c < nclusters; ++c) {
        const uint32_t ncl = cl0 + (uint32_t)c;
        const uint4 sw = philox4x32_10(0u, ncl, (uint32_t)grep, ST_WSEED, key);
        const int64_t seed = pos_of[(int64_t)(((uint64_t)sw.x * (uint64_t)N) >> 32)];
        bool glob = vlim == 0;  // membership in the global bitmap (block-uniform)
        int axis = 0;
        float3 rd = make_float3(0.f, 0.f, 0.f);
        if (KIND == CK_CART) {
            axis = (int)(((uint64_t)sw.y * 3ull) >> 32);
        } else {
            rd = rnddir_c(sw.y, sw.z);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            {
                const float4 v = creflect<KIND>(sp4[seed * ps], axis, rd);
                sp4[seed * ps] = v;
                s_proj[0] = cproj<KIND>(v, axis, rd);
                if (glob) atomicOr(&vis[seed >> 5], 1u << (seed & 31));
                else vset_insert(s_vset, (uint32_t)seed);
                q[0] = (int32_t)seed;
            }
            s_kb[0] = rp[seed];
            s_deg[0] = rp[seed + 1] - rp[seed];
            s_tail = 1;
        }
        __syncthreads();
        int head = 0, hprev = 0;
        while (true) {
            const int tail = s_tail;
            __syncthreads();
            if (head >= tail) break;
            if (!glob && (int64_t)tail + (int64_t)(tail - head) * maxdeg > (int64_t)vlim) {
                // this level could take the cluster past vlim sites: move the members
                // (all listed in q) to the global bitmap and claim there from now on
                for (int idx = threadIdx.x; idx < tail; idx += blockDim.x) {
                    const uint32_t m = (uint32_t)q[idx];
                    atomicOr(&vis[m >> 5], 1u << (m & 31));
                }
                glob = true;
                __syncthreads();
            }
            // The level's edges are spread over the whole workgroup: chunks of up to
            // blockDim frontier entries are copied to LDS with a block-wide scan of their
            // degrees, and lane t of the workgroup takes the chunk's edges t, t + blockDim,
            // ...: it finds its entry by a binary search over the scan, loads the far end,
            // draws the edge's Philox uniform and claims.  A level of W entries of mean
            // degree d then costs about W*d/blockDim rounds of two dependent round trips
            // (neighbour index, claim) instead of W*d/64 rounds in one wavefront.  New
            // entries are placed with one LDS atomic per wavefront.
            const int lane = (int)(threadIdx.x & 63u);
            const int wv = (int)(threadIdx.x >> 6);
            const int nwv = (int)(blockDim.x >> 6);
            for (int c0 = head; c0 < tail; c0 += (int)blockDim.x) {
                const int n = min((int)blockDim.x, tail - c0);
                const int idx = c0 + (int)threadIdx.x;
                int kb = 0, deg = 0;
                float pc = 0.f;
                if ((int)threadIdx.x < n) {
                    if (idx - hprev < CAP) {
                        const int sl = idx & (CAP - 1);
                        kb = s_kb[sl];
                        deg = s_deg[sl];
                        pc = s_proj[sl];
                    } else {
                        const int32_t cur = q[idx];
                        kb = rp[cur];
                        deg = rp[cur + 1] - kb;
                        pc = cproj<KIND>(sp4[(int64_t)cur * ps], axis, rd);
                    }
                }
                // block-wide exclusive scan of the degrees
                int inc = deg;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int v = __shfl_up(inc, o);
                    if (lane >= o) inc += v;
                }
                if (lane == 63) s_wsum[wv] = inc;
                __syncthreads();
                int wpre = 0;
                for (int w = 0; w < wv; ++w) wpre += s_wsum[w];
                const int pre = wpre + inc - deg;
                if ((int)threadIdx.x < n) {
                    s_cpre[threadIdx.x] = pre;
                    s_ckb[threadIdx.x] = kb;
                    s_cproj[threadIdx.x] = pc;
                }
                int total = 0;
                for (int w = 0; w < nwv; ++w) total += s_wsum[w];
                __syncthreads();
                for (int t0 = 0; t0 < total; t0 += (int)blockDim.x) {  // workgroup-uniform
                    const int t = t0 + (int)threadIdx.x;
                    bool ok = false;
                    int32_t j = 0, jk0 = 0, jk1 = 0;
                    float pnew = 0.f;
                    if (t < total) {
                        // owner: the last entry whose scan start is <= t (its degree is then > t - start)
                        int lo = 0, hi = n - 1;
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (s_cpre[mid] <= t) lo = mid;
                            else hi = mid - 1;
                        }
                        const int k = s_ckb[lo] + (t - s_cpre[lo]);
                        const float epc = s_cproj[lo];
                        j = ci[k];
                        const uint32_t eid = (uint32_t)edge[k];
                        // j's row, kept with its ring entry if j joins: in flight with the
                        // spin load and the claim
                        jk0 = rp[j];
                        jk1 = rp[j + 1];
                        {
                            const float4 vj = sp4[(int64_t)j * ps];  // in flight during the Philox rounds
                            const uint4 bw = philox4x32_10(eid >> 2, ncl, (uint32_t)grep, ST_WBOND, key);
                            const float cpl = J * epc * cproj<KIND>(vj, axis, rd);
                            if (cpl > 0.f && u01f(u4_get(bw, (int)(eid & 3u))) < -expm1f(-2.0f * beta * cpl)) {
                                if (glob) {
                                    const uint32_t ob = atomicOr(&vis[j >> 5], 1u << (j & 31));
                                    ok = !(ob & (1u << (j & 31)));
                                } else {
                                    ok = vset_insert(s_vset, (uint32_t)j);
                                }
                                if (ok) {
                                    // the claim succeeded, so nobody changed vj since it was read
                                    const float4 v = creflect<KIND>(vj, axis, rd);
                                    sp4[(int64_t)j * ps] = v;
                                    pnew = cproj<KIND>(v, axis, rd);
                                }
                            }
                        }
                    }
                    const int pos = wave_reserve(&s_tail, ok ? 1u : 0u);
                    if (ok) {
                        if (pos - head < CAP) {
                            const int sl = pos & (CAP - 1);
                            s_kb[sl] = jk0;
                            s_deg[sl] = jk1 - jk0;
                            s_proj[sl] = pnew;
                        }
                        q[pos] = j;  // every member, for the bitmap
                    }
                }
                __syncthreads();  // the chunk's LDS copy is read until here
            }
            __syncthreads();
            hprev = head;
            head = tail;
        }
        const int size = s_tail;
        if (glob)
            for (int idx = threadIdx.x; idx < size; idx += blockDim.x) vis[q[idx] >> 5] = 0u;
        if (vlim > 0)
            for (int i = threadIdx.x; i < MQ_VSET; i += blockDim.x) s_vset[i] = 0xFFFFFFFFu;
        __syncthreads();
        if (threadIdx.x == 0 && csize) csize[(int64_t)grep * nclusters + c] = size;
    }
}

// Ising / Potts Wolff on a CSR graph with inline neighbours.  As k_csr_wolff,
// plus: each frontier entry of a narrow level carries its first EW
// (neighbour, original slot) pairs in LDS, written by the lane that claimed the
// entry, which loaded them from the lattice's padded row table (ell[pos*EW + t],
// one 128-byte line per vertex) in the same round trip as the claim.  A level of
// at most ECAP entries, all of degree <= EW, whose entries were written by a
// level of at most ESPEC entries, then needs no degree scan, no search and no
// neighbour-index load: lane t takes pair t % EW of entry t / EW and waits on
// one memory round trip (the claim).  Other levels run the block-wide scan of
// k_csr_wolff.  Same tests, same clusters.  On the 2^20-vertex RGG (64 thermalised
// replicas, the same clusters timed) this lifts Ising 0.324 -> 0.360 and Potts-3
// 0.230 -> 0.268 sites/ns; inline buffers of 64 or 256 entries, or loading the
// pairs in levels of up to 256 entries, measured 3-5% below 128 / 128.
constexpr int EW = 16;
constexpr int ECAP = 128;
constexpr int ESPEC = 128;

// Test of graph site j over the edge with original slot eid: the test's
// Philox uniform against the add threshold; for Potts also whether j still
// holds the cluster's old value (its word is returned in seen for the claim).
template <int KIND>
__device__ __forceinline__ bool bond_try(const int8_t* sj, uint32_t eid, uint32_t ncl, int grep, uint32_t tw, int val,
                                         Key2 key, uint32_t& seen)
{
    if (KIND == CK_ISING) {
        const uint4 bw = philox4x32_10(eid >> 2, ncl, (uint32_t)grep, ST_WBOND, key);
        return u4_get(bw, (int)(eid & 3u)) < tw;  // the claim itself tests alignment
    }
    uint32_t sh;
    seen = peek_word(const_cast<int8_t*>(sj), sh);  // in flight during the Philox rounds
    const uint4 bw = philox4x32_10(eid >> 2, ncl, (uint32_t)grep, ST_WBOND, key);
    return (int8_t)(seen >> sh) == (int8_t)val && u4_get(bw, (int)(eid & 3u)) < tw;
}

template <int KIND>
__device__ __forceinline__ bool claim8(int8_t* sj, int val, int newq, uint32_t seen)
{
    if (KIND == CK_ISING) return claim_ising(sj, val);
    uint32_t sh;
    uint32_t* w = spin_word(sj, sh);
    const uint32_t prev = atomicCAS(w, seen, (seen & ~(0xFFu << sh)) | ((uint32_t)(uint8_t)newq << sh));
    // the word changed under the first swap: retry from what it held
    return (prev == seen) || claim_potts(sj, prev, val, newq);
}

// One test of the int8 graph Wolff: the uniform, then (row, inline pairs, claim)
// issued together so that they share one memory round trip.
// inline pairs are moved as clang vector types, which stay in registers (the HIP
// vector structs do not when copied through an array)
typedef int mq_v4i __attribute__((ext_vector_type(4)));

#define MQ_BOND(J, EID)                                                                                          \
    do {                                                                                                         \
        int8_t* sj_ = sp8 + (int64_t)(J) * ps;                                                                   \
        uint32_t seen_ = 0u;                                                                                     \
        if (bond_try<KIND>(sj_, (EID), ncl, grep, tw, val, key, seen_)) {                                        \
            jk0 = rp[(J)];                                                                                       \
            jk1 = rp[(J) + 1];                                                                                   \
            if (spec) {                                                                                          \
                const mq_v4i* rw_ = reinterpret_cast<const mq_v4i*>(ell + (int64_t)(J) * EW);                  \
                row[0] = rw_[0], row[1] = rw_[1], row[2] = rw_[2], row[3] = rw_[3];                              \
                row[4] = rw_[4], row[5] = rw_[5], row[6] = rw_[6], row[7] = rw_[7];                              \
            }                                                                                                    \
            ok = claim8<KIND>(sj_, val, newq, seen_);                                                            \
        }                                                                                                        \
    } while (0)

template <int KIND>
__global__ void __launch_bounds__(1024) k_csr_wolff8(int8_t* __restrict__ base, int64_t rs, int64_t ps, int64_t N,
                                                    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                    const int32_t* __restrict__ edge, const int2* __restrict__ ell,
                                                    const int32_t* __restrict__ pos_of, int nclusters, uint32_t cl0,
                                                    int rep0, int Q, const uint32_t* __restrict__ wthr,
                                                    int32_t* __restrict__ queue,
                                                    const int64_t* __restrict__ rep_site_off,
                                                    int64_t* __restrict__ csize, Key2 key)
{
    constexpr int CAP = CWCAP / 2;  // with the inline pairs: 57 KB of LDS
    __shared__ int s_tail;
    __shared__ int32_t s_kb[CAP];   // ring: row start and degree of every entry (scan path)
    __shared__ int32_t s_deg[CAP];
    __shared__ int2 s_en[2][ECAP][EW];  // inline pairs, one buffer per level parity
    __shared__ int32_t s_edeg[2][ECAP];
    // largest degree among the entries with inline pairs of level k: s_emax[k % 3]; the
    // counter of level k + 2 is cleared during level k, after its last read (level k - 1)
    __shared__ int s_emax[3];
    __shared__ int32_t s_ckb[1024];
    __shared__ int32_t s_cpre[1024];
    __shared__ int s_wsum[16];
    const int r = blockIdx.x;
    const int grep = rep0 + r;
    int8_t* sp8 = base + r * rs;
    int32_t* q = queue + rep_site_off[grep];
    const uint32_t tw = wthr[grep];
    for (int c = 0; c < nclusters; ++c) {
        const uint32_t ncl = cl0 + (uint32_t)c;
        const uint4 sw = philox4x32_10(0u, ncl, (uint32_t)grep, ST_WSEED, key);
        const int64_t seed = pos_of[(int64_t)(((uint64_t)sw.x * (uint64_t)N) >> 32)];
        uint32_t sh0;
        const int val = (int8_t)(peek_word(sp8 + seed * ps, sh0) >> sh0);  // earlier clusters wrote with L2 atomics
        const int newq = (KIND == CK_POTTS) ? (val + 1 + (int)(((uint64_t)sw.y * (uint64_t)(Q - 1)) >> 32)) % Q : -val;
        __syncthreads();  // the seed's old value is read and the previous cluster's LDS use is over
        if (threadIdx.x == 0) {
            if (KIND == CK_ISING) {
                claim_ising(sp8 + seed * ps, val);
            } else {
                uint32_t sh;
                claim_potts(sp8 + seed * ps, peek_word(sp8 + seed * ps, sh), val, newq);
            }
            const int k0 = rp[seed], d0 = rp[seed + 1] - k0;
            s_kb[0] = k0;
            s_deg[0] = d0;
            s_edeg[0][0] = d0;
            for (int t = 0; t < EW; ++t) s_en[0][0][t] = ell[seed * EW + t];
            s_emax[0] = d0;
            s_emax[1] = 0;
            s_emax[2] = 0;
            s_tail = 1;
        }
        __syncthreads();
        int head = 0, hprev = 0, lvl = 0, nprev = 1;
        while (true) {
            const int tail = s_tail;
            const int emax = s_emax[lvl % 3];
            __syncthreads();
            if (head >= tail) break;
            if (threadIdx.x == 0) s_emax[(lvl + 2) % 3] = 0;
            const int n = tail - head;
            const int eb = lvl & 1;
            const bool spec = n <= ESPEC;  // this level's claims load their inline pairs
            // reserves the new entry's position and writes its ring slot (or queue
            // entry) and, in a level that loaded them, its inline pairs
#define MQ_PLACE(OK, J, JK0, JK1, ROW)                                                  \
    do {                                                                                \
        const int pos_ = wave_reserve(&s_tail, (OK) ? 1u : 0u);                         \
        if (OK) {                                                                       \
            if (pos_ - head < CAP) {                                                    \
                s_kb[pos_ & (CAP - 1)] = (JK0);                                         \
                s_deg[pos_ & (CAP - 1)] = (JK1) - (JK0);                                \
            } else {                                                                    \
                q[pos_] = (J);                                                          \
            }                                                                           \
            const int re_ = pos_ - tail;                                                \
            if (spec && re_ < ECAP) {                                                   \
                s_edeg[eb ^ 1][re_] = (JK1) - (JK0);                                    \
                mq_v4i* dst_ = reinterpret_cast<mq_v4i*>(&s_en[eb ^ 1][re_][0]);       \
                dst_[0] = ROW[0], dst_[1] = ROW[1], dst_[2] = ROW[2], dst_[3] = ROW[3];   \
                dst_[4] = ROW[4], dst_[5] = ROW[5], dst_[6] = ROW[6], dst_[7] = ROW[7];   \
                atomicMax(&s_emax[(lvl + 1) % 3], (JK1) - (JK0));                       \
            }                                                                           \
        }                                                                               \
    } while (0)
            if (n <= ECAP && emax <= EW && nprev <= ESPEC) {
                // inline level: lane t takes pair t % EW of entry t / EW (workgroup-uniform trip count)
                for (int t0 = 0; t0 < n * EW; t0 += (int)blockDim.x) {
                    const int t = t0 + (int)threadIdx.x;
                    const int e = t / EW, k = t % EW;
                    bool ok = false;
                    int32_t j = 0;
                    int jk0 = 0, jk1 = 0;
                    mq_v4i row[EW / 2] = {};
                    if (t < n * EW && k < s_edeg[eb][e]) {
                        const int2 pr = s_en[eb][e][k];
                        j = pr.x;
                        MQ_BOND(j, (uint32_t)pr.y);
                    }
                    MQ_PLACE(ok, j, jk0, jk1, row);
                }
            } else {
                // The level's edges are spread over the whole workgroup: chunks of up to
                // blockDim entries are copied to LDS with a block-wide scan of their degrees,
                // and lane t takes the chunk's edges t, t + blockDim, ...: it finds its entry
                // by a binary search over the scan, loads the far end and claims.
                const int lane = (int)(threadIdx.x & 63u);
                const int wv = (int)(threadIdx.x >> 6);
                const int nwv = (int)(blockDim.x >> 6);
                for (int c0 = head; c0 < tail; c0 += (int)blockDim.x) {
                    const int nch = min((int)blockDim.x, tail - c0);
                    const int idx = c0 + (int)threadIdx.x;
                    int kb = 0, deg = 0;
                    if ((int)threadIdx.x < nch) {
                        if (idx - hprev < CAP) {
                            kb = s_kb[idx & (CAP - 1)];
                            deg = s_deg[idx & (CAP - 1)];
                        } else {
                            const int32_t cur = q[idx];
                            kb = rp[cur];
                            deg = rp[cur + 1] - kb;
                        }
                    }
                    // block-wide exclusive scan of the degrees
                    int inc = deg;
#pragma unroll
                    for (int o = 1; o < 64; o <<= 1) {
                        const int v = __shfl_up(inc, o);
                        if (lane >= o) inc += v;
                    }
                    if (lane == 63) s_wsum[wv] = inc;
                    __syncthreads();
                    int wpre = 0;
                    for (int w = 0; w < wv; ++w) wpre += s_wsum[w];
                    if ((int)threadIdx.x < nch) {
                        s_cpre[threadIdx.x] = wpre + inc - deg;
                        s_ckb[threadIdx.x] = kb;
                    }
                    int total = 0;
                    for (int w = 0; w < nwv; ++w) total += s_wsum[w];
                    __syncthreads();
                    for (int t0 = 0; t0 < total; t0 += (int)blockDim.x) {  // workgroup-uniform
                        const int t = t0 + (int)threadIdx.x;
                        bool ok = false;
                        int32_t j = 0;
                        int jk0 = 0, jk1 = 0;
                        mq_v4i row[EW / 2] = {};
                        if (t < total) {
                            // owner: the last entry whose scan start is <= t
                            int lo = 0, hi = nch - 1;
                            while (lo < hi) {
                                const int mid = (lo + hi + 1) >> 1;
                                if (s_cpre[mid] <= t) lo = mid;
                                else hi = mid - 1;
                            }
                            const int k = s_ckb[lo] + (t - s_cpre[lo]);
                            j = ci[k];
                            MQ_BOND(j, (uint32_t)edge[k]);
                        }
                        MQ_PLACE(ok, j, jk0, jk1, row);
                    }
                    __syncthreads();  // the chunk's LDS copy is read until here
                }
            }
#undef MQ_PLACE
            __syncthreads();
            hprev = head;
            head = tail;
            nprev = n;
            ++lvl;
        }
        const int size = s_tail;
        __syncthreads();
        if (threadIdx.x == 0 && csize) csize[(int64_t)grep * nclusters + c] = size;
    }
}

#undef MQ_BOND

// Padded row table of a graph: ell[p*EW + t] = (position, original slot) of the
// t-th neighbour of position p, (-1, -1) past its degree.
__global__ void __launch_bounds__(256) k_build_ell(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                   const int32_t* __restrict__ edge, int64_t N, int2* __restrict__ ell)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * EW) return;
    const int64_t p = i / EW;
    const int t = (int)(i - p * EW);
    const int a = rp[p], d = rp[p + 1] - a;
    ell[i] = t < d ? make_int2(ci[a + t], edge[a + t]) : make_int2(-1, -1);
}

inline dim3 grid_for(int64_t work, int nrep) { return dim3((unsigned)((work + 255) / 256), (unsigned)nrep, 1); }

}  // namespace

// ------------------------------------------------------------------ host
extern "C" int mq_lattice_csr(mq_ctx* ctx, int64_t N, const int32_t* row_ptr, const int32_t* col_idx, mq_lattice** out)
{
    if (!ctx) return MQ_ERR_ARG;
    if (!out || !row_ptr || !col_idx) return mq_fail(ctx, MQ_ERR_ARG, "mq_lattice_csr: NULL argument");
    if (N <= 0) return mq_fail(ctx, MQ_ERR_ARG, "mq_lattice_csr: N must be > 0");
    *out = nullptr;
    if (N >= ((int64_t)1 << 31)) return mq_fail(ctx, MQ_ERR_ARG, "graph too large");
    if (row_ptr[0] != 0) return mq_fail(ctx, MQ_ERR_ARG, "row_ptr[0] must be 0");
    const int64_t nnz = row_ptr[N];
    for (int64_t i = 0; i < N; ++i) {
        if (row_ptr[i + 1] < row_ptr[i]) return mq_fail(ctx, MQ_ERR_ARG, "row_ptr not monotone");
        for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
            const int j = col_idx[k];
            if (j < 0 || j >= N || j == i) return mq_fail(ctx, MQ_ERR_ARG, "bad neighbour index or self loop");
        }
    }
    // symmetry and mirrored slots: slot of (u -> v) mirrored to (v -> u), in O(nnz)
    // with local sorts: the incoming slots of every vertex are listed by a counting
    // transpose (in increasing source), each row's own slots are sorted by column, and
    // the two lists of a vertex must then name the same neighbours in the same order
    std::vector<int32_t> mirror(nnz, -1);
    {
        std::vector<int32_t> by_col(nnz);  // row u's slots sorted by column
        for (int64_t u = 0; u < N; ++u) {
            const int a = row_ptr[u], b = row_ptr[u + 1];
            for (int k = a; k < b; ++k) by_col[k] = k;
            std::sort(by_col.begin() + a, by_col.begin() + b,
                      [&](int32_t x, int32_t y) { return col_idx[x] < col_idx[y]; });
            for (int k = a + 1; k < b; ++k)
                if (col_idx[by_col[k]] == col_idx[by_col[k - 1]]) return mq_fail(ctx, MQ_ERR_ARG, "duplicate edge");
        }
        std::vector<int64_t> tp(N + 1, 0);  // incoming slots per vertex
        for (int64_t k = 0; k < nnz; ++k) tp[col_idx[k] + 1]++;
        for (int64_t v = 0; v < N; ++v) tp[v + 1] += tp[v];
        std::vector<int32_t> tin(nnz), src(nnz);
        {
            std::vector<int64_t> fill(tp.begin(), tp.end() - 1);
            for (int64_t u = 0; u < N; ++u)
                for (int k = row_ptr[u]; k < row_ptr[u + 1]; ++k) {
                    tin[fill[col_idx[k]]++] = k;  // sources in increasing order
                    src[k] = (int32_t)u;
                }
        }
        for (int64_t v = 0; v < N; ++v) {
            const int a = row_ptr[v], deg = row_ptr[v + 1] - a;
            if (tp[v + 1] - tp[v] != deg) return mq_fail(ctx, MQ_ERR_ARG, "adjacency is not symmetric");
            for (int t = 0; t < deg; ++t) {
                const int32_t ko = by_col[a + t];       // v -> u
                const int32_t ki = tin[tp[v] + t];      // u' -> v
                if (col_idx[ko] != src[ki]) return mq_fail(ctx, MQ_ERR_ARG, "adjacency is not symmetric");
                mirror[ko] = ki;
            }
        }
    }
    // greedy colouring in vertex order: the smallest colour no earlier neighbour has
    // (at most degree + 1 colours)
    std::vector<int32_t> colour(N, -1);
    std::vector<int64_t> seen_by;  // seen_by[c] == u: colour c is taken by a neighbour of u
    int ncol = 0;
    for (int64_t u = 0; u < N; ++u) {
        const int deg = row_ptr[u + 1] - row_ptr[u];
        if ((int)seen_by.size() < deg + 1) seen_by.resize(deg + 1, -1);
        for (int k = row_ptr[u]; k < row_ptr[u + 1]; ++k) {
            const int c = colour[col_idx[k]];
            if (c >= 0 && c <= deg) seen_by[c] = u;
        }
        int c = 0;
        while (seen_by[c] == u) ++c;
        colour[u] = c;
        ncol = std::max(ncol, c + 1);
    }
    uint32_t stride = 64;
    while (stride < (uint32_t)ncol) stride <<= 1;
    if (stride > MQ_MAX_COLOR_STRIDE)
        return mq_fail(ctx, MQ_ERR_UNSUPPORTED, "greedy colouring needs more than 65536 colours");
    auto* l = new mq_lattice();
    l->ctx = ctx;
    l->kind = LAT_CSR;
    l->N = N;
    l->nnz = nnz;
    l->ncolors = ncol;
    l->color_stride = stride;
    l->color_off.assign(ncol + 1, 0);
    for (int64_t u = 0; u < N; ++u) l->color_off[colour[u] + 1]++;
    for (int c = 0; c < ncol; ++c) l->color_off[c + 1] += l->color_off[c];
    l->pos_of.resize(N);
    l->vert_of.resize(N);
    l->colour.assign(colour.begin(), colour.end());
    {
        // Within a class, vertices are placed in breadth-first order (every
        // component, seeds in id order): graph neighbours get nearby ranks, so
        // the neighbours of a run of consecutive positions fall in a few short
        // windows of the other classes and the gathers hit L2.  Results do not
        // depend on this order (random numbers are keyed by original ids).
        std::vector<int32_t> order;
        order.reserve(N);
        std::vector<char> seen(N, 0);
        for (int64_t s0 = 0; s0 < N; ++s0) {
            if (seen[s0]) continue;
            seen[s0] = 1;
            size_t head = order.size();
            order.push_back((int32_t)s0);
            while (head < order.size()) {
                const int32_t u = order[head++];
                for (int k = row_ptr[u]; k < row_ptr[u + 1]; ++k) {
                    const int32_t v = col_idx[k];
                    if (!seen[v]) {
                        seen[v] = 1;
                        order.push_back(v);
                    }
                }
            }
        }
        std::vector<int64_t> next(l->color_off.begin(), l->color_off.end() - 1);
        for (int64_t t = 0; t < N; ++t) {
            const int32_t u = order[t];
            const int64_t p = next[colour[u]]++;
            l->pos_of[u] = (int32_t)p;
            l->vert_of[p] = u;
        }
    }
    l->h_rp.assign(row_ptr, row_ptr + N + 1);
    l->h_ci.assign(col_idx, col_idx + nnz);
    std::vector<int32_t> rp(N + 1), ci(nnz), edge(nnz), slot_of_orig(nnz);
    rp[0] = 0;
    int maxdeg = 0;
    for (int64_t p = 0; p < N; ++p) {
        const int64_t u = l->vert_of[p];
        const int deg = row_ptr[u + 1] - row_ptr[u];
        maxdeg = std::max(maxdeg, deg);
        rp[p + 1] = rp[p] + deg;
        for (int t = 0; t < deg; ++t) {
            const int k = row_ptr[u] + t;
            const int64_t v = col_idx[k];
            ci[rp[p] + t] = l->pos_of[v];
            edge[rp[p] + t] = k;  // the test from u to v keys its uniform by its own slot
            slot_of_orig[k] = rp[p] + t;
        }
    }
    l->maxdeg = maxdeg;
    l->h_slot_of_orig = slot_of_orig;
    l->h_mirror = mirror;
    auto cleanup = [&](hipError_t e) {
        ctx->err = std::string("csr upload: ") + hipGetErrorString(e);
        mq_lattice_destroy(l);
        return MQ_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipMalloc(&l->d_rp, sizeof(int32_t) * (N + 1))) != hipSuccess) return cleanup(e);
    if ((e = hipMalloc(&l->d_ci, sizeof(int32_t) * std::max<int64_t>(nnz, 1))) != hipSuccess) return cleanup(e);
    if ((e = hipMalloc(&l->d_edge, sizeof(int32_t) * std::max<int64_t>(nnz, 1))) != hipSuccess) return cleanup(e);
    if ((e = hipMalloc(&l->d_pos_of, sizeof(int32_t) * N)) != hipSuccess) return cleanup(e);
    if ((e = hipMalloc(&l->d_vert_of, sizeof(int32_t) * N)) != hipSuccess) return cleanup(e);
    if ((e = hipMemcpy(l->d_rp, rp.data(), sizeof(int32_t) * (N + 1), hipMemcpyHostToDevice)) != hipSuccess) return cleanup(e);
    if (nnz) {
        if ((e = hipMemcpy(l->d_ci, ci.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice)) != hipSuccess) return cleanup(e);
        if ((e = hipMemcpy(l->d_edge, edge.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice)) != hipSuccess) return cleanup(e);
    }
    if ((e = hipMemcpy(l->d_pos_of, l->pos_of.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice)) != hipSuccess) return cleanup(e);
    if ((e = hipMemcpy(l->d_vert_of, l->vert_of.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice)) != hipSuccess) return cleanup(e);
    *out = l;
    return MQ_OK;
}

extern "C" int mq_lattice_color_of(const mq_lattice* lat, int32_t* colour)
{
    if (!lat) return MQ_ERR_ARG;
    if (!colour) return mq_fail(lat->ctx, MQ_ERR_ARG, "mq_lattice_color_of: colour is NULL");
    if (lat->kind == LAT_CSR) {
        std::copy(lat->colour.begin(), lat->colour.end(), colour);
    } else {
        for (int64_t i = 0; i < lat->N; ++i) colour[i] = regular_colour(lat->dim, lat->L, i);
    }
    return MQ_OK;
}

hipError_t launch_init_csr(mq_state* st, const Group& g, Key2 key)
{
    hipLaunchKernelGGL(k_csr_init, grid_for(g.lat->N, g.nrep), dim3(256), 0, st->ctx->stream, st->d_spins + g.off,
                       (int64_t)1, (int64_t)g.nrep, g.lat->d_vert_of, g.lat->N, st->p.model, st->p.q, g.rep0, key);
    return hipGetLastError();
}

hipError_t launch_sweep_csr(mq_state* st, const Group& g, int colour, uint32_t phase, Key2 key)
{
    const mq_lattice* l = g.lat;
    const int64_t c0 = l->color_off[colour], nc = l->color_off[colour + 1] - c0;
    if (nc == 0) return hipSuccess;
    const int model = st->p.model;
    const bool bonds = st->d_bonds != nullptr;
    hipStream_t s = st->ctx->stream;
    const int64_t work = nc * g.nrep;
    const int blk = st->tune_block;
    const dim3 grid((unsigned)((work + blk - 1) / blk));
    if (model == MQ_HEISENBERG || model == MQ_PHI4) {
        float4* b = reinterpret_cast<float4*>(st->d_spins + g.off);
        const float J = (float)st->p.J, d = (float)st->p.delta, lam = (float)st->p.lambda, m = (float)st->p.mass;
        if (bonds)
            hipLaunchKernelGGL(k_csr_sweep_o3<true>, grid, dim3(blk), 0, s, b, (int64_t)1, (int64_t)g.nrep, g.nrep, l->d_rp,
                               l->d_ci, l->d_vert_of, st->d_bonds, st->d_rep_bond_off, st->d_betaf, c0, nc, model,
                               st->p.proposal, J, d, lam, m, phase, st->cur_dyn, g.rep0, key);
        else
            hipLaunchKernelGGL(k_csr_sweep_o3<false>, grid, dim3(blk), 0, s, b, (int64_t)1, (int64_t)g.nrep, g.nrep, l->d_rp,
                               l->d_ci, l->d_vert_of, st->d_bonds, st->d_rep_bond_off, st->d_betaf, c0, nc, model,
                               st->p.proposal, J, d, lam, m, phase, st->cur_dyn, g.rep0, key);
    } else {
        int8_t* b = reinterpret_cast<int8_t*>(st->d_spins + g.off);
        bool oct_ok = g.nrep % 8 == 0 && g.rep0 % 8 == 0 && l->maxdeg <= 127 &&
                      st->kernel_choice != 2;
        if (oct_ok && bonds)  // the 8 replicas of a lane must share one coupling sample
            for (int r = g.rep0; r < g.rep0 + g.nrep && oct_ok; ++r)
                oct_ok = st->rep_sample[r] == st->rep_sample[r - r % 8];
        if (oct_ok) {
            const int64_t w8 = nc * (g.nrep / 8);
            const dim3 g8((unsigned)((w8 + 255) / 256));
            if (model == MQ_POTTS)
                hipLaunchKernelGGL(k_csr_sweep_potts8, g8, dim3(256), 0, s, b, g.nrep, l->d_rp, l->d_ci, l->d_vert_of,
                                   st->d_thr, st->tbl, c0, nc, st->p.q, phase, st->cur_dyn, g.rep0, key);
            else if (bonds)
                hipLaunchKernelGGL(k_csr_sweep_ising8<true>, g8, dim3(256), 0, s, b, g.nrep, l->d_rp, l->d_ci,
                                   l->d_vert_of, st->d_bonds, st->d_rep_bond_off, st->d_thr, st->tbl, c0, nc, phase,
                                   st->cur_dyn, g.rep0, key);
            else
                hipLaunchKernelGGL(k_csr_sweep_ising8<false>, g8, dim3(256), 0, s, b, g.nrep, l->d_rp, l->d_ci,
                                   l->d_vert_of, st->d_bonds, st->d_rep_bond_off, st->d_thr, st->tbl, c0, nc, phase,
                                   st->cur_dyn, g.rep0, key);
            return hipGetLastError();
        }
        if (bonds)
            hipLaunchKernelGGL(k_csr_sweep_int8<true>, grid, dim3(blk), 0, s, b, (int64_t)1, (int64_t)g.nrep, g.nrep, l->d_rp,
                               l->d_ci, l->d_vert_of, st->d_bonds, st->d_rep_bond_off, st->d_thr, st->tbl, c0, nc,
                               model, st->p.q, phase, st->cur_dyn, g.rep0, key);
        else
            hipLaunchKernelGGL(k_csr_sweep_int8<false>, grid, dim3(blk), 0, s, b, (int64_t)1, (int64_t)g.nrep, g.nrep, l->d_rp,
                               l->d_ci, l->d_vert_of, st->d_bonds, st->d_rep_bond_off, st->d_thr, st->tbl, c0, nc,
                               model, st->p.q, phase, st->cur_dyn, g.rep0, key);
    }
    return hipGetLastError();
}

hipError_t launch_measure_csr(mq_state* st, const Group& g, long long* isum, double* fsum)
{
    const mq_lattice* l = g.lat;
    const bool vec = st->p.model == MQ_HEISENBERG || st->p.model == MQ_PHI4;
    const bool bonds = st->d_bonds != nullptr;
    hipStream_t s = st->ctx->stream;
    const int64_t per_rep = std::max<int64_t>(1, std::min<int64_t>((l->N + 255) / 256,
                                                                   (8ll * st->ctx->num_cus + g.nrep - 1) / g.nrep));
    const dim3 grid((unsigned)per_rep, (unsigned)g.nrep, 1);
#define MQ_M(B, V)                                                                                                  \
    hipLaunchKernelGGL((k_csr_measure<B, V>), grid, dim3(256), 0, s, st->d_spins + g.off, (int64_t)1,             \
                       (int64_t)g.nrep, l->d_rp,                                                                 \
                       l->d_ci, st->d_bonds, st->d_rep_bond_off, l->N, st->p.model, st->p.q, g.rep0, isum,       \
                       fsum)
    if (vec) {
        if (bonds) MQ_M(true, true);
        else MQ_M(false, true);
    } else {
        if (bonds) MQ_M(true, false);
        else MQ_M(false, false);
    }
#undef MQ_M
    return hipGetLastError();
}

hipError_t launch_wolff_csr(mq_state* st, const Group& g, int nclusters, uint32_t cl0, Key2 key)
{
    const mq_lattice* l = g.lat;
    const int model = st->p.model;
    hipStream_t s = st->ctx->stream;
    const float J = (float)st->p.J;
#define MQ_W(K)                                                                                                   \
    hipLaunchKernelGGL((k_csr_wolff<K>), dim3(g.nrep), dim3(wolff_block(st, g.nrep)), 0, s, st->d_spins + g.off, (int64_t)1,          \
                       (int64_t)g.nrep, l->N,                                                                    \
                       l->d_rp, l->d_ci, l->d_edge, l->d_pos_of, nclusters, cl0, g.rep0, st->p.q, st->d_wthr,    \
                       st->d_betaf, J, st->d_queue, st->d_visit, st->d_rep_site_off, st->d_csize, l->maxdeg,   \
                       st->wolff_vset, key)
    if (model == MQ_ISING || model == MQ_POTTS) {
        mq_lattice* lm = const_cast<mq_lattice*>(l);
        if (!lm->d_ell) {  // the padded row table, built on first use (N * EW * 8 bytes)
            hipError_t e = hipMalloc(&lm->d_ell, sizeof(int2) * (size_t)(l->N * EW));
            if (e != hipSuccess) return e;
            const int64_t work = l->N * EW;
            hipLaunchKernelGGL(k_build_ell, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, l->d_rp, l->d_ci,
                               l->d_edge, l->N, reinterpret_cast<int2*>(lm->d_ell));
        }
#define MQ_W8(K)                                                                                                  \
    hipLaunchKernelGGL((k_csr_wolff8<K>), dim3(g.nrep), dim3(wolff_block(st, g.nrep)), 0, s,                     \
                       reinterpret_cast<int8_t*>(st->d_spins + g.off), (int64_t)1, (int64_t)g.nrep, l->N, l->d_rp, \
                       l->d_ci, l->d_edge, reinterpret_cast<const int2*>(l->d_ell), l->d_pos_of, nclusters, cl0,  \
                       g.rep0, st->p.q, st->d_wthr, st->d_queue, st->d_rep_site_off, st->d_csize, key)
        if (model == MQ_ISING) MQ_W8(CK_ISING);
        else MQ_W8(CK_POTTS);
#undef MQ_W8
    } else if (model == MQ_HEISENBERG && st->p.embedding == MQ_EMBED_CARTESIAN) MQ_W(CK_CART);
    else if (model == MQ_HEISENBERG || model == MQ_PHI4) MQ_W(CK_RDIR);
    else return hipErrorInvalidValue;
#undef MQ_W
    return hipGetLastError();
}
