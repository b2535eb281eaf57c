// Multi-spin-coded (1 bit per spin) 3D Ising: storage, random start,
// checkerboard Metropolis sweep and observables.
//
// Layout per replica: for each colour the half-lattice in the int8 layout's
// order (half index h = i >> 1, rows of L/2 sites), packed 32 consecutive
// half-sites per uint32, bit j of word w = site h = 32w + j; a set bit is a
// down spin (-1).  Requires L a power of two with L >= 64 (whole words per row).
//
// Sweep (one lane per word of 32 sites):
//   * the six neighbour words of the other colour are the words at the same
//     position in rows y+-1, z+-1, the same row, and the same row shifted by
//     one site (v_alignbit with the adjacent word);
//   * the number c of anti-aligned neighbours (0..6) is summed bit-sliced with
//     two full adders (xor3 / majority) and a final adder into three bit planes;
//   * a site is flipped if its class c is "always" (dE <= 0) or it wins a
//     Bernoulli(p_c) draw.  The draw compares the site's 32-bit uniform U with
//     T = floor(p_c 2^32) bit-serially, MSB first: level k uses one random word
//     whose bit j is bit 31-k of U_j; the first level where U's bit differs from
//     T's bit decides the site (U < T accepts), and a lane stops drawing once
//     all its 32 sites are decided.  The decision is exactly U < T, the rule of
//     the int8 kernels, with U assembled from per-level words:
//       level k word = Philox(word index, 2*sweep + colour, replica, 8 + k/4)[k % 4].
#include "mq_internal.hpp"

namespace {

constexpr uint32_t ST_MSC = 8;  // streams 8..15: bit levels 0..31

// full adder on bit planes
__device__ __forceinline__ void fadd(uint32_t a, uint32_t b, uint32_t c, uint32_t& s, uint32_t& cy)
{
    s = a ^ b ^ c;
    cy = (a & b) | (a & c) | (b & c);
}

// Bernoulli classes are passed as (class id, threshold, flags) for at most 3
// classes; FERRO selects the classes {0,1,2} (J < 0) or {4,5,6} (J > 0).
struct MscClasses {
    uint32_t T[3];
    uint32_t always[3];  // class accepted without a draw (p >= 1)
};

template <bool FERRO>
__global__ void __launch_bounds__(512) k_sweep_msc(uint32_t* __restrict__ base, int64_t stride_words,
                                                   const MscClasses* __restrict__ cls, int lgL, int colour,
                                                   uint32_t phase, const uint32_t* __restrict__ dyn, int rep0,
                                                   int lgbpr, int lgblk, Key2 key)
{
    if (dyn) phase += dyn[0];
    const uint32_t lb = blockIdx.x;
    const uint32_t r = lb >> lgbpr;
    const uint32_t grep = (uint32_t)rep0 + r;
    const MscClasses C = cls[grep];
    const uint32_t L = 1u << lgL;
    const uint32_t lgwpr = (uint32_t)lgL - 6u;  // words per row: (L/2)/32
    const uint32_t wpr = 1u << lgwpr;
    const uint32_t plane_w = (1u << (2 * lgL - 1)) >> 5;  // words per z-plane
    const uint32_t half_w = plane_w << lgL;
    const uint32_t widx = ((lb - (r << lgbpr)) << lgblk) + threadIdx.x;  // word index in the colour array
    const uint32_t row = widx >> lgwpr;
    const uint32_t wi = widx & (wpr - 1u);
    const uint32_t y = row & (L - 1u);
    const uint32_t z = row >> lgL;
    const uint32_t p = (colour + y + z) & 1u;
    uint32_t* own = base + (int64_t)r * stride_words + (int64_t)colour * half_w;
    const uint32_t* oth = base + (int64_t)r * stride_words + (int64_t)(1 - colour) * half_w;
    const uint32_t rb = row << lgwpr;

    const uint32_t S = own[widx];
    const uint32_t O = oth[widx];
    const uint32_t X = p ? oth[rb + ((wi + 1u) & (wpr - 1u))] : oth[rb + ((wi - 1u) & (wpr - 1u))];
    const uint32_t Ym = oth[(y == 0u) ? widx + (L - 1u) * wpr : widx - wpr];
    const uint32_t Yp = oth[(y == L - 1u) ? widx - (L - 1u) * wpr : widx + wpr];
    const uint32_t Zm = oth[(z == 0u) ? widx + (L - 1u) * plane_w : widx - plane_w];
    const uint32_t Zp = oth[(z == L - 1u) ? widx - (L - 1u) * plane_w : widx + plane_w];
    // x-shifted neighbour word: p = 1 needs site k+1, p = 0 needs site k-1
    const uint32_t Xs = p ? __builtin_amdgcn_alignbit(X, O, 1) : __builtin_amdgcn_alignbit(O, X, 31);

    // anti-aligned count c = sum_k (S ^ n_k), bit-sliced; S ^ maj(a,b,c) = maj(S^a, S^b, S^c)
    uint32_t s1, c1, s2, c2;
    fadd(O, Xs, Ym, s1, c1);
    fadd(Yp, Zm, Zp, s2, c2);
    s1 ^= S;
    c1 ^= S;
    s2 ^= S;
    c2 ^= S;
    const uint32_t b0 = s1 ^ s2;
    const uint32_t k0 = s1 & s2;
    const uint32_t b1 = c1 ^ c2 ^ k0;
    const uint32_t b2 = (c1 & c2) | (c1 & k0) | (c2 & k0);
    // class masks (c = b2 b1 b0)
    uint32_t m0, m1, m2, acc;
    if (FERRO) {  // dE > 0 for c in {0,1,2}
        m0 = ~b2 & ~b1 & ~b0;
        m1 = ~b2 & ~b1 & b0;
        m2 = ~b2 & b1 & ~b0;
        acc = b2 | (b1 & b0);  // c >= 3
    } else {      // dE > 0 for c in {4,5,6}
        m0 = b2 & ~b1 & ~b0;
        m1 = b2 & ~b1 & b0;
        m2 = b2 & b1 & ~b0;
        acc = ~b2 | (b1 & b0);  // c <= 3, c = 7 cannot occur
    }
    if (C.always[0]) acc |= m0;
    if (C.always[1]) acc |= m1;
    if (C.always[2]) acc |= m2;
    uint32_t D = (C.always[0] ? 0u : m0) | (C.always[1] ? 0u : m1) | (C.always[2] ? 0u : m2);
    uint4 w = make_uint4(0, 0, 0, 0);
    for (int k = 0; k < 32 && D != 0u; ++k) {
        if ((k & 3) == 0) w = philox4x32_10(widx, phase, grep, ST_MSC + (uint32_t)(k >> 2), key);
        const uint32_t R = u4_get(w, k & 3);
        const uint32_t sh = 31u - (uint32_t)k;
        const uint32_t t = (((C.T[0] >> sh) & 1u) ? m0 : 0u) | (((C.T[1] >> sh) & 1u) ? m1 : 0u) |
                           (((C.T[2] >> sh) & 1u) ? m2 : 0u);
        acc |= D & ~R & t;
        D &= ~(R ^ t);
    }
    own[widx] = S ^ acc;
}

// random start: the same spins as the int8 layout (Philox on the natural index)
__global__ void __launch_bounds__(256) k_init_msc(uint32_t* __restrict__ base, int64_t stride_words, int lgL, int rep0,
                                                  Key2 key)
{
    const int r = blockIdx.y;
    const uint32_t L = 1u << lgL;
    const uint32_t H = L >> 1;
    const uint32_t half = (H * L) << lgL;
    const uint32_t widx_all = blockIdx.x * blockDim.x + threadIdx.x;  // word over both colours
    if (widx_all >= (2u * half) >> 5) return;
    const uint32_t colour = widx_all / (half >> 5);
    const uint32_t widx = widx_all - colour * (half >> 5);
    uint32_t word = 0u;
    for (int j = 0; j < 32; ++j) {
        const uint32_t h = widx * 32u + (uint32_t)j;
        const uint32_t row = h / H;
        const uint32_t k = h - row * H;
        const uint32_t y = row & (L - 1u), z = row >> lgL;
        const uint32_t x0 = 2u * k + ((colour + y + z) & 1u);
        const uint32_t i = x0 + L * row;
        const uint4 q = philox4x32_10(i >> 2, 0u, (uint32_t)(rep0 + r), ST_INIT, key);
        const uint32_t up = u4_get(q, (int)(i & 3u)) >> 31;
        word |= (up ? 0u : 1u) << j;
    }
    base[(int64_t)r * stride_words + widx_all] = word;
}

// isum[r*8+0] = sum over bonds s_i s_j; isum[r*8+1] = sum s
__global__ void __launch_bounds__(256) k_measure_msc(const uint32_t* __restrict__ base, int64_t stride_words, int lgL,
                                                     int rep0, unsigned long long* __restrict__ isum)
{
    __shared__ long long red[2][4];
    const int r = blockIdx.y;
    const uint32_t L = 1u << lgL;
    const uint32_t lgwpr = (uint32_t)lgL - 6u;
    const uint32_t wpr = 1u << lgwpr;
    const uint32_t plane_w = (1u << (2 * lgL - 1)) >> 5;
    const uint32_t half_w = plane_w << lgL;
    const uint32_t widx = blockIdx.x * blockDim.x + threadIdx.x;
    long long bs = 0, ms = 0;
    if (widx < half_w) {
        const uint32_t* c0 = base + (int64_t)r * stride_words;
        const uint32_t* c1 = c0 + half_w;
        const uint32_t row = widx >> lgwpr, wi = widx & (wpr - 1u);
        const uint32_t y = row & (L - 1u), z = row >> lgL;
        const uint32_t p = (y + z) & 1u;  // colour 0
        const uint32_t rb = row << lgwpr;
        const uint32_t S = c0[widx], O = c1[widx];
        const uint32_t X = p ? c1[rb + ((wi + 1u) & (wpr - 1u))] : c1[rb + ((wi - 1u) & (wpr - 1u))];
        const uint32_t Xs = p ? __builtin_amdgcn_alignbit(X, O, 1) : __builtin_amdgcn_alignbit(O, X, 31);
        const uint32_t n[6] = {O, Xs, c1[(y == 0u) ? widx + (L - 1u) * wpr : widx - wpr],
                               c1[(y == L - 1u) ? widx - (L - 1u) * wpr : widx + wpr],
                               c1[(z == 0u) ? widx + (L - 1u) * plane_w : widx - plane_w],
                               c1[(z == L - 1u) ? widx - (L - 1u) * plane_w : widx + plane_w]};
        int anti = 0;
        for (int k = 0; k < 6; ++k) anti += (int)__popc(S ^ n[k]);  // __popc returns unsigned
        bs = 6 * 32 - 2 * anti;
        ms = 64 - 2 * ((int)__popc(S) + (int)__popc(c1[widx]));
    }
    for (int o = 32; o > 0; o >>= 1) {
        bs += __shfl_xor(bs, o, 64);
        ms += __shfl_xor(ms, o, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wid] = bs;
        red[1][wid] = ms;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        long long t = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[threadIdx.x][i];
        atomicAdd(isum + (int64_t)(rep0 + r) * 8 + threadIdx.x, (unsigned long long)t);
    }
}

int lg2(int L)
{
    int l = 0;
    while ((1 << l) < L) ++l;
    return l;
}

}  // namespace

bool msc_supported(const mq_lattice* lat)
{
    return lat->kind == LAT_REGULAR && lat->dim == 3 && (lat->L & (lat->L - 1)) == 0 && lat->L >= 64 && lat->L <= 1024;
}

// thresholds -> per-replica Bernoulli classes (host)
hipError_t msc_upload_classes(mq_state* st, const std::vector<uint32_t>& thr)
{
    const bool ferro = st->p.J < 0.0;
    std::vector<MscClasses> c(st->nrep);
    for (int r = 0; r < st->nrep; ++r) {
        for (int j = 0; j < 3; ++j) {
            const int cls = ferro ? j : 4 + j;
            const int e = 12 - 2 * cls;  // table index e = q + 6 = 2 * aligned = 12 - 2c
            const uint32_t hi = thr[((size_t)r * st->tbl + e) * 2], lo = thr[((size_t)r * st->tbl + e) * 2 + 1];
            c[r].always[j] = hi > 0xFFFFu ? 1u : 0u;
            c[r].T[j] = hi > 0xFFFFu ? 0xFFFFFFFFu : ((hi << 16) | lo);
        }
    }
    if (!st->d_msc_cls) {
        hipError_t e = hipMalloc(&st->d_msc_cls, sizeof(MscClasses) * st->nrep);
        if (e != hipSuccess) return e;
    }
    return hipMemcpy(st->d_msc_cls, c.data(), sizeof(MscClasses) * st->nrep, hipMemcpyHostToDevice);
}

hipError_t launch_sweep_msc(mq_state* st, const Group& g, int colour, uint32_t phase, Key2 key)
{
    const int lgL = lg2(g.lat->L);
    int lgblk = 0;
    while ((1 << lgblk) < st->tune_block) ++lgblk;
    const int lgbpr = 3 * lgL - 1 - 5 - lgblk;  // words per colour / block
    if (lgbpr < 0) return hipErrorInvalidValue;
    const int64_t blocks = ((int64_t)1 << lgbpr) * g.nrep;
    uint32_t* b = reinterpret_cast<uint32_t*>(st->d_spins + g.off);
    const MscClasses* cls = reinterpret_cast<const MscClasses*>(st->d_msc_cls);
    if (st->p.J < 0.0)
        hipLaunchKernelGGL(k_sweep_msc<true>, dim3((unsigned)blocks), dim3(1 << lgblk), 0, st->ctx->stream, b,
                           g.stride / 4, cls, lgL, colour, phase, st->cur_dyn, g.rep0, lgbpr, lgblk, key);
    else
        hipLaunchKernelGGL(k_sweep_msc<false>, dim3((unsigned)blocks), dim3(1 << lgblk), 0, st->ctx->stream, b,
                           g.stride / 4, cls, lgL, colour, phase, st->cur_dyn, g.rep0, lgbpr, lgblk, key);
    return hipGetLastError();
}

hipError_t launch_init_msc(mq_state* st, const Group& g, Key2 key)
{
    const int lgL = lg2(g.lat->L);
    const int64_t words = g.lat->N / 32;
    hipLaunchKernelGGL(k_init_msc, dim3((unsigned)((words + 255) / 256), (unsigned)g.nrep), dim3(256), 0,
                       st->ctx->stream, reinterpret_cast<uint32_t*>(st->d_spins + g.off), g.stride / 4, lgL, g.rep0,
                       key);
    return hipGetLastError();
}

hipError_t launch_measure_msc(mq_state* st, const Group& g, long long* isum)
{
    const int lgL = lg2(g.lat->L);
    const int64_t words = g.lat->N / 64;
    hipLaunchKernelGGL(k_measure_msc, dim3((unsigned)((words + 255) / 256), (unsigned)g.nrep), dim3(256), 0,
                       st->ctx->stream, reinterpret_cast<const uint32_t*>(st->d_spins + g.off), g.stride / 4, lgL,
                       g.rep0, reinterpret_cast<unsigned long long*>(isum));
    return hipGetLastError();
}
