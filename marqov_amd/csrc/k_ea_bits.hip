// Bit-sliced +-J Edwards-Anderson batches: 32 replicas of one disorder sample
// (typically 32 temperatures) share one uint32 per site.
//
// Layout: replicas come in word groups of 32 consecutive replicas on one lattice
// and one sample; word group w stores, per colour and half index h (the int8
// layout's order), one uint32 whose bit j is the spin of replica 32w + j (set =
// down).  The sample's couplings ([colour][d][h] int8) are read once per 32
// replicas.
//
// Sweep (one lane per site word): neighbour words n_k are folded with their bond
// sign (n_k ^ ~0 for b = -1), the anti-aligned count c = sum_k (s ^ n_k) is summed
// bit-sliced into three planes, and each bit j is accepted if its class is
// "always" (dE <= 0) or it wins the bit-serial draw U_j < T_j[c] (MSB first):
// level k uses one random word (bit j = bit 31-k of U_j) and the per-bit
// threshold word W[k][c] whose bit j is bit 31-k of replica (32w+j)'s 32-bit
// threshold for class c.  W is rebuilt from the per-replica tables whenever
// temperatures move (set_beta, parallel-tempering swaps).
//   level k word = Philox(h, 2*sweep + colour, w, 8 + k/4)[k % 4]
#include "mq_internal.hpp"

namespace {

constexpr uint32_t ST_BITS = 8;

struct WTab {
    uint32_t W[32][3];  // per level, per Bernoulli class: bit j = threshold bit of replica j
    uint32_t A[3];      // per class: bit j set when replica j accepts without a draw
    uint32_t pad;
};

// planes (b2 b1 b0) -> mask of words whose count equals V
template <int V>
__device__ __forceinline__ uint32_t cls_mask(uint32_t b0, uint32_t b1, uint32_t b2)
{
    return ((V & 4) ? b2 : ~b2) & ((V & 2) ? b1 : ~b1) & ((V & 1) ? b0 : ~b0);
}

template <int DIM>
__device__ __forceinline__ void nbrs_h(int L, int colour, uint32_t h, uint32_t nb[2 * DIM])
{
    // N < 2^24 for these lattices: 32-bit indices and reciprocal division
    const uint32_t H = (uint32_t)L >> 1;
    const uint32_t plane = H * (uint32_t)L;
    uint32_t k, x1;
    const uint32_t row = fdiv_u24(h, H, 1.0f / (float)H, k);
    uint32_t x2 = 0;
    if (DIM == 3) x2 = fdiv_u24(row, (uint32_t)L, 1.0f / (float)L, x1);
    else x1 = row;
    const uint32_t p = ((uint32_t)colour + x1 + x2) & 1u;
    const uint32_t rb = h - k;
    nb[0] = rb + (p ? ((k + 1 == H) ? 0u : k + 1) : k);
    nb[1] = rb + (p ? k : ((k == 0) ? H - 1 : k - 1));
    nb[2] = (x1 == (uint32_t)L - 1) ? h - ((uint32_t)L - 1) * H : h + H;
    nb[3] = (x1 == 0) ? h + ((uint32_t)L - 1) * H : h - H;
    if (DIM == 3) {
        nb[4] = (x2 == (uint32_t)L - 1) ? h - ((uint32_t)L - 1) * plane : h + plane;
        nb[5] = (x2 == 0) ? h + ((uint32_t)L - 1) * plane : h - plane;
    }
}

// bit-sliced sum of 2*DIM one-bit inputs a[k] = s ^ n[k] -> planes (b0, b1, b2)
template <int DIM>
__device__ __forceinline__ void anti_count(uint32_t s, const uint32_t (&n)[2 * DIM], uint32_t& b0, uint32_t& b1,
                                           uint32_t& b2)
{
    // full adders on the raw neighbour words, then fold s (maj(a^s,b^s,c^s) = maj(a,b,c)^s)
    const uint32_t s1 = n[0] ^ n[1] ^ n[2] ^ s;
    const uint32_t c1 = ((n[0] & n[1]) | (n[0] & n[2]) | (n[1] & n[2])) ^ s;
    if (DIM == 3) {
        const uint32_t s2 = n[3] ^ n[4] ^ n[5] ^ s;
        const uint32_t c2 = ((n[3] & n[4]) | (n[3] & n[5]) | (n[4] & n[5])) ^ s;
        b0 = s1 ^ s2;
        const uint32_t k0 = s1 & s2;
        b1 = c1 ^ c2 ^ k0;
        b2 = (c1 & c2) | (c1 & k0) | (c2 & k0);
    } else {
        const uint32_t a3 = n[3] ^ s;  // 4 inputs: (s1 + 2 c1) + a3
        b0 = s1 ^ a3;
        const uint32_t k0 = s1 & a3;
        b1 = c1 ^ k0;
        b2 = c1 & k0;
    }
}

// one launch covers every lattice size: block ranges per lattice group
struct BitsGroup {
    int64_t off;     // byte offset of the group's first word group
    int64_t wstride; // words between word groups
    int L, gw0, bpw, block0;
};

template <int DIM, bool FERRO>
__global__ void __launch_bounds__(256) k_sweep_eabits(char* __restrict__ spins, const BitsGroup* __restrict__ groups,
                                                      int ngroups, const int8_t* __restrict__ bonds,
                                                      const int64_t* __restrict__ rep_bond_off,
                                                      const WTab* __restrict__ wtab, int colour, uint32_t phase,
                                                      const uint32_t* __restrict__ dyn, Key2 key)
{
    __shared__ WTab T;
    if (dyn) phase += dyn[0];
    int gi = 0;
    while (gi + 1 < ngroups && (int)blockIdx.x >= groups[gi + 1].block0) ++gi;
    const BitsGroup G = groups[gi];
    const int L = G.L, bpw = G.bpw;
    const int lb = (int)blockIdx.x - G.block0;
    const int wl = lb / bpw;  // word group within this lattice group
    const int gw = G.gw0 + wl;
    uint32_t* base = reinterpret_cast<uint32_t*>(spins + G.off);
    const int64_t wstride = G.wstride;
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(wtab + gw);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&T);
        for (int i = threadIdx.x; i < (int)(sizeof(WTab) / 4); i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    int64_t N = (int64_t)L * L;
    if (DIM == 3) N *= L;
    const int64_t half = N >> 1;
    const uint32_t h = (uint32_t)(lb - wl * bpw) * blockDim.x + threadIdx.x;
    if (h >= (uint32_t)half) return;
    const uint32_t hf = (uint32_t)half;
    uint32_t* own = base + (int64_t)wl * wstride + colour * half;
    const uint32_t* oth = base + (int64_t)wl * wstride + (1 - colour) * half;
    const int8_t* bs = bonds + rep_bond_off[32 * gw];
    const int8_t* bown = bs + (int64_t)colour * DIM * half;
    const int8_t* both = bs + (int64_t)(1 - colour) * DIM * half;
    uint32_t nb[2 * DIM];
    nbrs_h<DIM>(L, colour, h, nb);
    uint32_t n[2 * DIM];
#pragma unroll
    for (int sl = 0; sl < 2 * DIM; ++sl) {
        const uint32_t d = (uint32_t)(sl >> 1);
        const int b = (sl & 1) ? both[d * hf + nb[sl]] : bown[d * hf + h];
        n[sl] = oth[nb[sl]] ^ (b < 0 ? 0xFFFFFFFFu : 0u);
    }
    const uint32_t S = own[h];
    uint32_t b0, b1, b2;
    anti_count<DIM>(S, n, b0, b1, b2);
    // Bernoulli classes: ferro c in [0, DIM), antiferro c in (DIM, 2 DIM]
    constexpr int C0 = FERRO ? 0 : DIM + 1;
    const uint32_t m0 = cls_mask<C0>(b0, b1, b2);
    const uint32_t m1 = cls_mask<C0 + 1>(b0, b1, b2);
    const uint32_t m2 = (DIM == 3) ? cls_mask<C0 + 2>(b0, b1, b2) : 0u;
    // classes with dE <= 0: everything that is not a Bernoulli class
    uint32_t acc = ~(m0 | m1 | m2);
    acc |= (m0 & T.A[0]) | (m1 & T.A[1]) | (m2 & T.A[2]);
    uint32_t D = (m0 & ~T.A[0]) | (m1 & ~T.A[1]) | (m2 & ~T.A[2]);
    uint4 w = make_uint4(0, 0, 0, 0);
    for (int k = 0; k < 32 && D != 0u; ++k) {
        if ((k & 3) == 0) w = philox4x32_10(h, phase, (uint32_t)gw, ST_BITS + (uint32_t)(k >> 2), key);
        const uint32_t R = u4_get(w, k & 3);
        const uint32_t t = (T.W[k][0] & m0) | (T.W[k][1] & m1) | (T.W[k][2] & m2);
        acc |= D & ~R & t;
        D &= ~(R ^ t);
    }
    own[h] = S ^ acc;
}

// W tables from the per-replica acceptance tables (entry e = q + 2*DIM, q = 2*DIM - 2c)
__global__ void k_eabits_wtab(const uint32_t* __restrict__ thr, int tbl, int dim, int ferro, int nw,
                              WTab* __restrict__ wtab)
{
    const int gw = blockIdx.x;
    if (gw >= nw) return;
    const int t = threadIdx.x;  // 0..98: (level, class) or the 3 "always" words
    const int z = 2 * dim;
    if (t < 96) {
        const int k = t / 3, j3 = t % 3;
        if (dim == 2 && j3 == 2) {
            wtab[gw].W[k][j3] = 0u;
            return;
        }
        const int c = ferro ? j3 : dim + 1 + j3;
        const int e = 2 * z - 2 * c;
        uint32_t word = 0u;
        for (int j = 0; j < 32; ++j) {
            const uint32_t hi = thr[((int64_t)(32 * gw + j) * tbl + e) * 2];
            const uint32_t lo = thr[((int64_t)(32 * gw + j) * tbl + e) * 2 + 1];
            const uint32_t T32 = hi > 0xFFFFu ? 0xFFFFFFFFu : ((hi << 16) | lo);
            word |= ((T32 >> (31 - k)) & 1u) << j;
        }
        wtab[gw].W[k][j3] = word;
    } else if (t < 99) {
        const int j3 = t - 96;
        if (dim == 2 && j3 == 2) {
            wtab[gw].A[j3] = 0u;
            return;
        }
        const int c = ferro ? j3 : dim + 1 + j3;
        const int e = 2 * z - 2 * c;
        uint32_t word = 0u;
        for (int j = 0; j < 32; ++j)
            if (thr[((int64_t)(32 * gw + j) * tbl + e) * 2] > 0xFFFFu) word |= 1u << j;
        wtab[gw].A[j3] = word;
    }
}

template <int DIM>
__global__ void __launch_bounds__(256) k_init_eabits(uint32_t* __restrict__ base, int64_t wstride, int L, int gw0,
                                                     Key2 key)
{
    const int wl = blockIdx.y;
    int64_t N = (int64_t)L * L;
    if (DIM == 3) N *= L;
    const int64_t half = N >> 1;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int64_t x0 = i % L, rr = i / L;
    int cs = (int)x0 + (int)(rr % L);
    if (DIM == 3) cs += (int)(rr / L);
    const int colour = cs & 1;
    uint32_t word = 0u;
    for (int j = 0; j < 32; ++j) {
        const uint4 q = philox4x32_10((uint32_t)(i >> 2), 0u, (uint32_t)(32 * (gw0 + wl) + j), ST_INIT, key);
        if (!(u4_get(q, (int)(i & 3)) >> 31)) word |= 1u << j;
    }
    base[(int64_t)wl * wstride + colour * half + (i >> 1)] = word;
}

// add the 3-bit per-lane counts (b2 b1 b0) into an NP-plane bit-sliced accumulator
template <int NP>
__device__ __forceinline__ void acc_add3(uint32_t (&A)[NP], uint32_t b0, uint32_t b1, uint32_t b2)
{
    uint32_t carry = 0u;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const uint32_t in = (p == 0) ? b0 : (p == 1) ? b1 : (p == 2) ? b2 : 0u;
        const uint32_t sm = A[p] ^ in ^ carry;
        carry = (A[p] & in) | (A[p] & carry) | (in & carry);
        A[p] = sm;
    }
}

// Per replica: isum[r*8+0] += sum over bonds b s_i s_j, isum[r*8+1] += sum s.
// One launch over every lattice size; a wavefront takes up to 64*MSPL colour-0
// sites of one word of 32 replicas (table entries ngroups.. of the group
// table: bpw = wavefronts per word, block0 = first wavefront).  Lanes fold
// their sites into bit-sliced per-bit counters of anti-aligned bonds and down
// spins, the wavefront reduce-scatters them so lane j holds replica j's
// totals, and lanes 0..31 add them with one atomic each.
constexpr int MSPL = 8;  // sites per lane
template <int DIM>
__global__ void __launch_bounds__(256) k_measure_eabits(const char* __restrict__ spins,
                                                        const BitsGroup* __restrict__ groups, int ngroups, int nwaves,
                                                        const int8_t* __restrict__ bonds,
                                                        const int64_t* __restrict__ rep_bond_off,
                                                        unsigned long long* __restrict__ isum)
{
    constexpr int NE = 7 + 5;  // anti-aligned: <= 2*DIM*MSPL per lane (6 bits), +5 for the exchange steps
    constexpr int ND = 5 + 5;  // down spins: <= 2*MSPL per lane
    const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (wave >= nwaves) return;  // whole wavefront; nothing below synchronises the workgroup
    const int lane = (int)(threadIdx.x & 63);
    int gi = 0;
    while (gi + 1 < ngroups && groups[ngroups + gi + 1].block0 <= wave) ++gi;
    const BitsGroup G = groups[ngroups + gi];
    const int local = wave - G.block0;
    const int wl = local / G.bpw;
    const int chunk = local - wl * G.bpw;
    const int gw = G.gw0 + wl;
    const int L = G.L;
    uint32_t N = (uint32_t)L * (uint32_t)L;
    if (DIM == 3) N *= (uint32_t)L;
    const uint32_t half = N >> 1;
    const uint32_t* c0 = reinterpret_cast<const uint32_t*>(spins + G.off) + (int64_t)wl * G.wstride;
    const uint32_t* c1 = c0 + half;
    const int8_t* bs = bonds + rep_bond_off[32 * gw];
    const uint32_t first = (uint32_t)chunk * 64u * MSPL;
    uint32_t EA[NE], DA[ND];
#pragma unroll
    for (int p = 0; p < NE; ++p) EA[p] = 0u;
#pragma unroll
    for (int p = 0; p < ND; ++p) DA[p] = 0u;
    // All MSPL sites' words are loaded before any is counted, so a lane waits for one
    // memory round trip rather than one per site; sites past the end of the colour
    // read site 0 and are not counted.
    uint32_t n[MSPL][2 * DIM], S[MSPL], T2[MSPL];
    bool in[MSPL];
#pragma unroll
    for (int t = 0; t < MSPL; ++t) {
        const uint32_t hr = first + (uint32_t)t * 64u + (uint32_t)lane;  // coalesced across the wavefront
        in[t] = hr < half;
        const uint32_t h = in[t] ? hr : 0u;
        uint32_t nb[2 * DIM];
        nbrs_h<DIM>(L, 0, h, nb);
#pragma unroll
        for (int sl = 0; sl < 2 * DIM; ++sl) {
            const uint32_t d = (uint32_t)(sl >> 1);
            const int b = (sl & 1) ? bs[(uint32_t)DIM * half + d * half + nb[sl]] : bs[d * half + h];
            n[t][sl] = c1[nb[sl]] ^ (b < 0 ? 0xFFFFFFFFu : 0u);
        }
        S[t] = c0[h];
        T2[t] = c1[h];
    }
#pragma unroll
    for (int t = 0; t < MSPL; ++t) {
        if (!in[t]) continue;
        uint32_t b0, b1, b2;
        anti_count<DIM>(S[t], n[t], b0, b1, b2);
        acc_add3<NE - 5>(reinterpret_cast<uint32_t(&)[NE - 5]>(EA), b0, b1, b2);
        acc_add3<ND - 5>(reinterpret_cast<uint32_t(&)[ND - 5]>(DA), S[t] ^ T2[t], S[t] & T2[t], 0u);  // S + T2 per bit
    }
    const int anti = bs_reduce_scatter<NE>(EA);
    const int down = bs_reduce_scatter<ND>(DA);
    if (lane < 32) {
        const int64_t rem = (int64_t)half - first;
        const int64_t nsites = rem < 64 * MSPL ? rem : 64 * MSPL;  // colour-0 sites of this wavefront
        const int64_t r = 32ll * gw + lane;
        atomicAdd(isum + r * 8 + 0, (unsigned long long)(nsites * 2 * DIM - 2ll * anti));
        atomicAdd(isum + r * 8 + 1, (unsigned long long)(2ll * nsites - 2ll * down));
    }
}

}  // namespace

hipError_t launch_wtab_eabits(mq_state* st)
{
    const int nw = st->nrep / 32;
    hipLaunchKernelGGL(k_eabits_wtab, dim3(nw), dim3(128), 0, st->ctx->stream, st->d_thr, st->tbl, st->lat[0]->dim,
                       st->p.J < 0.0 ? 1 : 0, nw, reinterpret_cast<WTab*>(st->d_wtab));
    return hipGetLastError();
}

size_t wtab_bytes() { return sizeof(WTab); }

hipError_t launch_init_eabits(mq_state* st, const Group& g, Key2 key)
{
    const int L = g.lat->L;
    const int nwl = g.nrep / 32;
    uint32_t* b = reinterpret_cast<uint32_t*>(st->d_spins + g.off);
    const dim3 grid((unsigned)((g.lat->N + 255) / 256), (unsigned)nwl);
    if (g.lat->dim == 3)
        hipLaunchKernelGGL(k_init_eabits<3>, grid, dim3(256), 0, st->ctx->stream, b, g.stride / 4, L, g.rep0 / 32, key);
    else
        hipLaunchKernelGGL(k_init_eabits<2>, grid, dim3(256), 0, st->ctx->stream, b, g.stride / 4, L, g.rep0 / 32, key);
    return hipGetLastError();
}

hipError_t prepare_eabits(mq_state* st)
{
    std::vector<BitsGroup> gs;
    int blocks = 0;
    for (const auto& g : st->groups) {
        BitsGroup d;
        d.off = g.off;
        d.wstride = g.stride / 4;
        d.L = g.lat->L;
        d.gw0 = g.rep0 / 32;
        d.bpw = (int)((g.lat->N / 2 + 255) / 256);
        d.block0 = blocks;
        blocks += d.bpw * (g.nrep / 32);
        gs.push_back(d);
    }
    if (gs.size() > 64) return hipErrorNotSupported;
    const size_t ng = gs.size();
    // measurement entries: bpw = wavefronts per word, block0 = first wavefront
    int waves = 0;
    for (size_t i = 0; i < ng; ++i) {
        const Group& g = st->groups[i];
        BitsGroup d = gs[i];
        d.bpw = (int)((g.lat->N / 2 + 64 * MSPL - 1) / (64 * MSPL));
        d.block0 = waves;
        waves += d.bpw * (g.nrep / 32);
        gs.push_back(d);
    }
    hipError_t e = hipMalloc(&st->d_bits_groups, gs.size() * sizeof(BitsGroup));
    if (e != hipSuccess) return e;
    st->bits_ngroups = (int)ng;
    st->bits_blocks = blocks;
    st->bits_mwaves = waves;
    return hipMemcpy(st->d_bits_groups, gs.data(), gs.size() * sizeof(BitsGroup), hipMemcpyHostToDevice);
}

// one launch per colour over all lattice groups
hipError_t launch_sweep_eabits_all(mq_state* st, int colour, uint32_t phase, Key2 key)
{
    const WTab* wt = reinterpret_cast<const WTab*>(st->d_wtab);
    const BitsGroup* dg = reinterpret_cast<const BitsGroup*>(st->d_bits_groups);
    const dim3 grid((unsigned)st->bits_blocks);
    const bool ferro = st->p.J < 0.0;
#define MQ_E(D, F)                                                                                              \
    hipLaunchKernelGGL((k_sweep_eabits<D, F>), grid, dim3(256), 0, st->ctx->stream, st->d_spins, dg,            \
                       st->bits_ngroups, st->d_bonds, st->d_rep_bond_off, wt, colour, phase, st->cur_dyn, key)
    if (st->lat[0]->dim == 3) {
        if (ferro) MQ_E(3, true);
        else MQ_E(3, false);
    } else {
        if (ferro) MQ_E(2, true);
        else MQ_E(2, false);
    }
#undef MQ_E
    return hipGetLastError();
}

// one fused launch over every lattice group
hipError_t launch_measure_eabits_all(mq_state* st, long long* isum)
{
    const BitsGroup* dg = reinterpret_cast<const BitsGroup*>(st->d_bits_groups);
    const dim3 grid((unsigned)((st->bits_mwaves + 3) / 4));
    unsigned long long* is = reinterpret_cast<unsigned long long*>(isum);
    if (st->lat[0]->dim == 3)
        hipLaunchKernelGGL(k_measure_eabits<3>, grid, dim3(256), 0, st->ctx->stream, st->d_spins, dg, st->bits_ngroups,
                           st->bits_mwaves, st->d_bonds, st->d_rep_bond_off, is);
    else
        hipLaunchKernelGGL(k_measure_eabits<2>, grid, dim3(256), 0, st->ctx->stream, st->d_spins, dg, st->bits_ngroups,
                           st->bits_mwaves, st->d_bonds, st->d_rep_bond_off, is);
    return hipGetLastError();
}
