// Regular-lattice kernels: random start, checkerboard Metropolis sweeps and
// FP64 observable sums for Ising / +-J EA / Potts (int8 spins) and
// Heisenberg / Phi4 (FP32 float4 spins).
//
// Layout in HBM (per replica): the two checkerboard colours are stored as two
// contiguous half-lattices, colour c of site i = (x0+x1+x2)&1, half index
// h = i >> 1 (x0 fastest, so h = (x0>>1) + (L/2)*(x1 + L*x2)).  A colour-c
// site at (row, k) has its four in-plane neighbours at half index h +- L/2
// (x1) and h +- L*L/2 (x2) in the other colour, and its two x0 neighbours at k
// and k-1+2p (p = x0 & 1) of the same row.  One sweep therefore streams two
// reads and one write of the lattice.
//
// EA couplings per disorder sample: [colour][d][h] int8, the +x_d bond of each
// site stored with that site; the -x_d bond of a site is the +x_d bond of its
// -x_d neighbour.
#include "mq_internal.hpp"

namespace {

// natural index -> (colour, half index)
template <int DIM>
__device__ __forceinline__ int reg_colour_of(int L, int64_t i)
{
    const int64_t x0 = i % L;
    const int64_t r = i / L;
    int s = (int)x0 + (int)(r % L);
    if (DIM == 3) s += (int)(r / L);
    return s & 1;
}

// ----------------------------------------------------------------- init
template <int DIM>
__global__ void __launch_bounds__(256) k_init_int8(int8_t* __restrict__ base, int64_t stride, int L, int model, int Q,
                                                   int rep0, Key2 key)
{
    const int r = blockIdx.y;
    const RegGeom G = reg_geom(DIM, L);
    const int64_t N = 2 * G.half;
    const int64_t q4 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q4 * 4 >= N) return;
    const uint4 w = philox4x32_10((uint32_t)q4, 0u, (uint32_t)(rep0 + r), ST_INIT, key);
    int8_t* s = base + r * stride;
    for (int j = 0; j < 4; ++j) {
        const int64_t i = q4 * 4 + j;
        if (i >= N) break;
        const uint32_t x = u4_get(w, j);
        int8_t v;
        if (model == MQ_POTTS) v = (int8_t)(((uint64_t)x * (uint64_t)Q) >> 32);
        else v = (x >> 31) ? (int8_t)1 : (int8_t)-1;
        s[reg_colour_of<DIM>(L, i) * G.half + (i >> 1)] = v;
    }
}


template <int DIM>
__global__ void __launch_bounds__(256) k_init_f4(float4* __restrict__ base, int64_t stride4, int L, int rep0, Key2 key)
{
    const int r = blockIdx.y;
    const RegGeom G = reg_geom(DIM, L);
    const int64_t N = 2 * G.half;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint4 w = philox4x32_10((uint32_t)i, 0u, (uint32_t)(rep0 + r), ST_INIT, key);
    const float3 v = rnddir_f(w.x, w.y);
    base[r * stride4 + reg_colour_of<DIM>(L, i) * G.half + (i >> 1)] = make_float4(v.x, v.y, v.z, 0.0f);
}

// ----------------------------------------------------------- Ising / EA sweep
// One thread updates 8 consecutive half-sites of the active colour with one
// Philox block: 8 x 16-bit uniforms compared against the high half of the
// 32-bit acceptance threshold; an equal high half (probability 2^-16) draws the
// low 16 bits from the tie stream, so the decision equals a full 32-bit compare.
template <int DIM, bool BONDS>
__global__ void __launch_bounds__(512) k_sweep_ising(int8_t* __restrict__ base, int64_t stride,
                                                     const int8_t* __restrict__ bonds,
                                                     const int64_t* __restrict__ rep_bond_off,
                                                     const uint32_t* __restrict__ thr, int tbl, int L, int colour,
                                                     uint32_t phase, const uint32_t* __restrict__ dyn, int rep0, Key2 key)
{
    __shared__ uint32_t s_thr[2 * 16];
    const int r = blockIdx.y;
    const int grep = rep0 + r;
    if (threadIdx.x < 2 * tbl) s_thr[threadIdx.x] = thr[(int64_t)grep * 2 * tbl + threadIdx.x];
    if (dyn) phase += dyn[0];
    __syncthreads();
    const RegGeom G = reg_geom(DIM, L);
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t h0 = g << 3;
    if (h0 >= G.half) return;
    int8_t* own = base + r * stride + colour * G.half;
    const int8_t* oth = base + r * stride + (1 - colour) * G.half;
    const int8_t* bown = nullptr;
    const int8_t* both = nullptr;
    if (BONDS) {
        const int8_t* bs = bonds + rep_bond_off[grep];
        bown = bs + (int64_t)colour * DIM * G.half;
        both = bs + (int64_t)(1 - colour) * DIM * G.half;
    }
    const uint4 w = philox4x32_10((uint32_t)g, phase, (uint32_t)grep, ST_METRO, key);
    uint4 tie = make_uint4(0, 0, 0, 0);
    bool have_tie = false;
    constexpr int QMAX = 2 * DIM;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int64_t h = h0 + m;
        if (h >= G.half) break;
        int64_t nb[2 * DIM];
        reg_nbrs<DIM>(G, colour, h, nb);
        int hs = 0;
#pragma unroll
        for (int sl = 0; sl < 2 * DIM; ++sl) {
            int b = 1;
            if (BONDS) {
                const int d = sl >> 1;
                b = (sl & 1) ? both[(int64_t)d * G.half + nb[sl]] : bown[(int64_t)d * G.half + h];
            }
            hs += b * oth[nb[sl]];
        }
        const int s = own[h];
        const int e = s * hs + QMAX;
        const uint32_t hi16 = (u4_get(w, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
        const uint32_t thi = s_thr[2 * e];
        bool acc = hi16 < thi;
        if (hi16 == thi) {
            if (!have_tie) {
                tie = philox4x32_10((uint32_t)g, phase, (uint32_t)grep, ST_TIE, key);
                have_tie = true;
            }
            acc = ((u4_get(tie, m >> 1) >> (16 * (m & 1))) & 0xFFFFu) < s_thr[2 * e + 1];
        }
        if (acc) own[h] = (int8_t)(-s);
    }
}


// ------------------------------------------- 3D Ising, row-vectorized sweep
// Specialisation for 3D Ising without couplings and L % 32 == 0.  A thread
// owns 16 consecutive half-sites of one row (one 16-byte load/store); its six
// neighbour chunks are the other colour's chunks at the same position in rows
// y+-1 and z+-1, the chunk itself, and the chunk shifted by one byte toward
// -x0 (p = 0) or +x0 (p = 1), completed by one dword from the adjacent chunk.
// Spins are +-1 bytes (0x01 / 0xFF); bit 1 of a byte is set iff the spin is
// -1, so the doubled count of down neighbours is a byte-wise sum of
// (word & 0x02020202), and the table index e = q + 6 (q = s_i h_i) is
// 2*aligned = own_down ? 2n : 12 - 2n per byte.  Random numbers and
// acceptance are those of k_sweep_ising (group g = h >> 3).
// The block -> tile mapping gives each XCD a contiguous slab of rows so the
// y+-1 / z+-1 rows are re-read from the same L2.
__global__ void __launch_bounds__(512) k_sweep_ising3_row16(int8_t* __restrict__ base, int64_t stride,
                                                            const uint32_t* __restrict__ thr, int lgL, int colour,
                                                            uint32_t phase, const uint32_t* __restrict__ dyn, int rep0,
                                                            int lgbpr, int lgblk, int xcd, Key2 key)
{
    if (dyn) phase += dyn[0];
    __shared__ uint32_t s_thi[13];
    __shared__ uint32_t s_tlo[13];
    const uint32_t B = gridDim.x;
    const uint32_t b = blockIdx.x;
    const uint32_t lb0 = (xcd && (B & 7u) == 0u) ? (b & 7u) * (B >> 3) + (b >> 3) : b;
    // colour 1 walks the blocks backwards, so every pass starts where the previous
    // one ended: those lines are still in the 256 MB last-level cache
    const uint32_t lb = colour ? B - 1u - lb0 : lb0;
    const uint32_t r = lb >> lgbpr;
    const uint32_t grep = (uint32_t)rep0 + r;
    if (threadIdx.x < 13) {
        const uint32_t hi = thr[grep * 26u + 2u * threadIdx.x];
        const uint32_t lo = thr[grep * 26u + 2u * threadIdx.x + 1u];
        // "always" (hi = 65536) becomes hi = 0xFFFF, lo = 0x10000: same decision
        s_thi[threadIdx.x] = hi > 0xFFFFu ? 0xFFFFu : hi;
        s_tlo[threadIdx.x] = hi > 0xFFFFu ? 0x10000u : lo;
    }
    __syncthreads();
    const uint32_t chunk = ((lb - (r << lgbpr)) << lgblk) + threadIdx.x;
    row16_chunk(base + (int64_t)r * stride, chunk, lgL, colour, phase, grep, s_thi, s_tlo, key);
}


// ------------------------------------------- LDS-resident multi-sweep kernel
// Small lattices (whole replica plus its couplings fit in LDS): one workgroup
// per replica loads the spins (and EA couplings) into LDS once, runs nsweeps
// checkerboard sweeps there with a barrier between colours, and writes the
// spins back.  One launch covers the replicas of every lattice size through a
// small descriptor table.  Random numbers and acceptance are those of
// k_sweep_ising (group g = h >> 3, phase 2*sweep + colour), so the result is
// bit-identical to nsweeps launches of the streaming kernel.
struct LdsGroup {
    int64_t off;     // byte offset of the group's first replica
    int64_t stride;  // bytes per replica
    int L;
    // ceil(2^32 / (L/2)), ceil(2^32 / L): exact division for indices < 2^16;
    // L = 2 has no 32-bit magic for L/2 = 1 (magic_h = 0 there, row = h)
    uint32_t magic_h, magic_l;
    int rep0, nrep, block0;
};

template <int DIM>
__device__ __forceinline__ void lds_nbrs(int L, int H, uint32_t mh, uint32_t ml, int colour, int h, int nb[2 * DIM])
{
    const int row = (H == 1) ? h : (int)__umulhi((uint32_t)h, mh);
    const int k = h - row * H;
    int x1 = row, x2 = 0;
    if (DIM == 3) {
        x2 = (int)__umulhi((uint32_t)row, ml);
        x1 = row - x2 * L;
    }
    const int plane = H * L;
    const int p = (colour + x1 + x2) & 1;
    const int rb = h - k;
    nb[0] = rb + (p ? ((k + 1 == H) ? 0 : k + 1) : k);
    nb[1] = rb + (p ? k : ((k == 0) ? H - 1 : k - 1));
    nb[2] = (x1 == L - 1) ? h - (L - 1) * H : h + H;
    nb[3] = (x1 == 0) ? h + (L - 1) * H : h - H;
    if (DIM == 3) {
        nb[4] = (x2 == L - 1) ? h - (L - 1) * plane : h + plane;
        nb[5] = (x2 == 0) ? h + (L - 1) * plane : h - plane;
    }
}

template <int DIM, bool BONDS>
__global__ void __launch_bounds__(512) k_sweep_lds(char* __restrict__ spins, const LdsGroup* __restrict__ groups,
                                                   int ngroups, const int8_t* __restrict__ bonds,
                                                   const int64_t* __restrict__ rep_bond_off,
                                                   const uint32_t* __restrict__ thr, int tbl, int nsweeps,
                                                   uint32_t sweep0, const uint32_t* __restrict__ dyn, Key2 key)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t s_thr[2 * 16];
    int gi = 0;
    while (gi + 1 < ngroups && (int)blockIdx.x >= groups[gi + 1].block0) ++gi;
    const LdsGroup G = groups[gi];
    const int r = (int)blockIdx.x - G.block0;
    const int grep = G.rep0 + r;
    const int L = G.L, H = L >> 1;
    int N = L * L;
    if (DIM == 3) N *= L;
    const int half = N >> 1;
    int8_t* S = reinterpret_cast<int8_t*>(smem);
    int8_t* Bd = S + ((N + 15) & ~15);
    int8_t* gs = reinterpret_cast<int8_t*>(spins + G.off + (int64_t)r * G.stride);
    if (threadIdx.x < 2 * tbl) s_thr[threadIdx.x] = thr[(int64_t)grep * 2 * tbl + threadIdx.x];
    for (int i = threadIdx.x; i < N; i += blockDim.x) S[i] = gs[i];
    if (BONDS) {
        const int8_t* gb = bonds + rep_bond_off[grep];
        for (int i = threadIdx.x; i < DIM * N; i += blockDim.x) Bd[i] = gb[i];
    }
    uint32_t phase0 = 2u * sweep0;
    if (dyn) phase0 += dyn[0];
    __syncthreads();
    constexpr int QMAX = 2 * DIM;
    const int ngr = (half + 7) >> 3;
    for (int t = 0; t < nsweeps; ++t) {
        for (int colour = 0; colour < 2; ++colour) {
            const uint32_t phase = phase0 + 2u * (uint32_t)t + (uint32_t)colour;
            int8_t* own = S + colour * half;
            const int8_t* oth = S + (1 - colour) * half;
            const int8_t* bown = Bd + colour * DIM * half;
            const int8_t* both = Bd + (1 - colour) * DIM * half;
            for (int g = threadIdx.x; g < ngr; g += blockDim.x) {
                const uint4 w = philox4x32_10((uint32_t)g, phase, (uint32_t)grep, ST_METRO, key);
                bool have_tie = false;
                uint4 tie = make_uint4(0, 0, 0, 0);
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const int h = 8 * g + m;
                    if (h >= half) break;
                    int nb[2 * DIM];
                    lds_nbrs<DIM>(L, H, G.magic_h, G.magic_l, colour, h, nb);
                    int hs = 0;
#pragma unroll
                    for (int sl = 0; sl < 2 * DIM; ++sl) {
                        int b = 1;
                        if (BONDS) b = (sl & 1) ? both[(sl >> 1) * half + nb[sl]] : bown[(sl >> 1) * half + h];
                        hs += b * oth[nb[sl]];
                    }
                    const int s = own[h];
                    const int e = s * hs + QMAX;
                    const uint32_t hi16 = (u4_get(w, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
                    const uint32_t thi = s_thr[2 * e];
                    bool acc = hi16 < thi;
                    if (hi16 == thi) {
                        if (!have_tie) {
                            tie = philox4x32_10((uint32_t)g, phase, (uint32_t)grep, ST_TIE, key);
                            have_tie = true;
                        }
                        acc = ((u4_get(tie, m >> 1) >> (16 * (m & 1))) & 0xFFFFu) < s_thr[2 * e + 1];
                    }
                    if (acc) own[h] = (int8_t)(-s);
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < N; i += blockDim.x) gs[i] = S[i];
}

// ----------------------------------------------------------------- Potts sweep
// One Philox block per site: word 0 -> proposal in [0,Q) (Potts.h:171-177),
// word 1 -> 32-bit uniform against the threshold of dE = k*(#new - #old).
template <int DIM>
__global__ void __launch_bounds__(512) k_sweep_potts(int8_t* __restrict__ base, int64_t stride,
                                                     const uint32_t* __restrict__ thr, int tbl, int L, int Q,
                                                     int colour, uint32_t phase, const uint32_t* __restrict__ dyn,
                                                     int rep0, Key2 key)
{
    __shared__ uint32_t s_thr[2 * 16];
    if (dyn) phase += dyn[0];
    const int r = blockIdx.y;
    const int grep = rep0 + r;
    if (threadIdx.x < 2 * tbl) s_thr[threadIdx.x] = thr[(int64_t)grep * 2 * tbl + threadIdx.x];
    __syncthreads();
    const RegGeom G = reg_geom(DIM, L);
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= G.half) return;
    int8_t* own = base + r * stride + colour * G.half;
    const int8_t* oth = base + r * stride + (1 - colour) * G.half;
    int64_t nb[2 * DIM];
    reg_nbrs<DIM>(G, colour, h, nb);
    const uint4 w = philox4x32_10((uint32_t)h, phase, (uint32_t)grep, ST_METRO, key);
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

// ------------------------------------------------------- Heisenberg / Phi4
// One Philox block per site: words 0,1 -> direction, word 2 -> acceptance.
// Local arithmetic in FP32; acceptance u < exp(-beta*dE).
template <int DIM, bool BONDS>
__global__ void __launch_bounds__(512) k_sweep_o3(float4* __restrict__ base, int64_t stride4,
                                                  const int8_t* __restrict__ bonds,
                                                  const int64_t* __restrict__ rep_bond_off,
                                                  const float* __restrict__ betaf, int L, int model, int proposal,
                                                  float J, float delta, float lambda, float mass, int colour,
                                                  uint32_t phase, const uint32_t* __restrict__ dyn, int rep0, int xcd,
                                                  Key2 key)
{
    if (dyn) phase += dyn[0];
    // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs in launch
    // order; giving each XCD a contiguous run of tiles keeps the z+-1 planes of
    // a tile in the same L2
    uint32_t bx = blockIdx.x, by = blockIdx.y;
    const uint32_t nblk = gridDim.x * gridDim.y;
    if (xcd && (nblk & 7u) == 0u) {
        const uint32_t lin = blockIdx.x + blockIdx.y * gridDim.x;
        const uint32_t lb = (lin & 7u) * (nblk >> 3) + (lin >> 3);
        by = lb / gridDim.x;
        bx = lb - by * gridDim.x;
    }
    if (colour) {  // backwards, so the pass starts on what the previous one left in the last-level cache
        const uint32_t lin = nblk - 1u - (bx + by * gridDim.x);
        by = lin / gridDim.x;
        bx = lin - by * gridDim.x;
    }
    const int r = (int)by;
    const int grep = rep0 + r;
    const RegGeom G = reg_geom(DIM, L);
    const int64_t h = (int64_t)bx * blockDim.x + threadIdx.x;
    if (h >= G.half) return;
    float4* own = base + r * stride4 + colour * G.half;
    const float4* oth = base + r * stride4 + (1 - colour) * G.half;
    const int8_t* bs = BONDS ? bonds + rep_bond_off[grep] : nullptr;
    o3_site_update<DIM, BONDS>(own, oth, bs, G, colour, h, phase, (uint32_t)grep, betaf[grep], model, proposal, J, delta,
                               lambda, mass, key);
}

// --------------------------------------------------------------- measure
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ void atomic_add_any(double* p, double v) { atomicAdd(p, v); }
__device__ __forceinline__ void atomic_add_any(long long* p, long long v)
{
    atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}

template <typename T, int K>
__device__ __forceinline__ void block_sum_atomic(T (&v)[K], T* out)
{
    __shared__ T red[K][4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        T x = wave_sum(v[k]);
        if (lane == 0) red[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        T s = 0;
        for (int wv = 0; wv < (int)(blockDim.x >> 6); ++wv) s += red[threadIdx.x][wv];
        atomic_add_any(out + threadIdx.x, s);
    }
}

// isum[r*8 + 0] = sum over bonds of b s_i s_j (Ising/EA) or #equal bonds (Potts)
// isum[r*8 + 1] = sum s (Ising/EA); Potts: isum[r*8 + 1 + q] = count of colour q
// for q < 7 (with q = 8 colours, colour 7 is N minus the others)
template <int DIM, bool BONDS>
__global__ void __launch_bounds__(256) k_measure_int8(const int8_t* __restrict__ base, int64_t stride,
                                                      const int8_t* __restrict__ bonds,
                                                      const int64_t* __restrict__ rep_bond_off, int L, int model,
                                                      int Q, int rep0, unsigned long long* __restrict__ isum)
{
    const int r = blockIdx.y;
    const int grep = rep0 + r;
    const RegGeom G = reg_geom(DIM, L);
    // grid-stride over the colour-0 half: a few blocks per replica, so the
    // per-block atomics stay few
    long long acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int8_t* c0 = base + r * stride;
    const int8_t* c1 = c0 + G.half;
    const int8_t* bs = BONDS ? bonds + rep_bond_off[grep] : nullptr;
    for (int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; h < G.half; h += (int64_t)gridDim.x * blockDim.x) {
        int64_t nb[2 * DIM];
        reg_nbrs<DIM>(G, 0, h, nb);
        const int s = c0[h];
        if (model == MQ_POTTS) {
            int eq = 0;
#pragma unroll
            for (int sl = 0; sl < 2 * DIM; ++sl) eq += (c1[nb[sl]] == s);
            acc[0] += eq;
            const int t = c1[h];
#pragma unroll
            for (int q = 0; q < 7; ++q)  // static indices keep acc in registers
                if (q < Q) acc[1 + q] += (s == q) + (t == q);
        } else {
            int hs = 0;
#pragma unroll
            for (int sl = 0; sl < 2 * DIM; ++sl) {
                int b = 1;
                if (BONDS) {
                    const int d = sl >> 1;
                    b = (sl & 1) ? bs[(int64_t)DIM * G.half + (int64_t)d * G.half + nb[sl]] : bs[(int64_t)d * G.half + h];
                }
                hs += b * c1[nb[sl]];
            }
            acc[0] += s * hs;
            acc[1] += s + c1[h];
        }
    }
    block_sum_atomic<long long, 8>(acc, reinterpret_cast<long long*>(isum) + (int64_t)grep * 8);
}

// Observables of 3D Ising without couplings, L a power of two >= 32, with the
// loads of k_sweep_ising3_row16 for colour 0: a thread takes 16 colour-0 sites
// of one row, forms 2*aligned per byte from the same SWAR neighbour sums (every
// bond is seen once, from its colour-0 end), and counts the down spins of its
// chunk in both colours.  isum[r*8] += sum over bonds s_i s_j, isum[r*8+1] +=
// sum s.
__global__ void __launch_bounds__(256) k_measure_ising3_row16(const int8_t* __restrict__ base, int64_t stride, int lgL,
                                                              int rep0, int lgbpr, int lgit,
                                                              unsigned long long* __restrict__ isum)
{
    const uint32_t lb = blockIdx.x;
    const uint32_t r = lb >> lgbpr;
    const uint32_t grep = (uint32_t)rep0 + r;
    const uint32_t L = 1u << lgL;
    const uint32_t H = L >> 1;
    const uint32_t lgcpr = (uint32_t)lgL - 5u;
    const uint32_t lgplane = 2u * (uint32_t)lgL - 1u;
    const uint32_t plane = 1u << lgplane;
    const uint32_t half = plane << lgL;
    const int8_t* own = base + (int64_t)r * stride;
    const int8_t* oth = own + half;
    const uint32_t M2 = 0x02020202u;
    long long v[2] = {0, 0};
    // the block takes 2^lgit runs of 256 consecutive chunks, one atomic pair in the end
    for (uint32_t it = 0; it < (1u << lgit); ++it) {
        const uint32_t chunk = ((((lb - (r << lgbpr)) << lgit) + it) << 8) + threadIdx.x;
        const uint32_t row = chunk >> lgcpr;
        const uint32_t c = chunk & ((1u << lgcpr) - 1u);
        const uint32_t y = row & (L - 1u);
        const uint32_t z = row >> lgL;
        const uint32_t p = (y + z) & 1u;
        const uint32_t rb = row * H;
        const uint32_t h0 = rb + (c << 4);
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
        auto aligned2 = [&](uint32_t s, uint32_t o, uint32_t xs, uint32_t ym, uint32_t yp, uint32_t zm, uint32_t zp) {
            const uint32_t n2 = (o & M2) + (xs & M2) + (ym & M2) + (yp & M2) + (zm & M2) + (zp & M2);
            const uint32_t mask = ((s & M2) >> 1) * 0xFFu;
            return (n2 & mask) | ((0x0C0C0C0Cu - n2) & ~mask);  // 2 * aligned neighbours, 0..12 per byte
        };
        const uint32_t a4 = aligned2(S.x, O.x, xs0, Ym.x, Yp.x, Zm.x, Zp.x) +
                            aligned2(S.y, O.y, xs1, Ym.y, Yp.y, Zm.y, Zp.y) +
                            aligned2(S.z, O.z, xs2, Ym.z, Yp.z, Zm.z, Zp.z) +
                            aligned2(S.w, O.w, xs3, Ym.w, Yp.w, Zm.w, Zp.w);
        const uint32_t t2 = (a4 & 0x00FF00FFu) + ((a4 >> 8) & 0x00FF00FFu);
        // s_i h_i = aligned - anti-aligned = 2 aligned - 6 per site
        v[0] += (long long)((t2 & 0xFFFFu) + (t2 >> 16)) - 96;
        const int down = __popc(S.x & M2) + __popc(S.y & M2) + __popc(S.z & M2) + __popc(S.w & M2) +
                         __popc(O.x & M2) + __popc(O.y & M2) + __popc(O.z & M2) + __popc(O.w & M2);
        v[1] += 32 - 2 * down;
    }
    block_sum_atomic<long long, 2>(v, reinterpret_cast<long long*>(isum) + (int64_t)grep * 8);
}


// fsum[r*8 + 0] = sum over bonds b s_i.s_j, [1..3] = M, [4] = sum |s|^2, [5] = sum (|s|^2-1)^2
template <int DIM, bool BONDS>
__global__ void __launch_bounds__(256) k_measure_f4(const float4* __restrict__ base, int64_t stride4,
                                                    const int8_t* __restrict__ bonds,
                                                    const int64_t* __restrict__ rep_bond_off, int L, int rep0,
                                                    double* __restrict__ fsum)
{
    const int r = blockIdx.y;
    const int grep = rep0 + r;
    const RegGeom G = reg_geom(DIM, L);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const float4* c0 = base + r * stride4;
    const float4* c1 = c0 + G.half;
    const int8_t* bb = BONDS ? bonds + rep_bond_off[grep] : nullptr;
    for (int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; h < G.half; h += (int64_t)gridDim.x * blockDim.x) {
        int64_t nb[2 * DIM];
        reg_nbrs<DIM>(G, 0, h, nb);
        const float4 s = c0[h];
        const float4 t = c1[h];
        double bs = 0;
#pragma unroll
        for (int sl = 0; sl < 2 * DIM; ++sl) {
            const float4 v = c1[nb[sl]];
            double b = 1.0;
            if (BONDS) {
                const int d = sl >> 1;
                b = (sl & 1) ? bb[(int64_t)DIM * G.half + (int64_t)d * G.half + nb[sl]] : bb[(int64_t)d * G.half + h];
            }
            bs += b * ((double)s.x * v.x + (double)s.y * v.y + (double)s.z * v.z);
        }
        acc[0] += bs;
        acc[1] += (double)s.x + t.x;
        acc[2] += (double)s.y + t.y;
        acc[3] += (double)s.z + t.z;
        const double n2s = (double)s.x * s.x + (double)s.y * s.y + (double)s.z * s.z;
        const double n2t = (double)t.x * t.x + (double)t.y * t.y + (double)t.z * t.z;
        acc[4] += n2s + n2t;
        acc[5] += (n2s - 1) * (n2s - 1) + (n2t - 1) * (n2t - 1);
    }
    block_sum_atomic<double, 6>(acc, fsum + (int64_t)grep * 8);
}

// MagFTComp sums (observables.h:104-170, with the reference's site coordinates
// (x + 1/2)/L, regular_hypercubic.h:73-85): ft[r*18 + (d*3 + j)*2 + {0, 1}] =
// sum_i s_i[j] * {cos, sin}(2 pi (x_d + 1/2) / L) for axes d < DIM and spin
// components j < NC (1 for int8 spins, 3 for float4), over both colours, in FP64.
template <int DIM, bool VEC>
__global__ void __launch_bounds__(256) k_measure_ft(const char* __restrict__ base, int64_t stride, int L, int rep0,
                                                    double* __restrict__ ft)
{
    constexpr int NC = VEC ? 3 : 1;
    const int r = blockIdx.y;
    const int grep = rep0 + r;
    const RegGeom G = reg_geom(DIM, L);
    double acc[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) acc[k] = 0.0;
    const double twoL = 2.0 / (double)L;
    const char* rb = base + r * stride;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < 2 * G.half;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int c = t >= G.half ? 1 : 0;
        const int64_t h = t - c * G.half;
        const int64_t row = h / G.H;
        const int k = (int)(h - row * G.H);
        const int x1 = (int)(row % G.L);
        const int x2 = (DIM == 3) ? (int)(row / G.L) : 0;
        const int x[3] = {2 * k + ((c + x1 + x2) & 1), x1, x2};
        double v[3] = {0.0, 0.0, 0.0};
        if (VEC) {
            const float4 s = reinterpret_cast<const float4*>(rb)[t];
            v[0] = s.x;
            v[1] = s.y;
            v[2] = s.z;
        } else {
            v[0] = (double)reinterpret_cast<const int8_t*>(rb)[t];
        }
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            double sn, cs;
            sincospi(((double)x[d] + 0.5) * twoL, &sn, &cs);
#pragma unroll
            for (int j = 0; j < NC; ++j) {
                acc[(d * 3 + j) * 2] += v[j] * cs;
                acc[(d * 3 + j) * 2 + 1] += v[j] * sn;
            }
        }
    }
    block_sum_atomic<double, 18>(acc, ft + (int64_t)grep * 18);
}

// EA couplings for one sample: natural bond b = i*dim + d drawn from
// (seed, sample, ST_DISORDER), +1 with probability T/2^32.
template <int DIM>
__global__ void __launch_bounds__(256) k_bonds(int8_t* __restrict__ out, int L, uint32_t T, int sample, Key2 key)
{
    const RegGeom G = reg_geom(DIM, L);
    const int64_t N = 2 * G.half;
    const int64_t q4 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q4 * 4 >= N * DIM) return;
    const uint4 w = philox4x32_10((uint32_t)q4, 0u, (uint32_t)sample, ST_DISORDER, key);
    for (int j = 0; j < 4; ++j) {
        const int64_t b = q4 * 4 + j;
        if (b >= N * DIM) break;
        const int64_t i = b / DIM;
        const int d = (int)(b - i * DIM);
        out[(int64_t)reg_colour_of<DIM>(L, i) * DIM * G.half + (int64_t)d * G.half + (i >> 1)] =
            (u4_get(w, j) < T) ? (int8_t)1 : (int8_t)-1;
    }
}

inline dim3 grid_for(int64_t work, int block, int nrep)
{
    return dim3((unsigned)((work + block - 1) / block), (unsigned)nrep, 1);
}

inline bool is_vec(int model) { return model == MQ_HEISENBERG || model == MQ_PHI4; }

}  // namespace

hipError_t launch_init_regular(mq_state* st, const Group& g, Key2 key)
{
    const int L = g.lat->L, dim = g.lat->dim;
    const int64_t N = g.lat->N;
    hipStream_t s = st->ctx->stream;
    if (is_vec(st->p.model)) {
        float4* b = reinterpret_cast<float4*>(st->d_spins + g.off);
        const int64_t s4 = g.stride / 16;
        if (dim == 3) hipLaunchKernelGGL(k_init_f4<3>, grid_for(N, 256, g.nrep), dim3(256), 0, s, b, s4, L, g.rep0, key);
        else hipLaunchKernelGGL(k_init_f4<2>, grid_for(N, 256, g.nrep), dim3(256), 0, s, b, s4, L, g.rep0, key);
    } else {
        int8_t* b = reinterpret_cast<int8_t*>(st->d_spins + g.off);
        if (dim == 3)
            hipLaunchKernelGGL(k_init_int8<3>, grid_for((N + 3) / 4, 256, g.nrep), dim3(256), 0, s, b, g.stride, L,
                               st->p.model, st->p.q, g.rep0, key);
        else
            hipLaunchKernelGGL(k_init_int8<2>, grid_for((N + 3) / 4, 256, g.nrep), dim3(256), 0, s, b, g.stride, L,
                               st->p.model, st->p.q, g.rep0, key);
    }
    return hipGetLastError();
}

hipError_t launch_sweep_regular(mq_state* st, const Group& g, int colour, uint32_t phase, Key2 key)
{
    const int L = g.lat->L, dim = g.lat->dim;
    const int64_t half = g.lat->N / 2;
    hipStream_t s = st->ctx->stream;
    const int model = st->p.model;
    const bool bonds = (model == MQ_EA_ISING) || (st->d_bonds != nullptr);
    if (model == MQ_ISING || model == MQ_EA_ISING) {
        int8_t* b = reinterpret_cast<int8_t*>(st->d_spins + g.off);
        const int blk = st->tune_block;
        const dim3 grid = grid_for((half + 7) / 8, blk, g.nrep);
        const bool pow2 = (L & (L - 1)) == 0;
        if (dim == 3 && !bonds && pow2 && L >= 32 && L <= 1024 && st->tbl == 13 && st->kernel_choice != 2) {
            int lgL = 0;
            while ((1 << lgL) < L) ++lgL;
            int lgblk = 0;
            while ((1 << lgblk) < st->tune_block) ++lgblk;
            const int lgbpr = 3 * lgL - 5 - lgblk;  // (N/2)/16 chunks per replica / block
            const int64_t nblk = ((int64_t)1 << std::max(lgbpr, 0)) * g.nrep;
            if (lgbpr >= 0 && nblk < (1ll << 31)) {
                hipLaunchKernelGGL(k_sweep_ising3_row16, dim3((unsigned)nblk), dim3(1 << lgblk), 0, s, b, g.stride,
                                   st->d_thr, lgL, colour, phase, st->cur_dyn, g.rep0, lgbpr, lgblk, st->tune_xcd, key);
                return hipGetLastError();
            }
        }
        if (dim == 3) {
            if (bonds)
                hipLaunchKernelGGL((k_sweep_ising<3, true>), grid, dim3(blk), 0, s, b, g.stride, st->d_bonds,
                                   st->d_rep_bond_off, st->d_thr, st->tbl, L, colour, phase, st->cur_dyn, g.rep0, key);
            else
                hipLaunchKernelGGL((k_sweep_ising<3, false>), grid, dim3(blk), 0, s, b, g.stride, st->d_bonds,
                                   st->d_rep_bond_off, st->d_thr, st->tbl, L, colour, phase, st->cur_dyn, g.rep0, key);
        } else {
            if (bonds)
                hipLaunchKernelGGL((k_sweep_ising<2, true>), grid, dim3(blk), 0, s, b, g.stride, st->d_bonds,
                                   st->d_rep_bond_off, st->d_thr, st->tbl, L, colour, phase, st->cur_dyn, g.rep0, key);
            else
                hipLaunchKernelGGL((k_sweep_ising<2, false>), grid, dim3(blk), 0, s, b, g.stride, st->d_bonds,
                                   st->d_rep_bond_off, st->d_thr, st->tbl, L, colour, phase, st->cur_dyn, g.rep0, key);
        }
    } else if (model == MQ_POTTS) {
        int8_t* b = reinterpret_cast<int8_t*>(st->d_spins + g.off);
        const int blk = st->tune_block;
        const dim3 grid = grid_for(half, blk, g.nrep);
        if (dim == 3)
            hipLaunchKernelGGL(k_sweep_potts<3>, grid, dim3(blk), 0, s, b, g.stride, st->d_thr, st->tbl, L, st->p.q,
                               colour, phase, st->cur_dyn, g.rep0, key);
        else
            hipLaunchKernelGGL(k_sweep_potts<2>, grid, dim3(blk), 0, s, b, g.stride, st->d_thr, st->tbl, L, st->p.q,
                               colour, phase, st->cur_dyn, g.rep0, key);
    } else {
        float4* b = reinterpret_cast<float4*>(st->d_spins + g.off);
        const int64_t s4 = g.stride / 16;
        const int blk = st->tune_block;
        const dim3 grid = grid_for(half, blk, g.nrep);
        const float J = (float)st->p.J, delta = (float)st->p.delta, lam = (float)st->p.lambda, mass = (float)st->p.mass;
        if (dim == 3) {
            if (bonds)
                hipLaunchKernelGGL((k_sweep_o3<3, true>), grid, dim3(blk), 0, s, b, s4, st->d_bonds, st->d_rep_bond_off,
                                   st->d_betaf, L, model, st->p.proposal, J, delta, lam, mass, colour, phase, st->cur_dyn, g.rep0,
                                   st->tune_xcd, key);
            else
                hipLaunchKernelGGL((k_sweep_o3<3, false>), grid, dim3(blk), 0, s, b, s4, st->d_bonds,
                                   st->d_rep_bond_off, st->d_betaf, L, model, st->p.proposal, J, delta, lam, mass,
                                   colour, phase, st->cur_dyn, g.rep0, st->tune_xcd, key);
        } else {
            if (bonds)
                hipLaunchKernelGGL((k_sweep_o3<2, true>), grid, dim3(blk), 0, s, b, s4, st->d_bonds, st->d_rep_bond_off,
                                   st->d_betaf, L, model, st->p.proposal, J, delta, lam, mass, colour, phase, st->cur_dyn, g.rep0,
                                   st->tune_xcd, key);
            else
                hipLaunchKernelGGL((k_sweep_o3<2, false>), grid, dim3(blk), 0, s, b, s4, st->d_bonds,
                                   st->d_rep_bond_off, st->d_betaf, L, model, st->p.proposal, J, delta, lam, mass,
                                   colour, phase, st->cur_dyn, g.rep0, st->tune_xcd, key);
        }
    }
    return hipGetLastError();
}

hipError_t launch_measure_ft_regular(mq_state* st, const Group& g)
{
    const int L = g.lat->L, dim = g.lat->dim;
    const int64_t N = g.lat->N;
    const int64_t per_rep = std::max<int64_t>(1, std::min<int64_t>((N + 255) / 256,
                                                                   (8ll * st->ctx->num_cus + g.nrep - 1) / g.nrep));
    const dim3 grid((unsigned)per_rep, (unsigned)g.nrep, 1);
    const char* b = st->d_spins + g.off;
    hipStream_t s = st->ctx->stream;
    const bool vec = st->site_bytes == 16;
    if (dim == 3) {
        if (vec) hipLaunchKernelGGL((k_measure_ft<3, true>), grid, dim3(256), 0, s, b, g.stride, L, g.rep0, st->d_ft);
        else hipLaunchKernelGGL((k_measure_ft<3, false>), grid, dim3(256), 0, s, b, g.stride, L, g.rep0, st->d_ft);
    } else {
        if (vec) hipLaunchKernelGGL((k_measure_ft<2, true>), grid, dim3(256), 0, s, b, g.stride, L, g.rep0, st->d_ft);
        else hipLaunchKernelGGL((k_measure_ft<2, false>), grid, dim3(256), 0, s, b, g.stride, L, g.rep0, st->d_ft);
    }
    return hipGetLastError();
}

hipError_t launch_measure_regular(mq_state* st, const Group& g, long long* isum, double* fsum)
{
    const int L = g.lat->L, dim = g.lat->dim;
    const int64_t half = g.lat->N / 2;
    hipStream_t s = st->ctx->stream;
    const int model = st->p.model;
    const bool bonds = (model == MQ_EA_ISING) || (st->d_bonds != nullptr);
    // about 8 blocks per CU in total, at least one per replica
    const int64_t per_rep = std::max<int64_t>(1, std::min<int64_t>((half + 255) / 256,
                                                                   (8ll * st->ctx->num_cus + g.nrep - 1) / g.nrep));
    const dim3 grid((unsigned)per_rep, (unsigned)g.nrep, 1);
    if (is_vec(model)) {
        const float4* b = reinterpret_cast<const float4*>(st->d_spins + g.off);
        const int64_t s4 = g.stride / 16;
        if (dim == 3) {
            if (bonds) hipLaunchKernelGGL((k_measure_f4<3, true>), grid, dim3(256), 0, s, b, s4, st->d_bonds, st->d_rep_bond_off, L, g.rep0, fsum);
            else hipLaunchKernelGGL((k_measure_f4<3, false>), grid, dim3(256), 0, s, b, s4, st->d_bonds, st->d_rep_bond_off, L, g.rep0, fsum);
        } else {
            if (bonds) hipLaunchKernelGGL((k_measure_f4<2, true>), grid, dim3(256), 0, s, b, s4, st->d_bonds, st->d_rep_bond_off, L, g.rep0, fsum);
            else hipLaunchKernelGGL((k_measure_f4<2, false>), grid, dim3(256), 0, s, b, s4, st->d_bonds, st->d_rep_bond_off, L, g.rep0, fsum);
        }
    } else {
        const int8_t* b = reinterpret_cast<const int8_t*>(st->d_spins + g.off);
        unsigned long long* is = reinterpret_cast<unsigned long long*>(isum);
        const bool pow2 = (L & (L - 1)) == 0;
        if (model == MQ_ISING && dim == 3 && !bonds && pow2 && L >= 32 && L <= 1024 && st->kernel_choice != 2) {
            int lgL = 0;
            while ((1 << lgL) < L) ++lgL;
            // (N/2)/16 chunks per replica = 2^(3 lgL - 5): blocks of 256 threads, up to
            // 8 chunks per thread
            const int lgit = std::min(3, 3 * lgL - 13);
            const int lgbpr = 3 * lgL - 13 - lgit;
            hipLaunchKernelGGL(k_measure_ising3_row16, dim3((unsigned)(((int64_t)1 << lgbpr) * g.nrep)), dim3(256), 0, s,
                               b, g.stride, lgL, g.rep0, lgbpr, lgit, is);
            return hipGetLastError();
        }
        if (dim == 3) {
            if (bonds) hipLaunchKernelGGL((k_measure_int8<3, true>), grid, dim3(256), 0, s, b, g.stride, st->d_bonds, st->d_rep_bond_off, L, model, st->p.q, g.rep0, is);
            else hipLaunchKernelGGL((k_measure_int8<3, false>), grid, dim3(256), 0, s, b, g.stride, st->d_bonds, st->d_rep_bond_off, L, model, st->p.q, g.rep0, is);
        } else {
            if (bonds) hipLaunchKernelGGL((k_measure_int8<2, true>), grid, dim3(256), 0, s, b, g.stride, st->d_bonds, st->d_rep_bond_off, L, model, st->p.q, g.rep0, is);
            else hipLaunchKernelGGL((k_measure_int8<2, false>), grid, dim3(256), 0, s, b, g.stride, st->d_bonds, st->d_rep_bond_off, L, model, st->p.q, g.rep0, is);
        }
    }
    return hipGetLastError();
}

hipError_t launch_bonds_regular(mq_state* st, const Group& g, int sample, double p_plus, Key2 key)
{
    const int L = g.lat->L, dim = g.lat->dim;
    const int64_t N = g.lat->N;
    double t = std::floor(std::ldexp(p_plus, 32));
    uint32_t T = (p_plus <= 0) ? 0u : (t >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)t);
    int8_t* out = st->d_bonds + st->sample_bond_off[sample];
    const dim3 grid = grid_for((N * dim + 3) / 4, 256, 1);
    if (dim == 3) hipLaunchKernelGGL(k_bonds<3>, grid, dim3(256), 0, st->ctx->stream, out, L, T, sample, key);
    else hipLaunchKernelGGL(k_bonds<2>, grid, dim3(256), 0, st->ctx->stream, out, L, T, sample, key);
    return hipGetLastError();
}

// Descriptor table of the LDS multi-sweep kernel, built once when the state is
// created (no allocation or copy may happen while a graph is being captured).
hipError_t prepare_lds(mq_state* st)
{
    st->lds_ok = false;
    const int model = st->p.model;
    if (model != MQ_ISING && model != MQ_EA_ISING) return hipSuccess;
    if (st->msc || st->lat[0]->kind != LAT_REGULAR) return hipSuccess;
    const int dim = st->lat[0]->dim;
    std::vector<LdsGroup> gs;
    int blocks = 0;
    size_t need0 = 0, need1 = 0;
    for (const auto& g : st->groups) {
        const int64_t N = g.lat->N;
        if (N >= 65536) return hipSuccess;
        need0 = std::max(need0, (size_t)((N + 15) & ~15ll));
        need1 = std::max(need1, (size_t)((N + 15) & ~15ll) + (size_t)(N * dim));
        LdsGroup d;
        d.off = g.off;
        d.stride = g.stride;
        d.L = g.lat->L;
        d.magic_h = g.lat->L == 2 ? 0u : (uint32_t)((((uint64_t)1 << 32) + (uint64_t)(g.lat->L / 2) - 1) / (uint64_t)(g.lat->L / 2));
        d.magic_l = (uint32_t)((((uint64_t)1 << 32) + (uint64_t)g.lat->L - 1) / (uint64_t)g.lat->L);
        d.rep0 = g.rep0;
        d.nrep = g.nrep;
        d.block0 = blocks;
        blocks += g.nrep;
        gs.push_back(d);
    }
    if (gs.size() > 64 || need0 > 65536 - 256) return hipSuccess;
    hipError_t e = hipMalloc(&st->d_lds_groups, gs.size() * sizeof(LdsGroup));
    if (e != hipSuccess) return e;
    e = hipMemcpy(st->d_lds_groups, gs.data(), gs.size() * sizeof(LdsGroup), hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    st->lds_ok = true;
    st->lds_ngroups = (int)gs.size();
    st->lds_blocks = blocks;
    st->lds_need[0] = need0;
    st->lds_need[1] = need1;
    return hipSuccess;
}

// Multi-sweep launch over all groups when every replica (and its couplings) fits
// in 64 KiB of LDS.  Returns hipErrorNotSupported when the state does not qualify.
hipError_t launch_sweeps_lds(mq_state* st, int nsweeps, uint32_t sweep0, Key2 key)
{
    if (!st->lds_ok) return hipErrorNotSupported;
    const bool bonds = (st->p.model == MQ_EA_ISING) || (st->d_bonds != nullptr);
    const size_t lds = st->lds_need[bonds ? 1 : 0];
    if (lds > 65536 - 256) return hipErrorNotSupported;
    const int dim = st->lat[0]->dim;
    const int blocks = st->lds_blocks;
    const LdsGroup* dg = reinterpret_cast<const LdsGroup*>(st->d_lds_groups);
    const int ng = st->lds_ngroups;
    hipStream_t s = st->ctx->stream;
#define MQ_L(D, B)                                                                                               \
    hipLaunchKernelGGL((k_sweep_lds<D, B>), dim3(blocks), dim3(st->lds_block), lds, s, st->d_spins, dg, ng,     \
                       st->d_bonds, st->d_rep_bond_off, st->d_thr, st->tbl, nsweeps, sweep0, st->cur_dyn, key)
    if (dim == 3) {
        if (bonds) MQ_L(3, true);
        else MQ_L(3, false);
    } else {
        if (bonds) MQ_L(2, true);
        else MQ_L(2, false);
    }
#undef MQ_L
    return hipGetLastError();
}
