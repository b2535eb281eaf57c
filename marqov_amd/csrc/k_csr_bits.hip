// Replica-packed Ising on graphs (spin_bits = 1 on a CSR lattice): 32 replicas
// of one graph share a uint32 per vertex, so a neighbour gather brings 32
// replicas' spins in 4 bytes instead of 32.
//
// Layout: a group of R replicas on one graph (R a multiple of 32) has W = R/32
// word groups; the word of word group w at device position p is p*W + w, and
// its bit j is the spin of replica 32w + j (set = down).
//
// Sweep (two lanes per (vertex of the colour class, word group), 16 replicas
// each, on large colour classes; four lanes of 8 replicas on small ones): the
// anti-aligned neighbour counts c_j of a lane's replicas are summed bit-sliced
// into P planes (degrees below 2^P), spread to one byte per replica, and turned
// into the acceptance-table index e = q + qmax with q = deg - 2c.  Each replica is
// then decided exactly as by the int8 graph sweep: replica r's 16-bit uniform is
// half (r & 7) of Philox(vertex, 64*sweep + colour, r >> 3, METRO), compared with
// the high half of its 32-bit threshold, and on a tie the low half decides with
// Philox(vertex, phase, r >> 3, TIE).  The chains are therefore those of int8
// storage bit for bit, and the CPU replay of the int8 graph sweep replays them.
// Reference semantics: metropolis.h:86-126 with the Ising energy of
// Ising.h:63-84 on the graph lattices of graph_from_csv.h:32-57.
#include "mq_internal.hpp"

namespace {

// bit k of a 4-bit value to bit 0 of byte k
__device__ __forceinline__ uint32_t spread4(uint32_t nib) { return (nib * 0x00204081u) & 0x01010101u; }

// LPW lanes per (vertex, word group), each deciding NO = 4 / LPW octets of 8 replicas;
// nwork is a multiple of LPW, so the lanes of a word are active together
template <int P, int LPW>
__global__ void __launch_bounds__(256) k_csr_sweep_bits(uint32_t* __restrict__ base, int W,
                                                        const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                        const int32_t* __restrict__ vert_of,
                                                        const uint32_t* __restrict__ thr, int tbl, int64_t c0,
                                                        int nwork, uint32_t phase, const uint32_t* __restrict__ dyn,
                                                        int rep0, Key2 key)
{
    constexpr int NO = 4 / LPW;        // octets per lane
    constexpr int NB = 8 * NO;         // replica bits per lane
    constexpr uint32_t BM = NB == 32 ? 0xFFFFFFFFu : (1u << NB) - 1u;
    if (dyn) phase += dyn[0];
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= nwork) return;
    const int ol = t % LPW;  // lane within the word
    const int vw = t / LPW;
    const int vl = vw / W;
    const int w = vw - vl * W;
    const int64_t pos = c0 + vl;
    const int qmax = (tbl - 1) >> 1;
    const int grep0 = rep0 + 32 * w + NB * ol;  // first replica of this lane's bits
    const int sh = NB * ol;
    uint32_t* ownp = base + pos * W + w;
    const uint32_t S = *ownp;
    const uint32_t vid = (uint32_t)vert_of[pos];
    // the METRO blocks of the lane's octets depend only on the vertex: drawn while
    // the neighbour loads are in flight
    uint4 mw[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) mw[o] = philox4x32_10(vid, phase, ((uint32_t)grep0 >> 3) + (uint32_t)o, ST_METRO, key);
    const int a = rp[pos], b = rp[pos + 1];
    const int deg = b - a;
    // anti-aligned neighbour counts of the lane's replicas, bit-sliced into P planes
    const uint32_t sl = (S >> sh) & BM;
    uint32_t pl[P];
#pragma unroll
    for (int p = 0; p < P; ++p) pl[p] = 0u;
    for (int k0 = a; k0 < b; k0 += NBCH) {
        int nb[NBCH];
        chunk_cols(ci, k0, b, nb);
        uint32_t v[NBCH];
#pragma unroll
        for (int k = 0; k < NBCH; ++k) v[k] = base[(int64_t)nb[k] * W + w];
#pragma unroll
        for (int k = 0; k < NBCH; ++k) {
            uint32_t c = (k0 + k < b) ? (sl ^ ((v[k] >> sh) & BM)) : 0u;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const uint32_t n = pl[p] & c;
                pl[p] ^= c;
                c = n;
            }
        }
    }
    // table index e = deg + qmax - 2c, one byte per replica (E[k] byte i: replica
    // grep0 + 4k + i); at least qmax - deg >= 0 and at most 2 qmax <= 254
    const uint32_t ebase = (uint32_t)(deg + qmax) * 0x01010101u;
    uint32_t E[2 * NO];
#pragma unroll
    for (int k = 0; k < 2 * NO; ++k) {
        uint32_t C = 0u;
#pragma unroll
        for (int p = 0; p < P; ++p) C |= spread4((pl[p] >> (4 * k)) & 15u) << p;
        E[k] = ebase - 2u * C;
    }
    // rows of the lane's replicas (the table has nrep * tbl * 2 < 2^32 words)
    const uint32_t row0 = (uint32_t)grep0 * (uint32_t)tbl * 2u;
    uint32_t thv[NB];  // all table reads in flight together
#pragma unroll
    for (int j = 0; j < NB; ++j)
        thv[j] = thr[row0 + ((uint32_t)j * (uint32_t)tbl + ((E[j >> 2] >> (8 * (j & 3))) & 0xFFu)) * 2u];
    uint32_t acc = 0u;
#pragma unroll
    for (int o = 0; o < NO; ++o) {
        uint32_t oct = 0u;
        bool tie = false;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const uint32_t u = (u4_get(mw[o], m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
            const uint32_t th = thv[8 * o + m];
            const uint32_t tt = th > 0xFFFFu ? 0xFFFFu : th;  // "always" (65536) as 0xFFFF + lo 0x10000
            oct |= (u < tt ? 1u : 0u) << m;
            tie |= (u == tt);
        }
        if (tie) {
            const uint4 tw = philox4x32_10(vid, phase, ((uint32_t)grep0 >> 3) + (uint32_t)o, ST_TIE, key);
            oct = 0u;
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const int j = 8 * o + m;
                const uint32_t e = (E[j >> 2] >> (8 * (j & 3))) & 0xFFu;
                const uint32_t u = (u4_get(mw[o], m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
                const uint32_t lo = (u4_get(tw, m >> 1) >> (16 * (m & 1))) & 0xFFFFu;
                const uint32_t th = thv[j];
                const uint32_t tt = th > 0xFFFFu ? 0xFFFFu : th;
                const uint32_t tl = th > 0xFFFFu ? 0x10000u : thr[row0 + ((uint32_t)j * (uint32_t)tbl + e) * 2u + 1u];
                oct |= ((u < tt) || (u == tt && lo < tl) ? 1u : 0u) << m;
            }
        }
        acc |= oct << (8 * o);
    }
    acc <<= sh;
    // the lanes of a word are adjacent in the wavefront: their bits meet in lane ol = 0
    if (LPW >= 2) acc |= (uint32_t)__shfl_xor((int)acc, 1, 64);
    if (LPW >= 4) acc |= (uint32_t)__shfl_xor((int)acc, 2, 64);
    if (ol == 0) *ownp = S ^ acc;
}

// Random start: replica r's spin at vertex v is +1 when bit 31 of word (v & 3) of
// Philox(v >> 2, 0, r, INIT) is set, as for int8 storage.
__global__ void __launch_bounds__(256) k_csr_init_bits(uint32_t* __restrict__ base, int W,
                                                       const int32_t* __restrict__ vert_of, int64_t N, int rep0,
                                                       Key2 key)
{
    const int w = blockIdx.y;
    const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= N) return;
    const uint32_t v = (uint32_t)vert_of[pos];
    uint32_t word = 0u;
    for (int j = 0; j < 32; ++j) {
        const uint4 q = philox4x32_10(v >> 2, 0u, (uint32_t)(rep0 + 32 * w + j), ST_INIT, key);
        if (!(u4_get(q, (int)(v & 3u)) >> 31)) word |= 1u << j;
    }
    base[pos * W + w] = word;
}

// Per replica: isum[r*8+0] += sum over edges s_i s_j (each edge once, from its
// lower position), isum[r*8+1] += sum s.  A wavefront takes 64*MPL positions of
// one word group (grid y), folds anti-aligned edges and down spins into
// bit-sliced per-bit counters, reduce-scatters them so that lane j holds
// replica j's totals, and lanes 0..31 add them with one atomic each.
constexpr int MPL = 8;  // positions per lane
__global__ void __launch_bounds__(256) k_csr_measure_bits(const uint32_t* __restrict__ base, int W,
                                                          const int32_t* __restrict__ rp,
                                                          const int32_t* __restrict__ ci, int64_t N, int rep0,
                                                          unsigned long long* __restrict__ isum)
{
    constexpr int NE = 10 + 5;  // anti-aligned edges per lane <= MPL * 127 < 2^10, +5 for the exchange steps
    constexpr int ND = 4 + 5;   // down spins per lane <= MPL
    const int w = blockIdx.y;
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = (int)(threadIdx.x & 63);
    const int64_t first = wave * 64 * MPL;
    if (first >= N) return;  // whole wavefront; nothing below synchronises the workgroup
    uint32_t EA[NE], DA[ND];
#pragma unroll
    for (int p = 0; p < NE; ++p) EA[p] = 0u;
#pragma unroll
    for (int p = 0; p < ND; ++p) DA[p] = 0u;
    int nedge = 0, npos = 0;
    for (int k = 0; k < MPL; ++k) {
        const int64_t pos = first + (int64_t)k * 64 + lane;  // coalesced across the wavefront
        if (pos >= N) break;
        const uint32_t S = base[pos * W + w];
        uint32_t c = S;
#pragma unroll
        for (int p = 0; p < ND; ++p) {
            const uint32_t n = DA[p] & c;
            DA[p] ^= c;
            c = n;
        }
        ++npos;
        for (int e = rp[pos]; e < rp[pos + 1]; ++e) {
            const int j = ci[e];
            if ((int64_t)j <= pos) continue;
            uint32_t x = S ^ base[(int64_t)j * W + w];
#pragma unroll
            for (int p = 0; p < NE; ++p) {
                const uint32_t n = EA[p] & x;
                EA[p] ^= x;
                x = n;
            }
            ++nedge;
        }
    }
    const int anti = bs_reduce_scatter<NE>(EA);
    const int down = bs_reduce_scatter<ND>(DA);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        nedge += __shfl_xor(nedge, 32 >> k, 64);
        npos += __shfl_xor(npos, 32 >> k, 64);
    }
    if (lane < 32) {
        const int64_t r = (int64_t)rep0 + 32 * w + lane;
        atomicAdd(isum + r * 8 + 0, (unsigned long long)(long long)(nedge - 2 * anti));
        atomicAdd(isum + r * 8 + 1, (unsigned long long)(long long)(npos - 2 * down));
    }
}

}  // namespace

bool csrbits_supported(const mq_lattice* lat) { return lat->kind == LAT_CSR && lat->maxdeg <= 127; }

hipError_t launch_init_csrbits(mq_state* st, const Group& g, Key2 key)
{
    const int W = g.nrep / 32;
    const dim3 grid((unsigned)((g.lat->N + 255) / 256), (unsigned)W);
    hipLaunchKernelGGL(k_csr_init_bits, grid, dim3(256), 0, st->ctx->stream,
                       reinterpret_cast<uint32_t*>(st->d_spins + g.off), W, g.lat->d_vert_of, g.lat->N, g.rep0, key);
    return hipGetLastError();
}

hipError_t launch_sweep_csrbits(mq_state* st, const Group& g, int colour, uint32_t phase, Key2 key)
{
    const mq_lattice* l = g.lat;
    const int64_t c0 = l->color_off[colour], nc = l->color_off[colour + 1] - c0;
    if (nc == 0) return hipSuccess;
    const int W = g.nrep / 32;
    if (nc * W * 4 >= (1ll << 31)) return hipErrorInvalidValue;  // checked at state creation (N * R / 8)
    if ((int64_t)st->nrep * st->tbl * 2 >= (1ll << 32)) return hipErrorInvalidValue;  // 32-bit table offsets
    // lanes per word: two (16 replicas each) on all but the smallest colour classes,
    // four on those, whose launches are latency floors that a shorter chain per lane
    // (fewer Philox blocks and table reads) shortens.  Measured on the 2^20-vertex
    // RGG with 64 replicas, per class of 254k / 106k / 9.9k / 0.8k vertices: one lane
    // 22.2 / 13.1 / 7.1 / 7.6 us, two 20.6 / 11.6 / 5.6 / 5.5, four 27.6 / 15.0 / 5.7 / 5.0
    const int lpw = nc >= 4096 ? 2 : 4;
    const int nwork = (int)(nc * W * lpw);
    const dim3 grid((unsigned)((nwork + 255) / 256));
    uint32_t* b = reinterpret_cast<uint32_t*>(st->d_spins + g.off);
#define MQ_B(P)                                                                                                    \
    do {                                                                                                           \
        if (lpw == 2)                                                                                              \
            hipLaunchKernelGGL((k_csr_sweep_bits<P, 2>), grid, dim3(256), 0, st->ctx->stream, b, W, l->d_rp,      \
                               l->d_ci, l->d_vert_of, st->d_thr, st->tbl, c0, nwork, phase, st->cur_dyn, g.rep0,  \
                               key);                                                                               \
        else                                                                                                       \
            hipLaunchKernelGGL((k_csr_sweep_bits<P, 4>), grid, dim3(256), 0, st->ctx->stream, b, W, l->d_rp,      \
                               l->d_ci, l->d_vert_of, st->d_thr, st->tbl, c0, nwork, phase, st->cur_dyn, g.rep0,  \
                               key);                                                                               \
    } while (0)
    if (l->maxdeg < 8) MQ_B(3);
    else if (l->maxdeg < 16) MQ_B(4);
    else if (l->maxdeg < 32) MQ_B(5);
    else if (l->maxdeg < 64) MQ_B(6);
    else MQ_B(7);
#undef MQ_B
    return hipGetLastError();
}

hipError_t launch_measure_csrbits(mq_state* st, const Group& g, long long* isum)
{
    const int W = g.nrep / 32;
    const int64_t waves = (g.lat->N + 64 * MPL - 1) / (64 * MPL);
    const dim3 grid((unsigned)((waves + 3) / 4), (unsigned)W);
    hipLaunchKernelGGL(k_csr_measure_bits, grid, dim3(256), 0, st->ctx->stream,
                       reinterpret_cast<const uint32_t*>(st->d_spins + g.off), W, g.lat->d_rp, g.lat->d_ci, g.lat->N,
                       g.rep0, reinterpret_cast<unsigned long long*>(isum));
    return hipGetLastError();
}
