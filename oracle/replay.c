/*
 * oracle/replay.c -- CPU replay of the GPU update schedule (test infrastructure).
 *
 * Only tests/, the smoke check and the bench's cpu_baseline leg may load this
 * library, and only to check results; the product path never calls it.
 *
 * What it restates, per function:
 *  - orc_thresholds_*: the integer acceptance tables the library builds on the
 *    host.  Acceptance follows the reference rule "accept if dE <= 0, else with
 *    probability exp(-beta*dE)" (src/libmarqov/metropolishelpers.h:109-142,
 *    metropolis.h:60-64), with the probability quantised to 32 bits.
 *  - orc_ising_sweeps: a checkerboard Metropolis sweep of Ising / +-J
 *    Edwards-Anderson spins on a periodic hypercubic lattice.  dE follows
 *    compute_interactionenergydiff (metropolishelpers.h:598-623) with the
 *    bond-weighted neighbourhood of nbrhoodloop (metropolishelpers.h:456-487):
 *    dE = J * (s_new - s_old) * sum_j b_ij s_j.  Site order differs from the
 *    reference's random-site order (emcs.h:93-106) by design: all sites of one
 *    colour, then all of the other.
 *  - orc_wolff_ising: single-cluster Wolff for Ising (Ising.h:122-169); an
 *    aligned neighbour not yet in the cluster joins with p = -expm1(2*beta*J)
 *    (Ising.h:139).  As in the reference, every such test draws its own
 *    uniform (rng.real() per candidate, Ising.h:150-158); here the uniform is
 *    keyed by the test (site, neighbour slot), so the cluster is the set of
 *    sites reachable from the seed along passing tests and does not depend on
 *    visit order.
 *  - orc_potts_*: Potts Metropolis (Potts.h:87-107,165-178) and Wolff
 *    (Potts.h:189-322) on the same schedule.
 *  - orc_energy_* / orc_mag_*: observables (hamiltonian/util/observables.h:32-66,
 *    173-231; Potts.h:40-65).
 * Lattice indexing follows FixedDimRegularHypercubic / RegularLattice
 * (src/lattice/regular_hypercubic.h:31-66, src/lattice/util/regular_lattice.h:54-89):
 * site index sum_j x_j L^j with x_0 fastest, neighbour order +x0,-x0,+x1,-x1,...
 *
 * Random-number layout (identical on the GPU):
 *   key = 64-bit seed; counter = (c0, c1, replica, stream)
 *   STREAM_INIT  : c0 = site>>2, c1 = 0, word site&3
 *   STREAM_METRO : c0 = (site>>1)>>3, c1 = 2*sweep+colour, 16-bit half (site>>1)&7
 *   STREAM_TIE   : same counter as METRO, low 16 bits of the 32-bit uniform
 *   STREAM_WSEED : c0 = 0, c1 = cluster number
 *   STREAM_WBOND : c0 = e>>2 with e = site*2*dim+slot (the test from site to its
 *                  neighbour over slot), c1 = cluster number, word e&3; on a graph
 *                  e is the CSR slot of the test's edge (row of the tested-from site)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "philox.h"

enum { ST_INIT = 0, ST_METRO = 1, ST_TIE = 2, ST_WSEED = 3, ST_WBOND = 4, ST_PT = 5, ST_DISORDER = 6, ST_PROPOSE = 7 };

#define EXPORT __attribute__((visibility("default")))

/* ---------------------------------------------------------------- lattice */

static void coords(int dim, int L, int64_t i, int x[3])
{
    for (int d = 0; d < dim; ++d) { x[d] = (int)(i % L); i /= L; }
}

static int64_t nbr(int dim, int L, int64_t i, int d, int sign)
{
    int64_t pw = 1;
    for (int k = 0; k < d; ++k) pw *= L;
    int x[3];
    coords(dim, L, i, x);
    int xn = x[d] + sign;
    if (xn == L) xn = 0;
    if (xn < 0) xn = L - 1;
    return i + (int64_t)(xn - x[d]) * pw;
}

EXPORT int64_t orc_neighbour(int dim, int L, int64_t i, int slot)
{
    /* slot order +x0,-x0,+x1,-x1,... as in regular_hypercubic.h:31-66 */
    return nbr(dim, L, i, slot >> 1, (slot & 1) ? -1 : +1);
}

static int colour_of(int dim, int L, int64_t i)
{
    int x[3];
    coords(dim, L, i, x);
    int s = 0;
    for (int d = 0; d < dim; ++d) s += x[d];
    return s & 1;
}

/* ------------------------------------------------------------- thresholds */

static void split_threshold(double p, uint32_t* hi, uint32_t* lo)
{
    if (!(p < 1.0)) { *hi = 65536u; *lo = 0u; return; }
    double t = floor(ldexp(p, 32));
    uint64_t T = (t <= 0.0) ? 0u : (uint64_t)t;
    if (T > 0xFFFFFFFFull) T = 0xFFFFFFFFull;
    *hi = (uint32_t)(T >> 16);
    *lo = (uint32_t)(T & 0xFFFFu);
}

/* Ising / EA: entry e = q + qmax, q = s_i * sum_j b_ij s_j, dE = -2 J q. */
EXPORT void orc_thresholds_ising(double beta, double J, int qmax, uint32_t* thi, uint32_t* tlo)
{
    for (int e = 0; e <= 2 * qmax; ++e) {
        int q = e - qmax;
        double dE = -2.0 * J * (double)q;
        double p = (dE <= 0.0) ? 1.0 : exp(-beta * dE);
        split_threshold(p, &thi[e], &tlo[e]);
    }
}

/* Potts: entry e = (n_new - n_old) + zmax, dE = k * (n_new - n_old). */
EXPORT void orc_thresholds_potts(double beta, double k, int zmax, uint32_t* thi, uint32_t* tlo)
{
    for (int e = 0; e <= 2 * zmax; ++e) {
        double dE = k * (double)(e - zmax);
        double p = (dE <= 0.0) ? 1.0 : exp(-beta * dE);
        split_threshold(p, &thi[e], &tlo[e]);
    }
}

/* Wolff add probability as a 32-bit threshold: add iff u32 < T. */
EXPORT uint32_t orc_wolff_threshold(double p)
{
    if (!(p > 0.0)) return 0u;
    double t = floor(ldexp(p, 32));
    if (t >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)t;
}

/* ------------------------------------------------------------------- init */

EXPORT void orc_init_ising(int64_t N, int8_t* s, uint64_t seed, uint32_t rep)
{
    for (int64_t i = 0; i < N; ++i) {
        uint32_t w[4];
        orc_philox_u64key((uint32_t)(i >> 2), 0u, rep, ST_INIT, seed, w);
        s[i] = (w[i & 3] >> 31) ? (int8_t)1 : (int8_t)-1;
    }
}

EXPORT void orc_init_potts(int64_t N, int8_t* s, int Q, uint64_t seed, uint32_t rep)
{
    for (int64_t i = 0; i < N; ++i) {
        uint32_t w[4];
        orc_philox_u64key((uint32_t)(i >> 2), 0u, rep, ST_INIT, seed, w);
        s[i] = (int8_t)(((uint64_t)w[i & 3] * (uint64_t)Q) >> 32);
    }
}

/* One colour half-sweep of the FP32 O(3)/Phi4 Metropolis kernel restated in
 * FP64 (the GPU schedule; the reference's move is metropolis.h:86-126 with the
 * proposals of initializers.h:61-81 / Phi4.h:20-34 and the on-site terms of
 * Phi4.h:69-74,110-124).  For each site i of the colour: counter
 * (i >> 1, phase, rep, STREAM_METRO) gives words w0..w2; the direction is
 * (1 - a0, rho sin(pi a1), rho cos(pi a1)) with a_k = 2 (w_k >> 8) 2^-24 and
 * rho = sqrt(a0 (2 - a0)) (randomdir.h:45-65); the proposal is that direction
 * (fresh), s + delta r normalised (small angle) or s + 0.5 r (Phi4); dE =
 * J (s_new - s_old) . sum of neighbours plus, for Phi4, (mass/beta) d|s|^2 +
 * (lambda/beta) d(|s|^2 - 1)^2; accept iff dE <= 0 or u < exp(-beta dE) with
 * u = (w2 >> 8) 2^-24.  margin[i] = p - u (p = 1 when dE <= 0), so a GPU
 * decision made in FP32 can differ only where |margin| is tiny; other sites'
 * margin is 0.  s is float[3] per site in natural site order. */
/* proposal, energy difference and decision of one O(3)/Phi4 site: so = old spin,
 * nsum = neighbour sum, w = its Philox words; writes sn, returns p - u */
static double o3_site(const double* so, const double nsum[3], const uint32_t w[4], int phi4, int small_angle,
                      double J, double delta, double lambda, double mass, double beta, double sn[3])
{
    const double a0 = 2.0 * (double)(w[0] >> 8) * 0x1.0p-24;
    const double a1 = 2.0 * (double)(w[1] >> 8) * 0x1.0p-24;
    const double rho = sqrt(fmax(a0 * (2.0 - a0), 0.0));
    const double rd[3] = {1.0 - a0, rho * sin(3.14159265358979323846 * a1), rho * cos(3.14159265358979323846 * a1)};
    if (phi4) {
        for (int k = 0; k < 3; ++k) sn[k] = so[k] + 0.5 * rd[k];
    } else if (small_angle) {
        double t[3], n2 = 0.0;
        for (int k = 0; k < 3; ++k) {
            t[k] = so[k] + delta * rd[k];
            n2 += t[k] * t[k];
        }
        for (int k = 0; k < 3; ++k) sn[k] = t[k] / sqrt(n2);
    } else {
        for (int k = 0; k < 3; ++k) sn[k] = rd[k];
    }
    double dE = J * ((sn[0] - so[0]) * nsum[0] + (sn[1] - so[1]) * nsum[1] + (sn[2] - so[2]) * nsum[2]);
    if (phi4) {
        const double n2 = sn[0] * sn[0] + sn[1] * sn[1] + sn[2] * sn[2];
        const double o2 = so[0] * so[0] + so[1] * so[1] + so[2] * so[2];
        dE += (mass / beta) * (n2 - o2) + (lambda / beta) * ((n2 - 1.0) * (n2 - 1.0) - (o2 - 1.0) * (o2 - 1.0));
    }
    const double u = (double)(w[2] >> 8) * 0x1.0p-24;
    const double p = dE <= 0.0 ? 1.0 : exp(-beta * dE);
    return p - u;
}

EXPORT void orc_o3_halfsweep(int dim, int L, const double* s_in, double* s_out, double* margin, int phi4,
                             int small_angle, double J, double delta, double lambda, double mass, double beta,
                             int colour, uint32_t phase, uint32_t rep, uint64_t seed)
{
    int64_t N = 1;
    for (int d = 0; d < dim; ++d) N *= L;
    memcpy(s_out, s_in, sizeof(double) * 3 * (size_t)N);
    for (int64_t i = 0; i < N; ++i) {
        margin[i] = 0.0;
        int x[3] = {0, 0, 0};
        coords(dim, L, i, x);
        if (((x[0] + x[1] + x[2]) & 1) != colour) continue;
        double ns[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d)
            for (int sg = 1; sg >= -1; sg -= 2) {
                const double* v = s_in + 3 * nbr(dim, L, i, d, sg);
                for (int k = 0; k < 3; ++k) ns[k] += v[k];
            }
        uint32_t w[4];
        orc_philox_u64key((uint32_t)(i >> 1), phase, rep, ST_METRO, seed, w);
        double sn[3];
        margin[i] = o3_site(s_in + 3 * i, ns, w, phi4, small_angle, J, delta, lambda, mass, beta, sn);
        if (margin[i] > 0.0)
            for (int k = 0; k < 3; ++k) s_out[3 * i + k] = sn[k];
    }
}

/* The same for one colour class of a CSR graph: counter (vertex, phase, rep,
 * STREAM_METRO) with phase = 64 * sweep + colour; colour_of[] from
 * mq_lattice_color_of. */
EXPORT void orc_csr_o3_colour(int64_t N, const int32_t* rp, const int32_t* ci, const int32_t* colour_of,
                              const double* s_in, double* s_out, double* margin, int phi4, int small_angle, double J,
                              double delta, double lambda, double mass, double beta, int colour, uint32_t phase,
                              uint32_t rep, uint64_t seed)
{
    memcpy(s_out, s_in, sizeof(double) * 3 * (size_t)N);
    for (int64_t v = 0; v < N; ++v) {
        margin[v] = 0.0;
        if (colour_of[v] != colour) continue;
        double ns[3] = {0, 0, 0};
        for (int64_t k = rp[v]; k < rp[v + 1]; ++k)
            for (int c = 0; c < 3; ++c) ns[c] += s_in[3 * (int64_t)ci[k] + c];
        uint32_t w[4];
        orc_philox_u64key((uint32_t)v, phase, rep, ST_METRO, seed, w);
        double sn[3];
        margin[v] = o3_site(s_in + 3 * v, ns, w, phi4, small_angle, J, delta, lambda, mass, beta, sn);
        if (margin[v] > 0.0)
            for (int c = 0; c < 3; ++c) s_out[3 * v + c] = sn[c];
    }
}

/* Uniform of the Wolff test from site i to its neighbour over slot (2d: +e_d,
 * 2d+1: -e_d): word e & 3 of counter (e >> 2, ncl, rep, STREAM_WBOND) with
 * e = i * 2 * dim + slot.  A site's 2*dim tests use two Philox blocks in 3D and
 * one in 2D. */
static uint32_t test_uniform(int dim, int64_t i, int slot, uint32_t ncl, uint32_t rep, uint64_t seed)
{
    const uint64_t e = (uint64_t)i * (uint64_t)(2 * dim) + (uint64_t)slot;
    uint32_t w[4];
    orc_philox_u64key((uint32_t)(e >> 2), ncl, rep, ST_WBOND, seed, w);
    return w[e & 3];
}

/* One O(3)/Phi4 Wolff cluster (Heisenberg.h:191-243 with a Cartesian axis;
 * Phi4.h:182-244 / the random-direction embedding: s -= 2 (s.r) r) restated in
 * FP64 on the GPU's random numbers: seed and embedding from counter
 * (0, ncl, rep, STREAM_WSEED) (seed = (w0 N) >> 32; axis = (w1 3) >> 32 or the
 * direction of words w1, w2 as in orc_o3_halfsweep); the test from in-cluster
 * a to its neighbour j over slot passes iff cpl = J p_a' p_j > 0 and
 * u < 1 - exp(-2 beta cpl), u = (word >> 8) 2^-24 of the word test_uniform
 * names (e = a*2*dim + slot), p_a' the projection of a after its flip and p_j
 * that of j before the cluster.  The cluster is the set reachable from the seed
 * along passing tests; *min_margin receives min |p - u| over every
 * (member, neighbour) test, so a cluster grown in FP32 in any order can differ
 * only if it is tiny.
 * Returns the cluster size; s_out = s_in with the cluster flipped. */
EXPORT int64_t orc_o3_wolff(int dim, int L, const double* s_in, double* s_out, int random_dir, double J, double beta,
                            uint32_t ncl, uint32_t rep, uint64_t seed, double* min_margin)
{
    int64_t N = 1;
    for (int d = 0; d < dim; ++d) N *= L;
    memcpy(s_out, s_in, sizeof(double) * 3 * (size_t)N);
    uint32_t sw[4];
    orc_philox_u64key(0u, ncl, rep, ST_WSEED, seed, sw);
    const int64_t s0 = (int64_t)(((uint64_t)sw[0] * (uint64_t)N) >> 32);
    const int axis = (int)(((uint64_t)sw[1] * 3ull) >> 32);
    double rd[3] = {0, 0, 0};
    if (random_dir) {
        const double a0 = 2.0 * (double)(sw[1] >> 8) * 0x1.0p-24;
        const double a1 = 2.0 * (double)(sw[2] >> 8) * 0x1.0p-24;
        const double rho = sqrt(fmax(a0 * (2.0 - a0), 0.0));
        rd[0] = 1.0 - a0;
        rd[1] = rho * sin(3.14159265358979323846 * a1);
        rd[2] = rho * cos(3.14159265358979323846 * a1);
    }
    double* proj0 = (double*)malloc(sizeof(double) * (size_t)N);  /* projections before the cluster */
    char* in = (char*)calloc((size_t)N, 1);
    int64_t* stack = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
    for (int64_t i = 0; i < N; ++i) {
        const double* v = s_in + 3 * i;
        proj0[i] = random_dir ? v[0] * rd[0] + v[1] * rd[1] + v[2] * rd[2] : v[axis];
    }
    /* flip: Cartesian negates the axis component; random direction reflects */
#define ORC_FLIP(i)                                                                            \
    do {                                                                                       \
        double* v_ = s_out + 3 * (i);                                                          \
        if (random_dir) {                                                                      \
            const double d_ = v_[0] * rd[0] + v_[1] * rd[1] + v_[2] * rd[2];                   \
            for (int k_ = 0; k_ < 3; ++k_) v_[k_] -= 2.0 * d_ * rd[k_];                        \
        } else {                                                                               \
            v_[axis] = -v_[axis];                                                              \
        }                                                                                      \
    } while (0)
    double mm = INFINITY;
    int64_t top = 0, size = 1;
    in[s0] = 1;
    ORC_FLIP(s0);
    stack[top++] = s0;
    while (top > 0) {
        const int64_t a = stack[--top];
        const double* va = s_out + 3 * a;
        const double pa = random_dir ? va[0] * rd[0] + va[1] * rd[1] + va[2] * rd[2] : va[axis];
        for (int d = 0; d < dim; ++d)
            for (int sg = 1; sg >= -1; sg -= 2) {
                const int64_t j = nbr(dim, L, a, d, sg);
                const double u = (double)(test_uniform(dim, a, 2 * d + (sg < 0), ncl, rep, seed) >> 8) * 0x1.0p-24;
                const double cpl = J * pa * proj0[j];
                const double p = cpl > 0.0 ? -expm1(-2.0 * beta * cpl) : 0.0;
                const double mg = fabs(p - u);
                if (mg < mm) mm = mg;
                if (!in[j] && u < p) {
                    in[j] = 1;
                    ORC_FLIP(j);
                    stack[top++] = j;
                    ++size;
                }
            }
    }
#undef ORC_FLIP
    *min_margin = mm;
    free(proj0);
    free(in);
    free(stack);
    return size;
}

/* Points of the device random geometric graph builder (mq_rgg_build): coordinate
 * d of point i is the 53-bit uniform of counter (i, d >> 1, 0, 8), words 2(d&1)
 * (high) and 2(d&1)+1 (low); out[i*dim + d]. */
EXPORT void orc_rgg_points(int dim, int64_t n, uint64_t seed, double* out)
{
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < dim; ++d) {
            uint32_t w[4];
            orc_philox_u64key((uint32_t)i, (uint32_t)(d >> 1), 0u, 8u, seed, w);
            const uint64_t u = ((uint64_t)w[2 * (d & 1)] << 32) | w[2 * (d & 1) + 1];
            out[i * dim + d] = (double)(u >> 11) * 0x1.0p-53;
        }
}

/* EA +-J couplings with P(+1) = p_plus: bonds[i*dim+d] couples i and i+e_d. */
EXPORT void orc_init_bonds(int dim, int64_t N, int8_t* bonds, double p_plus, uint64_t seed, uint32_t sample)
{
    uint32_t T = orc_wolff_threshold(p_plus);
    for (int64_t b = 0; b < N * dim; ++b) {
        uint32_t w[4];
        orc_philox_u64key((uint32_t)(b >> 2), 0u, sample, ST_DISORDER, seed, w);
        bonds[b] = (w[b & 3] < T) ? (int8_t)1 : (int8_t)-1;
    }
}

/* ------------------------------------------------------------ Metropolis */

static int8_t bond_of(const int8_t* bonds, int dim, int L, int64_t i, int slot)
{
    if (!bonds) return 1;
    int d = slot >> 1;
    if ((slot & 1) == 0) return bonds[i * dim + d];
    int64_t j = nbr(dim, L, i, d, -1);
    return bonds[j * dim + d];
}

static int accept16(uint32_t hi16, uint32_t thi, uint32_t tlo, uint32_t g, uint32_t phase,
                    uint32_t rep, uint64_t seed, int m)
{
    if (hi16 < thi) return 1;
    if (hi16 > thi) return 0;
    uint32_t t[4];
    orc_philox_u64key(g, phase, rep, ST_TIE, seed, t);
    uint32_t lo16 = (t[m >> 1] >> (16 * (m & 1))) & 0xFFFFu;
    return lo16 < tlo;
}

EXPORT void orc_ising_sweeps(int dim, int L, int8_t* s, const int8_t* bonds,
                             const uint32_t* thi, const uint32_t* tlo,
                             int nsweeps, uint64_t seed, uint32_t sweep0, uint32_t rep)
{
    int64_t N = 1;
    for (int d = 0; d < dim; ++d) N *= L;
    const int qmax = 2 * dim;
    for (int t = 0; t < nsweeps; ++t) {
        for (int c = 0; c < 2; ++c) {
            uint32_t phase = 2u * (sweep0 + (uint32_t)t) + (uint32_t)c;
            for (int64_t i = 0; i < N; ++i) {
                if (colour_of(dim, L, i) != c) continue;
                int64_t h = i >> 1;
                uint32_t g = (uint32_t)(h >> 3);
                int m = (int)(h & 7);
                int hsum = 0;
                for (int slot = 0; slot < 2 * dim; ++slot) {
                    int64_t j = orc_neighbour(dim, L, i, slot);
                    hsum += bond_of(bonds, dim, L, i, slot) * s[j];
                }
                int q = s[i] * hsum;
                int e = q + qmax;
                uint32_t w[4];
                orc_philox_u64key(g, phase, rep, ST_METRO, seed, w);
                uint32_t hi16 = (w[m >> 1] >> (16 * (m & 1))) & 0xFFFFu;
                if (accept16(hi16, thi[e], tlo[e], g, phase, rep, seed, m)) s[i] = (int8_t)-s[i];
            }
        }
    }
}

/* Potts Metropolis: proposal uniform in [0,Q) (may equal the old value, as in
 * Potts.h:171-177); dE = k*(#nbr==new - #nbr==old) (Potts.h:87-107).
 * Per site one Philox block: word 0 -> proposal, word 1 -> 32-bit uniform.
 * counter: c0 = site>>1 (colour-half index), c1 = 2*sweep+colour. */
EXPORT void orc_potts_sweeps(int dim, int L, int8_t* s, int Q, const uint32_t* thi, const uint32_t* tlo,
                             int nsweeps, uint64_t seed, uint32_t sweep0, uint32_t rep)
{
    int64_t N = 1;
    for (int d = 0; d < dim; ++d) N *= L;
    const int zmax = 2 * dim;
    for (int t = 0; t < nsweeps; ++t) {
        for (int c = 0; c < 2; ++c) {
            uint32_t phase = 2u * (sweep0 + (uint32_t)t) + (uint32_t)c;
            for (int64_t i = 0; i < N; ++i) {
                if (colour_of(dim, L, i) != c) continue;
                uint32_t w[4];
                orc_philox_u64key((uint32_t)(i >> 1), phase, rep, ST_METRO, seed, w);
                int snew = (int)(((uint64_t)w[0] * (uint64_t)Q) >> 32);
                int sold = s[i];
                int nn = 0, no = 0;
                for (int slot = 0; slot < 2 * dim; ++slot) {
                    int sj = s[orc_neighbour(dim, L, i, slot)];
                    nn += (sj == snew);
                    no += (sj == sold);
                }
                int e = nn - no + zmax;
                uint64_t T = ((uint64_t)thi[e] << 16) | tlo[e];
                if ((uint64_t)w[1] < T) s[i] = (int8_t)snew;
            }
        }
    }
}

/* ------------------------------------------------------------------ Wolff */

static int32_t* g_stack = NULL;
static int64_t g_stack_n = 0;
static uint8_t* g_mark = NULL;
static int64_t g_mark_n = 0;

static int ensure_work(int64_t N)
{
    if (g_stack_n < N) {
        free(g_stack);
        g_stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
        g_stack_n = N;
    }
    if (g_mark_n < N) {
        free(g_mark);
        g_mark = (uint8_t*)calloc((size_t)N, 1);
        g_mark_n = N;
    }
    return g_stack && g_mark;
}


/* Returns the cluster sizes in sizes[nclusters]. */
EXPORT int orc_wolff_ising(int dim, int L, int8_t* s, uint32_t tw, int nclusters,
                           uint64_t seed, uint32_t cl0, uint32_t rep, int64_t* sizes)
{
    int64_t N = 1;
    for (int d = 0; d < dim; ++d) N *= L;
    if (!ensure_work(N)) return -1;
    for (int c = 0; c < nclusters; ++c) {
        uint32_t ncl = cl0 + (uint32_t)c;
        uint32_t w[4];
        orc_philox_u64key(0u, ncl, rep, ST_WSEED, seed, w);
        int64_t seedsite = (int64_t)(((uint64_t)w[0] * (uint64_t)N) >> 32);
        int8_t val = s[seedsite];
        int64_t top = 0, size = 1;
        g_stack[top++] = (int32_t)seedsite;
        g_mark[seedsite] = 1;
        while (top > 0) {
            int64_t cur = g_stack[--top];
            for (int slot = 0; slot < 2 * dim; ++slot) {
                int64_t j = orc_neighbour(dim, L, cur, slot);
                if (g_mark[j] || s[j] != val) continue;
                if (test_uniform(dim, cur, slot, ncl, rep, seed) < tw) {
                    g_mark[j] = 1;
                    g_stack[top++] = (int32_t)j;
                    ++size;
                }
            }
        }
        /* flip the cluster and clear marks */
        for (int64_t i = 0; i < N; ++i)
            if (g_mark[i]) { s[i] = (int8_t)-s[i]; g_mark[i] = 0; }
        if (sizes) sizes[c] = size;
    }
    return 0;
}

/* Potts Wolff: new colour uniform among the Q-1 others (Potts.h:221-229),
 * drawn from word 1 of the seed block; neighbours equal to the seed's old
 * colour join with p = 1 - exp(-beta*|k|) (Potts.h:237-242 with cpl = k*(-1/2)
 * fed to the pref=2 rule of wolff.h:35-39). */
EXPORT int orc_wolff_potts(int dim, int L, int8_t* s, int Q, uint32_t tw, int nclusters,
                           uint64_t seed, uint32_t cl0, uint32_t rep, int64_t* sizes)
{
    int64_t N = 1;
    for (int d = 0; d < dim; ++d) N *= L;
    if (!ensure_work(N)) return -1;
    for (int c = 0; c < nclusters; ++c) {
        uint32_t ncl = cl0 + (uint32_t)c;
        uint32_t w[4];
        orc_philox_u64key(0u, ncl, rep, ST_WSEED, seed, w);
        int64_t seedsite = (int64_t)(((uint64_t)w[0] * (uint64_t)N) >> 32);
        int oldq = s[seedsite];
        int newq = (oldq + 1 + (int)(((uint64_t)w[1] * (uint64_t)(Q - 1)) >> 32)) % Q;
        int64_t top = 0, size = 1;
        g_stack[top++] = (int32_t)seedsite;
        g_mark[seedsite] = 1;
        while (top > 0) {
            int64_t cur = g_stack[--top];
            for (int slot = 0; slot < 2 * dim; ++slot) {
                int64_t j = orc_neighbour(dim, L, cur, slot);
                if (g_mark[j] || s[j] != oldq) continue;
                if (test_uniform(dim, cur, slot, ncl, rep, seed) < tw) {
                    g_mark[j] = 1;
                    g_stack[top++] = (int32_t)j;
                    ++size;
                }
            }
        }
        for (int64_t i = 0; i < N; ++i)
            if (g_mark[i]) { s[i] = (int8_t)newq; g_mark[i] = 0; }
        if (sizes) sizes[c] = size;
    }
    return 0;
}

/* ------------------------------------------------------------ observables */

/* sum over bonds (each once) of b_ij s_i s_j */
EXPORT int64_t orc_bond_sum_ising(int dim, int L, const int8_t* s, const int8_t* bonds)
{
    int64_t N = 1;
    for (int d = 0; d < dim; ++d) N *= L;
    int64_t acc = 0;
    for (int64_t i = 0; i < N; ++i)
        for (int d = 0; d < dim; ++d) {
            int64_t j = nbr(dim, L, i, d, +1);
            int b = bonds ? bonds[i * dim + d] : 1;
            acc += (int64_t)b * s[i] * s[j];
        }
    return acc;
}

EXPORT int64_t orc_mag_sum_ising(int64_t N, const int8_t* s)
{
    int64_t m = 0;
    for (int64_t i = 0; i < N; ++i) m += s[i];
    return m;
}

/* number of equal-colour bonds (each once) */
EXPORT int64_t orc_bond_sum_potts(int dim, int L, const int8_t* s)
{
    int64_t N = 1;
    for (int d = 0; d < dim; ++d) N *= L;
    int64_t acc = 0;
    for (int64_t i = 0; i < N; ++i)
        for (int d = 0; d < dim; ++d) acc += (s[i] == s[nbr(dim, L, i, d, +1)]);
    return acc;
}

/* Philox4x32-10 exported for the known-answer tests. */
EXPORT void orc_philox4x32_10_export(const uint32_t* ctr, const uint32_t* key, uint32_t* out)
{
    orc_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------ CSR graphs */
/* Colour-class sweeps on a general graph (original vertex numbering); the
 * colour of every vertex is an input (the library's greedy colouring).
 * The sweep's phase is stride*sweep + colour, with stride 64 for up to 64
 * colour classes and the next power of two of the class count above that.
 * Ising/EA: counter (vertex, phase, replica >> 3, METRO), 16-bit
 * half (replica & 7) against the threshold's high half, TIE stream on equality;
 * Potts: counter (vertex, phase, replica, METRO), proposes (word0*Q)>>32 and
 * accepts iff word 1 < T. */
EXPORT void orc_csr_sweeps(int64_t N, const int32_t* rp, const int32_t* ci, const int32_t* colour, int ncol,
                           int8_t* s, const int8_t* bonds, const uint32_t* thi, const uint32_t* tlo, int qmax,
                           int potts, int Q, int nsweeps, uint64_t seed, uint32_t sweep0, uint32_t rep)
{
    uint32_t stride = 64u;
    while (stride < (uint32_t)ncol) stride <<= 1;
    for (int t = 0; t < nsweeps; ++t)
        for (int c = 0; c < ncol; ++c) {
            uint32_t phase = stride * (sweep0 + (uint32_t)t) + (uint32_t)c;
            for (int64_t v = 0; v < N; ++v) {
                if (colour[v] != c) continue;
                uint32_t w[4];
                int si = s[v];
                if (potts) {
                    /* proposal: word (rep & 3) of the PROPOSE block of 4 replicas;
                     * acceptance: 16-bit half (rep & 7) of the METRO block of 8
                     * replicas against the high half of T, the TIE stream's half
                     * against the low half on equality (= the 32-bit test u < T) */
                    uint32_t pw[4];
                    orc_philox_u64key((uint32_t)v, phase, rep >> 2, ST_PROPOSE, seed, pw);
                    int snew = (int)(((uint64_t)pw[rep & 3u] * (uint64_t)Q) >> 32);
                    int dn = 0;
                    for (int k = rp[v]; k < rp[v + 1]; ++k) dn += (s[ci[k]] == snew) - (s[ci[k]] == si);
                    int e = dn + qmax;
                    orc_philox_u64key((uint32_t)v, phase, rep >> 3, ST_METRO, seed, w);
                    int m = (int)(rep & 7u);
                    uint32_t hi16 = (w[m >> 1] >> (16 * (m & 1))) & 0xFFFFu;
                    int acc;
                    if (hi16 != thi[e]) acc = hi16 < thi[e];
                    else {
                        uint32_t tw[4];
                        orc_philox_u64key((uint32_t)v, phase, rep >> 3, ST_TIE, seed, tw);
                        acc = ((tw[m >> 1] >> (16 * (m & 1))) & 0xFFFFu) < tlo[e];
                    }
                    if (acc) s[v] = (int8_t)snew;
                } else {
                    int h = 0;
                    for (int k = rp[v]; k < rp[v + 1]; ++k) h += (bonds ? bonds[k] : 1) * s[ci[k]];
                    int e = si * h + qmax;
                    /* one Philox block per vertex and 8 replicas: 16-bit half (rep & 7) */
                    orc_philox_u64key((uint32_t)v, phase, rep >> 3, ST_METRO, seed, w);
                    int m = (int)(rep & 7u);
                    uint32_t hi16 = (w[m >> 1] >> (16 * (m & 1))) & 0xFFFFu;
                    int acc;
                    if (hi16 != thi[e]) acc = hi16 < thi[e];
                    else {
                        uint32_t tw[4];
                        orc_philox_u64key((uint32_t)v, phase, rep >> 3, ST_TIE, seed, tw);
                        acc = ((tw[m >> 1] >> (16 * (m & 1))) & 0xFFFFu) < tlo[e];
                    }
                    if (acc) s[v] = (int8_t)(-si);
                }
            }
        }
}

/* Wolff on a graph: the test from u to v over CSR slot k (v = ci[k]) uses word
 * k & 3 of counter (k >> 2, ncl, rep, STREAM_WBOND).  potts = 0: Ising; 1:
 * Potts with Q states. */
EXPORT int orc_csr_wolff(int64_t N, const int32_t* rp, const int32_t* ci, int8_t* s, int potts, int Q, uint32_t tw,
                         int nclusters, uint64_t seed, uint32_t cl0, uint32_t rep, int64_t* sizes)
{
    if (!ensure_work(N)) return -1;
    for (int c = 0; c < nclusters; ++c) {
        uint32_t ncl = cl0 + (uint32_t)c;
        uint32_t w[4];
        orc_philox_u64key(0u, ncl, rep, ST_WSEED, seed, w);
        int64_t seedv = (int64_t)(((uint64_t)w[0] * (uint64_t)N) >> 32);
        int val = s[seedv];
        int newq = potts ? (val + 1 + (int)(((uint64_t)w[1] * (uint64_t)(Q - 1)) >> 32)) % Q : -val;
        int64_t top = 0, size = 1;
        g_stack[top++] = (int32_t)seedv;
        g_mark[seedv] = 1;
        while (top > 0) {
            int64_t u = g_stack[--top];
            for (int64_t k = rp[u]; k < rp[u + 1]; ++k) {
                int64_t v = ci[k];
                if (g_mark[v] || s[v] != val) continue;
                uint32_t b[4];
                orc_philox_u64key((uint32_t)(k >> 2), ncl, rep, ST_WBOND, seed, b);
                if (b[k & 3] < tw) {
                    g_mark[v] = 1;
                    g_stack[top++] = (int32_t)v;
                    ++size;
                }
            }
        }
        for (int64_t i = 0; i < N; ++i)
            if (g_mark[i]) { s[i] = (int8_t)newq; g_mark[i] = 0; }
        if (sizes) sizes[c] = size;
    }
    return 0;
}

/* One O(3)/Phi4 Wolff cluster on a CSR graph, FP64, as orc_o3_wolff: seed
 * vertex (w0 N) >> 32, test uniforms keyed by the CSR slot of the tested edge in
 * the tested-from vertex's row, as orc_csr_wolff. */
EXPORT int64_t orc_csr_o3_wolff(int64_t N, const int32_t* rp, const int32_t* ci, const double* s_in, double* s_out,
                                int random_dir, double J, double beta, uint32_t ncl, uint32_t rep, uint64_t seed,
                                double* min_margin)
{
    memcpy(s_out, s_in, sizeof(double) * 3 * (size_t)N);
    uint32_t sw[4];
    orc_philox_u64key(0u, ncl, rep, ST_WSEED, seed, sw);
    const int64_t s0 = (int64_t)(((uint64_t)sw[0] * (uint64_t)N) >> 32);
    const int axis = (int)(((uint64_t)sw[1] * 3ull) >> 32);
    double rd[3] = {0, 0, 0};
    if (random_dir) {
        const double a0 = 2.0 * (double)(sw[1] >> 8) * 0x1.0p-24;
        const double a1 = 2.0 * (double)(sw[2] >> 8) * 0x1.0p-24;
        const double rho = sqrt(fmax(a0 * (2.0 - a0), 0.0));
        rd[0] = 1.0 - a0;
        rd[1] = rho * sin(3.14159265358979323846 * a1);
        rd[2] = rho * cos(3.14159265358979323846 * a1);
    }
    double* proj0 = (double*)malloc(sizeof(double) * (size_t)N);
    char* in = (char*)calloc((size_t)N, 1);
    int64_t* stack = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
    for (int64_t i = 0; i < N; ++i) {
        const double* v = s_in + 3 * i;
        proj0[i] = random_dir ? v[0] * rd[0] + v[1] * rd[1] + v[2] * rd[2] : v[axis];
    }
#define ORC_FLIP(i)                                                                            \
    do {                                                                                       \
        double* v_ = s_out + 3 * (i);                                                          \
        if (random_dir) {                                                                      \
            const double d_ = v_[0] * rd[0] + v_[1] * rd[1] + v_[2] * rd[2];                   \
            for (int k_ = 0; k_ < 3; ++k_) v_[k_] -= 2.0 * d_ * rd[k_];                        \
        } else {                                                                               \
            v_[axis] = -v_[axis];                                                              \
        }                                                                                      \
    } while (0)
    double mm = INFINITY;
    int64_t top = 0, size = 1;
    in[s0] = 1;
    ORC_FLIP(s0);
    stack[top++] = s0;
    while (top > 0) {
        const int64_t a = stack[--top];
        const double* va = s_out + 3 * a;
        const double pa = random_dir ? va[0] * rd[0] + va[1] * rd[1] + va[2] * rd[2] : va[axis];
        for (int64_t k = rp[a]; k < rp[a + 1]; ++k) {
            const int64_t j = ci[k];
            uint32_t w[4];
            orc_philox_u64key((uint32_t)(k >> 2), ncl, rep, ST_WBOND, seed, w);
            const double u = (double)(w[k & 3] >> 8) * 0x1.0p-24;
            const double cpl = J * pa * proj0[j];
            const double p = cpl > 0.0 ? -expm1(-2.0 * beta * cpl) : 0.0;
            const double mg = fabs(p - u);
            if (mg < mm) mm = mg;
            if (!in[j] && u < p) {
                in[j] = 1;
                ORC_FLIP(j);
                stack[top++] = j;
                ++size;
            }
        }
    }
#undef ORC_FLIP
    *min_margin = mm;
    free(proj0);
    free(in);
    free(stack);
    return size;
}

/* ---------------------------------------------- multi-spin-coded 3D Ising */
/* Replays the 1-bit sweep: a site of anti-aligned count c is flipped if dE <= 0,
 * else iff U < T (T = thi<<16 | tlo) where U's bit 31-k is bit (h & 31) of the
 * level-k word Philox(h >> 5, 2*sweep + colour, rep, 8 + k/4)[k % 4]; an
 * "always" entry (thi = 65536) flips without a draw. */
EXPORT void orc_ising_msc_sweeps(int L, int8_t* s, const uint32_t* thi, const uint32_t* tlo, int nsweeps,
                                 uint64_t seed, uint32_t sweep0, uint32_t rep)
{
    const int dim = 3;
    int64_t N = (int64_t)L * L * L;
    for (int t = 0; t < nsweeps; ++t)
        for (int c = 0; c < 2; ++c) {
            uint32_t phase = 2u * (sweep0 + (uint32_t)t) + (uint32_t)c;
            for (int64_t i = 0; i < N; ++i) {
                if (colour_of(dim, L, i) != c) continue;
                int hs = 0;
                for (int slot = 0; slot < 6; ++slot) hs += s[orc_neighbour(dim, L, i, slot)];
                int e = s[i] * hs + 6;
                int flip;
                if (thi[e] > 0xFFFFu) {
                    flip = 1;
                } else {
                    uint32_t T = (thi[e] << 16) | tlo[e];
                    int64_t h = i >> 1;
                    uint32_t widx = (uint32_t)(h >> 5), j = (uint32_t)(h & 31);
                    flip = 0;
                    for (int k = 0; k < 32; ++k) {
                        uint32_t w[4];
                        orc_philox_u64key(widx, phase, rep, 8u + (uint32_t)(k >> 2), seed, w);
                        uint32_t rb = (w[k & 3] >> j) & 1u, tb = (T >> (31 - k)) & 1u;
                        if (rb != tb) { flip = (rb == 0u); break; }
                    }
                }
                if (flip) s[i] = (int8_t)-s[i];
            }
        }
}

/* -------------------------------------------- bit-sliced EA replica words */
/* Replays the 32-replica word sweep: replica j of word group gw sits in bit j.
 * Site h of colour c, replica j: q = s_j(i) sum_k b_k s_j(nbr_k); an "always"
 * entry flips, otherwise U_j < T_j with U_j's bit 31-k = bit j of
 * Philox(h, 2*sweep + colour, gw, 8 + k/4)[k % 4].  spins: 32 replicas x N,
 * replica-major, natural order; thi/tlo: 32 tables of 4*dim+1 entries. */
EXPORT void orc_ea_bits_sweeps(int dim, int L, int8_t* spins, const int8_t* bonds, const uint32_t* thi,
                               const uint32_t* tlo, int nsweeps, uint64_t seed, uint32_t sweep0, uint32_t gw)
{
    int64_t N = 1;
    for (int d = 0; d < dim; ++d) N *= L;
    const int z = 2 * dim, tbl = 2 * z + 1;
    for (int t = 0; t < nsweeps; ++t)
        for (int c = 0; c < 2; ++c) {
            uint32_t phase = 2u * (sweep0 + (uint32_t)t) + (uint32_t)c;
            for (int64_t i = 0; i < N; ++i) {
                if (colour_of(dim, L, i) != c) continue;
                uint32_t words[32];
                int have = 0; /* number of 4-level blocks computed */
                for (int j = 0; j < 32; ++j) {
                    int8_t* s = spins + (int64_t)j * N;
                    int hs = 0;
                    for (int slot = 0; slot < z; ++slot)
                        hs += bond_of(bonds, dim, L, i, slot) * s[orc_neighbour(dim, L, i, slot)];
                    int e = s[i] * hs + z;
                    uint32_t hi = thi[j * tbl + e], lo = tlo[j * tbl + e];
                    int flip;
                    if (hi > 0xFFFFu) {
                        flip = 1;
                    } else {
                        uint32_t T = (hi << 16) | lo;
                        flip = 0;
                        for (int k = 0; k < 32; ++k) {
                            while (have <= (k >> 2)) {
                                uint32_t w[4];
                                orc_philox_u64key((uint32_t)(i >> 1), phase, gw, 8u + (uint32_t)have, seed, w);
                                for (int m = 0; m < 4; ++m) words[4 * have + m] = w[m];
                                ++have;
                            }
                            uint32_t rb = (words[k] >> j) & 1u, tb = (T >> (31 - k)) & 1u;
                            if (rb != tb) { flip = (rb == 0u); break; }
                        }
                    }
                    if (flip) s[i] = (int8_t)-s[i];
                }
            }
        }
}
