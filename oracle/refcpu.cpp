/*
 * oracle/refcpu.cpp -- reference-faithful CPU restatement of MARQOV's update
 * hot path (test infrastructure; also the CPU baseline the bench times).
 *
 * Only tests/, the smoke check and the bench's cpu_baseline leg may load this
 * library, and only to check results or as the timed CPU baseline.
 *
 * Restated semantics, each following the cited reference code:
 *  - RNGCache over std::mt19937_64: 1024 cached 64-bit words per refill,
 *    integer(r) = x % r, real() = (x >> 11) * 2^-53
 *    (src/libmarqov/rngcache.h:156-219, 258-261; engine from core.h:372).
 *  - EMCS: ncluster Wolff moves seeded at integer(N), then nmetro sweeps of N
 *    random-site Metropolis attempts at integer(N) (src/libmarqov/emcs.h:69-109).
 *  - Metropolis move: propose, dE = interaction + on-site + flex, accept if
 *    dE <= 0 else if real() < exp(-beta*dE) (metropolis.h:86-126,
 *    metropolishelpers.h:109-142).  The reference evaluates exp(-beta*dE)
 *    through its own 2^x approximation; its test
 *    testsuite/components/5-metro_exp.cpp shows that rule and plain exp accept
 *    the same moves on 5e7 draws, so plain exp is used here.
 *  - Proposals: Ising sign flip (initializers.h:37-59); O(3) fresh uniform
 *    direction (initializers.h:61-81, randomdir.h:34-106); Potts uniform in
 *    [0,Q) (Potts.h:165-178); Phi4 phi + 0.5*r (Phi4.h:20-34,169).
 *  - dE terms: J*(s_new-s_old).sum_j b_ij s_j (metropolishelpers.h:456-623);
 *    Phi4 on-site (mass/beta)|phi|^2 + (lambda/beta)(|phi|^2-1)^2
 *    (Phi4.h:69-74,110-124; metropolishelpers.h:651-683); Potts flex term
 *    k*(#new - #old) (Potts.h:87-107; metropolishelpers.h:714-738).
 *  - Wolff: Ising (Ising.h:122-169), Heisenberg random Cartesian component
 *    (Heisenberg.h:191-243), Potts relabel (Potts.h:189-322), Phi4 random unit
 *    vector reflection via the generic move (wolff.h:48-125, Phi4.h:182-244);
 *    the add test is !update_accepted<2>(cpl) (wolff.h:35-39), i.e. add with
 *    probability 1 - exp(-2*beta*cpl) when cpl > 0.
 *  - Lattices: FixedDimRegularHypercubic neighbour arithmetic
 *    (regular_hypercubic.h:31-66) and CSR adjacency lists (graph_from_csv.h).
 *  - Observables: |sum s|/N (observables.h:32-66); E/N = J sum_i sum_nbr
 *    s_i.s_j / 2N (observables.h:173-231) plus on-site terms
 *    (observables.h:237-352); Potts (Q*max - N)/(N(Q-1)) (Potts.h:40-65).
 */
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#define EXPORT extern "C" __attribute__((visibility("default")))

namespace {

/* Planted mistakes for the tests' power checks (0 = none, the faithful
 * restatement; 1 = Phi4 Wolff renormalises phi to unit length after each
 * reflection, which Phi4.h:241-242 does not do). */
int g_plant = 0;

enum Model { ISING = 1, EA = 2, POTTS = 3, HEIS = 4, PHI4 = 5 };

struct Cache {
    std::mt19937_64 eng;
    uint64_t buf[1024];
    int pos = 1024;
    explicit Cache(uint64_t seed) : eng(seed) {}
    inline uint64_t integer()
    {
        if (pos >= 1024) {
            for (int i = 0; i < 1024; ++i) buf[i] = eng();
            pos = 0;
        }
        return buf[pos++];
    }
    inline uint64_t integer(uint64_t r) { return integer() % r; }
    inline double real() { return double(integer() >> 11) * 0x1.0p-53; }
};

/* uniform point on S^2, component order as randomdir.h:45-65 */
inline void rnddir(Cache& rng, double v[3])
{
    double a0 = 2.0 * rng.real();
    double a1 = 2.0 * rng.real();
    double r1 = a0 * (2.0 - a0);
    v[0] = 1.0 - a0;
    double rho = std::sqrt(r1);
    v[2] = rho * std::cos(M_PI * a1);
    v[1] = rho * std::sin(M_PI * a1);
}

struct Regular {
    int dim, L;
    int64_t N;
    int64_t pows[4];
    Regular(int d, int l) : dim(d), L(l), N(1)
    {
        for (int i = 0; i < d; ++i) { pows[i] = N; N *= l; }
    }
    int z(int64_t) const { return 2 * dim; }
    /* FixedDimRegularHypercubic::nbrs arithmetic */
    inline int64_t nb(int64_t i, int slot) const
    {
        int j = slot >> 1;
        int64_t pl = pows[j] * L;
        if ((slot & 1) == 0) {
            int64_t c = i + pows[j];
            return (i / pl) * pl + c % pl;
        }
        int64_t c = i - pows[j];
        if (c < (i / pl) * pl) return (i / pl) * pl + (c + pl) % pl;
        return c;
    }
};

struct Csr {
    int64_t N;
    const int* rp;
    const int* ci;
    int z(int64_t i) const { return rp[i + 1] - rp[i]; }
    inline int64_t nb(int64_t i, int slot) const { return ci[rp[i] + slot]; }
};

struct Params {
    int model;
    double J;
    int Q;
    double lambda, mass;
};

/* bonds: regular -> bonds[i*dim+d] for the +d bond; CSR -> per slot */
template <class Lat>
inline int bond(const Lat& lat, const int8_t* b, int64_t i, int slot);
template <>
inline int bond<Regular>(const Regular& lat, const int8_t* b, int64_t i, int slot)
{
    if (!b) return 1;
    int d = slot >> 1;
    if ((slot & 1) == 0) return b[i * lat.dim + d];
    return b[lat.nb(i, slot) * lat.dim + d];
}
template <>
inline int bond<Csr>(const Csr& lat, const int8_t* b, int64_t i, int slot)
{
    return b ? b[lat.rp[i] + slot] : 1;
}

inline bool metro_accept(double dE, double beta, Cache& rng)
{
    if (dE <= 0) return true;
    return rng.real() < std::exp(-beta * dE);
}

inline bool wolff_add(double cpl, double beta, Cache& rng)
{
    /* !update_accepted<2>(cpl): cpl <= 0 never adds */
    if (cpl <= 0) return false;
    return !(rng.real() < std::exp(-2.0 * beta * cpl));
}

template <class Lat>
struct Sim {
    const Lat& lat;
    Params p;
    const int8_t* bonds;
    int* is;      /* Ising/EA/Potts spins */
    double* ds;   /* O(3)/Phi4, 3 per site */
    Cache rng;
    std::vector<int> stack;
    Sim(const Lat& l, Params pp, const int8_t* b, int* i_s, double* d_s, uint64_t seed)
        : lat(l), p(pp), bonds(b), is(i_s), ds(d_s), rng(seed), stack(4096 / sizeof(int))
    {}

    void init()
    {
        for (int64_t i = 0; i < lat.N; ++i) {
            switch (p.model) {
            case ISING:
            case EA: is[i] = (rng.real() > 0.5) ? 1 : -1; break; /* Ising.h:99-107 */
            case POTTS: is[i] = int(rng.integer(p.Q)); break;     /* Potts.h:149-156 */
            default: rnddir(rng, ds + 3 * i); break;              /* core.h:774-782 */
            }
        }
    }

    int metropolis(double beta, int64_t i)
    {
        const int z = lat.z(i);
        switch (p.model) {
        case ISING:
        case EA: {
            double h = 0;
            for (int s = 0; s < z; ++s) h += bond(lat, bonds, i, s) * is[lat.nb(i, s)];
            double dE = p.J * double(-2 * is[i]) * h;
            if (metro_accept(dE, beta, rng)) { is[i] = -is[i]; return 1; }
            return 0;
        }
        case POTTS: {
            int snew = int(rng.integer(p.Q));
            int sold = is[i];
            int en = 0, eo = 0;
            for (int s = 0; s < z; ++s) {
                int v = is[lat.nb(i, s)];
                eo += (v == sold);
                en += (v == snew);
            }
            double dE = p.J * double(en - eo);
            if (metro_accept(dE, beta, rng)) { is[i] = snew; return 1; }
            return 0;
        }
        case HEIS:
        case PHI4: {
            double* so = ds + 3 * i;
            double sn[3];
            if (p.model == HEIS) {
                rnddir(rng, sn);
            } else {
                double r[3];
                rnddir(rng, r);
                for (int k = 0; k < 3; ++k) sn[k] = so[k] + 0.5 * r[k];
            }
            double nb[3] = {0, 0, 0};
            for (int s = 0; s < z; ++s) {
                const double* v = ds + 3 * lat.nb(i, s);
                double b = bond(lat, bonds, i, s);
                nb[0] += b * v[0]; nb[1] += b * v[1]; nb[2] += b * v[2];
            }
            double dE = p.J * ((sn[0] - so[0]) * nb[0] + (sn[1] - so[1]) * nb[1] + (sn[2] - so[2]) * nb[2]);
            if (p.model == PHI4) {
                double n2 = sn[0] * sn[0] + sn[1] * sn[1] + sn[2] * sn[2];
                double o2 = so[0] * so[0] + so[1] * so[1] + so[2] * so[2];
                dE += (p.mass / beta) * (n2 - o2);
                dE += (p.lambda / beta) * ((n2 - 1) * (n2 - 1) - (o2 - 1) * (o2 - 1));
            }
            if (metro_accept(dE, beta, rng)) { so[0] = sn[0]; so[1] = sn[1]; so[2] = sn[2]; return 1; }
            return 0;
        }
        }
        return 0;
    }

    void push(int q, int64_t v)
    {
        if (q >= int(stack.size())) stack.resize(2 * stack.size());
        stack[q] = int(v);
    }

    int64_t wolff(double beta, int64_t seed)
    {
        int q = 0;
        int64_t size = 1;
        stack[0] = int(seed);
        switch (p.model) {
        case ISING:
        case EA: {
            const int val = is[seed];
            is[seed] = -val;
            const double prob = -std::expm1(2.0 * beta * p.J); /* Ising.h:139 */
            while (q >= 0) {
                int64_t cur = stack[q--];
                for (int s = 0; s < lat.z(cur); ++s) {
                    int64_t j = lat.nb(cur, s);
                    int lval = (p.model == EA) ? val * bond(lat, bonds, cur, s) : val;
                    if (p.model == EA) {
                        /* generic embedding for bonded Ising: cpl = J*b*s_cur*s_j with s_cur flipped */
                        double cpl = p.J * bond(lat, bonds, cur, s) * is[cur] * is[j];
                        if (wolff_add(cpl, beta, rng)) { push(++q, j); ++size; is[j] = -is[j]; }
                    } else if (is[j] == lval) {
                        if (rng.real() < prob) { push(++q, j); ++size; is[j] = -is[j]; }
                    }
                }
            }
            return size;
        }
        case POTTS: {
            const int oldq = is[seed];
            int newq;
            do { newq = int(rng.integer(p.Q)); } while (newq == oldq);
            is[seed] = newq;
            while (q >= 0) {
                int64_t cur = stack[q--];
                for (int s = 0; s < lat.z(cur); ++s) {
                    int64_t j = lat.nb(cur, s);
                    double cpl = p.J * ((is[j] == oldq) ? -0.5 : 0.0);
                    if (wolff_add(cpl, beta, rng)) { push(++q, j); ++size; is[j] = newq; }
                }
            }
            return size;
        }
        case HEIS: {
            const int r = int(rng.integer(3));
            ds[3 * seed + r] *= -1;
            while (q >= 0) {
                int64_t cur = stack[q--];
                for (int s = 0; s < lat.z(cur); ++s) {
                    int64_t j = lat.nb(cur, s);
                    double cpl = p.J * bond(lat, bonds, cur, s) * ds[3 * cur + r] * ds[3 * j + r];
                    if (wolff_add(cpl, beta, rng)) { push(++q, j); ++size; ds[3 * j + r] *= -1; }
                }
            }
            return size;
        }
        case PHI4: {
            double rd[3];
            rnddir(rng, rd);
            auto flip = [&](int64_t i) {
                double* v = ds + 3 * i;
                double d = v[0] * rd[0] + v[1] * rd[1] + v[2] * rd[2];
                for (int k = 0; k < 3; ++k) v[k] -= 2 * d * rd[k];
                if (g_plant == 1) {
                    const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                    for (int k = 0; k < 3; ++k) v[k] /= n;
                }
            };
            auto proj = [&](int64_t i) {
                const double* v = ds + 3 * i;
                return v[0] * rd[0] + v[1] * rd[1] + v[2] * rd[2];
            };
            flip(seed);
            while (q >= 0) {
                int64_t cur = stack[q--];
                for (int s = 0; s < lat.z(cur); ++s) {
                    int64_t j = lat.nb(cur, s);
                    double cpl = p.J * bond(lat, bonds, cur, s) * proj(cur) * proj(j);
                    if (wolff_add(cpl, beta, rng)) { push(++q, j); ++size; flip(j); }
                }
            }
            return size;
        }
        }
        return size;
    }

    /* one EMCS (emcs.h:69-109); returns mean cluster size (NaN if ncluster=0) */
    double emcs(double beta, int ncluster, int nmetro, int64_t* nacc)
    {
        double avg = 0;
        for (int j = 0; j < ncluster; ++j) {
            int64_t s = int64_t(rng.integer(uint64_t(lat.N)));
            avg += double(wolff(beta, s));
        }
        int64_t acc = 0;
        for (int j = 0; j < nmetro; ++j)
            for (int64_t i = 0; i < lat.N; ++i) acc += metropolis(beta, int64_t(rng.integer(uint64_t(lat.N))));
        if (nacc) *nacc += acc;
        return avg / ncluster;
    }

    /* obs[0] = E/N, obs[1] = |m| (Potts: the Potts magnetisation), obs[2] = m^2,
     * obs[3] = m^4 (the moments Binder's cumulant 1 - <m^4>/(3<m^2>^2) is formed
     * from), obs[4] = mean |phi|^2 per site (1 for unit spins) */
    void measure(double beta, double* obs) const
    {
        moments(beta, obs);
        obs[2] = obs[1] * obs[1];
        obs[3] = obs[2] * obs[2];
    }

    void moments(double beta, double* obs) const
    {
        obs[4] = 1.0;
        const int64_t N = lat.N;
        double e = 0;
        if (p.model == POTTS) {
            std::vector<int64_t> cnt(p.Q, 0);
            int64_t same = 0;
            for (int64_t i = 0; i < N; ++i) {
                cnt[is[i]]++;
                for (int s = 0; s < lat.z(i); ++s) same += (is[lat.nb(i, s)] == is[i]);
            }
            int64_t mx = *std::max_element(cnt.begin(), cnt.end());
            obs[0] = p.J * double(same) / double(2 * N);
            obs[1] = double(p.Q * mx - N) / double(N * (p.Q - 1));
            return;
        }
        if (p.model == ISING || p.model == EA) {
            int64_t m = 0, acc = 0;
            for (int64_t i = 0; i < N; ++i) {
                m += is[i];
                for (int s = 0; s < lat.z(i); ++s) acc += int64_t(bond(lat, bonds, i, s)) * is[i] * is[lat.nb(i, s)];
            }
            obs[0] = p.J * double(acc) / double(2 * N);
            obs[1] = std::fabs(double(m)) / double(N);
            return;
        }
        double M[3] = {0, 0, 0};
        double on = 0, phi2 = 0;
        for (int64_t i = 0; i < N; ++i) {
            const double* v = ds + 3 * i;
            for (int k = 0; k < 3; ++k) M[k] += v[k];
            phi2 += v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            for (int s = 0; s < lat.z(i); ++s) {
                const double* w = ds + 3 * lat.nb(i, s);
                e += bond(lat, bonds, i, s) * (v[0] * w[0] + v[1] * w[1] + v[2] * w[2]);
            }
            if (p.model == PHI4) {
                double n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                on += (p.mass / beta) * n2 + (p.lambda / beta) * (n2 - 1) * (n2 - 1);
            }
        }
        obs[0] = p.J * e / double(2 * N) + on / double(N);
        obs[1] = std::sqrt(M[0] * M[0] + M[1] * M[1] + M[2] * M[2]) / double(N);
        obs[4] = phi2 / double(N);
    }
};

template <class Lat>
int run_impl(const Lat& lat, Params p, const int8_t* bonds, double beta, int ncluster, int nmetro,
             int nwarm, int nmeas, uint64_t seed, int init, int* is, double* ds, double* obs,
             double* clusters, int64_t* nacc)
{
    Sim<Lat> sim(lat, p, bonds, is, ds, seed);
    if (init) sim.init();
    for (int t = 0; t < nwarm; ++t) sim.emcs(beta, ncluster, nmetro, nullptr);
    for (int t = 0; t < nmeas; ++t) {
        double c = sim.emcs(beta, ncluster, nmetro, nacc);
        if (clusters) clusters[t] = c;
        if (obs) sim.measure(beta, obs + 5 * t);
    }
    return 0;
}

}  // namespace

/*
 * Run one simulation: optional init, nwarm + nmeas EMCS, one measurement per
 * measured EMCS (obs[5*t .. 5*t+4] = E/N, |m|, m^2, m^4, mean |phi|^2).  Integer models use is[N],
 * O(3)/Phi4 use ds[3N].  bonds may be NULL (uniform coupling).
 */
EXPORT int ref_run_regular(int model, int dim, int L, double J, int Q, double lambda, double mass,
                           const int8_t* bonds, double beta, int ncluster, int nmetro, int nwarm,
                           int nmeas, uint64_t seed, int init, int* is, double* ds, double* obs,
                           double* clusters, int64_t* nacc)
{
    Regular lat(dim, L);
    Params p{model, J, Q, lambda, mass};
    return run_impl(lat, p, bonds, beta, ncluster, nmetro, nwarm, nmeas, seed, init, is, ds, obs, clusters, nacc);
}

EXPORT int ref_run_csr(int model, int64_t N, const int* rp, const int* ci, double J, int Q, double lambda,
                       double mass, const int8_t* bonds, double beta, int ncluster, int nmetro, int nwarm,
                       int nmeas, uint64_t seed, int init, int* is, double* ds, double* obs,
                       double* clusters, int64_t* nacc)
{
    Csr lat{N, rp, ci};
    Params p{model, J, Q, lambda, mass};
    return run_impl(lat, p, bonds, beta, ncluster, nmetro, nwarm, nmeas, seed, init, is, ds, obs, clusters, nacc);
}

/* test hook: plant a known mistake (see g_plant) to show a check can fail */
EXPORT void ref_set_plant(int which) { g_plant = which; }

EXPORT uint64_t ref_mt19937_64_nth(uint64_t seed, int64_t n)
{
    std::mt19937_64 e(seed);
    uint64_t v = 0;
    for (int64_t i = 0; i < n; ++i) v = e();
    return v;
}

/* RNGCache draws for pinning: kind 0 = integer(), 1 = real() bits, 2 = integer(range) */
EXPORT void ref_rngcache_draws(uint64_t seed, int kind, uint64_t range, int64_t n, uint64_t* out)
{
    Cache c(seed);
    for (int64_t i = 0; i < n; ++i) {
        if (kind == 0) out[i] = c.integer();
        else if (kind == 1) { double r = c.real(); std::memcpy(&out[i], &r, 8); }
        else out[i] = c.integer(range);
    }
}

/*
 * CPU baseline timing: nthreads independent replicas (one per thread, as
 * CXX11Scheduler runs one simulation per worker), each doing nwarm + nsweeps
 * Metropolis sweeps (ncluster = 0) on a regular lattice from a random start.
 * out[0] = aggregate attempts/ns over all threads, out[1] = per-thread mean
 * attempts/ns, out[2] = wall seconds of the timed part.
 * With wolff != 0 the replicas instead grow nsweeps Wolff clusters each and
 * out[] reports cluster sites/ns.
 */
EXPORT int ref_bench_regular(int model, int dim, int L, double J, int Q, double beta, int nwarm,
                             int nsweeps, int nthreads, uint64_t seed, int wolff, double* out)
{
    Regular lat(dim, L);
    Params p{model, J, Q, 4.5, 1.0};
    const bool vec = (model == HEIS || model == PHI4);
    std::vector<std::vector<int>> is(nthreads);
    std::vector<std::vector<double>> ds(nthreads);
    std::vector<double> secs(nthreads, 0), work(nthreads, 0);
    std::atomic<int> ready{0};
    std::vector<std::thread> th;
    auto t0 = std::chrono::steady_clock::now();
    std::atomic<int64_t> t_start_ns{0};
    for (int t = 0; t < nthreads; ++t) {
        th.emplace_back([&, t]() {
            if (vec) ds[t].assign(size_t(3 * lat.N), 0.0);
            else is[t].assign(size_t(lat.N), 0);
            Sim<Regular> sim(lat, p, nullptr, is[t].data(), ds[t].data(), seed + uint64_t(t));
            sim.init();
            for (int w = 0; w < nwarm; ++w) sim.emcs(beta, wolff ? 1 : 0, wolff ? 0 : 1, nullptr);
            ready++;
            while (ready.load() < nthreads) std::this_thread::yield();
            auto a = std::chrono::steady_clock::now();
            double sites = 0;
            if (wolff) {
                for (int c = 0; c < nsweeps; ++c) sites += double(sim.wolff(beta, int64_t(sim.rng.integer(uint64_t(lat.N)))));
            } else {
                for (int s = 0; s < nsweeps; ++s)
                    for (int64_t i = 0; i < lat.N; ++i) sim.metropolis(beta, int64_t(sim.rng.integer(uint64_t(lat.N))));
                sites = double(nsweeps) * double(lat.N);
            }
            auto b = std::chrono::steady_clock::now();
            secs[t] = std::chrono::duration<double>(b - a).count();
            work[t] = sites;
        });
    }
    for (auto& x : th) x.join();
    (void)t0;
    (void)t_start_ns;
    double wall = *std::max_element(secs.begin(), secs.end());
    double total = 0, per = 0;
    for (int t = 0; t < nthreads; ++t) {
        total += work[t];
        per += work[t] / (secs[t] * 1e9);
    }
    out[0] = total / (wall * 1e9);
    out[1] = per / nthreads;
    out[2] = wall;
    return 0;
}

/*
 * CPU baseline of the Wolff update on a given state (the restatement's stack
 * Wolff, Ising.h:122-169, Heisenberg.h:191-243, Potts.h:262-322, wolff.h:48-125):
 * nthreads threads, each on its own copy of the state with its own RNGCache
 * (seed + thread), grow single clusters at uniformly drawn seed sites until
 * `seconds` of wall time have passed, as CXX11Scheduler runs one simulation per
 * worker thread.  Regular lattice when rp is NULL (dim, L), else the CSR graph
 * (N, rp, ci).  Integer models start from is0[N], O(3)/Phi4 from ds0[3N].
 * out[0] = aggregate cluster sites/ns, out[1] = mean per-thread sites/ns,
 * out[2] = wall seconds, out[3] = clusters grown, out[4] = mean cluster size.
 */
EXPORT int ref_bench_wolff(int model, int dim, int L, int64_t N, const int* rp, const int* ci, double J, int Q,
                           double beta, const int* is0, const double* ds0, double seconds, int nthreads,
                           uint64_t seed, double* out)
{
    const bool vec = (model == HEIS || model == PHI4);
    Params p{model, J, Q, 4.5, 1.0};
    std::vector<double> secs(nthreads, 0), sites(nthreads, 0), ncl(nthreads, 0);
    std::atomic<int> ready{0};
    auto work = [&](int t, auto& lat) {
        std::vector<int> is;
        std::vector<double> ds;
        if (vec) ds.assign(ds0, ds0 + 3 * lat.N);
        else is.assign(is0, is0 + lat.N);
        using Lat = std::decay_t<decltype(lat)>;
        Sim<Lat> sim(lat, p, nullptr, is.data(), ds.data(), seed + uint64_t(t));
        ready++;
        while (ready.load() < nthreads) std::this_thread::yield();
        const auto a = std::chrono::steady_clock::now();
        double s = 0, n = 0, el = 0;
        do {
            s += double(sim.wolff(beta, int64_t(sim.rng.integer(uint64_t(lat.N)))));
            n += 1;
            el = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
        } while (el < seconds);
        secs[t] = el;
        sites[t] = s;
        ncl[t] = n;
    };
    std::vector<std::thread> th;
    if (rp) {
        Csr lat{N, rp, ci};
        for (int t = 0; t < nthreads; ++t) th.emplace_back([&, t]() { work(t, lat); });
        for (auto& x : th) x.join();
    } else {
        Regular lat(dim, L);
        for (int t = 0; t < nthreads; ++t) th.emplace_back([&, t]() { work(t, lat); });
        for (auto& x : th) x.join();
    }
    double wall = *std::max_element(secs.begin(), secs.end());
    double total = 0, per = 0, clusters = 0;
    for (int t = 0; t < nthreads; ++t) {
        total += sites[t];
        per += sites[t] / (secs[t] * 1e9);
        clusters += ncl[t];
    }
    out[0] = total / (wall * 1e9);
    out[1] = per / nthreads;
    out[2] = wall;
    out[3] = clusters;
    out[4] = total / clusters;
    return 0;
}
