/*
 * oracle/ref_metropolis_driver.cpp -- drives the reference's own Metropolis
 * move (test infrastructure; built only into oracle/_ref/).
 *
 * This program is ours; it includes the reference headers where they lie
 * (src/libmarqov/metropolis.h, src/libmarqov/rngcache.h,
 * src/hamiltonian/util/termcollection.h, src/lattice/util/regular_lattice.h)
 * and instantiates MARQOV::Metropolis<H, RegularLattice>::move exactly as
 * EMCS does (src/libmarqov/emcs.h:93-106), with the RNGCache<std::mt19937_64>
 * the reference Core owns (core.h:372).  The Hamiltonians below are declared
 * the way a MARQOV user declares one (an `interactions` array of
 * Standard_Interaction terms, as src/hamiltonian/Ising.h:62-80 does); the
 * initial states follow Ising::initstatespace (Ising.h:99-107) and
 * Core::init_hot (core.h:774-782).
 *
 * The reference's Wolff/EMCS/Core headers need MARQOV::Space from the
 * HDF5-bound core.h and are not built here.
 *
 * usage: ref_metropolis ising|heis DIM L BETA J NSWEEPS SEED
 *   prints the final state, one site per line, in site-index order.
 * usage: ref_metropolis ea DIM L BETA J NSWEEPS SEED
 *   the +-J Edwards-Anderson model: Ising spins on the reference's
 *   RegularRandomBond<BimodalPDF> lattice (src/lattice/regular_random_bond.h,
 *   bonds drawn and symmetrised there, from std::random_device), whose bnds the
 *   Metropolis move weights the neighbour sum with (metropolishelpers.h:456-537);
 *   prints the N x 2*DIM bonds in neighbour-slot order, then the final state.
 * usage: ref_metropolis phi4 DIM L BETA LAMBDA MASS NSWEEPS SEED
 *   the same for the Phi4 model (J = -1, Phi4.h:93), declared as a user model:
 *   the reference's Standard_Interaction and Onsite_Quadratic(mass/beta) terms,
 *   a user on-site term lambda/beta (|phi|^2 - 1)^2 (the model Phi4.h:69-74 and
 *   :110-124 define) and the radial proposal phi + 0.5 r as the Initializer
 *   specialisation (Phi4.h:20-34, :169), so the reference's Metropolis move
 *   assembles the on-site energy difference (metropolishelpers.h:651-683).
 * usage: ref_metropolis bench ising|heis DIM L BETA J NWARM NSWEEPS SEED NTHREADS
 *   times the reference's move as the CPU baseline: NTHREADS independent
 *   simulations, one per thread (the CXX11Scheduler default of one simulation per
 *   hardware thread, marqovscheduler.h:281-286), NWARM untimed then NSWEEPS
 *   timed sweeps of N random-site moves each; prints one JSON line with the
 *   aggregate and per-thread attempts/ns.
 */
#include <array>
#include <cmath>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include <algorithm>  // used, not included, by lattice/regular_random_bond.h

#include "util/svmath.h"
#include "metropolis.h"
#include "hamiltonian/util/termcollection.h"
#include "lattice/util/regular_lattice.h"
#include "lattice/regular_random_bond.h"

template <class SV>
struct UserHam {
    double J;
    static constexpr int SymD = std::tuple_size<SV>::value;
    typedef SV StateVector;
    std::array<Standard_Interaction<SV>*, 1> interactions;
    explicit UserHam(double j) : J(j), interactions{{new Standard_Interaction<SV>(j)}} {}
    ~UserHam() { delete interactions[0]; }
};

static void run_ea(int dim, int L, double beta, double J, int nsweeps, unsigned long long seed)
{
    typedef std::array<int, 1> SV;
    RegularRandomBond<BimodalPDF> lat(L, dim);
    UserHam<SV> ham(J);
    RNGCache<std::mt19937_64> rng(seed);
    std::vector<SV> s(lat.size());
    for (std::size_t i = 0; i < lat.size(); ++i) s[i][0] = (rng.real() > 0.5) ? 1 : -1;
    for (int t = 0; t < nsweeps; ++t)
        for (std::size_t i = 0; i < lat.size(); ++i) {
            const int rsite = rng.integer(lat.size());
            MARQOV::Metropolis<UserHam<SV>, RegularRandomBond<BimodalPDF>>::move(ham, lat, s, rng, beta, rsite);
        }
    for (std::size_t i = 0; i < lat.size(); ++i) {
        const auto b = lat.bnds(1, (int)i);
        for (std::size_t k = 0; k < b.size(); ++k) std::printf(k ? " %d" : "%d", b[k]);
        std::printf("\n");
    }
    for (std::size_t i = 0; i < lat.size(); ++i) std::printf("%d\n", s[i][0]);
}

typedef std::array<double, 3> Phi4SV;

class UserFourthMinusOne : public OnSite<Phi4SV, double>
{
public:
    explicit UserFourthMinusOne(double h) : OnSite<Phi4SV, double>(h) {}
    double get(const Phi4SV& phi) override { return std::pow(dot(phi, phi) - 1.0, 2); }
};

struct UserPhi4 {
    static constexpr int SymD = 3;
    typedef Phi4SV StateVector;
    std::array<Standard_Interaction<Phi4SV>*, 1> interactions;
    Onsite_Quadratic<Phi4SV> quadratic;
    UserFourthMinusOne fourth;
    std::vector<OnSite<Phi4SV, double>*> onsite;
    UserPhi4(double beta, double lambda, double mass)
        : interactions{{new Standard_Interaction<Phi4SV>(-1.0)}}, quadratic(mass / beta), fourth(lambda / beta)
    {
        onsite.push_back(&quadratic);
        onsite.push_back(&fourth);
    }
    ~UserPhi4() { delete interactions[0]; }
};

template <>
class Initializer<UserPhi4>
{
public:
    template <class RNG>
    static Phi4SV newsv(const Phi4SV& osv, RNG& rng)
    {
        Phi4SV dir;
        set_to_rnddir(rng, dir);
        return osv + mult(0.5, dir);
    }
};

static void run_phi4(int dim, int L, double beta, double lambda, double mass, int nsweeps, unsigned long long seed)
{
    RegularLattice lat(L, dim);
    UserPhi4 ham(beta, lambda, mass);
    RNGCache<std::mt19937_64> rng(seed);
    std::vector<Phi4SV> s(lat.size());
    for (std::size_t i = 0; i < lat.size(); ++i) set_to_rnddir(rng, s[i]);  // Core::init_hot, core.h:774-782
    for (int t = 0; t < nsweeps; ++t)
        for (std::size_t i = 0; i < lat.size(); ++i) {
            const int rsite = rng.integer(lat.size());
            MARQOV::Metropolis<UserPhi4, RegularLattice>::move(ham, lat, s, rng, beta, rsite);
        }
    for (std::size_t i = 0; i < lat.size(); ++i) std::printf("%.17g %.17g %.17g\n", s[i][0], s[i][1], s[i][2]);
}

template <class SV>
static void run(int dim, int L, double beta, double J, int nsweeps, unsigned long long seed, bool ising)
{
    RegularLattice lat(L, dim);
    UserHam<SV> ham(J);
    RNGCache<std::mt19937_64> rng(seed);
    std::vector<SV> s(lat.size());
    for (std::size_t i = 0; i < lat.size(); ++i) {
        if (ising) s[i][0] = (rng.real() > 0.5) ? 1 : -1;
        else set_to_rnddir(rng, s[i]);
    }
    for (int t = 0; t < nsweeps; ++t)
        for (std::size_t i = 0; i < lat.size(); ++i) {
            const int rsite = rng.integer(lat.size());
            MARQOV::Metropolis<UserHam<SV>, RegularLattice>::move(ham, lat, s, rng, beta, rsite);
        }
    for (std::size_t i = 0; i < lat.size(); ++i) {
        for (int k = 0; k < UserHam<SV>::SymD; ++k) std::printf(k ? " %.17g" : "%.17g", double(s[i][k]));
        std::printf("\n");
    }
}

template <class SV>
static void bench(int dim, int L, double beta, double J, int nwarm, int nsweeps, unsigned long long seed, int nth,
                  bool ising)
{
    RegularLattice lat(L, dim);
    const std::size_t N = lat.size();
    std::vector<double> secs(nth, 0.0);
    std::atomic<int> ready{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nth; ++t)
        th.emplace_back([&, t]() {
            UserHam<SV> ham(J);
            RNGCache<std::mt19937_64> rng(seed + (unsigned long long)t);
            std::vector<SV> s(N);
            for (std::size_t i = 0; i < N; ++i) {
                if (ising) s[i][0] = (rng.real() > 0.5) ? 1 : -1;
                else set_to_rnddir(rng, s[i]);
            }
            auto sweep = [&]() {
                for (std::size_t i = 0; i < N; ++i)
                    MARQOV::Metropolis<UserHam<SV>, RegularLattice>::move(ham, lat, s, rng, beta, rng.integer(N));
            };
            for (int w = 0; w < nwarm; ++w) sweep();
            ready++;
            while (ready.load() < nth) std::this_thread::yield();
            const auto a = std::chrono::steady_clock::now();
            for (int w = 0; w < nsweeps; ++w) sweep();
            secs[t] = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
        });
    for (auto& x : th) x.join();
    double wall = 0.0, per = 0.0;
    for (double x : secs) {
        wall = x > wall ? x : wall;
        per += (double)N * nsweeps / (x * 1e9);
    }
    std::printf("{\"attempts_per_ns\": %.6g, \"per_thread\": %.6g, \"threads\": %d, \"seconds\": %.3f, \"sites\": %zu}\n",
                (double)N * nsweeps * nth / (wall * 1e9), per / nth, nth, wall, N);
}

int main(int argc, char** argv)
{
    if (argc == 11 && std::strcmp(argv[1], "bench") == 0) {
        const bool ising = std::strcmp(argv[2], "ising") == 0;
        const int dim = std::atoi(argv[3]), L = std::atoi(argv[4]);
        const double beta = std::atof(argv[5]), J = std::atof(argv[6]);
        const int nwarm = std::atoi(argv[7]), nsweeps = std::atoi(argv[8]);
        const unsigned long long seed = std::strtoull(argv[9], nullptr, 10);
        const int nth = std::atoi(argv[10]) > 0 ? std::atoi(argv[10]) : 1;
        if (ising) bench<std::array<int, 1>>(dim, L, beta, J, nwarm, nsweeps, seed, nth, true);
        else bench<std::array<double, 3>>(dim, L, beta, J, nwarm, nsweeps, seed, nth, false);
        return 0;
    }
    if (argc == 8 && std::strcmp(argv[1], "ea") == 0) {
        run_ea(std::atoi(argv[2]), std::atoi(argv[3]), std::atof(argv[4]), std::atof(argv[5]), std::atoi(argv[6]),
               std::strtoull(argv[7], nullptr, 10));
        return 0;
    }
    if (argc == 9 && std::strcmp(argv[1], "phi4") == 0) {
        run_phi4(std::atoi(argv[2]), std::atoi(argv[3]), std::atof(argv[4]), std::atof(argv[5]), std::atof(argv[6]),
                 std::atoi(argv[7]), std::strtoull(argv[8], nullptr, 10));
        return 0;
    }
    if (argc != 8) {
        std::fprintf(stderr, "usage: %s ising|heis DIM L BETA J NSWEEPS SEED\n", argv[0]);
        return 2;
    }
    const bool ising = std::strcmp(argv[1], "ising") == 0;
    const int dim = std::atoi(argv[2]);
    const int L = std::atoi(argv[3]);
    const double beta = std::atof(argv[4]);
    const double J = std::atof(argv[5]);
    const int nsweeps = std::atoi(argv[6]);
    const unsigned long long seed = std::strtoull(argv[7], nullptr, 10);
    if (ising) run<std::array<int, 1>>(dim, L, beta, J, nsweeps, seed, true);
    else run<std::array<double, 3>>(dim, L, beta, J, nsweeps, seed, false);
    return 0;
}
