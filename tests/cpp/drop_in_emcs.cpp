// The EMCS specialisation a MARQOV maintainer adds (the code shown in the
// integration notes), compiled and run here.  MARQOV's own headers cannot be
// compiled in this image (core.h needs HDF5), so the few names the
// specialisation binds to are declared below with the signatures it uses:
//   FixedDimRegularHypercubic<dim>(L) with size()   (lattice/regular_hypercubic.h:15-100)
//   Ising<int> with member J                        (hamiltonian/Ising.h:62-109)
//   RNGCache<RNG>::integer()                        (libmarqov/rngcache.h:156-179)
//   template <class H, class L> struct EMCS         (libmarqov/emcs.h:62-67)
// They are interface shapes for the compiler, not MARQOV's implementation.
//
// The program runs a Core-like loop (warm-up EMCS, then game-loop EMCS each
// followed by a host-side measurement of the statespace, core.h:1001-1050)
// through the specialisation, and the same steps through Batch/GpuEMCS
// directly, and prints one JSON line: whether the two final states are equal,
// and how many state downloads and temperature uploads the specialisation made.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <random>
#include <vector>

#include "marqov_amd.hpp"

// ---- interface shapes (see above)
template <int dim>
struct FixedDimRegularHypercubic {
    int L;
    explicit FixedDimRegularHypercubic(int l) : L(l) {}
    std::size_t size() const
    {
        std::size_t n = 1;
        for (int d = 0; d < dim; ++d) n *= (std::size_t)L;
        return n;
    }
};
template <class S>
struct Ising {
    double J = -1.0;
};
template <class RNG>
struct RNGCache {
    RNG eng;
    explicit RNGCache(uint64_t seed) : eng(seed) {}
    uint64_t integer() { return eng(); }
};
namespace MARQOV {
template <class H, class L>
struct EMCS;
}

// ---- the maintainer's code
template <int dim>
struct GpuRegular : public FixedDimRegularHypercubic<dim> {
    marqov_amd::Context& ctx;
    marqov_amd::RegularLattice dev;
    mutable std::unique_ptr<marqov_amd::Batch> batch;  // one replica per Core
    mutable std::unique_ptr<marqov_amd::GpuEMCS> step;
    mutable double beta_on_device = NAN;
    mutable bool host_current = true;  // the Core's statespace holds the device state
    bool sync_host = false;            // set by the game loop: Core measures after this step
    mutable int downloads = 0, beta_uploads = 0;
    GpuRegular(marqov_amd::Context& c, int L) : FixedDimRegularHypercubic<dim>(L), ctx(c), dev(c, dim, L) {}
};

namespace MARQOV {
template <int dim>
struct EMCS<Ising<int>, GpuRegular<dim>> {
    template <class RNGType, class StateSpace, class Timer>
    static double move(const Ising<int>& ham, const GpuRegular<dim>& grid, StateSpace& ss, RNGCache<RNGType>& rng,
                       double beta, int ncluster, int nmetro, Timer&)
    {
        if (!grid.batch) {  // first step: upload Core's statespace, Philox key from its RNG
            std::vector<const marqov_amd::RegularLattice*> l{&grid.dev};
            grid.batch.reset(new marqov_amd::Batch(grid.ctx, marqov_amd::Ising{ham.J}, l));
            grid.step.reset(new marqov_amd::GpuEMCS(*grid.batch, rng.integer()));
            std::vector<int8_t> h(grid.size());
            for (std::size_t i = 0; i < h.size(); ++i) h[i] = (int8_t)ss[i][0];
            grid.batch->upload(0, h);
        }
        if (beta != grid.beta_on_device) {  // temperature tables only when beta changes
            grid.batch->set_beta({beta});
            grid.beta_on_device = beta;
            ++grid.beta_uploads;
        }
        double mean = NAN;
        if (grid.sync_host) {
            mean = grid.step->move(ncluster, nmetro)[0];  // NaN when ncluster == 0, as emcs.h:108
            auto h = grid.batch->template download<int8_t>(0, grid.size(), 1);
            for (std::size_t i = 0; i < h.size(); ++i) ss[i][0] = h[i];
            ++grid.downloads;
            grid.host_current = true;
        } else {
            grid.step->advance(ncluster, nmetro);  // no copy, no synchronisation
            grid.host_current = false;
        }
        return mean;
    }
};
}  // namespace MARQOV

// ---- a Core-like loop over it
int main(int argc, char** argv)
{
    const int L = 32, nwarm = argc > 1 ? std::atoi(argv[1]) : 50, nmeas = argc > 2 ? std::atoi(argv[2]) : 20;
    const double beta = 0.4406868;
    try {
        marqov_amd::Context ctx(0);
        GpuRegular<2> grid(ctx, L);
        Ising<int> ham;
        RNGCache<std::mt19937_64> rng(7);
        std::vector<std::array<int, 1>> ss(grid.size());
        std::mt19937_64 init(11);
        for (auto& v : ss) v[0] = (init() >> 63) ? 1 : -1;
        const std::vector<std::array<int, 1>> start = ss;
        struct NoTimer {} timer;
        double esum = 0;
        for (int t = 0; t < nwarm; ++t) MARQOV::EMCS<Ising<int>, GpuRegular<2>>::move(ham, grid, ss, rng, beta, 4, 1, timer);
        grid.sync_host = true;
        for (int t = 0; t < nmeas; ++t) {
            MARQOV::EMCS<Ising<int>, GpuRegular<2>>::move(ham, grid, ss, rng, beta, 4, 1, timer);
            long long bonds = 0;  // E/N = J sum_<ij> s_i s_j / N from the host statespace
            for (int y = 0; y < L; ++y)
                for (int x = 0; x < L; ++x)
                    bonds += (long long)ss[y * L + x][0] * (ss[y * L + (x + 1) % L][0] + ss[((y + 1) % L) * L + x][0]);
            esum += ham.J * (double)bonds / (double)(L * L);
        }
        // the same steps through the C++ API directly
        marqov_amd::RegularLattice lat(ctx, 2, L);
        std::vector<const marqov_amd::RegularLattice*> l{&lat};
        marqov_amd::Batch direct(ctx, marqov_amd::Ising{-1.0}, l);
        std::vector<int8_t> h(start.size());
        for (std::size_t i = 0; i < h.size(); ++i) h[i] = (int8_t)start[i][0];
        direct.upload(0, h);
        direct.set_beta({beta});
        RNGCache<std::mt19937_64> rng2(7);
        marqov_amd::GpuEMCS step(direct, rng2.integer());
        for (int t = 0; t < nwarm + nmeas; ++t) step.advance(4, 1);
        auto d = direct.download<int8_t>(0, start.size(), 1);
        bool equal = true;
        for (std::size_t i = 0; i < d.size(); ++i) equal = equal && d[i] == ss[i][0];
        std::printf("{\"equal\": %s, \"downloads\": %d, \"beta_uploads\": %d, \"e_mean\": %.6f}\n",
                    equal ? "true" : "false", grid.downloads, grid.beta_uploads, esum / nmeas);
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "error: %s\n", ex.what());
        return 1;
    }
    return 0;
}
