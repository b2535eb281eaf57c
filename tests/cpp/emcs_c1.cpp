// C++ driver over include/marqov_amd.hpp: the EMCS schedule of 2D Ising L=64 (2D Ising,
// L=64, J=-1, beta_c, 10 Wolff clusters + 1 checkerboard sweep per EMCS) on a
// batch of replicas, as Core::wrmploop/gameloop run it (core.h:1001-1050).
// Prints one JSON line with the mean energy per site, its standard error over
// replicas, the Binder cumulant and the mean cluster size.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "marqov_amd.hpp"

int main(int argc, char** argv)
{
    const int nrep = argc > 1 ? std::atoi(argv[1]) : 128;
    const int nwarm = argc > 2 ? std::atoi(argv[2]) : 200;
    const int nmeas = argc > 3 ? std::atoi(argv[3]) : 400;
    try {
        marqov_amd::Context ctx(0);
        marqov_amd::RegularLattice lat(ctx, 2, 64);
        std::vector<const marqov_amd::RegularLattice*> lats(nrep, &lat);
        marqov_amd::Batch batch(ctx, marqov_amd::Ising{-1.0}, lats);
        batch.set_beta(std::vector<double>(nrep, 0.4406868));
        batch.init(1);
        marqov_amd::GpuEMCS emcs(batch, 1);
        for (int t = 0; t < nwarm; ++t) emcs.move(10, 1);
        std::vector<double> esum(nrep, 0.0), m2(nrep, 0.0), m4(nrep, 0.0);
        double cl = 0;
        for (int t = 0; t < nmeas; ++t) {
            auto mc = emcs.move(10, 1);
            auto o = batch.measure();
            for (int r = 0; r < nrep; ++r) {
                esum[r] += o[r].e;
                m2[r] += o[r].m2;
                m4[r] += o[r].m4;
                cl += mc[r];
            }
        }
        double mean = 0, var = 0, M2 = 0, M4 = 0;
        for (int r = 0; r < nrep; ++r) mean += esum[r] / nmeas / nrep;
        for (int r = 0; r < nrep; ++r) {
            const double d = esum[r] / nmeas - mean;
            var += d * d / (nrep - 1);
            M2 += m2[r] / nmeas / nrep;
            M4 += m4[r] / nmeas / nrep;
        }
        std::printf("{\"e\": %.10f, \"se\": %.3e, \"binder\": %.6f, \"mean_cluster\": %.2f}\n", mean,
                    std::sqrt(var / nrep), 1.0 - M4 / (3 * M2 * M2), cl / ((double)nmeas * nrep));
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "error: %s\n", ex.what());
        return 1;
    }
    return 0;
}
