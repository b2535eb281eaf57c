// Writes a two-step run file through runfile.hpp (the driver's HDF5 writer) from
// synthetic values, reads its last step back with runfile::read_last and prints
// what it read as one JSON line; tests/test_runfile.py checks the file itself
// with the HDF5 project's h5ls / h5dump.
//
//   runfile_check <file.h5> ising|heisenberg
#include <cstdio>
#include <cstring>
#include <string>

#include "../../marqov_amd/driver/runfile.hpp"

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: runfile_check <file.h5> ising|heisenberg\n");
        return 2;
    }
    const std::string path = argv[1], kind = argv[2];
    const bool vec = kind == "heisenberg";
    const int L = 4, N = L * L, nmeas = 6;
    try {
        for (int step = 0; step < 2; ++step) {
            runfile::Meta m;
            m.beta = 0.4406868;
            m.id = 5;
            m.repid = 2;
            m.seed = 0xFEEDFACECAFEull;
            m.nsteps = (step ? 0 : 20) + nmeas;
            m.warmupsteps = step ? 0 : 20;
            m.gameloopsteps = nmeas;
            m.ncluster = 3;
            m.nmetro = 1;
            m.dim = 2;
            m.npoints = N;
            m.model = vec ? "Heisenberg" : "Ising";
            m.params = {{"param0", -1.0}};
            m.version = "runfile_check";
            std::vector<runfile::Series> obs;
            runfile::Series ms{"m", "magnetization", 1, {}};
            for (int t = 0; t < nmeas; ++t) ms.values.push_back(0.125 * (t + 1) + step + 1.0 / 3.0);
            obs.push_back(ms);
            if (!vec) {
                runfile::Series es{"e", "", 1, {}};
                for (int t = 0; t < nmeas; ++t) es.values.push_back(-1.5 - 0.01 * t - step);
                obs.push_back(es);
                runfile::Series d{"dummy", "testing vector-valued observables ...", 5, {}};
                for (int t = 0; t < nmeas; ++t)
                    for (int k = 0; k < 5; ++k) d.values.push_back(42 + 0.1 * k);
                obs.push_back(d);
            }
            runfile::State st;
            st.rng = "Philox4x32-10";
            st.rngstate = {0xFEEDFACECAFEull, 0x0123456789ABCDEFull, (uint64_t)(20 + nmeas * (step + 1))};
            st.per_site = vec ? 3 : 1;
            st.vector = vec;
            for (int i = 0; i < N; ++i) {
                if (vec) {
                    st.reals.push_back(0.6);
                    st.reals.push_back(-0.8 + 1e-3 * i);
                    st.reals.push_back(step ? 1.0 / 7.0 : 0.0);
                } else {
                    st.ints.push_back(((i * 7 + step) % 3) ? 1 : -1);
                }
            }
            runfile::write_step(path, step, m, obs, st);
        }
        const runfile::Last l = runfile::read_last(path, N, vec ? 3 : 1, vec);
        std::printf("{\"step\": %d, \"rng\": \"%s\", \"rngstate\": [%llu, %llu, %llu], \"model\": \"%s\", "
                    "\"lattice\": \"%s\", \"npoints\": %lld, \"beta\": %.17g, \"repid\": %d, \"seed\": %llu, \"spins\": [",
                    l.step, l.state.rng.c_str(), (unsigned long long)l.state.rngstate[0],
                    (unsigned long long)l.state.rngstate[1], (unsigned long long)l.state.rngstate[2],
                    l.meta.model.c_str(), l.meta.lattice.c_str(), (long long)l.meta.npoints, l.meta.beta,
                    l.meta.repid, (unsigned long long)l.meta.seed);
        const std::size_t n = vec ? l.state.reals.size() : l.state.ints.size();
        for (std::size_t k = 0; k < n; ++k)
            std::printf(k ? ", %.17g" : "%.17g", vec ? l.state.reals[k] : (double)l.state.ints[k]);
        std::printf("]}\n");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "runfile_check: %s\n", e.what());
        return 1;
    }
    return 0;
}
