// Writes a two-step run file through runfile.hpp (the driver's HDF5 writer) from
// synthetic values, reads its last step back with runfile::read_last and prints
// what it read as one JSON line; tests/test_runfile.py checks the file itself
// with the HDF5 project's h5ls / h5dump.
//
//   runfile_check <file.h5> ising|heisenberg
//   runfile_check <file.h5> series      (the time series of the reference's cache-container
//                                        component tests, testsuite/components/1-3)
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
    if (kind == "series") {
        // 1-cachecontainer.cpp: 2,000,000 doubles and ints; 2-cachecontainer-long-sim.cpp:
        // 500,000 records of a 7-vector; 3-cachecontainer-long-obs.cpp: 5 records of a
        // 600,000-vector (longer than the 4 MiB cache)
        try {
            using namespace h5store;
            Id file = create_file(path);
            {
                Id g = group(file.get(), "obs1");
                const std::vector<double> v(2000000, 42.0);
                series<double>(g.get(), "testobs", v.data(), v.size(), 1, false, chunk_of(sizeof(double)), false,
                               "a test description");
            }
            {
                Id g = group(file.get(), "obs2");
                const std::vector<int32_t> v(2000000, 42);
                series<int32_t>(g.get(), "testint", v.data(), v.size(), 1, false, chunk_of(sizeof(int32_t)), false,
                                "a test int description");
            }
            {
                Id g = group(file.get(), "obs3");
                std::vector<double> v;
                for (int i = 0; i < 500000; ++i)
                    for (int k = 1; k <= 7; ++k) v.push_back((double)k);
                series<double>(g.get(), "name", v.data(), 500000, 7, true, vector_chunk(sizeof(double), 7), false,
                               "desc");
            }
            {
                Id g = group(file.get(), "obs4");
                const std::vector<double> v((std::size_t)600000 * 5, 2.0);
                series<double>(g.get(), "name", v.data(), 5, 600000, true, vector_chunk(sizeof(double), 600000),
                               false, "desc");
            }
        } catch (const std::exception& e) {
            std::fprintf(stderr, "runfile_check: %s\n", e.what());
            return 1;
        }
        std::printf("{}\n");
        return 0;
    }
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
