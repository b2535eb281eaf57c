// runfile.hpp -- one simulation's HDF5 run file in MARQOV's layout: a new
// /step<N> per run (config, environment, observable series, final state), and
// the last step's state read back to continue a run.
//
// Mirrors the reference's Core: setupHDF5Container / createstep (core.h:650-768),
// Config::dumpparamstoH5 (core.h:180-191), dumphamparamstoH5 (core.h:696-711),
// writelat (regular_hypercubic.h:151-156, graph_from_csv.h:65-69, core.h:63-68),
// dumpEnvironmenttoHDF5Group (marqov.cpp:30-52), dumprng / dumpstatespace
// (core.h:841-896) and the state read of setupstatespace (core.h:543-606).
#pragma once

#include <unistd.h>

#include <climits>
#include <ctime>
#include <string>
#include <utility>
#include <vector>

#include "h5store.hpp"

namespace runfile {

// Everything a step records about the run besides observables and state.
struct Meta {
    // config/marqovconfig
    double beta = 0.0;
    int32_t id = 0, repid = 0;
    uint64_t seed = 0;
    int32_t nsteps = 0, warmupsteps = 0, gameloopsteps = 0, ncluster = 0, nmetro = 0;
    // config/lattice: RegularHypercubic writes name, dim, npoints (int); GraphFromCSV
    // name, npoints (int); a lattice without its own writer name "unknown" and
    // npoints (size_t)
    std::string lattice = "RegularHypercubic";
    int32_t dim = 0;
    int64_t npoints = 0;
    // config/hamiltonian: Model and the constructor arguments in order
    std::string model;
    std::vector<std::pair<std::string, double>> params;
    // environment
    std::string code = "marqov_amd", version;
};

// One observable's time series: n records of `width` values (width > 1: the
// reference's vector-valued observables, stored as fixed-size arrays).
struct Series {
    std::string name;
    std::string desc;  // dataset comment; empty for none
    std::size_t width = 1;
    std::vector<double> values;
};

// The state that continues a run: the generator's words and the spins in the
// lattice's site order, per_site values per site (int32 for Ising/Potts/EA spins,
// double for O(3)/Phi4 vectors).
struct State {
    std::string rng;
    std::vector<uint64_t> rngstate;
    int per_site = 1;
    bool vector = false;  // double components (else int32)
    std::vector<int32_t> ints;
    std::vector<double> reals;
};

inline std::string step_name(int s) { return "step" + std::to_string(s); }

inline void write_lattice(hid_t g, const Meta& m)
{
    h5store::string(g, "name", m.lattice);
    if (m.lattice == "RegularHypercubic") {
        h5store::scalar<int32_t>(g, "dim", m.dim);
        h5store::scalar<int32_t>(g, "npoints", (int32_t)m.npoints);
    } else if (m.lattice == "GraphFromCSV") {
        h5store::scalar<int32_t>(g, "npoints", (int32_t)m.npoints);
    } else {
        h5store::scalar<uint64_t>(g, "npoints", (uint64_t)m.npoints);
    }
}

inline void write_environment(hid_t g, const Meta& m)
{
    h5store::string(g, "Code", m.code);
    h5store::string(g, "Version", m.version);
    char host[HOST_NAME_MAX + 1] = {};
    if (gethostname(host, sizeof(host) - 1) == 0) h5store::string(g, "hostname", host);
    const std::time_t now = std::time(nullptr);
    std::tm tmv{};
    localtime_r(&now, &tmv);
    char stamp[64];
    std::strftime(stamp, sizeof(stamp), "%A %Y-%m-%d %H:%M:%S", &tmv);
    h5store::string(g, "startingdate", stamp);
}

// Writes /step<step> of the run file: step 0 creates (truncates) the file, a later
// step is added to it.
inline void write_step(const std::string& path, int step, const Meta& m, const std::vector<Series>& obs,
                       const State& st)
{
    using namespace h5store;
    Id file = step == 0 ? create_file(path) : open_file(path, true);
    if (step == 0)
        check(H5Oset_comment(file.get(), "A calculation is a series of steps; each one can continue from the "
                                         "state the previous one ended in."),
              "comment on the file");
    Id sg = group(file.get(), step_name(step), "One run: its configuration, its observable time series and its "
                                               "final state.");
    {
        Id env = group(sg.get(), "environment", "The host system and code of the run.");
        write_environment(env.get(), m);
    }
    {
        Id cfg = group(sg.get(), "config", "What is needed to start the run.");
        Id mc = group(cfg.get(), "marqovconfig", "The method's parameters.");
        scalar<double>(mc.get(), "beta", m.beta);
        scalar<int32_t>(mc.get(), "id", m.id);
        scalar<int32_t>(mc.get(), "repid", m.repid);
        scalar<uint64_t>(mc.get(), "seed", m.seed);
        scalar<int32_t>(mc.get(), "nsteps", m.nsteps);
        scalar<int32_t>(mc.get(), "warmupsteps", m.warmupsteps);
        scalar<int32_t>(mc.get(), "gameloopsteps", m.gameloopsteps);
        scalar<int32_t>(mc.get(), "ncluster", m.ncluster);
        scalar<int32_t>(mc.get(), "nmetro", m.nmetro);
        Id lat = group(cfg.get(), "lattice", "The lattice's parameters.");
        write_lattice(lat.get(), m);
        Id ham = group(cfg.get(), "hamiltonian", "The Hamiltonian's parameters.");
        string(ham.get(), "Model", m.model);
        for (const auto& p : m.params) scalar<double>(ham.get(), p.first, p.second);
    }
    {
        Id sgrp = group(sg.get(), "state", "The final state: the spins and the random number generator.");
        string(sgrp.get(), "RNG", st.rng);
        words(sgrp.get(), "rngstate", st.rngstate);
        const std::size_t n = st.vector ? st.reals.size() / st.per_site : st.ints.size() / st.per_site;
        if (st.vector)
            series<double>(sgrp.get(), "hamiltonianstatespace", st.reals.data(), n, st.per_site, true,
                           chunk_of(st.per_site * sizeof(double)), true, nullptr);
        else
            series<int32_t>(sgrp.get(), "hamiltonianstatespace", st.ints.data(), n, st.per_site, true,
                            chunk_of(st.per_site * sizeof(int32_t)), true, nullptr);
    }
    {
        Id og = group(sg.get(), "observables", "The observables' time series, one record per measurement.");
        for (const Series& s : obs) {
            const std::size_t n = s.values.size() / s.width;
            // scalar series: 4 MiB chunks; vector series: 4 MiB / record + 1 (cachecontainer.h:403)
            const std::size_t chunk = s.width == 1 ? chunk_of(sizeof(double)) : vector_chunk(sizeof(double), s.width);
            series<double>(og.get(), s.name, s.values.data(), n, s.width, s.width > 1, chunk, false,
                           s.desc.empty() ? nullptr : s.desc.c_str());
        }
    }
}

// The last step of a run file: its number, what its config recorded and its state.
struct Last {
    int step = -1;
    Meta meta;
    State state;
};

// Reads the last step of `path` for a lattice of nsites sites with per_site values
// per site; throws if the file has no complete step.
inline Last read_last(const std::string& path, std::size_t nsites, int per_site, bool vector)
{
    using namespace h5store;
    Id file = open_file(path, false);
    Last out;
    out.step = last_step(file.get());
    if (out.step < 0) throw Error(path + " holds no step");
    Id sg = open_group(file.get(), step_name(out.step));
    {
        Id mc = open_group(sg.get(), "config/marqovconfig");
        out.meta.beta = read_scalar<double>(mc.get(), "beta");
        out.meta.id = read_scalar<int32_t>(mc.get(), "id");
        out.meta.repid = read_scalar<int32_t>(mc.get(), "repid");
        out.meta.seed = read_scalar<uint64_t>(mc.get(), "seed");
        out.meta.ncluster = read_scalar<int32_t>(mc.get(), "ncluster");
        out.meta.nmetro = read_scalar<int32_t>(mc.get(), "nmetro");
        Id lat = open_group(sg.get(), "config/lattice");
        out.meta.lattice = read_string(lat.get(), "name");
        out.meta.npoints = read_scalar<int64_t>(lat.get(), "npoints");
        Id ham = open_group(sg.get(), "config/hamiltonian");
        out.meta.model = read_string(ham.get(), "Model");
    }
    Id st = open_group(sg.get(), "state");
    out.state.rng = read_string(st.get(), "RNG");
    out.state.rngstate = read_words(st.get(), "rngstate");
    out.state.per_site = per_site;
    out.state.vector = vector;
    if (vector)
        out.state.reals = read_series<double>(st.get(), "hamiltonianstatespace", nsites, per_site);
    else
        out.state.ints = read_series<int32_t>(st.get(), "hamiltonianstatespace", nsites, per_site);
    return out;
}

// The Hamiltonian parameters of the last step, by name.
inline double read_param(const std::string& path, int step, const std::string& name)
{
    using namespace h5store;
    Id file = open_file(path, false);
    Id ham = open_group(file.get(), step_name(step) + "/config/hamiltonian");
    return read_scalar<double>(ham.get(), name);
}

}  // namespace runfile
