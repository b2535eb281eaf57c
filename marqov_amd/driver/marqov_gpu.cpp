// marqov_gpu -- runs a MARQOV .ini configuration on one MI355X.
//
//   marqov_gpu <Hamiltonian> [--config DIR] [--out DIR] [--seed S] [--device D] [--dry-run] [--fresh]
//              [--graph FILE.csv [--idxoffs K]]
//
// Reads <Hamiltonian>.ini from the configuration directory (default ./config,
// as check_registry_availability, src/libmarqov/util/startup.h:69-80) with the
// reference's registry rules and runs what the reference's drivers run for it
// (demo/{Ising,Heisenberg,Phi4,Potts,IsingRGG}.cpp with RegularLatticeLoop,
// src/libmarqov/util/regularlatticeloop.h:11-136):
//   * parameter sets are the Cartesian product of the model's lists, the
//     first list varying fastest (cart_prod, util/cartprod.h); Ising,
//     Heisenberg, Potts<q>, EdwardsAnderson-Ising and IsingRGG take (beta, J),
//     Phi4 takes (beta, lambda, mass) with J = -1 (demo/Phi4.cpp:70-79);
//   * for each L of the list: ncluster = int(nclusteramp * L^nclusterexp), and
//     every parameter set runs `rep` replicas (one value, or one per L) with
//     repid 0..rep-1 (replicator, util/replicate.h:34-52);
//   * a simulation runs 10 * (warmupsteps/10) EMCS, then 10 * (measuresteps/10)
//     EMCS each followed by one measurement (Core::wrmploop / gameloop,
//     core.h:1001-1050); an EMCS is ncluster Wolff moves and nmetro sweeps
//     (emcs.h:62-109).
// "Potts" picks q from PottsSelect.ini [Which] q as demo/Potts.cpp does.
//
// All replicas of one L and one set of non-temperature parameters run as one
// batch on the GPU (one mq_state): each EMCS step is a handful of launches for
// the whole batch.  Periodic lattices with dimension 2 or 3 and even L use the
// checkerboard kernels; odd L and dimension 4 use the graph kernels on the
// same lattice as a CSR graph.  IsingRGG builds a periodic random geometric
// graph of L^dim points with mean degree 6 per repid (demo/IsingRGG.cpp).
// --graph runs the configuration on a graph read in GraphFromCSV's format
// (src/lattice/graph_from_csv.h:32-57, lattice/util/io.h:57-105) instead of the
// L list: one batch per parameter group on that graph, output under <out>/graph/,
// ncluster = int(nclusteramp * n^(nclusterexp/dim)) for n vertices.
// EdwardsAnderson-Ising draws +-1 couplings with probability 1/2 each
// (BimodalPDF, src/lattice/util/prob_density.h:25-35), one disorder sample per
// repid shared by all temperatures.
//
// Run files: every simulation has the HDF5 file <out>/<L>/<name>.h5 in MARQOV's
// layout (runfile.hpp; core.h:650-917): each run of the configuration adds a step
// /step<N> with the configuration, the observable time series of its game loop
// and the final state (spins and the generator's words).  When every simulation of
// a GPU batch has such a file, the next run continues from their last step, as
// MARQOV does (core.h:543-607, 650-700): it loads the spins, skips the warm-up
// (wrmploop runs only at step 0, core.h:1034-1050) and records its game loop as
// the next step.  Philox counters are keyed by the EMCS number, which the state
// stores, so a continued trajectory is exactly the one an uninterrupted run would
// have made.  --fresh starts again and overwrites the files.
//
// Observables in the run files are the reference tuples' (Ising.h:82-94: m, e,
// magft0, magft1 and the constant vector test series "dummy"; Heisenberg.h:88-90
// and Potts.h:143-144: m; Phi4.h:133-140: m and one magft<d> per lattice axis),
// MagFTComp on regular lattices only (graphs carry no coordinates here).  The
// Edwards-Anderson tuple (EdwardsAndersonIsing.h:313-317: running averages of
// two-site correlations over the whole run) is not computed on this path; its
// files hold e and m.
//
// Text output: one file per simulation, <out>/<L>/<name>.txt (the reference's
// outpath and defaultfilter name, util/filters.h:28-37), holding a commented
// header and one row per measurement: step, E/N, |m|, m^2, m^4, mean cluster
// size of that EMCS, and for Ising / Phi4 on regular lattices the MagFTComp
// columns (mq_measure_ft); a continued run appends its rows as the next block.
// <out>/<L>/summary.txt holds per-simulation means.
#include <sys/stat.h>

#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "lattices.hpp"
#include "marqov_amd.hpp"
#include "registry.hpp"
#include "runfile.hpp"

#ifndef MQ_DRIVER_VERSION
#define MQ_DRIVER_VERSION "unversioned"
#endif

using namespace marqov_amd;

namespace {

struct Options {
    std::string ham;
    std::string config = "./config";
    std::string out;
    uint64_t seed = 1;
    int device = 0;
    bool dry = false;
    bool fresh = false;
    std::string graph;  // GraphFromCSV file replacing the hypercubes
    int idxoffs = 0;
};

struct Spec {
    std::string block;  // configuration block and file stem, e.g. "Potts3"
    int model = MQ_ISING;
    int q = 2;
    bool rgg = false;
};

struct ParamSet {
    double beta;
    double J, lambda, mass;
};

struct Sim {
    ParamSet p;
    int repid;
    int id;
};

struct Group {  // one mq_state: sims sharing everything but beta and repid
    std::vector<Sim> sims;
};

uint64_t mix(uint64_t a, uint64_t b)
{
    uint64_t z = a ^ (b + 0x9E3779B97F4A7C15ull + (a << 6) + (a >> 2));
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void make_dir(const std::string& path)
{
    std::string acc;
    for (std::size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!acc.empty() && mkdir(acc.c_str(), 0755) != 0 && errno != EEXIST)
                throw std::runtime_error("cannot create directory " + acc);
        }
        if (i < path.size()) acc += path[i];
    }
}

// beta<beta>_<repid>, beta formatted by std::to_string as the reference's
// defaultfilter does (util/filters.h:28-37); with several groups of
// non-temperature parameters (the reference would overwrite one file with the
// next) the group index is appended.
std::string sim_name(double beta, int repid, int group, int ngroups)
{
    return "beta" + std::to_string(beta) + "_" + std::to_string(repid) +
           (ngroups > 1 ? "_g" + std::to_string(group) : "");
}

Spec spec_of(const std::string& ham, const Registry& reg)
{
    Spec s;
    s.block = ham;
    if (ham == "Ising" || ham == "Ising4d") {
        s.model = MQ_ISING;
    } else if (ham == "IsingRGG") {
        s.model = MQ_ISING;
        s.rgg = true;
    } else if (ham == "Heisenberg") {
        s.model = MQ_HEISENBERG;
    } else if (ham == "Phi4") {
        s.model = MQ_PHI4;
    } else if (ham == "EdwardsAnderson-Ising") {
        s.model = MQ_EA_ISING;
    } else if (ham == "Potts") {
        s.q = reg.get<int>("PottsSelect.ini", "Which", "q");
        s.block = "Potts" + std::to_string(s.q);
        s.model = MQ_POTTS;
    } else if (ham.rfind("Potts", 0) == 0) {
        s.model = MQ_POTTS;
        s.q = reg.get<int>(ham + ".ini", ham, "q");
    } else {
        throw std::runtime_error("no GPU update path for Hamiltonian '" + ham +
                                 "' (Ising, Ising4d, IsingRGG, Heisenberg, Phi4, Potts, Potts<q>, EdwardsAnderson-Ising)");
    }
    if (s.model == MQ_POTTS && (s.q < 2 || s.q > 8)) throw std::runtime_error("Potts q must be in 2..8");
    return s;
}

// cart_prod: the first list varies fastest
std::vector<ParamSet> parameter_sets(const Spec& s, const Registry& reg)
{
    const std::string f = s.block + ".ini";
    const auto beta = reg.get<std::vector<double>>(f, s.block, "beta");
    std::vector<ParamSet> out;
    if (s.model == MQ_PHI4) {
        const auto lambda = reg.get<std::vector<double>>(f, s.block, "lambda");
        const auto mass = reg.get<std::vector<double>>(f, s.block, "mass");
        for (double m : mass)
            for (double l : lambda)
                for (double b : beta) out.push_back({b, -1.0, l, m});
    } else {
        const auto J = reg.get<std::vector<double>>(f, s.block, "J");
        for (double j : J)
            for (double b : beta) out.push_back({b, j, 0.0, 0.0});
    }
    return out;
}

mq_model_params model_params(const Spec& s, const ParamSet& p)
{
    mq_model_params m{};
    m.model = s.model;
    m.J = p.J;
    m.q = s.q;
    m.lambda = p.lambda;
    m.mass = p.mass;
    m.proposal = MQ_PROPOSE_FRESH;      // SVInitializer<double>: fresh uniform direction
    m.delta = 0.5;
    m.embedding = MQ_EMBED_CARTESIAN;   // Wolff<Heisenberg>: random Cartesian component
    m.spin_bits = 8;
    return m;
}

struct Plan {
    int L, dim, nrep, ncluster, nmetro, nwarm, nmeas;
    int64_t nsites;
    std::string lattice;  // "regular", "csr-hypercube", "rgg", "csv"
    std::string source;   // graph file of a "csv" plan
    std::vector<Group> groups;
    std::string outpath;
};

void print_usage()
{
    std::fprintf(stderr,
                 "usage: marqov_gpu <Hamiltonian> [--config DIR] [--out DIR] [--seed S] [--device D] [--dry-run] [--fresh]\n"
                 "                  [--graph FILE.csv [--idxoffs K]]\n"
                 "  Hamiltonian: Ising, Ising4d, IsingRGG, Heisenberg, Phi4, Potts (q from PottsSelect.ini),\n"
                 "               Potts<q>, EdwardsAnderson-Ising\n");
}

// MagFTComp columns of the reference's observable tuples: Ising axes 0 and 1
// (Ising.h:82-94), Phi4 every axis of the lattice (Phi4.h:116-140); regular
// lattices only.  The reference's Phi4 tuple always holds three (axes 0, 1, 2);
// on a 2D lattice its axis-2 component reads crds(i)[2] past the end of a
// two-coordinate site (undefined behaviour), so 2D Phi4 writes two columns here.
int ft_columns(const Plan& pl, const Spec& s)
{
    if (pl.lattice != "regular") return 0;
    return s.model == MQ_ISING ? 2 : (s.model == MQ_PHI4 ? pl.dim : 0);
}

void write_row(FILE* fp, unsigned long long step, const double* o, double cl, const double* ft, int nft)
{
    std::fprintf(fp, "%llu %.17g %.17g %.17g %.17g %.17g", step, o[0], o[1], o[2], o[3], cl);
    for (int d = 0; d < nft; ++d) std::fprintf(fp, " %.17g", ft[d]);
    std::fprintf(fp, "\n");
}

void write_sim(const Plan& pl, const Spec& s, const Sim& sim, uint64_t seed, const std::vector<double>& obs,
               const std::vector<double>& cl, const std::vector<double>& ft, int nrep, int r, const std::string& path,
               int block, uint64_t step0)
{
    const int nft = ft_columns(pl, s);
    FILE* fp = std::fopen(path.c_str(), block > 0 ? "a" : "w");
    if (!fp) throw std::runtime_error("cannot write " + path);
    if (block > 0) {  // a resumed run appends its game loop as the next block
        std::fprintf(fp, "# block %d (resumed after %llu EMCS)\n", block, (unsigned long long)step0);
        for (int t = 0; t < pl.nmeas; ++t)
            write_row(fp, (unsigned long long)(step0 + t), &obs[((size_t)t * nrep + r) * MQ_NOBS], cl[(size_t)t * nrep + r],
                      &ft[((size_t)t * nrep + r) * 3], nft);
        std::fclose(fp);
        return;
    }
    std::fprintf(fp, "# marqov_gpu %s L=%d dim=%d lattice=%s sites=%lld%s%s\n", s.block.c_str(), pl.L, pl.dim,
                 pl.lattice.c_str(), (long long)pl.nsites, pl.source.empty() ? "" : " graph=", pl.source.c_str());
    std::fprintf(fp, "# beta=%.17g J=%.17g", sim.p.beta, sim.p.J);
    if (s.model == MQ_PHI4) std::fprintf(fp, " lambda=%.17g mass=%.17g", sim.p.lambda, sim.p.mass);
    if (s.model == MQ_POTTS) std::fprintf(fp, " q=%d", s.q);
    std::fprintf(fp, " repid=%d id=%d seed=%llu\n", sim.repid, sim.id, (unsigned long long)seed);
    std::fprintf(fp, "# ncluster=%d nmetro=%d warmup=%d measure=%d\n", pl.ncluster, pl.nmetro, pl.nwarm, pl.nmeas);
    std::fprintf(fp, "# step E M M2 M4 clustersize");
    for (int d = 0; d < nft; ++d) std::fprintf(fp, " magft%d", d);
    std::fprintf(fp, "\n");
    for (int t = 0; t < pl.nmeas; ++t)
        write_row(fp, (unsigned long long)(pl.nwarm + t), &obs[((size_t)t * nrep + r) * MQ_NOBS], cl[(size_t)t * nrep + r],
                  &ft[((size_t)t * nrep + r) * 3], nft);
    std::fclose(fp);
}

// ---- run files (runfile.hpp)
constexpr const char* RNG_NAME = "Philox4x32-10";

// the reference Hamiltonians' `name` members
std::string model_name(const Spec& s)
{
    switch (s.model) {
    case MQ_ISING: return "Ising";
    case MQ_HEISENBERG: return "Heisenberg";
    case MQ_PHI4: return "Phi4";
    case MQ_POTTS: return "Potts";
    default: return "EdwardsAndersonIsing";
    }
}

runfile::Meta run_meta(const Plan& pl, const Spec& s, const Sim& sim, uint64_t base_seed, int nwarm_run)
{
    runfile::Meta m;
    m.beta = sim.p.beta;
    m.id = sim.id;
    m.repid = sim.repid;
    m.seed = base_seed;
    m.nsteps = nwarm_run + pl.nmeas;  // the EMCS this run makes
    m.warmupsteps = nwarm_run;
    m.gameloopsteps = pl.nmeas;
    m.ncluster = pl.ncluster;
    m.nmetro = pl.nmetro;
    // the reference's lattice writers: RegularHypercubic for hypercubes (whichever
    // kernels run them), GraphFromCSV for --graph files, "unknown" otherwise
    m.lattice = (pl.lattice == "regular" || pl.lattice == "csr-hypercube") ? "RegularHypercubic"
                : (pl.lattice == "csv" ? "GraphFromCSV" : "unknown");
    m.dim = pl.dim;
    m.npoints = pl.nsites;
    m.model = model_name(s);
    if (s.model == MQ_PHI4)  // Phi4::paramname (Phi4.h:152-163)
        m.params = {{"beta", sim.p.beta}, {"lambda", sim.p.lambda}, {"mass", sim.p.mass}};
    else
        m.params = {{"param0", sim.p.J}};
    m.version = MQ_DRIVER_VERSION;
    return m;
}

// the observable series of simulation r of a batch over the game loop
std::vector<runfile::Series> run_series(const Plan& pl, const Spec& s, const std::vector<double>& obs,
                                        const std::vector<double>& ft, int nrep, int r)
{
    auto col = [&](int k) {
        std::vector<double> v(pl.nmeas);
        for (int t = 0; t < pl.nmeas; ++t) v[t] = obs[((size_t)t * nrep + r) * MQ_NOBS + k];
        return v;
    };
    std::vector<runfile::Series> out;
    out.push_back({"m", "magnetization", 1, col(1)});
    if (s.model == MQ_ISING || s.model == MQ_EA_ISING) out.push_back({"e", "", 1, col(0)});
    const int nft = ft_columns(pl, s);
    for (int d = 0; d < nft; ++d) {
        std::vector<double> v(pl.nmeas);
        for (int t = 0; t < pl.nmeas; ++t) v[t] = ft[((size_t)t * nrep + r) * 3 + d];
        out.push_back({"magft" + std::to_string(d), "Fourier Component of Magnetization", 1, v});
    }
    if (s.model == MQ_ISING) {  // IsingGenericVectorValuedObs (Ising.h:35-48): 42 + 0.1 k, k < 5
        runfile::Series d{"dummy", "testing vector-valued observables ...", 5, {}};
        for (int t = 0; t < pl.nmeas; ++t)
            for (int k = 0; k < 5; ++k) d.values.push_back(42 + 0.1 * k);
        out.push_back(d);
    }
    return out;
}

}  // namespace

int main(int argc, char** argv)
{
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](const char* what) -> std::string {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "marqov_gpu: %s needs a value\n", what);
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--config") o.config = next("--config");
        else if (a == "--out") o.out = next("--out");
        else if (a == "--seed") o.seed = std::strtoull(next("--seed").c_str(), nullptr, 10);
        else if (a == "--device") o.device = std::atoi(next("--device").c_str());
        else if (a == "--dry-run") o.dry = true;
        else if (a == "--fresh") o.fresh = true;
        else if (a == "--graph") o.graph = next("--graph");
        else if (a == "--idxoffs") o.idxoffs = std::atoi(next("--idxoffs").c_str());
        else if (a == "-h" || a == "--help") {
            print_usage();
            return 0;
        } else if (!a.empty() && a[0] == '-') {
            std::fprintf(stderr, "marqov_gpu: unknown option %s\n", a.c_str());
            print_usage();
            return 2;
        } else if (o.ham.empty()) o.ham = a;
        else {
            print_usage();
            return 2;
        }
    }
    if (o.ham.empty()) {
        print_usage();
        return 2;
    }
    try {
        Registry reg(o.config, ".ini");
        const Spec s = spec_of(o.ham, reg);
        const std::string f = s.block + ".ini";
        if (!reg.has_file(f)) throw RegistryError("no " + f + " in " + o.config);
        const auto nL = reg.get<std::vector<int>>(f, s.block, "L");
        const int dim = reg.get<int>(f, s.block, "dim");
        auto nrep = reg.get<std::vector<int>>(f, s.block, "rep");
        if (nrep.size() == 1) nrep.resize(nL.size(), nrep[0]);
        if (nrep.size() != nL.size()) throw RegistryError("rep needs one value or one per L in " + f);
        const double amp = reg.get<double>(f, "MC", "nclusteramp");
        const int cexp = reg.get<int>(f, "MC", "nclusterexp");
        const int nmetro = reg.get<int>(f, "MC", "nmetro");
        const int warm = reg.get<int>(f, "MC", "warmupsteps");
        const int meas = reg.get<int>(f, "MC", "measuresteps");
        const std::string outdir = o.out.empty() ? reg.get<std::string>(f, "IO", "outdir") : o.out;
        if (dim < 1) throw RegistryError("dim must be >= 1 in " + f);
        for (int r : nrep)
            if (r < 1) throw RegistryError("rep must be >= 1 in " + f);
        if (nmetro < 0 || warm < 0 || meas < 0 || amp < 0.0)
            throw RegistryError("nmetro, warmupsteps, measuresteps and nclusteramp must be >= 0 in " + f);
        const auto psets = parameter_sets(s, reg);

        // group parameter sets by everything but beta (one mq_state each)
        std::vector<std::vector<int>> by_group;
        {
            std::map<std::vector<double>, int> gid;
            for (int k = 0; k < (int)psets.size(); ++k) {
                const std::vector<double> key = {psets[k].J, psets[k].lambda, psets[k].mass};
                auto it = gid.find(key);
                if (it == gid.end()) {
                    it = gid.emplace(key, (int)by_group.size()).first;
                    by_group.emplace_back();
                }
                by_group[it->second].push_back(k);
            }
        }

        Csr csv;
        if (!o.graph.empty()) {
            if (s.rgg) throw std::runtime_error("--graph replaces the random geometric graphs of IsingRGG; use Ising");
            csv = read_csv_graph(o.graph, o.idxoffs);
        }
        std::vector<Plan> plans;
        int id = 0;
        const std::size_t nplans = o.graph.empty() ? nL.size() : 1;
        for (std::size_t li = 0; li < nplans; ++li) {
            Plan pl;
            pl.dim = dim;
            pl.nrep = nrep[li];
            pl.nmetro = nmetro;
            pl.nwarm = 10 * (warm / 10);
            pl.nmeas = 10 * (meas / 10);
            if (o.graph.empty()) {
                pl.L = nL[li];
                pl.nsites = (int64_t)std::llround(std::pow((double)pl.L, (double)dim));
                pl.ncluster = (int)(amp * std::pow((double)pl.L, (double)cexp));
                pl.lattice = s.rgg ? "rgg" : ((dim == 2 || dim == 3) && pl.L % 2 == 0 && pl.L >= 2) ? "regular" : "csr-hypercube";
                pl.outpath = outdir + "/" + std::to_string(pl.L) + "/";
            } else {
                pl.L = 0;
                pl.nsites = csv.nsites();
                pl.ncluster = (int)(amp * std::pow((double)pl.nsites, (double)cexp / (double)std::max(dim, 1)));
                pl.lattice = "csv";
                pl.source = o.graph;
                pl.outpath = outdir + "/graph/";
            }
            // simulation ids in the reference's order: parameter sets, then replicas
            std::vector<int> gof(psets.size());
            for (std::size_t g = 0; g < by_group.size(); ++g)
                for (int k : by_group[g]) gof[k] = (int)g;
            pl.groups.assign(by_group.size(), Group{});
            for (std::size_t k = 0; k < psets.size(); ++k)
                for (int r = 0; r < nrep[li]; ++r) pl.groups[gof[k]].sims.push_back({psets[k], r, id++});
            plans.push_back(pl);
        }

        if (o.dry) {
            for (const auto& pl : plans) {
                std::printf("{\"L\": %d, \"dim\": %d, \"nsites\": %lld, \"lattice\": \"%s\", \"ncluster\": %d, "
                            "\"nmetro\": %d, \"warmup\": %d, \"measure\": %d, \"out\": \"%s\", \"groups\": [",
                            pl.L, pl.dim, (long long)pl.nsites, pl.lattice.c_str(), pl.ncluster, pl.nmetro, pl.nwarm,
                            pl.nmeas, pl.outpath.c_str());
                for (std::size_t g = 0; g < pl.groups.size(); ++g) {
                    std::printf("%s[", g ? ", " : "");
                    for (std::size_t k = 0; k < pl.groups[g].sims.size(); ++k) {
                        const Sim& x = pl.groups[g].sims[k];
                        std::printf("%s{\"id\": %d, \"beta\": %.17g, \"J\": %.17g, \"lambda\": %.17g, \"mass\": %.17g, "
                                    "\"repid\": %d, \"name\": \"%s\"}",
                                    k ? ", " : "", x.id, x.p.beta, x.p.J, x.p.lambda, x.p.mass, x.repid,
                                    sim_name(x.p.beta, x.repid, (int)g, (int)pl.groups.size()).c_str());
                    }
                    std::printf("]");
                }
                std::printf("], \"model\": \"%s\", \"q\": %d}\n", s.block.c_str(), s.q);
            }
            return 0;
        }

        Context ctx(o.device);
        make_dir(outdir);
        for (const auto& pl : plans) {
            make_dir(pl.outpath);
            std::printf("\nL = %d, %lld sites (%s, %d groups, ncluster %d, nmetro %d)\n", pl.L, (long long)pl.nsites,
                        pl.lattice.c_str(),
                        (int)pl.groups.size(), pl.ncluster, pl.nmetro);
            std::unique_ptr<RegularLattice> reglat;
            std::vector<std::unique_ptr<GraphLattice>> graphs;
            Csr hg;  // the plan's one graph (hypercube or file), kept for EA couplings
            if (pl.lattice == "regular") {
                reglat.reset(new RegularLattice(ctx, pl.dim, pl.L));
            } else if (pl.lattice == "csr-hypercube" || pl.lattice == "csv") {
                hg = pl.lattice == "csv" ? csv : hypercube_csr(pl.dim, pl.L);
                graphs.emplace_back(new GraphLattice(ctx, hg.row_ptr, hg.col_idx));
            } else {
                int64_t n = 1;
                for (int d = 0; d < pl.dim; ++d) n *= pl.L;
                const double r = rgg_radius(pl.dim, n, 6.0);
                for (int rep = 0; rep < pl.nrep; ++rep) {  // one graph per replica, built on the device
                    const CsrGraph g = random_geometric_graph(
                        ctx, pl.dim, n, r, mix(mix(o.seed, 0x5267u), (uint64_t)pl.L * 1000003u + rep));
                    graphs.emplace_back(new GraphLattice(ctx, g.row_ptr, g.col_idx));
                }
            }
            std::unique_ptr<FILE, int (*)(FILE*)> sumf(std::fopen((pl.outpath + "summary.txt").c_str(), "w"), &std::fclose);
            if (!sumf) throw std::runtime_error("cannot write " + pl.outpath + "summary.txt");
            FILE* sum = sumf.get();
            std::fprintf(sum, "# name beta J lambda mass repid <E> err(E) <|m|> <m^2> <m^4> <clustersize> block\n");
            for (std::size_t g = 0; g < pl.groups.size(); ++g) {
                // RGG: replicas of one graph must be adjacent, so order by repid, then beta
                std::vector<Sim> sims = pl.groups[g].sims;
                if (pl.lattice == "rgg")
                    std::stable_sort(sims.begin(), sims.end(), [](const Sim& a, const Sim& b) { return a.repid < b.repid; });
                const int nr = (int)sims.size();
                std::vector<const RegularLattice*> rl;
                std::vector<const GraphLattice*> gl;
                for (const Sim& x : sims) {
                    if (reglat) rl.push_back(reglat.get());
                    else gl.push_back(pl.lattice == "rgg" ? graphs[x.repid].get() : graphs[0].get());
                }
                const mq_model_params mp = model_params(s, sims[0].p);
                std::unique_ptr<Batch> b(reglat ? new Batch(ctx, mp, rl) : new Batch(ctx, mp, gl));
                std::vector<double> beta(nr);
                for (int r = 0; r < nr; ++r) beta[r] = sims[r].p.beta;
                b->set_beta(beta);
                const uint64_t seed = mix(mix(o.seed, (uint64_t)pl.L), g);
                const int per_site = (s.model == MQ_HEISENBERG || s.model == MQ_PHI4) ? 3 : 1;
                const std::size_t nsites = reglat ? reglat->size() : gl[0]->size();
                const std::size_t bytes = nsites * (per_site == 3 ? 3 * sizeof(float) : 1);
                std::vector<std::string> names(nr), h5paths(nr);
                int have = 0;
                for (int r = 0; r < nr; ++r) {
                    names[r] = sim_name(sims[r].p.beta, sims[r].repid, (int)g, (int)pl.groups.size());
                    h5paths[r] = pl.outpath + names[r] + ".h5";
                    if (!o.fresh && h5store::is_hdf5(h5paths[r])) ++have;
                }
                if (have > 0 && have < nr)
                    throw std::runtime_error("only " + std::to_string(have) + " of the " + std::to_string(nr) +
                                             " simulations of group " + std::to_string(g) + " in " + pl.outpath +
                                             " have run files; remove them or run with --fresh");
                const bool resume = have == nr;
                uint64_t step0 = 0;
                int block = 0;
                if (resume) {  // continue from the last step of every simulation's run file
                    for (int r = 0; r < nr; ++r) {
                        const std::string& path = h5paths[r];
                        auto mismatch = [&](const std::string& what) {
                            throw std::runtime_error(path + " does not belong to this configuration (" + what +
                                                     "); remove it or run with --fresh");
                        };
                        runfile::Last last;
                        try {
                            last = runfile::read_last(path, nsites, per_site, per_site == 3);
                        } catch (const std::exception& ex) {
                            mismatch(std::string("unreadable last step: ") + ex.what());
                        }
                        const runfile::Meta want = run_meta(pl, s, sims[r], o.seed, 0);
                        if (last.state.rng != RNG_NAME || last.state.rngstate.size() != 3)
                            mismatch("state written by another generator");
                        if (last.state.rngstate[0] != o.seed)
                            mismatch("written with --seed " + std::to_string(last.state.rngstate[0]));
                        if (last.state.rngstate[1] != seed) mismatch("batch key differs");
                        if (last.meta.model != want.model || last.meta.lattice != want.lattice ||
                            last.meta.npoints != want.npoints)
                            mismatch("model or lattice differ");
                        if (last.meta.beta != sims[r].p.beta || last.meta.repid != sims[r].repid)
                            mismatch("temperature or replica differ");
                        for (const auto& pr : want.params) {
                            double v = 0.0;
                            try {
                                v = runfile::read_param(path, last.step, pr.first);
                            } catch (const std::exception& ex) {
                                mismatch(std::string("unreadable Hamiltonian parameter: ") + ex.what());
                            }
                            if (v != pr.second) mismatch("Hamiltonian parameter " + pr.first + " differs");
                        }
                        if (r == 0) {
                            block = last.step + 1;
                            step0 = last.state.rngstate[2];
                        } else if (last.step + 1 != block || last.state.rngstate[2] != step0) {
                            mismatch("its last step differs from the other simulations' of the batch");
                        }
                        std::vector<char> sp(bytes);
                        if (per_site == 3) {
                            float* f = reinterpret_cast<float*>(sp.data());
                            for (std::size_t k = 0; k < nsites * 3; ++k) {
                                const double v = last.state.reals[k];
                                if (!std::isfinite(v)) mismatch("spins are not finite");
                                f[k] = (float)v;
                            }
                        } else {
                            const int lo = s.model == MQ_POTTS ? 0 : -1, hi = s.model == MQ_POTTS ? s.q - 1 : 1;
                            for (std::size_t k = 0; k < nsites; ++k) {
                                const int32_t v = last.state.ints[k];
                                if (v < lo || v > hi || (s.model != MQ_POTTS && v == 0)) mismatch("spin values out of range");
                                sp[k] = (char)(int8_t)v;
                            }
                        }
                        b->upload(r, sp);
                    }
                } else {
                    b->init(seed);
                }
                if (s.model == MQ_EA_ISING) {
                    std::vector<int32_t> sample(nr);
                    for (int r = 0; r < nr; ++r) sample[r] = sims[r].repid;
                    const uint64_t bseed = mix(o.seed, 0xEA00u + (uint64_t)pl.L);
                    if (reglat) {
                        b->random_bonds(sample, 0.5, bseed);
                    } else {
                        // graphs: +-1 per undirected edge from a hash of (sample, edge), so
                        // both CSR slots of an edge carry the same coupling
                        ctx.check(mq_state_set_samples(b->get(), sample.data()), "mq_state_set_samples");
                        const Csr& g = hg;
                        const uint64_t n = (uint64_t)g.nsites();
                        std::vector<int8_t> bonds(g.col_idx.size());
                        for (int sm = 0; sm < pl.nrep; ++sm) {
                            const uint64_t ks = mix(bseed, (uint64_t)sm);
                            for (uint64_t i = 0; i < n; ++i)
                                for (int32_t k = g.row_ptr[i]; k < g.row_ptr[i + 1]; ++k) {
                                    const uint64_t j = (uint64_t)g.col_idx[k];
                                    const uint64_t e = std::min(i, j) * n + std::max(i, j);
                                    bonds[k] = (mix(ks, e) >> 63) ? (int8_t)1 : (int8_t)-1;
                                }
                            ctx.check(mq_state_set_bonds(b->get(), sm, bonds.data()), "mq_state_set_bonds");
                        }
                    }
                }
                GpuEMCS emcs(*b, seed, step0);
                const auto t0 = std::chrono::steady_clock::now();
                if (!resume) {  // the warm-up runs only for a fresh start (wrmploop, core.h:1034-1050)
                    std::printf("|");
                    for (int k = 0; k < 10; ++k) {
                        emcs.advance(pl.ncluster, pl.nmetro, pl.nwarm / 10);
                        std::printf(".");
                        std::fflush(stdout);
                    }
                    std::printf("|");
                } else {
                    std::printf("resuming block %d after %llu EMCS |", block, (unsigned long long)step0);
                }
                const uint64_t meas0 = emcs.steps();
                std::vector<double> obs((size_t)pl.nmeas * nr * MQ_NOBS), cl((size_t)pl.nmeas * nr);
                std::vector<double> ft((size_t)pl.nmeas * nr * 3, 0.0);
                const bool want_ft = ft_columns(pl, s) > 0;
                for (int k = 0, t = 0; k < 10; ++k) {
                    for (int i = 0; i < pl.nmeas / 10; ++i, ++t) {
                        const std::vector<double> mc = emcs.move(pl.ncluster, pl.nmetro);
                        ctx.check(mq_measure(b->get(), &obs[(size_t)t * nr * MQ_NOBS]), "mq_measure");
                        if (want_ft) ctx.check(mq_measure_ft(b->get(), &ft[(size_t)t * nr * 3]), "mq_measure_ft");
                        for (int r = 0; r < nr; ++r) cl[(size_t)t * nr + r] = mc[r];
                    }
                    std::printf(".");
                    std::fflush(stdout);
                }
                const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                std::printf("| group %d: %d simulations, %llu EMCS each, %.2f s\n", (int)g, nr,
                            (unsigned long long)(emcs.steps() - step0), secs);
                for (int r = 0; r < nr; ++r) {  // the step of each simulation's run file
                    std::vector<char> sp(bytes);
                    ctx.check(mq_state_download(b->get(), r, sp.data(), (int64_t)bytes), "mq_state_download");
                    runfile::State st;
                    st.rng = RNG_NAME;
                    st.rngstate = {o.seed, seed, emcs.steps()};
                    st.per_site = per_site;
                    st.vector = per_site == 3;
                    if (st.vector) {
                        const float* f = reinterpret_cast<const float*>(sp.data());
                        st.reals.assign(f, f + nsites * 3);
                    } else {
                        st.ints.resize(nsites);
                        for (std::size_t k = 0; k < nsites; ++k) st.ints[k] = (int8_t)sp[k];
                    }
                    runfile::write_step(h5paths[r], block, run_meta(pl, s, sims[r], o.seed, resume ? 0 : pl.nwarm),
                                        run_series(pl, s, obs, ft, nr, r), st);
                }
                for (int r = 0; r < nr; ++r) {
                    const Sim& x = sims[r];
                    const std::string& name = names[r];
                    write_sim(pl, s, x, seed, obs, cl, ft, nr, r, pl.outpath + name + ".txt", block, meas0);
                    double se = 0, se2 = 0, sm = 0, sm2 = 0, sm4 = 0, sc = 0;
                    for (int t = 0; t < pl.nmeas; ++t) {
                        const double* v = &obs[((size_t)t * nr + r) * MQ_NOBS];
                        se += v[0];
                        se2 += v[0] * v[0];
                        sm += v[1];
                        sm2 += v[2];
                        sm4 += v[3];
                        sc += cl[(size_t)t * nr + r];
                    }
                    const double n = pl.nmeas > 0 ? pl.nmeas : 1;
                    const double me = se / n;
                    const double err = pl.nmeas > 1 ? std::sqrt(std::max(0.0, se2 / n - me * me) / (n - 1)) : 0.0;
                    std::fprintf(sum, "%s %.10g %.10g %.10g %.10g %d %.10g %.3g %.10g %.10g %.10g %.6g %d\n", name.c_str(),
                                 x.p.beta, x.p.J, x.p.lambda, x.p.mass, x.repid, me, err, sm / n, sm2 / n, sm4 / n,
                                 sc / n, block);
                }
            }
        }
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "marqov_gpu: %s\n", e.what());
        return 1;
    }
}
