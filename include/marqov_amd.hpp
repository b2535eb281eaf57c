// marqov_amd.hpp -- C++ host API over the C ABI of marqov_amd.h.
//
// RAII owners for contexts, lattices and replica batches, errors as
// std::runtime_error (the reference's convention, src/libmarqov/core.h:338-339),
// and `GpuEMCS`, the per-batch counterpart of MARQOV's elementary Monte Carlo
// step EMCS<H,L>::move (src/libmarqov/emcs.h:62-109): ncluster Wolff moves,
// then nmetro sweeps, returning the mean cluster size (NaN for ncluster = 0).
// The model tags mirror the reference's Hamiltonians
// (src/hamiltonian/{Ising,EdwardsAndersonIsing,Potts,Heisenberg,Phi4}.h) with
// their parameters and sign convention (J = -1 is ferromagnetic).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "marqov_amd.h"

namespace marqov_amd {

class Error : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

class Context {
public:
    explicit Context(int device = 0)
    {
        if (mq_ctx_create(device, &h_) != MQ_OK) throw Error("marqov_amd: no usable HIP device");
    }
    ~Context() { mq_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    mq_ctx* get() const { return h_; }
    void check(int rc, const char* what) const
    {
        if (rc != MQ_OK) throw Error(std::string(what) + ": " + mq_last_error(h_));
    }
    void sync() const { check(mq_sync(h_), "mq_sync"); }

private:
    mq_ctx* h_ = nullptr;
};

// Periodic hypercube: FixedDimRegularHypercubic<dim> (src/lattice/regular_hypercubic.h:15-100)
class RegularLattice {
public:
    RegularLattice(Context& ctx, int dim, int L) : ctx_(ctx), dim_(dim), L_(L)
    {
        ctx.check(mq_lattice_regular(ctx.get(), dim, L, &h_), "mq_lattice_regular");
    }
    ~RegularLattice() { mq_lattice_destroy(h_); }
    RegularLattice(const RegularLattice&) = delete;
    RegularLattice& operator=(const RegularLattice&) = delete;
    mq_lattice* get() const { return h_; }
    std::size_t size() const { return (std::size_t)mq_lattice_nsites(h_); }
    int dim() const { return dim_; }
    int length() const { return L_; }

private:
    Context& ctx_;
    int dim_, L_;
    mq_lattice* h_ = nullptr;
};

// General graph from CSR adjacency: GraphFromCSV (src/lattice/graph_from_csv.h:32-57)
class GraphLattice {
public:
    GraphLattice(Context& ctx, const std::vector<int32_t>& row_ptr, const std::vector<int32_t>& col_idx)
    {
        ctx.check(mq_lattice_csr(ctx.get(), (int64_t)row_ptr.size() - 1, row_ptr.data(), col_idx.data(), &h_),
                  "mq_lattice_csr");
    }
    ~GraphLattice() { mq_lattice_destroy(h_); }
    GraphLattice(const GraphLattice&) = delete;
    GraphLattice& operator=(const GraphLattice&) = delete;
    mq_lattice* get() const { return h_; }
    std::size_t size() const { return (std::size_t)mq_lattice_nsites(h_); }

private:
    mq_lattice* h_ = nullptr;
};

// Random geometric graph built on the device (mq_rgg_build), as CSR for GraphLattice.
struct CsrGraph {
    std::vector<int32_t> row_ptr, col_idx;
};
inline CsrGraph random_geometric_graph(Context& ctx, int dim, int64_t n, double radius, uint64_t seed)
{
    CsrGraph g;
    g.row_ptr.resize((size_t)n + 1);
    int64_t nnz = 0;
    ctx.check(mq_rgg_build(ctx.get(), dim, n, radius, seed, &nnz, g.row_ptr.data(), nullptr, 0, nullptr),
              "mq_rgg_build");
    g.col_idx.resize((size_t)nnz);
    ctx.check(mq_rgg_build(ctx.get(), dim, n, radius, seed, &nnz, g.row_ptr.data(), g.col_idx.data(), nnz, nullptr),
              "mq_rgg_build");
    return g;
}

// Model tags with the reference's parameters.
struct Ising { double J = -1.0; int spin_bits = 8; };
struct EdwardsAndersonIsing { double J = -1.0; };
template <int Q> struct Potts { double J = -1.0; };
struct Heisenberg { double J = -1.0; int proposal = MQ_PROPOSE_FRESH; double delta = 0.5; int embedding = MQ_EMBED_CARTESIAN; };
struct Phi4 { double lambda = 4.5, mass = 1.0; };  // J = -1 fixed as Phi4.h:93

inline mq_model_params params_of(const Ising& m) { mq_model_params p{}; p.model = MQ_ISING; p.J = m.J; p.spin_bits = m.spin_bits; return p; }
inline mq_model_params params_of(const EdwardsAndersonIsing& m) { mq_model_params p{}; p.model = MQ_EA_ISING; p.J = m.J; return p; }
template <int Q> inline mq_model_params params_of(const Potts<Q>& m) { mq_model_params p{}; p.model = MQ_POTTS; p.J = m.J; p.q = Q; return p; }
inline mq_model_params params_of(const Heisenberg& m)
{
    mq_model_params p{};
    p.model = MQ_HEISENBERG;
    p.J = m.J;
    p.proposal = m.proposal;
    p.delta = m.delta;
    p.embedding = m.embedding;
    return p;
}
// A model chosen at run time (q, couplings and proposals read from a configuration).
inline mq_model_params params_of(const mq_model_params& p) { return p; }
inline mq_model_params params_of(const Phi4& m)
{
    mq_model_params p{};
    p.model = MQ_PHI4;
    p.J = -1.0;
    p.lambda = m.lambda;
    p.mass = m.mass;
    return p;
}

struct Observables {
    double e, m, m2, m4;  // E/N, |m|, m^2, m^4
};

// A batch of replicas of one model: the device counterpart of nrep Core
// statespaces (src/libmarqov/core.h:249-342).
class Batch {
public:
    template <class Model, class Lattice>
    Batch(Context& ctx, const Model& model, const std::vector<const Lattice*>& lattices) : ctx_(ctx)
    {
        std::vector<mq_lattice*> ls;
        for (auto* l : lattices) ls.push_back(l->get());
        nrep_ = (int)ls.size();
        const mq_model_params p = params_of(model);
        model_ = p.model;
        ctx.check(mq_state_create(ctx.get(), &p, nrep_, ls.data(), &h_), "mq_state_create");
    }
    ~Batch() { mq_state_destroy(h_); }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;
    mq_state* get() const { return h_; }
    int nrep() const { return nrep_; }
    int model() const { return model_; }
    void check(int rc, const char* what) const { ctx_.check(rc, what); }
    // FP64 renormalisation of Heisenberg spins to unit length (FP32 reflections
    // about a random direction drift |s| by rounding)
    void renormalize() { ctx_.check(mq_renormalize(h_), "mq_renormalize"); }

    void set_beta(const std::vector<double>& beta) { ctx_.check(mq_state_set_beta(h_, beta.data()), "mq_state_set_beta"); }
    void init(uint64_t seed) { ctx_.check(mq_state_init(h_, seed), "mq_state_init"); }
    void random_bonds(const std::vector<int32_t>& rep_sample, double p_plus, uint64_t seed)
    {
        ctx_.check(mq_state_random_bonds(h_, rep_sample.data(), p_plus, seed), "mq_state_random_bonds");
    }
    void metropolis(int nsweeps, uint64_t seed, uint64_t sweep0) { ctx_.check(mq_metropolis(h_, nsweeps, seed, sweep0), "mq_metropolis"); }
    std::vector<int64_t> wolff(int nclusters, uint64_t seed, uint64_t cluster0)
    {
        std::vector<int64_t> sizes((size_t)nrep_ * nclusters);
        ctx_.check(mq_wolff(h_, nclusters, seed, cluster0, sizes.data()), "mq_wolff");
        return sizes;
    }
    std::vector<Observables> measure()
    {
        std::vector<double> raw((size_t)nrep_ * MQ_NOBS);
        ctx_.check(mq_measure(h_, raw.data()), "mq_measure");
        std::vector<Observables> o(nrep_);
        for (int r = 0; r < nrep_; ++r) o[r] = {raw[4 * r], raw[4 * r + 1], raw[4 * r + 2], raw[4 * r + 3]};
        return o;
    }
    // MagFTComp per replica and axis (observables.h:104-170): [r*3 + d], NaN for d >= dim
    std::vector<double> measure_ft()
    {
        std::vector<double> ft((size_t)nrep_ * 3);
        ctx_.check(mq_measure_ft(h_, ft.data()), "mq_measure_ft");
        return ft;
    }
    std::vector<int32_t> pt_swap(int ntemp, int parity, uint64_t seed, uint64_t step)
    {
        std::vector<int32_t> acc(nrep_ / ntemp);
        ctx_.check(mq_pt_swap(h_, ntemp, parity, seed, step, acc.data()), "mq_pt_swap");
        return acc;
    }
    // one step of a PT plan: disjoint temperature-slot pairs, offered in every group
    std::vector<int32_t> pt_swap_pairs(int ntemp, const std::vector<std::pair<int32_t, int32_t>>& pairs, uint64_t seed,
                                       uint64_t step)
    {
        std::vector<int32_t> flat;
        for (const auto& p : pairs) {
            flat.push_back(p.first);
            flat.push_back(p.second);
        }
        std::vector<int32_t> acc(nrep_ / ntemp);
        ctx_.check(mq_pt_swap_pairs(h_, ntemp, (int)pairs.size(), flat.data(), seed, step, acc.data()),
                   "mq_pt_swap_pairs");
        return acc;
    }
    template <class T>
    void upload(int rep, const std::vector<T>& v)
    {
        ctx_.check(mq_state_upload(h_, rep, v.data(), (int64_t)(v.size() * sizeof(T))), "mq_state_upload");
    }
    template <class T>
    std::vector<T> download(int rep, std::size_t nsites, int per_site)
    {
        std::vector<T> v(nsites * per_site);
        ctx_.check(mq_state_download(h_, rep, v.data(), (int64_t)(v.size() * sizeof(T))), "mq_state_download");
        return v;
    }

private:
    Context& ctx_;
    mq_state* h_ = nullptr;
    int nrep_ = 0;
    int model_ = 0;
};

// One elementary Monte Carlo step of every replica (emcs.h:69-109).
// `step` numbers the EMCS so that every step draws fresh Philox counters; a
// run continued from a saved state passes the number of EMCS already done, and
// then continues the same trajectory.
// Heisenberg spins are renormalised to unit length in FP64 after every
// `renorm_every`-th EMCS (by EMCS number, so a resumed run renormalises at the
// same steps; 0 = never): Wolff reflections about a random direction are FP32
// and drift |s| by rounding, while the reference's doubles keep it at 1.
class GpuEMCS {
public:
    static constexpr int kRenormEvery = 1000;
    GpuEMCS(Batch& batch, uint64_t seed, uint64_t step0 = 0, int renorm_every = kRenormEvery)
        : b_(batch), seed_(seed), step_(step0), renorm_every_(renorm_every)
    {}
    // one step, returning each replica's mean cluster size (one device-to-host
    // copy and a stream synchronisation, as Core reports it per EMCS)
    std::vector<double> move(int ncluster, int nmetro)
    {
        std::vector<double> mean(b_.nrep());
        b_.check(mq_emcs(b_.get(), ncluster, nmetro, seed_, step_++, mean.data()), "mq_emcs");
        after_step();
        return mean;
    }
    // the same step without reading anything back (warm-up): no copy, no synchronisation
    void advance(int ncluster, int nmetro)
    {
        b_.check(mq_emcs(b_.get(), ncluster, nmetro, seed_, step_++, nullptr), "mq_emcs");
        after_step();
    }
    // nsteps such steps: persistent launches (mq_emcs_run, each replica at its own
    // pace) where they apply, split at the renormalisation cadence; else step by step
    void advance(int ncluster, int nmetro, int nsteps)
    {
        while (nsteps > 0) {
            int n = nsteps;
            const bool renorm = renorm_every_ > 0 && b_.model() == MQ_HEISENBERG;
            if (renorm) n = std::min<uint64_t>((uint64_t)n, (uint64_t)renorm_every_ - step_ % (uint64_t)renorm_every_);
            const int rc = mq_emcs_run(b_.get(), n, ncluster, nmetro, seed_, step_, nullptr);
            if (rc == MQ_ERR_UNSUPPORTED) {
                for (int k = 0; k < nsteps; ++k) advance(ncluster, nmetro);
                return;
            }
            b_.check(rc, "mq_emcs_run");
            step_ += (uint64_t)n;
            nsteps -= n;
            if (renorm && step_ % (uint64_t)renorm_every_ == 0) b_.renormalize();
        }
    }
    uint64_t steps() const { return step_; }

private:
    void after_step()
    {
        if (renorm_every_ > 0 && b_.model() == MQ_HEISENBERG && step_ % (uint64_t)renorm_every_ == 0) b_.renormalize();
    }
    Batch& b_;
    uint64_t seed_;
    uint64_t step_;
    int renorm_every_;
};

}  // namespace marqov_amd
