// lattices.hpp -- host-side graph builders for configurations the
// checkerboard path cannot take.
//
// * hypercube_csr: the periodic hypercube of any dimension and any L >= 3 as a
//   CSR graph, in the reference's site order (index sum_j x_j L^j, x0 fastest)
//   and neighbour order (+x0, -x0, +x1, -x1, ...), as RegularHypercubic builds
//   it (src/lattice/regular_hypercubic.h:102-143).  Odd L is not bipartite and
//   dimension 4 has no regular-lattice kernel, so these run on the graph path.
// * rgg_radius: the connection radius of a random geometric graph of mean
//   degree k; the graph itself is built on the device (mq_rgg_build,
//   random_geometric_graph in include/marqov_amd.hpp), which replaces the
//   reference's PCL kd-tree construction (src/lattice/random_geometric_graph.h:29-168).
// * read_csv_graph: GraphFromCSV's file format (src/lattice/graph_from_csv.h:32-57,
//   lattice/util/io.h:57-105): first line the number of coordinates per
//   vertex, second line skipped, then one line per vertex with that many
//   coordinates and the vertex's neighbours, tab-separated; neighbour indices
//   are shifted by idxoffs.  Every line up to the end of the file is a vertex;
//   an empty line is a vertex without neighbours.
#pragma once

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <sstream>
#include <string>
#include <stdexcept>
#include <vector>

namespace marqov_amd {

struct Csr {
    std::vector<int32_t> row_ptr, col_idx;
    int64_t nsites() const { return (int64_t)row_ptr.size() - 1; }
};

inline Csr hypercube_csr(int dim, int L)
{
    if (dim < 1 || L < 3) throw std::invalid_argument("hypercube_csr: dim >= 1 and L >= 3");
    int64_t N = 1;
    for (int d = 0; d < dim; ++d) N *= L;
    if (N * 2 * dim >= ((int64_t)1 << 31)) throw std::invalid_argument("hypercube_csr: lattice too large");
    Csr g;
    g.row_ptr.resize(N + 1);
    g.col_idx.resize(N * 2 * dim);
    for (int64_t i = 0; i < N; ++i) {
        g.row_ptr[i] = (int32_t)(i * 2 * dim);
        int64_t pw = 1;
        for (int d = 0; d < dim; ++d) {
            const int64_t x = (i / pw) % L;
            const int64_t up = (x == L - 1) ? i - (L - 1) * pw : i + pw;
            const int64_t dn = (x == 0) ? i + (L - 1) * pw : i - pw;
            g.col_idx[i * 2 * dim + 2 * d] = (int32_t)up;
            g.col_idx[i * 2 * dim + 2 * d + 1] = (int32_t)dn;
            pw *= L;
        }
    }
    g.row_ptr[N] = (int32_t)(N * 2 * dim);
    return g;
}

// Connection radius for mean degree k with n points in the unit square/cube:
// n * (volume of the r-ball) = k.  For dim 3 and n = L^3 this is the reference's
// cbrt(3/(4 pi) * k) / L (demo/IsingRGG.cpp).
inline double rgg_radius(int dim, int64_t n, double mean_degree)
{
    if (dim == 2) return std::sqrt(mean_degree / (M_PI * (double)n));
    return std::cbrt(3.0 / (4.0 * M_PI) * mean_degree / (double)n);
}

inline Csr read_csv_graph(const std::string& path, int idxoffs = 0)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot read graph file " + path);
    std::string row;
    if (!std::getline(in, row)) throw std::runtime_error("empty graph file " + path);
    int ncoords = 0;
    try {
        ncoords = std::stoi(row);
    } catch (const std::exception&) {
        throw std::runtime_error("graph file " + path + ": first line must be the number of coordinates");
    }
    std::getline(in, row);  // second header line
    // every further line is a vertex, an empty one a vertex without coordinates or
    // neighbours (import_geometry, lattice/util/io.h:99-123)
    std::vector<std::vector<int32_t>> nb;
    while (std::getline(in, row)) {
        if (!row.empty() && row.back() == '\r') row.pop_back();
        std::istringstream ss(row);
        std::string w;
        int k = 0;
        std::vector<int32_t> v;
        while (std::getline(ss, w, '\t')) {
            if (k++ < ncoords || w.empty()) continue;
            std::size_t used = 0;
            long long x = 0;
            try {
                x = std::stoll(w, &used);
            } catch (const std::exception&) {
                used = 0;
            }
            if (used != w.size() || x - idxoffs < INT32_MIN || x - idxoffs > INT32_MAX)
                throw std::runtime_error("graph file " + path + ": bad neighbour index '" + w + "' on vertex " +
                                         std::to_string(nb.size()));
            v.push_back((int32_t)(x - idxoffs));
        }
        nb.push_back(std::move(v));
    }
    const int64_t n = (int64_t)nb.size();
    Csr g;
    g.row_ptr.assign(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        for (int32_t j : nb[i])
            if (j < 0 || j >= n) throw std::runtime_error("graph file " + path + ": neighbour index out of range");
        g.row_ptr[i + 1] = g.row_ptr[i] + (int32_t)nb[i].size();
    }
    g.col_idx.reserve(g.row_ptr[n]);
    for (const auto& v : nb) g.col_idx.insert(g.col_idx.end(), v.begin(), v.end());
    return g;
}

}  // namespace marqov_amd
