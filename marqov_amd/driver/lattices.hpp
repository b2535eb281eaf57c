// lattices.hpp -- host-side graph builders for configurations the
// checkerboard path cannot take.
//
// * hypercube_csr: the periodic hypercube of any dimension and any L >= 3 as a
//   CSR graph, in the reference's site order (index sum_j x_j L^j, x0 fastest)
//   and neighbour order (+x0, -x0, +x1, -x1, ...), as RegularHypercubic builds
//   it (src/lattice/regular_hypercubic.h:102-143).  Odd L is not bipartite and
//   dimension 4 has no regular-lattice kernel, so these run on the graph path.
// * rgg_periodic: a random geometric graph of n uniform points in the periodic
//   unit square or cube, vertices joined when their minimum-image distance is
//   below r; the reference builds the same graph in 3D with a PCL kd-tree and
//   image searches across faces, edges and corners
//   (src/lattice/random_geometric_graph.h:29-168).  Here a cell list with
//   cells of side >= r finds the same pairs; neighbours are sorted by index.
// * read_csv_graph: GraphFromCSV's file format (src/lattice/graph_from_csv.h:32-57,
//   lattice/util/io.h:57-105): first line the number of coordinates per
//   vertex, second line skipped, then one line per vertex with that many
//   coordinates and the vertex's neighbours, tab-separated; neighbour indices
//   are shifted by idxoffs.  Reading stops at the end of the file or at an
//   empty line.
#pragma once

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <random>
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

inline Csr rgg_periodic(int dim, int64_t n, double r, uint64_t seed)
{
    if (dim != 2 && dim != 3) throw std::invalid_argument("rgg_periodic: dim 2 or 3");
    if (n < 2 || r <= 0.0 || r >= 0.5) throw std::invalid_argument("rgg_periodic: n >= 2 and 0 < r < 1/2");
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<double> p((size_t)n * dim);
    for (auto& v : p) v = U(gen);
    const int m = std::max(1, (int)std::floor(1.0 / r));  // cells per axis, side 1/m >= r
    int64_t ncell = m;
    for (int d = 1; d < dim; ++d) ncell *= m;
    auto cell_of = [&](int64_t i) {
        int64_t c = 0, pw = 1;
        for (int d = 0; d < dim; ++d) {
            int x = (int)(p[i * dim + d] * m);
            if (x >= m) x = m - 1;
            c += x * pw;
            pw *= m;
        }
        return c;
    };
    std::vector<int64_t> start(ncell + 1, 0), order(n);
    for (int64_t i = 0; i < n; ++i) ++start[cell_of(i) + 1];
    for (int64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
    {
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < n; ++i) order[fill[cell_of(i)]++] = i;
    }
    const double r2 = r * r;
    std::vector<std::vector<int32_t>> nb(n);
    const int reach = (m >= 3) ? 1 : 0;  // with fewer than 3 cells per axis every cell is a neighbour
    std::vector<int64_t> cells;
    for (int64_t i = 0; i < n; ++i) {
        int xc[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) {
            xc[d] = (int)(p[i * dim + d] * m);
            if (xc[d] >= m) xc[d] = m - 1;
        }
        cells.clear();
        const int lo = reach ? -1 : 0, hi = reach ? 1 : m - 1;
        for (int a = lo; a <= hi; ++a)
            for (int b = lo; b <= hi; ++b)
                for (int c = (dim == 3 ? lo : 0); c <= (dim == 3 ? hi : 0); ++c) {
                    const int off[3] = {a, b, c};
                    int64_t cell = 0, pw = 1;
                    for (int d = 0; d < dim; ++d) {
                        const int x = reach ? ((xc[d] + off[d]) % m + m) % m : off[d];
                        cell += x * pw;
                        pw *= m;
                    }
                    cells.push_back(cell);
                }
        std::sort(cells.begin(), cells.end());
        cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
        for (int64_t cell : cells)
            for (int64_t k = start[cell]; k < start[cell + 1]; ++k) {
                const int64_t j = order[k];
                if (j == i) continue;
                double s = 0.0;
                for (int d = 0; d < dim; ++d) {
                    double dx = std::fabs(p[i * dim + d] - p[j * dim + d]);
                    if (dx > 0.5) dx = 1.0 - dx;
                    s += dx * dx;
                }
                if (s < r2) nb[i].push_back((int32_t)j);
            }
        std::sort(nb[i].begin(), nb[i].end());
    }
    Csr g;
    g.row_ptr.resize(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) g.row_ptr[i + 1] = g.row_ptr[i] + (int32_t)nb[i].size();
    g.col_idx.reserve(g.row_ptr[n]);
    for (int64_t i = 0; i < n; ++i) g.col_idx.insert(g.col_idx.end(), nb[i].begin(), nb[i].end());
    return g;
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
    std::vector<std::vector<int32_t>> nb;
    while (std::getline(in, row)) {
        if (!row.empty() && row.back() == '\r') row.pop_back();
        if (row.empty()) break;
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
