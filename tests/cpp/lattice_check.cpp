// Prints the graphs built by marqov_amd/driver/lattices.hpp for the tests:
//   lattice_check cube <dim> <L>           -> "n nnz", row_ptr, col_idx
//   lattice_check rgg <dim> <n> <r> <seed> -> "n nnz", the points, row_ptr, col_idx
//   lattice_check csv <path> [idxoffs]      -> "n nnz", row_ptr, col_idx
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../../marqov_amd/driver/lattices.hpp"

using namespace marqov_amd;

static void dump(const Csr& g)
{
    std::printf("%lld %lld\n", (long long)g.nsites(), (long long)g.col_idx.size());
    for (auto v : g.row_ptr) std::printf("%d ", v);
    std::printf("\n");
    for (auto v : g.col_idx) std::printf("%d ", v);
    std::printf("\n");
}

int main(int argc, char** argv)
{
    if (argc >= 4 && !std::strcmp(argv[1], "cube")) {
        dump(hypercube_csr(std::atoi(argv[2]), std::atoi(argv[3])));
        return 0;
    }
    if (argc >= 6 && !std::strcmp(argv[1], "rgg")) {
        const int dim = std::atoi(argv[2]);
        const long long n = std::atoll(argv[3]);
        const double r = std::atof(argv[4]);
        const unsigned long long seed = std::strtoull(argv[5], nullptr, 10);
        const Csr g = rgg_periodic(dim, n, r, seed);
        std::printf("%lld %lld\n", n, (long long)g.col_idx.size());
        // the same points rgg_periodic drew
        std::mt19937_64 gen(seed);
        std::uniform_real_distribution<double> U(0.0, 1.0);
        for (long long i = 0; i < n * dim; ++i) std::printf("%.17g ", U(gen));
        std::printf("\n");
        for (auto v : g.row_ptr) std::printf("%d ", v);
        std::printf("\n");
        for (auto v : g.col_idx) std::printf("%d ", v);
        std::printf("\n");
        return 0;
    }
    if (argc >= 3 && !std::strcmp(argv[1], "csv")) {
        dump(read_csv_graph(argv[2], argc >= 4 ? std::atoi(argv[3]) : 0));
        return 0;
    }
    std::fprintf(stderr, "usage: lattice_check cube <dim> <L> | rgg <dim> <n> <r> <seed> | csv <path> [idxoffs]\n");
    return 2;
}
