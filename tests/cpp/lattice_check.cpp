// Prints the graphs built by marqov_amd/driver/lattices.hpp for the tests:
//   lattice_check cube <dim> <L>           -> "n nnz", row_ptr, col_idx
//   lattice_check csv <path> [idxoffs]      -> "n nnz", row_ptr, col_idx
#include <cstdio>
#include <cstdlib>
#include <cstring>

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
    if (argc >= 3 && !std::strcmp(argv[1], "csv")) {
        dump(read_csv_graph(argv[2], argc >= 4 ? std::atoi(argv[3]) : 0));
        return 0;
    }
    std::fprintf(stderr, "usage: lattice_check cube <dim> <L> | csv <path> [idxoffs]\n");
    return 2;
}
