// Random geometric graphs built on the GPU: n points uniform in the periodic
// unit square (dim 2) or cube (dim 3), vertices i != j joined when their
// minimum-image distance is below r.  The reference builds the 3D graph on the
// host with a PCL kd-tree and image searches across faces, edges and corners
// (src/lattice/random_geometric_graph.h:29-168); here every step runs on the
// device:
//
//   1. point i's coordinates are 53-bit Philox uniforms, counter
//      (i, d >> 1, 0, ST_GRAPH), words 2(d&1) (high) and 2(d&1)+1 (low), so the
//      graph is a function of (dim, n, r, seed) alone;
//   2. points are binned into m^dim cells of side 1/m >= r (m = floor(1/r)),
//      counted with atomics, the counts scanned, the points scattered;
//   3. one thread per point counts, then lists, the points of its 3^dim
//      neighbour cells (all m^dim when m < 3) within distance r, computing
//      dx = |x_i - x_j|, dx = 1 - dx when dx > 1/2, s = sum dx*dx with
//      correctly rounded multiplies and adds in dimension order (no
//      contraction), and accepting s < r*r;
//   4. the degrees are scanned into row pointers and every row is sorted, so
//      the CSR is canonical whatever order the cells were filled in.
//
// The layout and rounding let a host restatement reproduce the edge set
// exactly.  The result is the CSR that mq_lattice_csr consumes.
#include <hipcub/hipcub.hpp>

#include "mq_internal.hpp"

namespace {

constexpr uint32_t ST_GRAPH = 8;

__device__ __forceinline__ double rgg_coord(uint32_t i, int d, Key2 key)
{
    const uint4 w = philox4x32_10(i, (uint32_t)(d >> 1), 0u, ST_GRAPH, key);
    const uint32_t hi = (d & 1) ? w.z : w.x;
    const uint32_t lo = (d & 1) ? w.w : w.y;
    return (double)((((uint64_t)hi << 32) | lo) >> 11) * 0x1.0p-53;
}

__device__ __forceinline__ int cell_coord(double x, int m)
{
    const int k = (int)(x * (double)m);
    return k < m ? k : m - 1;
}

template <int DIM>
__global__ void k_rgg_points(int64_t n, int m, Key2 key, double* __restrict__ xyz, int32_t* __restrict__ cell,
                             int32_t* __restrict__ count)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = 0, pw = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const double x = rgg_coord((uint32_t)i, d, key);
        xyz[i * DIM + d] = x;
        c += cell_coord(x, m) * pw;
        pw *= m;
    }
    cell[i] = c;
    atomicAdd(&count[c], 1);
}

// points in cell order: order[p] = vertex, sxyz[p] = its coordinates
template <int DIM>
__global__ void k_rgg_scatter(int64_t n, const double* __restrict__ xyz, const int32_t* __restrict__ cell,
                              const int32_t* __restrict__ start, int32_t* __restrict__ fill,
                              int32_t* __restrict__ order, double* __restrict__ sxyz)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cell[i];
    const int p = start[c] + atomicAdd(&fill[c], 1);
    order[p] = (int32_t)i;
#pragma unroll
    for (int d = 0; d < DIM; ++d) sxyz[(int64_t)p * DIM + d] = xyz[i * DIM + d];
}

// FILL = false: deg[i] = number of neighbours; FILL = true: list them at
// ci[rp[i]..] and sort the row
template <int DIM, bool FILL>
__global__ void k_rgg_nbrs(int64_t n, int m, double r2, const double* __restrict__ xyz,
                           const int32_t* __restrict__ start, const int32_t* __restrict__ order,
                           const double* __restrict__ sxyz, int32_t* __restrict__ deg, const int32_t* __restrict__ rp,
                           int32_t* __restrict__ ci)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double xi[DIM];
    int ki[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        xi[d] = xyz[i * DIM + d];
        ki[d] = cell_coord(xi[d], m);
    }
    const bool near = m >= 3;  // else every cell is a neighbour cell (each visited once)
    const int lo = near ? -1 : 0, hi = near ? 1 : m - 1;
    const int clo = DIM == 3 ? lo : 0, chi = DIM == 3 ? hi : 0;
    int32_t* row = FILL ? ci + rp[i] : nullptr;
    int cnt = 0;
    for (int a = lo; a <= hi; ++a)
        for (int b = lo; b <= hi; ++b)
            for (int c = clo; c <= chi; ++c) {
                const int off[3] = {a, b, c};
                int cc = 0, pw = 1;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    const int x = near ? (ki[d] + off[d] + m) % m : off[d];
                    cc += x * pw;
                    pw *= m;
                }
                for (int p = start[cc]; p < start[cc + 1]; ++p) {
                    const int32_t j = order[p];
                    if (j == (int32_t)i) continue;
                    double s = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        double dx = fabs(xi[d] - sxyz[(int64_t)p * DIM + d]);
                        if (dx > 0.5) dx = 1.0 - dx;
                        s = __dadd_rn(s, __dmul_rn(dx, dx));
                    }
                    if (s < r2) {
                        if (FILL) row[cnt] = j;
                        ++cnt;
                    }
                }
            }
    if (!FILL) {
        deg[i] = cnt;
        return;
    }
    for (int u = 1; u < cnt; ++u) {  // rows hold a few dozen entries
        const int32_t v = row[u];
        int w = u - 1;
        while (w >= 0 && row[w] > v) {
            row[w + 1] = row[w];
            --w;
        }
        row[w + 1] = v;
    }
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

struct DevBuf {
    void* p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int mq_rgg_build(mq_ctx* ctx, int dim, int64_t n, double radius, uint64_t seed, int64_t* nnz,
                            int32_t* row_ptr, int32_t* col_idx, int64_t col_cap, double* coords)
{
    if (!ctx) return MQ_ERR_ARG;
    if (!nnz) return mq_fail(ctx, MQ_ERR_ARG, "random geometric graph: nnz is NULL");
    if (dim != 2 && dim != 3) return mq_fail(ctx, MQ_ERR_ARG, "random geometric graph: dim must be 2 or 3");
    if (n < 2 || n > ((int64_t)1 << 30)) return mq_fail(ctx, MQ_ERR_ARG, "random geometric graph: 2 <= n <= 2^30");
    if (!(radius > 0.0 && radius < 0.5))
        return mq_fail(ctx, MQ_ERR_ARG, "random geometric graph: radius must lie in (0, 1/2)");
    // CSR slots are int32: refuse graphs whose expected edge count comes near 2^31
    const double ball = dim == 2 ? M_PI * radius * radius : 4.0 / 3.0 * M_PI * radius * radius * radius;
    if ((double)n * (double)(n - 1) * ball * 1.25 + 1e4 > 2147483647.0)
        return mq_fail(ctx, MQ_ERR_ARG, "random geometric graph: expected number of edge slots exceeds 2^31");
    const double md = std::max(1.0, std::floor(1.0 / radius));  // cells per axis, side 1/m >= radius
    if (std::pow(md, dim) >= 2147483647.0 - 1.0)
        return mq_fail(ctx, MQ_ERR_ARG, "random geometric graph: radius too small for the cell grid");
    const int m = (int)md;
    int64_t ncell = m;
    for (int d = 1; d < dim; ++d) ncell *= m;
    hipStream_t s = ctx->stream;
    const Key2 key{(uint32_t)seed, (uint32_t)(seed >> 32)};
    DevBuf xyz, cell, count, start, fill, order, sxyz, deg, rp, tmp;
    MQ_HIP(ctx, hipMalloc(&xyz.p, sizeof(double) * dim * n));
    MQ_HIP(ctx, hipMalloc(&cell.p, sizeof(int32_t) * n));
    MQ_HIP(ctx, hipMalloc(&count.p, sizeof(int32_t) * (ncell + 1)));
    MQ_HIP(ctx, hipMalloc(&start.p, sizeof(int32_t) * (ncell + 1)));
    MQ_HIP(ctx, hipMalloc(&fill.p, sizeof(int32_t) * ncell));
    MQ_HIP(ctx, hipMalloc(&order.p, sizeof(int32_t) * n));
    MQ_HIP(ctx, hipMalloc(&sxyz.p, sizeof(double) * dim * n));
    MQ_HIP(ctx, hipMalloc(&deg.p, sizeof(int32_t) * (n + 1)));
    MQ_HIP(ctx, hipMalloc(&rp.p, sizeof(int32_t) * (n + 1)));
    auto* d_xyz = static_cast<double*>(xyz.p);
    auto* d_cell = static_cast<int32_t*>(cell.p);
    auto* d_count = static_cast<int32_t*>(count.p);
    auto* d_start = static_cast<int32_t*>(start.p);
    auto* d_fill = static_cast<int32_t*>(fill.p);
    auto* d_order = static_cast<int32_t*>(order.p);
    auto* d_sxyz = static_cast<double*>(sxyz.p);
    auto* d_deg = static_cast<int32_t*>(deg.p);
    auto* d_rp = static_cast<int32_t*>(rp.p);

    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(ncell + 1)), dim3(256), 0, s, d_count, ncell + 1, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(ncell)), dim3(256), 0, s, d_fill, ncell, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n + 1)), dim3(256), 0, s, d_deg, n + 1, 0);
    if (dim == 2) hipLaunchKernelGGL(k_rgg_points<2>, dim3(blocks_for(n)), dim3(256), 0, s, n, m, key, d_xyz, d_cell, d_count);
    else hipLaunchKernelGGL(k_rgg_points<3>, dim3(blocks_for(n)), dim3(256), 0, s, n, m, key, d_xyz, d_cell, d_count);
    MQ_HIP(ctx, hipGetLastError());
    // exclusive scans (cell starts over ncell+1 entries, row pointers over n+1)
    size_t tbytes = 0, t2 = 0;
    MQ_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tbytes, d_count, d_start, (int)(ncell + 1), s));
    MQ_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, t2, d_deg, d_rp, (int)(n + 1), s));
    tbytes = std::max(tbytes, t2);
    MQ_HIP(ctx, hipMalloc(&tmp.p, tbytes));
    MQ_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp.p, tbytes, d_count, d_start, (int)(ncell + 1), s));
    if (dim == 2)
        hipLaunchKernelGGL(k_rgg_scatter<2>, dim3(blocks_for(n)), dim3(256), 0, s, n, d_xyz, d_cell, d_start, d_fill,
                           d_order, d_sxyz);
    else
        hipLaunchKernelGGL(k_rgg_scatter<3>, dim3(blocks_for(n)), dim3(256), 0, s, n, d_xyz, d_cell, d_start, d_fill,
                           d_order, d_sxyz);
    MQ_HIP(ctx, hipGetLastError());
    const double r2 = radius * radius;
    if (dim == 2)
        hipLaunchKernelGGL((k_rgg_nbrs<2, false>), dim3(blocks_for(n)), dim3(256), 0, s, n, m, r2, d_xyz, d_start,
                           d_order, d_sxyz, d_deg, nullptr, nullptr);
    else
        hipLaunchKernelGGL((k_rgg_nbrs<3, false>), dim3(blocks_for(n)), dim3(256), 0, s, n, m, r2, d_xyz, d_start,
                           d_order, d_sxyz, d_deg, nullptr, nullptr);
    MQ_HIP(ctx, hipGetLastError());
    MQ_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp.p, tbytes, d_deg, d_rp, (int)(n + 1), s));
    int32_t total = 0;
    MQ_HIP(ctx, hipMemcpyAsync(&total, d_rp + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MQ_HIP(ctx, hipStreamSynchronize(s));
    *nnz = total;
    if (col_idx && col_cap < total)
        return mq_fail(ctx, MQ_ERR_ARG, "random geometric graph: col_idx holds fewer entries than nnz");
    if (row_ptr) MQ_HIP(ctx, hipMemcpyAsync(row_ptr, d_rp, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, s));
    if (coords) MQ_HIP(ctx, hipMemcpyAsync(coords, d_xyz, sizeof(double) * dim * n, hipMemcpyDeviceToHost, s));
    if (col_idx && total > 0) {
        DevBuf ci;
        MQ_HIP(ctx, hipMalloc(&ci.p, sizeof(int32_t) * (size_t)total));
        auto* d_ci = static_cast<int32_t*>(ci.p);
        if (dim == 2)
            hipLaunchKernelGGL((k_rgg_nbrs<2, true>), dim3(blocks_for(n)), dim3(256), 0, s, n, m, r2, d_xyz, d_start,
                               d_order, d_sxyz, nullptr, d_rp, d_ci);
        else
            hipLaunchKernelGGL((k_rgg_nbrs<3, true>), dim3(blocks_for(n)), dim3(256), 0, s, n, m, r2, d_xyz, d_start,
                               d_order, d_sxyz, nullptr, d_rp, d_ci);
        MQ_HIP(ctx, hipGetLastError());
        MQ_HIP(ctx, hipMemcpyAsync(col_idx, d_ci, sizeof(int32_t) * (size_t)total, hipMemcpyDeviceToHost, s));
        MQ_HIP(ctx, hipStreamSynchronize(s));
    }
    MQ_HIP(ctx, hipStreamSynchronize(s));
    return MQ_OK;
}
