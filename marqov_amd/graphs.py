"""Graph construction for the CSR path (host side).

* rgg2d: random geometric graph on the periodic unit square -- N points uniform
  from a seed, an edge between points closer than `radius` (minimum-image
  distance), built with a cell list on the host (numpy).  The device builder,
  2D and 3D, is `gpu.Context.rgg` (mq_rgg_build).  The reference's RandomGeometricGraph is
  3D only and uses a PCL kd-tree with a random_device seed
  (src/lattice/random_geometric_graph.h:29-168); this one is 2D, seeded and
  library-free.  rgg_radius(N, k) gives the radius for mean degree k:
  pi r^2 N = k.
* read_csv_graph / write_csv_graph: the tab-separated geometry format of
  GraphFromCSV / save_geometry (src/lattice/util/io.h:35-126): line 1 = number
  of coordinates, line 2 = header (skipped), then one line per vertex with its
  coordinates followed by its neighbours' indices.
* regular_csr: the periodic hypercube as CSR, neighbour order +x0,-x0,+x1,...
  (src/lattice/regular_hypercubic.h:31-66), for cross-checks of the two paths.
"""
import numpy as np


def rgg_radius(n, mean_degree=6.0, dim=2):
    """Radius for mean degree k: pi r^2 n = k (2D), 4/3 pi r^3 n = k (3D)."""
    if dim == 3:
        return float(np.cbrt(3.0 / (4.0 * np.pi) * mean_degree / n))
    return float(np.sqrt(mean_degree / (np.pi * n)))


def rgg2d(n, radius, seed):
    """Returns (row_ptr int32[n+1], col_idx int32[nnz], points float64[n,2])."""
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 2))
    m = max(1, int(np.floor(1.0 / radius)))
    cell = np.minimum((pts * m).astype(np.int64), m - 1)
    cid = cell[:, 0] * m + cell[:, 1]
    order = np.argsort(cid, kind="stable")
    starts = np.searchsorted(cid[order], np.arange(m * m + 1))
    src_all, dst_all = [], []
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            # pair every point with every point of the neighbouring cell (dx, dy)
            nc = ((cell[:, 0] + dx) % m) * m + (cell[:, 1] + dy) % m
            cnt = starts[nc + 1] - starts[nc]
            tot = int(cnt.sum())
            if tot == 0:
                continue
            src = np.repeat(np.arange(n), cnt)
            off = np.repeat(starts[nc], cnt) + (np.arange(tot) - np.repeat(np.cumsum(cnt) - cnt, cnt))
            dst = order[off]
            d = pts[src] - pts[dst]
            d -= np.round(d)
            keep = (np.einsum("ij,ij->i", d, d) < radius * radius) & (src != dst)
            src_all.append(src[keep])
            dst_all.append(dst[keep])
    src = np.concatenate(src_all)
    dst = np.concatenate(dst_all)
    # m < 3 makes neighbouring cells coincide: drop duplicate pairs
    key = np.unique(src.astype(np.int64) * n + dst)
    src, dst = key // n, key % n
    row_ptr = np.zeros(n + 1, np.int64)
    np.add.at(row_ptr, src + 1, 1)
    row_ptr = np.cumsum(row_ptr)
    return row_ptr.astype(np.int32), dst.astype(np.int32), pts


def regular_csr(dim, L):
    N = L ** dim
    i = np.arange(N)
    cols = []
    for d in range(dim):
        pw = L ** d
        x = (i // pw) % L
        cols.append(i + (((x + 1) % L) - x) * pw)
        cols.append(i + (((x - 1) % L) - x) * pw)
    ci = np.stack(cols, 1).reshape(-1)
    return (np.arange(N + 1) * 2 * dim).astype(np.int32), ci.astype(np.int32)


def write_csv_graph(path, row_ptr, col_idx, coords=None):
    n = len(row_ptr) - 1
    nc = 0 if coords is None else coords.shape[1]
    with open(path, "w") as f:
        f.write(f"{nc}\n")
        f.write("coordinates and neighbours\n")
        for v in range(n):
            parts = [] if coords is None else [f"{c:.16f}" for c in coords[v]]
            parts += [str(int(j)) for j in col_idx[row_ptr[v]:row_ptr[v + 1]]]
            f.write("\t".join(parts) + "\n")


def read_csv_graph(path, idxoffs=0):
    with open(path) as f:
        ncoords = int(f.readline())
        f.readline()
        coords, nbrs = [], []
        # every further line is a vertex; an empty line is one without coordinates or
        # neighbours, as import_geometry reads it (lattice/util/io.h:99-123): its
        # coordinates are NaN here
        for line in f:
            line = line.rstrip("\n").rstrip("\r")
            w = [x for x in line.split("\t") if x != ""]
            c = [float(x) for x in w[:ncoords]]
            coords.append(c + [float("nan")] * (ncoords - len(c)))
            nbrs.append([int(x) - idxoffs for x in w[ncoords:]])
    row_ptr = np.zeros(len(nbrs) + 1, np.int64)
    row_ptr[1:] = np.cumsum([len(x) for x in nbrs])
    col_idx = np.array([j for x in nbrs for j in x], np.int32)
    return row_ptr.astype(np.int32), col_idx, np.array(coords, np.float64).reshape(len(nbrs), ncoords)
