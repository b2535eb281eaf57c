"""Host-side graph construction (CPU)."""
import os

import numpy as np

from marqov_amd import graphs


def test_rgg_matches_brute_force():
    n = 1500
    r = graphs.rgg_radius(n)
    rp, ci, pts = graphs.rgg2d(n, r, 11)
    d = pts[:, None, :] - pts[None, :, :]
    d -= np.round(d)
    A = np.einsum("ijk,ijk->ij", d, d) < r * r
    np.fill_diagonal(A, False)
    B = np.zeros_like(A)
    for v in range(n):
        B[v, ci[rp[v]:rp[v + 1]]] = True
    assert (A == B).all()
    assert abs(np.diff(rp).mean() - 6.0) < 0.5


def test_csv_roundtrip(tmp_path):
    rp, ci, pts = graphs.rgg2d(300, graphs.rgg_radius(300), 3)
    path = os.path.join(tmp_path, "g.csv")
    graphs.write_csv_graph(path, rp, ci, pts)
    rp2, ci2, pts2 = graphs.read_csv_graph(path)
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2) and np.allclose(pts, pts2)


def test_regular_csr_neighbour_order():
    import oracle_bind as ob
    rp, ci = graphs.regular_csr(3, 4)
    for i in range(64):
        assert list(ci[rp[i]:rp[i + 1]]) == [ob.lib().orc_neighbour(3, 4, i, s) for s in range(6)]
