"""Random geometric graphs built on the device (mq_rgg_build) against the oracle.

The oracle regenerates the points from the same Philox counters (oracle/replay.c,
orc_rgg_points; its Philox is pinned by the Random123 and torch vectors) and joins
them by an all-pairs minimum-image search, an algorithm independent of the device's
cell list; the CSR must be identical.  At the C5 size the graph is checked through
properties: symmetry, sorted rows, no self loops and the mean degree
(n - 1) * pi * r^2 = 6 within 0.05 (about 20 standard deviations of the mean)."""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from marqov_amd import gpu
    return gpu.Context(0)


def _edges(rp, ci):
    src = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return src.astype(np.int64), ci.astype(np.int64)


@pytest.mark.parametrize("dim,n,radius,seed", [
    (2, 3000, 0.0252, 3),     # mean degree 6, 39 x 39 cells
    (3, 3000, 0.0785, 4),     # mean degree 6, 12^3 cells
    (2, 2000, 1.0 / 3.0, 5),  # exactly 3 cells per axis: the smallest wrapped 3x3 search
    (2, 600, 0.45, 6),        # 2 cells per axis: every cell searched once
    (3, 500, 0.34, 7),        # 2 cells per axis in 3D
    (3, 800, 0.2, 8),         # 5^3 cells, a wide radius
])
def test_rgg_equals_all_pairs_oracle(ctx, dim, n, radius, seed):
    rp, ci, xyz = ctx.rgg(dim, n, radius, seed, coords=True)
    pts = ob.rgg_points(dim, n, seed)
    assert np.array_equal(xyz, pts)
    erp, eci = ob.rgg_brute(pts, radius)
    assert np.array_equal(rp, erp)
    assert np.array_equal(ci, eci)


def test_rgg_c5_size_properties(ctx):
    from marqov_amd import graphs, gpu
    n = 1 << 20
    r = graphs.rgg_radius(n)
    rp, ci = ctx.rgg(2, n, r, 5)
    assert rp[0] == 0 and len(rp) == n + 1 and rp[-1] == len(ci)
    src, dst = _edges(rp, ci)
    assert not np.any(src == dst)
    assert np.all(np.diff(ci)[np.diff(src) == 0] > 0)  # rows strictly increasing
    fwd = np.sort(src * n + dst)
    rev = np.sort(dst * n + src)
    assert np.array_equal(fwd, rev)  # symmetric
    mean_deg = len(ci) / n
    assert abs(mean_deg - (n - 1) * np.pi * r * r) < 0.05, mean_deg
    lat = gpu.CsrLattice(ctx, rp, ci)
    assert 6 <= len(lat.colors()) <= 20


def test_rgg_3d_reference_shape(ctx):
    """The reference's RGG case (3D, mean degree 6, demo/IsingRGG.cpp) at 32^3 points."""
    from marqov_amd import graphs
    n = 32 ** 3
    r = graphs.rgg_radius(n, 6.0, dim=3)
    rp, ci = ctx.rgg(3, n, r, 11)
    src, dst = _edges(rp, ci)
    assert np.array_equal(np.sort(src * n + dst), np.sort(dst * n + src))
    assert abs(len(ci) / n - (n - 1) * 4.0 / 3.0 * np.pi * r ** 3) < 0.15


@pytest.mark.parametrize("dim,n,radius", [(4, 100, 0.1), (2, 1, 0.1), (2, 100, 0.5), (2, 100, 0.0),
                                          (3, 100, -0.1), (2, 1 << 20, 0.2)])
def test_rgg_rejects_bad_arguments(ctx, dim, n, radius):
    from marqov_amd import gpu
    with pytest.raises(gpu.MarqovError):
        ctx.rgg(dim, n, radius, 1)


def test_rgg_refuses_short_col_idx(ctx):
    """col_idx smaller than nnz is refused with the count still reported."""
    import ctypes
    from marqov_amd import gpu
    n, r = 2000, 0.03
    rp_ref, ci_ref = ctx.rgg(2, n, r, 3)
    nnz = ctypes.c_int64()
    rp = np.zeros(n + 1, np.int32)
    short = np.zeros(len(ci_ref) - 1, np.int32)
    rc = ctx.lib.mq_rgg_build(ctx.h, 2, n, r, 3, ctypes.byref(nnz), gpu._ptr(rp), gpu._ptr(short), short.size, None)
    assert rc != 0
    assert nnz.value == len(ci_ref)
    assert b"fewer entries" in ctx.lib.mq_last_error(ctx.h)

