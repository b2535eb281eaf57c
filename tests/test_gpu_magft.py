"""MagFTComp on the GPU (mq_measure_ft) against the oracle's restatement of
observables.h:104-170 on the downloaded spins, for the models whose reference
observables include it (Ising, Phi4) and the other regular-lattice spins (EA,
Heisenberg); Potts, bit-packed storage and graphs are refused."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import observables  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from marqov_amd import gpu
    return gpu.Context(0)


@pytest.mark.parametrize("model,dim,L", [("ising", 2, 16), ("ising", 3, 8), ("ea", 3, 6), ("heisenberg", 3, 8),
                                         ("phi4", 2, 12), ("phi4", 3, 6), ("ising", 2, 2)])
def test_magft_matches_oracle(ctx, model, dim, L):
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, dim, L)
    lat2 = gpu.RegularLattice(ctx, dim, L + 2)
    st = gpu.State(ctx, model, [lat, lat, lat2], J=-1.0, proposal="small", delta=1.0)
    st.set_beta([0.3, 0.9, 0.6])
    if model == "ea":
        st.random_bonds(p_plus=0.5, seed=2)
    st.init(seed=4)
    st.metropolis(3, seed=5)
    if model == "ising":  # an ordered stripe along x0 in replica 1: a large, known component
        N = L ** dim
        st.upload(1, np.where(np.arange(N) % L < L // 2, 1, -1).astype(np.int8))
    ft = st.measure_ft()
    for r, lt in enumerate([L, L, L + 2]):
        s = st.download(r)
        for d in range(3):
            if d < dim:
                assert ft[r, d] == pytest.approx(observables.magft(s, dim, lt, d), rel=1e-9, abs=1e-14), (r, d)
            else:
                assert np.isnan(ft[r, d])
    if model == "ising":
        assert ft[1, 0] == pytest.approx((2.0 / (L * np.sin(np.pi / L))) ** 2, rel=1e-12)


def test_magft_refusals(ctx):
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, 2, 8)
    st = gpu.State(ctx, "potts", [lat], q=3)
    with pytest.raises(gpu.MarqovError, match="Potts"):
        st.measure_ft()
    rp = np.arange(0, 2 * 8 + 1, 2, dtype=np.int32)
    ci = np.empty(16, np.int32)
    ci[0::2] = (np.arange(8) + 1) % 8
    ci[1::2] = (np.arange(8) - 1) % 8
    g = gpu.State(ctx, "ising", [gpu.CsrLattice(ctx, rp, ci)])
    with pytest.raises(gpu.MarqovError, match="regular lattices"):
        g.measure_ft()
    m = gpu.State(ctx, "ising", [gpu.RegularLattice(ctx, 3, 64)], spin_bits=1)
    with pytest.raises(gpu.MarqovError, match="spin_bits"):
        m.measure_ft()
