"""Persistent EMCS (mq_emcs_run): nsteps EMCS of every replica in one launch, each
replica at its own pace, equal nsteps calls of mq_emcs bit for bit (states and
mean cluster sizes)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BETA = {("ising", 2): 0.4406868, ("ising", 3): 0.2216546, ("potts", 2): 1.0050525, ("potts", 3): 0.55,
        ("heisenberg", 3): 0.693, ("heisenberg", 2): 1.0, ("phi4", 3): 0.69, ("cart", 3): 0.693}


@pytest.mark.parametrize("model,dim,L,nrep,nc,nm", [
    ("ising", 2, 64, 8, 10, 1), ("ising", 3, 32, 8, 5, 2), ("ising", 3, 64, 4, 10, 1),
    ("potts", 2, 32, 8, 4, 1), ("potts", 3, 16, 8, 3, 2),
    ("ising", 2, 48, 4, 0, 3), ("ising", 3, 16, 4, 6, 0), ("ising", 3, 256, 2, 2, 1),
    ("heisenberg", 3, 16, 8, 5, 1), ("cart", 3, 24, 4, 3, 2), ("heisenberg", 2, 32, 4, 4, 1), ("phi4", 3, 12, 8, 4, 1),
    ("cart", 3, 32, 4, 5, 0)])
def test_emcs_run_equals_emcs_calls(ctx, model, dim, L, nrep, nc, nm):
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, dim, L)
    betas = BETA[(model, dim)] * np.linspace(0.9, 1.1, nrep)
    kw = dict(J=-1.0, q=3)
    if model == "cart":  # Heisenberg, small-angle proposal, Cartesian embedding
        model, kw = "heisenberg", dict(J=-1.0, proposal="small", delta=0.8, embedding="cartesian")
    elif model in ("heisenberg", "phi4"):
        kw = dict(J=-1.0, embedding="random")

    def fresh():
        st = gpu.State(ctx, model, [lat] * nrep, **kw)
        st.set_beta(betas)
        st.init(seed=6)
        return st

    a, b = fresh(), fresh()
    K, step0 = (2, 1) if L == 256 else (7, 3)
    ma = np.array([a.emcs(nc, nm, seed=11, step=step0 + k) for k in range(K)])
    mb = b.emcs_run(K, nc, nm, seed=11, step0=step0)
    for r in range(nrep):
        assert np.array_equal(a.download(r), b.download(r)), r
    np.testing.assert_array_equal(ma, mb)
    if nc > 0:
        assert np.all(mb >= 1)


def test_emcs_run_two_lattice_groups(ctx):
    from marqov_amd import gpu
    la, lb = gpu.RegularLattice(ctx, 3, 16), gpu.RegularLattice(ctx, 3, 24)
    lats = [la] * 4 + [lb] * 3

    def fresh():
        st = gpu.State(ctx, "ising", lats, J=-1.0)
        st.set_beta(np.linspace(0.2, 0.25, 7))
        st.init(seed=2)
        return st

    a, b = fresh(), fresh()
    ma = np.array([a.emcs(4, 1, seed=5, step=k) for k in range(5)])
    mb = b.emcs_run(5, 4, 1, seed=5, step0=0)
    assert all(np.array_equal(a.download(r), b.download(r)) for r in range(7))
    np.testing.assert_array_equal(ma, mb)


def test_emcs_run_refusals(ctx):
    from marqov_amd import gpu, graphs
    from marqov_amd.gpu import MarqovError
    lat = gpu.RegularLattice(ctx, 2, 16)
    ea = gpu.State(ctx, "ea", [lat] * 2, J=-1.0)
    ea.set_beta(0.5)
    ea.random_bonds()
    with pytest.raises(MarqovError, match="spin glass"):
        ea.emcs_run(1, 1, 1, seed=1, step0=0)
    g = gpu.CsrLattice(ctx, *graphs.regular_csr(2, 8))
    st = gpu.State(ctx, "ising", [g] * 2, J=-1.0)
    st.set_beta(0.4)
    with pytest.raises(MarqovError, match="regular lattices"):
        st.emcs_run(1, 1, 1, seed=1, step0=0)
    st = gpu.State(ctx, "ising", [lat] * 2, J=-1.0)
    st.set_beta(0.4)
    st.init(seed=1)
    with pytest.raises(MarqovError, match=">= 0"):
        st.emcs_run(1, -1, 1, seed=1, step0=0)
    with pytest.raises(MarqovError, match="cluster counter"):
        st.emcs_run(1, 1 << 30, 0, seed=1, step0=4)
    assert st.emcs_run(0, 1, 1, seed=1, step0=0).shape == (0, 2)
