"""Autotuner: every candidate launch shape gives the same chain, and the tuner
caches and re-applies its choice."""
import json

import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,L", [(3, 32), (2, 64)])
def test_all_shapes_bit_exact(ctx, dim, L):
    from marqov_amd import autotune, gpu
    lat = gpu.RegularLattice(ctx, dim, L)
    ref = None
    for shape in autotune.candidates(gpu.State(ctx, "ising", [lat, lat], J=-1.0)):
        st = gpu.State(ctx, "ising", [lat, lat], J=-1.0)
        st.set_tuning(**shape)
        st.set_beta([0.25, 0.5])
        st.init(seed=4)
        st.metropolis(3, seed=8, sweep0=0)
        out = [st.download(r) for r in range(2)]
        if ref is None:
            ref = out
            exp = ob.ising_sweeps(dim, L, ob.init_ising(L ** dim, 4, 0), 0.25, -1.0, 3, 8, 0, 0)
            assert np.array_equal(out[0], exp)
        assert all(np.array_equal(a, b) for a, b in zip(out, ref)), shape


@pytest.mark.parametrize("model", ["heisenberg", "potts"])
def test_all_shapes_bit_exact_other_models(ctx, model):
    """The FP32 and Potts sweeps give the same spins under every launch shape
    (threads per block, XCD-aware tile order)."""
    from marqov_amd import autotune, gpu
    lat = gpu.RegularLattice(ctx, 3, 32)
    kw = dict(J=-1.0, q=3, proposal="small", delta=0.6)
    ref = None
    for shape in autotune.candidates(gpu.State(ctx, model, [lat] * 8, **kw)):
        st = gpu.State(ctx, model, [lat] * 8, **kw)
        st.set_tuning(**shape)
        st.set_beta(0.6)
        st.init(seed=4)
        st.metropolis(3, seed=8, sweep0=0)
        out = [st.download(r) for r in (0, 7)]
        if ref is None:
            ref = out
        assert all(np.array_equal(a, b) for a, b in zip(out, ref)), shape


def test_tune_caches_choice(ctx, tmp_path):
    from marqov_amd import autotune, gpu
    lat = gpu.RegularLattice(ctx, 3, 64)
    st = gpu.State(ctx, "ising", [lat] * 4, J=-1.0)
    st.set_beta(0.22)
    st.init(seed=1)
    path = str(tmp_path / "t.json")
    shape, timings = autotune.tune(ctx, st, cache_path=path, sweeps=2)
    assert {k: st.get_tuning()[k] for k in shape} == shape
    assert len(timings) == len(autotune.candidates(st))
    data = json.load(open(path))
    assert data[autotune.key_of(st)]["shape"] == shape
    st2 = gpu.State(ctx, "ising", [lat] * 4, J=-1.0)
    shape2, _ = autotune.tune(ctx, st2, cache_path=path)
    assert shape2 == shape and {k: st2.get_tuning()[k] for k in shape} == shape


@pytest.mark.parametrize("model,lattice", [("ising", "3d16"), ("ising", "2d256"), ("potts", "2d32"),
                                           ("heisenberg", "3d16"), ("ising", "graph")])
def test_all_wolff_shapes_bit_exact(ctx, model, lattice, tmp_path):
    """Every Wolff launch shape the tuner tries (workgroup size, LDS membership, LDS ring)
    grows the same clusters; tune_wolff leaves the caller's state unchanged, and the chain
    after tuning equals the chain without it."""
    from marqov_amd import autotune, gpu, graphs
    if lattice == "graph":
        rp, ci = graphs.regular_csr(2, 40)
        lat = gpu.CsrLattice(ctx, rp, ci)
    else:
        dim, L = int(lattice[0]), int(lattice[2:])
        lat = gpu.RegularLattice(ctx, dim, L)
    kw = dict(J=-1.0, q=3, embedding="random")
    beta = {"ising": 0.3 if lattice.startswith("3d") else 0.44, "potts": 1.0, "heisenberg": 0.7}[model]

    def fresh():
        st = gpu.State(ctx, model, [lat] * 4, **kw)
        st.set_beta(beta)
        st.init(seed=3)
        st.metropolis(4, seed=5, sweep0=0)
        return st

    ref = None
    for shape in autotune.wolff_candidates(fresh()):
        st = fresh()
        st.set_tuning(**shape)
        sizes = st.wolff(6, seed=7, cluster0=0)
        out = (sizes, [st.download(r) for r in range(4)])
        if ref is None:
            ref = out
        assert np.array_equal(out[0], ref[0]), shape
        assert all(np.array_equal(a, b) for a, b in zip(out[1], ref[1])), shape
    st = fresh()
    before = [st.download(r) for r in range(4)]
    shape, timings = autotune.tune_wolff(ctx, st, nclusters=4, cache_path=str(tmp_path / "w.json"))
    assert len(timings) == len(autotune.wolff_candidates(st))
    assert all(np.array_equal(st.download(r), before[r]) for r in range(4))
    assert {k: st.get_tuning()[k] for k in shape} == shape
    sizes = st.wolff(6, seed=7, cluster0=0)
    assert np.array_equal(sizes, ref[0])
