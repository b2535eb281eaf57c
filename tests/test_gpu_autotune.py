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
    assert st.get_tuning() == shape
    assert len(timings) == len(autotune.candidates(st))
    data = json.load(open(path))
    assert data[autotune.key_of(st)]["shape"] == shape
    st2 = gpu.State(ctx, "ising", [lat] * 4, J=-1.0)
    shape2, _ = autotune.tune(ctx, st2, cache_path=path)
    assert shape2 == shape and st2.get_tuning() == shape
