"""The C ABI is re-entrant per handle: every context owns its own HIP stream, and
nothing else is shared, so simulations driven from several host threads at once
(as MARQOV's scheduler runs one Core per pool thread, marqovqueue.h:169-226) give
exactly the spins and cluster sizes they give one after another, and the same
observables (to the last-bit rounding of FP64 atomics for the FP32 models)."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(out, idx, model):
    from marqov_amd import gpu
    ctx = gpu.Context(0)
    if model == "graph":
        from marqov_amd import graphs
        rp, ci, _ = graphs.rgg2d(3000, graphs.rgg_radius(3000), 40 + idx)
        lat = gpu.CsrLattice(ctx, rp, ci)
        st = gpu.State(ctx, "potts", [lat] * 8, J=-1.0, q=3)
        st.set_beta(1.0)
    elif model == "heisenberg":
        lat = gpu.RegularLattice(ctx, 3, 16)
        st = gpu.State(ctx, "heisenberg", [lat] * 4, J=-1.0, proposal="small", delta=0.6)
        st.set_beta(0.693)
    else:
        lat = gpu.RegularLattice(ctx, 3, 32)
        st = gpu.State(ctx, "ising", [lat] * 4, J=-1.0)
        st.set_beta(0.2216546)
    st.init(seed=100 + idx)
    sizes = []
    for t in range(10):
        st.metropolis(2, seed=200 + idx, sweep0=2 * t)
        sizes.append(st.wolff(3, seed=300 + idx, cluster0=3 * t))
    out[idx] = ([st.download(r) for r in range(st.nrep)], st.measure(), np.concatenate(sizes, axis=1))
    del st
    del lat
    ctx.close()


@pytest.mark.parametrize("models", [("ising", "ising", "heisenberg", "graph")])
def test_contexts_in_threads_match_sequential_runs(models):
    seq, par = {}, {}
    for i, m in enumerate(models):
        _run(seq, i, m)
    errors = []

    def guarded(i, m):
        try:
            _run(par, i, m)
        except Exception as e:  # surfaced below
            errors.append(repr(e))

    threads = [threading.Thread(target=guarded, args=(i, m)) for i, m in enumerate(models)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not errors, errors
    assert all(not t.is_alive() for t in threads)
    for i in range(len(models)):
        for a, b in zip(seq[i][0], par[i][0]):
            assert np.array_equal(a, b), i
        # FP64 observables of FP32 models are sums of per-block atomics, whose order
        # (and last-bit rounding) follows the block schedule; integer models' are exact
        assert np.allclose(seq[i][1], par[i][1], rtol=1e-12, atol=1e-15), i
        assert np.array_equal(seq[i][2], par[i][2]), i
