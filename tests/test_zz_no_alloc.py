"""mq_measure_noalloc (the allocation-free measurement entry point) against mq_measure,
through the C ABI: same observables, bit for bit for integer models and to FP64
summation round-off (1e-12 relative) for FP32 spins, on two calls in a row, the second
on a smaller state so that what the first call left in the shared static scratch must
not leak into it.  Output goes to device memory from hipMalloc (the test's, not the
library's) and comes back with hipMemcpy after mq_sync."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        self.lib = ctypes.CDLL(os.path.join(rocm, "lib", "libamdhip64.so"))

    def malloc(self, n):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) == 0
        return p

    def to_host(self, p, count):
        out = np.zeros(count, np.float64)
        assert self.lib.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self, p):
        self.lib.hipFree(p)


def _noalloc(ctx, st, hip, dptr):
    rc = ctx.lib.mq_measure_noalloc(st.h, dptr)
    ctx.check(rc, "mq_measure_noalloc")
    ctx.sync()
    return hip.to_host(dptr, st.nrep * 4).reshape(st.nrep, 4)


def _state(ctx, model, lats, beta, **kw):
    from marqov_amd import gpu
    st = gpu.State(ctx, model, lats, **kw)
    st.set_beta(beta)
    st.init(seed=5)
    st.metropolis(4, seed=6)
    if model not in ("ea",):
        st.wolff(2, seed=6)
    return st


def test_measure_noalloc_equals_measure_two_calls(ctx):
    from marqov_amd import gpu
    hip = _Hip()
    dptr = hip.malloc(512 * 4 * 8)
    try:
        # first call: a larger integer state with two lattice sizes (two groups)
        la, lb = gpu.RegularLattice(ctx, 2, 32), gpu.RegularLattice(ctx, 2, 16)
        big = _state(ctx, "ising", [la] * 300 + [lb] * 200, np.linspace(0.2, 0.6, 500), J=-1.0)
        got, ref = _noalloc(ctx, big, hip, dptr), big.measure()
        assert np.array_equal(got, ref), np.abs(got - ref).max()
        # second call: a smaller state of another model; the first call's sums must not leak in
        small = _state(ctx, "potts", [gpu.RegularLattice(ctx, 2, 8)] * 6, 1.0, J=-1.0, q=3)
        got, ref = _noalloc(ctx, small, hip, dptr), small.measure()
        assert np.array_equal(got, ref), np.abs(got - ref).max()
        # FP32 spins, FP64 sums in atomic order: equal to round-off
        vec = _state(ctx, "phi4", [gpu.RegularLattice(ctx, 3, 8)] * 4, 0.69, J=-1.0, lam=4.5, mass=1.0,
                     embedding="random")
        got, ref = _noalloc(ctx, vec, hip, dptr), vec.measure()
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-14), np.abs(got - ref).max()
        # graph lattice and 1-bit storage take their own measurement kernels
        from marqov_amd import graphs
        rp, ci = graphs.regular_csr(2, 6)
        csr = _state(ctx, "ising", [gpu.CsrLattice(ctx, rp, ci)] * 3, 0.44, J=-1.0)
        got, ref = _noalloc(ctx, csr, hip, dptr), csr.measure()
        assert np.array_equal(got, ref)
        msc = gpu.State(ctx, "ising", [gpu.RegularLattice(ctx, 3, 64)] * 2, J=-1.0, spin_bits=1)
        msc.set_beta(0.22)
        msc.init(seed=3)
        msc.metropolis(2, seed=3)
        got, ref = _noalloc(ctx, msc, hip, dptr), msc.measure()
        assert np.array_equal(got, ref)
    finally:
        hip.free(dptr)


def test_measure_noalloc_refuses_more_replicas_than_its_scratch(ctx):
    from marqov_amd import gpu
    cap = 32768
    lats = [gpu.RegularLattice(ctx, 2, 2), gpu.RegularLattice(ctx, 2, 4)]
    st = gpu.State(ctx, "ising", [lats[0]] * (cap // 2) + [lats[1]] * (cap // 2 + 1), J=-1.0)
    st.set_beta(0.3)
    hip = _Hip()
    dptr = hip.malloc((cap + 1) * 4 * 8)
    try:
        rc = ctx.lib.mq_measure_noalloc(st.h, dptr)
        assert rc == -1  # MQ_ERR_ARG
        assert b"MQ_NOALLOC_MAX_REP" in ctx.lib.mq_last_error(ctx.h)
        assert ctx.lib.mq_measure_noalloc(st.h, None) != 0
    finally:
        hip.free(dptr)
