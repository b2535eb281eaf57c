"""Error behaviour of the C ABI: invalid arguments and unsupported requests are refused on
the host with a negative status and a message in mq_last_error (the reference throws
std::runtime_error or exits with a message, core.h:338-339, metropolishelpers.h:299-301);
nothing is launched, and the context stays usable afterwards."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from marqov_amd import gpu
    return gpu.Context(0)


def _ring(n):
    rp = np.arange(0, 2 * n + 1, 2, dtype=np.int32)
    ci = np.empty(2 * n, np.int32)
    ci[0::2] = (np.arange(n) + 1) % n
    ci[1::2] = (np.arange(n) - 1) % n
    return rp, ci


def _lattice_cases():
    from marqov_amd import gpu
    rp, ci = _ring(6)
    asym = ci.copy()
    asym[0] = 3  # 0 -> 3 without 3 -> 0
    selfl = ci.copy()
    selfl[0] = 0
    dup = ci.copy()
    dup[1] = dup[0]
    bad_rp = rp.copy()
    bad_rp[4] = bad_rp[3] - 1  # row 3 ends before it starts
    oob = ci.copy()
    oob[2] = 6
    return [
        ("dim must be 2 or 3", lambda c: gpu.RegularLattice(c, 4, 8)),
        ("L must be even", lambda c: gpu.RegularLattice(c, 3, 7)),
        ("not symmetric", lambda c: gpu.CsrLattice(c, rp, asym)),
        ("self loop", lambda c: gpu.CsrLattice(c, rp, selfl)),
        ("duplicate edge", lambda c: gpu.CsrLattice(c, rp, dup)),
        ("not monotone", lambda c: gpu.CsrLattice(c, bad_rp, ci)),
        ("bad neighbour index", lambda c: gpu.CsrLattice(c, rp, oob)),
    ]


@pytest.mark.parametrize("case", range(7))
def test_lattice_arguments_refused(ctx, case):
    from marqov_amd import gpu
    match, make = _lattice_cases()[case]
    with pytest.raises(gpu.MarqovError, match=match):
        make(ctx)


def test_state_arguments_refused(ctx):
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, 3, 8)
    lat2 = gpu.RegularLattice(ctx, 2, 8)
    rp, ci = _ring(8)
    g = gpu.CsrLattice(ctx, rp, ci)
    with pytest.raises(gpu.MarqovError, match="q must be in"):
        gpu.State(ctx, "potts", [lat], q=9)
    with pytest.raises(gpu.MarqovError, match="one kind and dim"):
        gpu.State(ctx, "ising", [lat, lat2])
    with pytest.raises(gpu.MarqovError, match="one kind and dim"):
        gpu.State(ctx, "ising", [lat, g])
    with pytest.raises(gpu.MarqovError, match="spin_bits"):
        gpu.State(ctx, "ising", [lat], spin_bits=4)
    with pytest.raises(gpu.MarqovError, match="power of two >= 64"):
        gpu.State(ctx, "ising", [lat], spin_bits=1)
    with pytest.raises(gpu.MarqovError, match="multiple of 32"):
        gpu.State(ctx, "ea", [lat] * 16, spin_bits=1)
    with pytest.raises(gpu.MarqovError, match="Ising model"):
        gpu.State(ctx, "potts", [gpu.RegularLattice(ctx, 3, 64)], spin_bits=1)


def test_calls_in_the_wrong_state_refused(ctx):
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, 3, 8)
    st = gpu.State(ctx, "ising", [lat] * 2)
    with pytest.raises(gpu.MarqovError, match="beta not set"):
        st.metropolis(1, seed=1)
    with pytest.raises(gpu.MarqovError, match="beta must be finite"):
        st.set_beta([0.3, np.nan])
    st.set_beta(0.3)
    with pytest.raises(gpu.MarqovError, match="expected one int8 per site"):
        st.upload(0, np.ones(100, np.int8))
    with pytest.raises(gpu.MarqovError, match="out of range"):
        st.upload(0, np.full(512, 3, np.int8))
    with pytest.raises(gpu.MarqovError, match="ntemp"):
        st.pt_swap(3, 0, 1, 0)
    with pytest.raises(gpu.MarqovError, match="renormalisation"):
        st.renormalize()
    with pytest.raises(gpu.MarqovError, match="tuning"):
        st.set_tuning(block=100)
    with pytest.raises(gpu.MarqovError, match="graph in 0..2"):
        st.set_tuning(graph=3)
    ea = gpu.State(ctx, "ea", [lat] * 2)
    ea.set_beta(0.5)
    with pytest.raises(gpu.MarqovError, match="EA couplings not set"):
        ea.metropolis(1, seed=1)
    ea.random_bonds(p_plus=0.5, seed=3)
    with pytest.raises(gpu.MarqovError, match="spin glass"):
        ea.wolff(1, seed=1)
    with pytest.raises(gpu.MarqovError, match="p_plus"):
        ea.random_bonds(p_plus=1.5, seed=3)
    # the context and the states still work after every refusal
    st.init(seed=2)
    st.metropolis(2, seed=1)
    ea.metropolis(2, seed=1)
    assert np.all(np.isfinite(st.measure())) and np.all(np.isfinite(ea.measure()))


def test_refusals_carry_their_own_message(ctx):
    """An argument refused on a valid handle leaves its own message, not the previous one."""
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, 2, 8)
    st = gpu.State(ctx, "ising", [lat] * 2)
    buf = np.zeros(64, np.int8)
    b = np.zeros(2, np.float64)
    lib = ctx.lib
    assert lib.mq_state_get_beta(st.h, gpu._ptr(b)) == -3
    assert b"beta not set" in lib.mq_last_error(ctx.h)
    st.set_beta(0.4)
    assert lib.mq_state_download(st.h, 5, gpu._ptr(buf), buf.nbytes) != 0
    assert b"replica index out of range" in lib.mq_last_error(ctx.h)
    assert lib.mq_state_upload(st.h, 0, None, buf.nbytes) != 0
    assert b"host buffer is NULL" in lib.mq_last_error(ctx.h)
    acc = np.zeros(1, np.int32)
    assert lib.mq_pt_swap_pairs(st.h, 2, -1, None, 1, 0, gpu._ptr(acc)) != 0
    assert b"npairs < 0" in lib.mq_last_error(ctx.h)
    assert lib.mq_lattice_csr(ctx.h, 0, gpu._ptr(np.zeros(1, np.int32)), gpu._ptr(np.zeros(1, np.int32)),
                              None) != 0
    assert b"NULL argument" in lib.mq_last_error(ctx.h)


def test_sweep_counter_limit_per_lattice_kind(ctx):
    """The Philox phase of a sweep is one 32-bit word: 2*sweep + colour on regular lattices,
    64*sweep + colour on graphs.  A graph state refuses sweep numbers from 2^26 on (otherwise
    sweeps s and s + 2^26 would draw the same numbers), through mq_metropolis, mq_emcs and
    mq_run_graph; a regular state still runs there and refuses from 2^31 on."""
    from marqov_amd import gpu
    rp, ci = _ring(8)
    g = gpu.State(ctx, "ising", [gpu.CsrLattice(ctx, rp, ci)] * 2)
    g.set_beta(0.4)
    g.init(seed=1)
    lim = 1 << 26
    g.metropolis(1, seed=1, sweep0=lim - 1)  # the last sweep number that fits
    with pytest.raises(gpu.MarqovError, match="2\\^26"):
        g.metropolis(1, seed=1, sweep0=lim)
    with pytest.raises(gpu.MarqovError, match="2\\^26"):
        g.metropolis(2, seed=1, sweep0=lim - 1)
    with pytest.raises(gpu.MarqovError, match="2\\^26"):
        g.emcs(0, 1, 1, lim)  # sweep0 = step * nmetro
    with pytest.raises(gpu.MarqovError, match="2\\^26"):
        g.run_graph(1, 1, 1, lim - 1, 0, 2, 1)
    r = gpu.State(ctx, "ising", [gpu.RegularLattice(ctx, 2, 8)] * 2)
    r.set_beta(0.4)
    r.init(seed=1)
    r.metropolis(1, seed=1, sweep0=lim)
    r.metropolis(1, seed=1, sweep0=(1 << 31) - 1)
    with pytest.raises(gpu.MarqovError, match="2\\^31"):
        r.metropolis(1, seed=1, sweep0=1 << 31)
    assert np.all(np.isfinite(g.measure())) and np.all(np.isfinite(r.measure()))
