"""Bit-exact parity of the 3D Ising sweep at the size it is benchmarked: L = 256
(16,777,216 sites per replica) in a batch of 32 replicas, int8 and 1 bit per spin,
against the oracle's CPU replay of the same colour schedule and Philox stream
(regular_hypercubic.h:31-66 neighbours, metropolis.h:86-126 acceptance).

Each replay takes about 2 s per sweep per replica on one host core, so the tests
replay a few replicas of the batch (the first, one inside, the last) in threads;
every launch shape of the row-vectorised kernel (threads per block, XCD-aware
block order on and off; colour 1 always walks its blocks backwards) must give the
same bits."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu

L = 256
N = L ** 3
NREP = 32
BETAS = np.linspace(0.20, 0.26, NREP)  # one table per replica: beta_c = 0.2216546 lies inside
CHECK = (0, 13, 31)
SEED, SWEEP0, NSW = 1234, 7, 2


def _replay(fn, reps):
    with ThreadPoolExecutor(max_workers=len(reps)) as ex:
        return dict(zip(reps, ex.map(fn, reps)))


@pytest.fixture(scope="module")
def c2_int8_expected():
    def one(r):
        return ob.ising_sweeps(3, L, ob.init_ising(N, 2, r), float(BETAS[r]), -1.0, NSW, SEED, SWEEP0, r)
    return _replay(one, CHECK)


@pytest.mark.parametrize("block,xcd", [(64, 1), (256, 1), (512, 1), (256, 0)])
def test_c2_int8_row16_batch_bit_exact(ctx, c2_int8_expected, block, xcd):
    """32 replicas of 3D L = 256 int8 Ising (the headline state) through k_sweep_ising3_row16
    with the given launch shape: replicas 0, 13 and 31 equal the replay after 2 sweeps, and the
    FP64 observables equal the oracle's bond and magnetisation sums of the downloaded states."""
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, 3, L)
    st = gpu.State(ctx, "ising", [lat] * NREP, J=-1.0)
    st.set_tuning(kernels=0, block=block, xcd_remap=xcd, lds_block=256)
    st.set_beta(BETAS)
    st.init(seed=2)
    st.metropolis(NSW, seed=SEED, sweep0=SWEEP0)
    obs = st.measure()
    for r in CHECK:
        got = st.download(r)
        exp = c2_int8_expected[r]
        assert np.array_equal(got, exp), (block, xcd, r, int((got != exp).sum()))
        e = -1.0 * ob.lib().orc_bond_sum_ising(3, L, ob.p(got), None) / N
        m = abs(int(got.astype(np.int64).sum())) / N
        assert obs[r, 0] == pytest.approx(e, abs=1e-12), r
        assert obs[r, 1] == pytest.approx(m, abs=1e-12), r


def test_c2_replica_in_batch_equals_replica_alone(ctx):
    """Replica 0 of the 32-replica L = 256 batch and a one-replica state (same seed, same
    temperature) end in identical bits: the batch changes the launch, not the arithmetic."""
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, 3, L)
    out = []
    for nrep in (NREP, 1):
        st = gpu.State(ctx, "ising", [lat] * nrep, J=-1.0)
        st.set_beta(BETAS[:nrep])
        st.init(seed=2)
        st.metropolis(3, seed=SEED, sweep0=SWEEP0)
        out.append(st.download(0))
        del st
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("Lm,nrep,reps", [(128, 2, (0, 1)), (256, NREP, (0, 31))])
def test_c2_msc_bit_exact_full_size(ctx, Lm, nrep, reps):
    """1-bit multi-spin-coded 3D Ising (k_sweep_msc) at L = 128 and at L = 256 in the
    32-replica batch: the checked replicas equal the replay of orc_ising_msc_sweeps."""
    from marqov_amd import gpu
    Nm = Lm ** 3
    betas = np.linspace(0.20, 0.26, nrep)
    lat = gpu.RegularLattice(ctx, 3, Lm)
    st = gpu.State(ctx, "ising", [lat] * nrep, J=-1.0, spin_bits=1)
    st.set_beta(betas)
    st.init(seed=6)
    st.metropolis(NSW, seed=41, sweep0=5)

    def one(r):
        return ob.ising_msc_sweeps(Lm, ob.init_ising(Nm, 6, r), float(betas[r]), -1.0, NSW, 41, 5, r)

    exp = _replay(one, reps)
    obs = st.measure()
    for r in reps:
        got = st.download(r)
        assert np.array_equal(got, exp[r]), (Lm, r, int((got != exp[r]).sum()))
        e = -1.0 * ob.lib().orc_bond_sum_ising(3, Lm, ob.p(got), None) / Nm
        assert obs[r, 0] == pytest.approx(e, abs=1e-12), r
