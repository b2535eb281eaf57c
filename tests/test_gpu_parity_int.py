"""Bit-exact parity of the HIP integer-spin path against the CPU replay oracle.

Same seed, same Philox counters, same colour schedule and the same integer
acceptance tables: every spin must match after every call.
"""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu


def _state(ctx, model, dim, L, nrep=1, **kw):
    from marqov_amd import gpu
    lat = gpu.RegularLattice(ctx, dim, L)
    st = gpu.State(ctx, model, [lat] * nrep, **kw)
    st._lat = lat
    return st


@pytest.mark.parametrize("dim,L", [(2, 4), (2, 6), (2, 16), (3, 4), (3, 8), (3, 10)])
def test_init_matches_oracle(ctx, dim, L):
    st = _state(ctx, "ising", dim, L, nrep=3)
    st.set_beta(0.3)
    st.init(seed=11)
    for r in range(3):
        assert np.array_equal(st.download(r), ob.init_ising(L ** dim, 11, r)), (dim, L, r)


@pytest.mark.parametrize("dim,L,beta", [(2, 8, 0.4406868), (2, 16, 0.3), (2, 6, 1.0), (3, 8, 0.2216546),
                                        (3, 6, 0.5), (3, 16, 0.1)])
def test_ising_sweeps_bit_exact(ctx, dim, L, beta):
    nrep = 2
    st = _state(ctx, "ising", dim, L, nrep=nrep, J=-1.0)
    st.set_beta([beta, beta * 0.5])
    st.init(seed=5)
    st.metropolis(7, seed=123, sweep0=3)
    for r in range(nrep):
        s0 = ob.init_ising(L ** dim, 5, r)
        exp = ob.ising_sweeps(dim, L, s0, [beta, beta * 0.5][r], -1.0, 7, 123, 3, r)
        got = st.download(r)
        assert np.array_equal(got, exp), (dim, L, r, int((got != exp).sum()))


def test_ising_upload_roundtrip_and_antiferro(ctx):
    dim, L = 3, 6
    st = _state(ctx, "ising", dim, L, J=+1.0)
    st.set_beta(0.7)
    rng = np.random.default_rng(0)
    s = np.where(rng.random(L ** 3) < 0.5, 1, -1).astype(np.int8)
    st.upload(0, s)
    assert np.array_equal(st.download(0), s)
    st.metropolis(5, seed=9, sweep0=0)
    exp = ob.ising_sweeps(dim, L, s, 0.7, 1.0, 5, 9, 0, 0)
    assert np.array_equal(st.download(0), exp)


@pytest.mark.parametrize("dim,L", [(2, 8), (3, 6)])
def test_ea_sweeps_bit_exact(ctx, dim, L):
    N = L ** dim
    nrep = 4
    st = _state(ctx, "ea", dim, L, nrep=nrep, J=-1.0)
    betas = [0.5, 1.0, 2.0, 3.3]
    st.set_beta(betas)
    st.random_bonds(rep_sample=[0, 0, 1, 1], p_plus=0.5, seed=44)
    st.init(seed=7)
    st.metropolis(6, seed=77, sweep0=0)
    for r in range(nrep):
        b = ob.init_bonds(dim, N, 0.5, 44, r // 2)
        exp = ob.ising_sweeps(dim, L, ob.init_ising(N, 7, r), betas[r], -1.0, 6, 77, 0, r, bonds=b)
        assert np.array_equal(st.download(r), exp), r


def test_ea_host_bonds(ctx):
    dim, L = 3, 4
    N = L ** dim
    st = _state(ctx, "ea", dim, L, J=-1.0)
    st.set_beta(1.5)
    rng = np.random.default_rng(3)
    b = np.where(rng.random(N * dim) < 0.5, 1, -1).astype(np.int8)
    st.set_bonds(0, b)
    st.init(seed=2)
    st.metropolis(4, seed=8, sweep0=10)
    exp = ob.ising_sweeps(dim, L, ob.init_ising(N, 2, 0), 1.5, -1.0, 4, 8, 10, 0, bonds=b)
    assert np.array_equal(st.download(0), exp)


@pytest.mark.parametrize("dim,L,Q", [(2, 8, 3), (2, 6, 4), (3, 4, 3)])
def test_potts_sweeps_bit_exact(ctx, dim, L, Q):
    N = L ** dim
    st = _state(ctx, "potts", dim, L, nrep=2, J=-1.0, q=Q)
    st.set_beta([1.0, 0.6])
    st.init(seed=3)
    for r in range(2):
        assert np.array_equal(st.download(r), ob.init_potts(N, Q, 3, r))
    st.metropolis(5, seed=31, sweep0=2)
    for r in range(2):
        exp = ob.potts_sweeps(dim, L, ob.init_potts(N, Q, 3, r), Q, [1.0, 0.6][r], -1.0, 5, 31, 2, r)
        assert np.array_equal(st.download(r), exp), r


@pytest.mark.parametrize("dim,L,beta", [(2, 16, 0.4406868), (2, 8, 0.6), (3, 8, 0.2216546), (3, 6, 0.4)])
def test_wolff_ising_bit_exact(ctx, dim, L, beta):
    N = L ** dim
    nrep = 3
    st = _state(ctx, "ising", dim, L, nrep=nrep, J=-1.0)
    st.set_beta(beta)
    st.init(seed=21)
    sizes = st.wolff(6, seed=99, cluster0=5)
    for r in range(nrep):
        exp, esz = ob.wolff_ising(dim, L, ob.init_ising(N, 21, r), beta, -1.0, 6, 99, 5, r)
        assert np.array_equal(sizes[r], esz), (r, sizes[r], esz)
        assert np.array_equal(st.download(r), exp), r


@pytest.mark.parametrize("dim,L,Q", [(2, 8, 3), (3, 4, 4)])
def test_wolff_potts_bit_exact(ctx, dim, L, Q):
    N = L ** dim
    st = _state(ctx, "potts", dim, L, nrep=2, J=-1.0, q=Q)
    st.set_beta(1.1)
    st.init(seed=4)
    sizes = st.wolff(5, seed=55, cluster0=0)
    for r in range(2):
        exp, esz = ob.wolff_potts(dim, L, ob.init_potts(N, Q, 4, r), Q, 1.1, -1.0, 5, 55, 0, r)
        assert np.array_equal(sizes[r], esz)
        assert np.array_equal(st.download(r), exp)


@pytest.mark.parametrize("model,dim,L", [("ising", 2, 8), ("ising", 3, 6), ("potts", 2, 6)])
def test_measure_matches_oracle(ctx, model, dim, L):
    N = L ** dim
    st = _state(ctx, model, dim, L, nrep=2, J=-1.0, q=3)
    st.set_beta(0.5)
    st.init(seed=6)
    st.metropolis(3, seed=1)
    obs = st.measure()
    for r in range(2):
        s = st.download(r)
        if model == "ising":
            e = -1.0 * ob.lib().orc_bond_sum_ising(dim, L, ob.p(s), None) / N
            m = abs(int(s.astype(np.int64).sum())) / N
        else:
            e = -1.0 * ob.lib().orc_bond_sum_potts(dim, L, ob.p(s)) / N
            cnt = np.bincount(s.astype(np.int64), minlength=3)
            m = (3 * cnt.max() - N) / (N * 2)
        assert obs[r, 0] == pytest.approx(e, abs=1e-12)
        assert obs[r, 1] == pytest.approx(m, abs=1e-12)


@pytest.mark.parametrize("L", [32, 64])
def test_measure_row16_matches_oracle(ctx, L):
    """3D Ising with L a power of two >= 32 is measured by the row-vectorised kernel:
    bond sum and magnetisation must equal the oracle's from the downloaded spins."""
    dim, N = 3, L ** 3
    st = _state(ctx, "ising", dim, L, nrep=3, J=-1.0)
    st.set_beta([0.1, 0.2216546, 0.5])
    st.init(seed=12)
    st.metropolis(4, seed=3)
    obs = st.measure()
    for r in range(3):
        s = st.download(r)
        e = -1.0 * ob.lib().orc_bond_sum_ising(dim, L, ob.p(s), None) / N
        m = abs(int(s.astype(np.int64).sum())) / N
        assert obs[r, 0] == pytest.approx(e, abs=1e-12), r
        assert obs[r, 1] == pytest.approx(m, abs=1e-12), r


@pytest.mark.parametrize("L,nsweeps", [(32, 6), (64, 2)])
def test_ising3_row16_kernel_bit_exact(ctx, L, nsweeps):
    """L % 32 == 0 selects the row-vectorised 3D kernel; it must replay bit-exactly."""
    dim = 3
    nrep = 2
    st = _state(ctx, "ising", dim, L, nrep=nrep, J=-1.0)
    betas = [0.2216546, 0.35]
    st.set_beta(betas)
    st.init(seed=9)
    st.metropolis(nsweeps, seed=1234, sweep0=17)
    for r in range(nrep):
        exp = ob.ising_sweeps(dim, L, ob.init_ising(L ** 3, 9, r), betas[r], -1.0, nsweeps, 1234, 17, r)
        got = st.download(r)
        assert np.array_equal(got, exp), (L, r, int((got != exp).sum()))


def test_ising3_row16_antiferro_and_hot(ctx):
    """Antiferromagnetic J and beta = 0 (every move accepted) through the fast kernel."""
    dim, L = 3, 32
    for J, beta in ((+1.0, 0.4), (-1.0, 0.0), (-1.0, 5.0)):
        st = _state(ctx, "ising", dim, L, J=J)
        st.set_beta(beta)
        st.init(seed=3)
        st.metropolis(3, seed=5, sweep0=0)
        exp = ob.ising_sweeps(dim, L, ob.init_ising(L ** 3, 3, 0), beta, J, 3, 5, 0, 0)
        assert np.array_equal(st.download(0), exp), (J, beta)


@pytest.mark.parametrize("kernels", ["0", "1"])
def test_mixed_L_ea_batch_bit_exact(ctx, kernels, monkeypatch):
    """Mixed lattice sizes in one state (LDS multi-sweep kernel when kernels=0,
    streaming kernels when kernels=1) against the CPU replay."""
    from marqov_amd import gpu
    monkeypatch.setenv("MARQOV_AMD_KERNELS", kernels)
    dim = 3
    Ls = [4, 6, 8]
    lats = {L: gpu.RegularLattice(ctx, dim, L) for L in Ls}
    reps = [lats[L] for L in Ls for _ in range(4)]
    st = gpu.State(ctx, "ea", reps, J=-1.0)
    betas = np.tile([0.5, 1.1, 2.0, 3.3], 3)
    st.set_beta(betas)
    sample = np.repeat(np.arange(3), 4)
    st.random_bonds(rep_sample=sample, p_plus=0.5, seed=4)
    st.init(seed=12)
    st.metropolis(5, seed=3, sweep0=1)
    for r, lat in enumerate(reps):
        N = lat.N
        b = ob.init_bonds(dim, N, 0.5, 4, int(sample[r]))
        exp = ob.ising_sweeps(dim, lat.L, ob.init_ising(N, 12, r), betas[r], -1.0, 5, 3, 1, r, bonds=b)
        assert np.array_equal(st.download(r), exp), (kernels, r)


@pytest.mark.parametrize("kernels", ["0", "1"])
def test_2d_ising_batch_bit_exact(ctx, kernels, monkeypatch):
    from marqov_amd import gpu
    monkeypatch.setenv("MARQOV_AMD_KERNELS", kernels)
    lat = gpu.RegularLattice(ctx, 2, 64)
    st = gpu.State(ctx, "ising", [lat] * 3, J=-1.0)
    st.set_beta([0.4406868, 0.3, 0.6])
    st.init(seed=2)
    st.metropolis(4, seed=9, sweep0=0)
    for r in range(3):
        exp = ob.ising_sweeps(2, 64, ob.init_ising(4096, 2, r), [0.4406868, 0.3, 0.6][r], -1.0, 4, 9, 0, r)
        assert np.array_equal(st.download(r), exp), (kernels, r)


@pytest.mark.parametrize("J,betas", [(-1.0, [0.2216546, 0.4]), (+1.0, [0.3, 0.0])])
def test_msc_sweeps_bit_exact(ctx, J, betas):
    """1-bit multi-spin-coded 3D Ising (L=64) against its CPU replay; the random start
    equals the int8 layout's."""
    from marqov_amd import gpu
    L = 64
    lat = gpu.RegularLattice(ctx, 3, L)
    st = gpu.State(ctx, "ising", [lat, lat], J=J, spin_bits=1)
    st.set_beta(betas)
    st.init(seed=6)
    for r in range(2):
        assert np.array_equal(st.download(r), ob.init_ising(L ** 3, 6, r))
    st.metropolis(2, seed=41, sweep0=5)
    for r in range(2):
        exp = ob.ising_msc_sweeps(L, ob.init_ising(L ** 3, 6, r), betas[r], J, 2, 41, 5, r)
        got = st.download(r)
        assert np.array_equal(got, exp), (r, int((got != exp).sum()))
    obs = st.measure()
    for r in range(2):
        s = st.download(r)
        e = J * ob.lib().orc_bond_sum_ising(3, L, ob.p(s), None) / L ** 3
        assert obs[r, 0] == pytest.approx(e, abs=1e-12)
        assert obs[r, 1] == pytest.approx(abs(int(s.astype(np.int64).sum())) / L ** 3, abs=1e-12)


def test_msc_upload_roundtrip(ctx):
    from marqov_amd import gpu
    L = 64
    lat = gpu.RegularLattice(ctx, 3, L)
    st = gpu.State(ctx, "ising", [lat], J=-1.0, spin_bits=1)
    st.set_beta(0.2)
    s = np.where(np.random.default_rng(1).random(L ** 3) < 0.3, 1, -1).astype(np.int8)
    st.upload(0, s)
    assert np.array_equal(st.download(0), s)


@pytest.mark.parametrize("kernels", ["0", "1"])
@pytest.mark.parametrize("model,dim,L", [("ising", 2, 64), ("ising", 3, 16), ("potts", 2, 32)])
def test_wolff_lds_and_streaming_bit_exact(ctx, kernels, model, dim, L, monkeypatch):
    """Small lattices grow clusters in LDS (kernels=0) or in global memory (kernels=1):
    both must equal the CPU replay."""
    from marqov_amd import gpu
    monkeypatch.setenv("MARQOV_AMD_KERNELS", kernels)
    N = L ** dim
    lat = gpu.RegularLattice(ctx, dim, L)
    st = gpu.State(ctx, model, [lat] * 3, J=-1.0, q=3)
    beta = {2: 0.4406868, 3: 0.2216546}[dim] if model == "ising" else 1.0
    st.set_beta(beta)
    st.init(seed=30)
    sizes = st.wolff(7, seed=31, cluster0=3)
    for r in range(3):
        if model == "ising":
            exp, esz = ob.wolff_ising(dim, L, ob.init_ising(N, 30, r), beta, -1.0, 7, 31, 3, r)
        else:
            exp, esz = ob.wolff_potts(dim, L, ob.init_potts(N, 3, 30, r), 3, beta, -1.0, 7, 31, 3, r)
        assert np.array_equal(sizes[r], esz), (r, sizes[r], esz)
        assert np.array_equal(st.download(r), exp), r


@pytest.mark.parametrize("model", ["ising", "potts"])
@pytest.mark.parametrize("kernels,ring,dim,L,block", [("0", "64", 2, 32, "0"), ("0", "64", 3, 16, "0"),
                                                      ("1", "2048", 3, 64, "256"), ("1", "2048", 3, 64, "1024")])
def test_wolff_frontier_spill_bit_exact(ctx, model, kernels, ring, dim, L, block, monkeypatch):
    """Ordered starts at large beta grow near-percolating clusters whose BFS fronts
    overflow the LDS frontier ring (64 entries forced in the small-lattice kernel, 4096 in
    the streaming kernel at 3D L=64, with 256- and 1024-thread workgroups): entries spilled
    to the global queue must leave the clusters equal to the CPU replay."""
    from marqov_amd import gpu
    monkeypatch.setenv("MARQOV_AMD_KERNELS", kernels)
    N = L ** dim
    lat = gpu.RegularLattice(ctx, dim, L)
    st = gpu.State(ctx, model, [lat] * 2, J=-1.0, q=3)
    st.set_tuning(wolff_block=int(block))
    st.set_tuning(wolff_ring=int(ring))
    beta = 1.0 if model == "ising" else 2.0  # bond probability 1 - e^-2 for both
    st.set_beta(beta)
    starts = []
    for r in range(2):
        s0 = (np.ones(N, np.int8) if model == "ising" else np.zeros(N, np.int8))
        s0[np.random.default_rng(40 + r).random(N) < 0.1] = -1 if model == "ising" else 2  # 10% defects
        st.upload(r, s0)
        starts.append(s0)
    sizes = st.wolff(3, seed=17, cluster0=1)
    for r in range(2):
        if model == "ising":
            exp, esz = ob.wolff_ising(dim, L, starts[r], beta, -1.0, 3, 17, 1, r)
        else:
            exp, esz = ob.wolff_potts(dim, L, starts[r], 3, beta, -1.0, 3, 17, 1, r)
        assert esz.max() > N // 2  # the fronts are far wider than the ring
        assert np.array_equal(sizes[r], esz), (r, sizes[r], esz)
        assert np.array_equal(st.download(r), exp), r


@pytest.mark.parametrize("model,dim,L", [("ising", 2, 64), ("ising", 3, 16), ("potts", 2, 32)])
def test_wolff_wave_chained_calls_bit_exact(ctx, model, dim, L):
    """Ordered starts with 10% defects at bond probability 1 - e^-1 grow large clusters
    in the small-lattice kernel; three calls of four clusters each must match the CPU
    replay cluster by cluster and spin by spin (state carried across launches)."""
    from marqov_amd import gpu
    N = L ** dim
    lat = gpu.RegularLattice(ctx, dim, L)
    st = gpu.State(ctx, model, [lat] * 3, J=-1.0, q=3)
    beta = 0.5 if model == "ising" else 1.0
    st.set_beta(beta)
    cur = []
    for r in range(3):
        s0 = (np.ones(N, np.int8) if model == "ising" else np.zeros(N, np.int8))
        s0[np.random.default_rng(70 + r).random(N) < 0.1] = -1 if model == "ising" else 2
        st.upload(r, s0)
        cur.append(s0)
    for call in range(3):
        sizes = st.wolff(4, seed=19, cluster0=4 * call)
        for r in range(3):
            if model == "ising":
                cur[r], esz = ob.wolff_ising(dim, L, cur[r], beta, -1.0, 4, 19, 4 * call, r)
            else:
                cur[r], esz = ob.wolff_potts(dim, L, cur[r], 3, beta, -1.0, 4, 19, 4 * call, r)
            assert np.array_equal(sizes[r], esz), (call, r, sizes[r], esz)
            assert np.array_equal(st.download(r), cur[r]), (call, r)
        if call == 0:
            assert sizes.max() * 8 > N


@pytest.mark.parametrize("model", ["ising", "potts"])
@pytest.mark.parametrize("vset", ["0", "16", "4096"])
@pytest.mark.parametrize("dim,L", [(2, 256), (3, 32)])
def test_wolff_membership_modes_bit_exact(ctx, model, vset, dim, L, monkeypatch):
    """The large-lattice Wolff kernel keeps a cluster's membership in an LDS set of packed
    sites while the cluster cannot outgrow the tuning's wolff_vset sites (the alignment test is
    then a plain load and the flip a byte store) and in the spin bytes after that (one global
    atomic per claim).  With the set off (0), left early by nearly every cluster (16) and at
    its largest (4096), three chained calls of four clusters each equal the CPU replay on
    ordered starts with 10% defects above the bond-percolation threshold (cluster sizes from a
    few sites to most of the lattice)."""
    from marqov_amd import gpu
    monkeypatch.setenv("MARQOV_AMD_KERNELS", "1")
    N = L ** dim
    lat = gpu.RegularLattice(ctx, dim, L)
    st = gpu.State(ctx, model, [lat] * 3, J=-1.0, q=3)
    st.set_tuning(wolff_vset=int(vset))
    # bond probability 1 - e^-0.92 = 0.60 (2D) and 1 - e^-0.36 = 0.30 (3D), above the thresholds
    # 0.5 and 0.25: seeds on the majority grow clusters of much of the lattice, seeds on defects
    # stay small
    beta = {2: 0.46, 3: 0.18}[dim] * (1 if model == "ising" else 2)
    st.set_beta(beta)
    cur = []
    for r in range(3):
        s0 = (np.ones(N, np.int8) if model == "ising" else np.zeros(N, np.int8))
        s0[np.random.default_rng(80 + r).random(N) < 0.1] = -1 if model == "ising" else 2
        st.upload(r, s0)
        cur.append(s0)
    allsz = []
    for call in range(3):
        sizes = st.wolff(4, seed=23, cluster0=4 * call)
        allsz.append(sizes)
        for r in range(3):
            if model == "ising":
                cur[r], esz = ob.wolff_ising(dim, L, cur[r], beta, -1.0, 4, 23, 4 * call, r)
            else:
                cur[r], esz = ob.wolff_potts(dim, L, cur[r], 3, beta, -1.0, 4, 23, 4 * call, r)
            assert np.array_equal(sizes[r], esz), (call, r, sizes[r], esz)
            assert np.array_equal(st.download(r), cur[r]), (call, r)
    allsz = np.concatenate(allsz, axis=1)
    print(f"[wolff membership] {model} {dim}D L={L} vset={vset}: sizes {np.sort(allsz.ravel()).tolist()}", flush=True)
    assert allsz.max() > 4096 and allsz.min() < 4096
