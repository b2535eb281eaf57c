"""FP32 O(3) / Phi4 Metropolis sweeps against an FP64 restatement of the same
schedule (oracle/replay.c, orc_o3_halfsweep: same Philox counters, proposals,
energy difference and acceptance rule).  One sweep from a downloaded state is
replayed colour by colour, the second colour from the GPU's own first-colour
result, so FP32 rounding cannot cascade.  Decisions must agree wherever the
acceptance margin |p - u| exceeds 1e-4 (FP32 energy and exp errors are orders of
magnitude smaller); accepted spins must equal the FP64 proposal to 2e-5 (FP32
rounding of the direction and normalisation) and rejected spins must not move."""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu

TOL_MARGIN = 1e-4


@pytest.fixture(scope="module")
def ctx():
    from marqov_amd import gpu
    return gpu.Context(0)


def _colour(dim, L):
    idx = np.arange(L ** dim)
    return sum((idx // L ** d) % L for d in range(dim)) & 1


@pytest.mark.parametrize("kernels", ["0", "1"])
@pytest.mark.parametrize("model,proposal,dim,L,beta", [
    ("heisenberg", "fresh", 3, 6, 0.693), ("heisenberg", "small", 3, 8, 0.693), ("heisenberg", "small", 2, 8, 1.2),
    ("phi4", "fresh", 3, 6, 0.69), ("phi4", "fresh", 2, 10, 0.9)])
def test_o3_sweep_decisions_match_fp64_restatement(ctx, model, proposal, dim, L, beta, kernels, monkeypatch):
    from marqov_amd import gpu
    monkeypatch.setenv("MARQOV_AMD_KERNELS", kernels)
    lat = gpu.RegularLattice(ctx, dim, L)
    nrep, seed, sweep0 = 3, 9, 3
    st = gpu.State(ctx, model, [lat] * nrep, J=-1.0, proposal=proposal, delta=0.6, lam=4.5, mass=1.0)
    st.set_beta(beta)
    st.init(seed=5)
    st.metropolis(sweep0, seed=seed, sweep0=0)
    s0 = [st.download(r).astype(np.float64) for r in range(nrep)]
    st.metropolis(1, seed=seed, sweep0=sweep0)
    s1 = [st.download(r).astype(np.float64) for r in range(nrep)]
    col = _colour(dim, L)
    kw = dict(phi4=(model == "phi4"), small_angle=(proposal == "small"), delta=0.6, lam=4.5, mass=1.0)
    n_checked = n_acc = n_near = 0
    for r in range(nrep):
        old = s0[r]
        for c in (0, 1):
            cpu, m = ob.o3_halfsweep(dim, L, old, c, 2 * sweep0 + c, r, seed, beta, J=-1.0, **kw)
            sites = col == c
            moved = np.any(s1[r] != old, axis=1)
            sure = sites & (np.abs(m) > TOL_MARGIN)
            n_near += int((sites & ~sure).sum())
            assert np.array_equal(moved[sure], (m > 0)[sure]), (r, c)
            acc = sure & (m > 0)
            assert np.abs(s1[r][acc] - cpu[acc]).max(initial=0.0) < 2e-5, (r, c)
            n_checked += int(sure.sum())
            n_acc += int(acc.sum())
            # the next colour sees the GPU's result for this one
            old = np.where(sites[:, None], s1[r], old)
    assert n_near <= 3, n_near
    assert 0.1 * n_checked < n_acc < 0.95 * n_checked, (n_acc, n_checked)
