"""Statistics shared by the statistical parity tests.

Every chain (a GPU replica or a CPU seed) is an independent Markov chain, so the
error of an equilibrium average is the spread of the per-chain time averages
over chains (autocorrelation inside a chain then needs no separate treatment).
The Binder cumulant U = 1 - <m^4> / (3 <m^2>^2) is a ratio of averages; its
error is the jackknife over chains.

The stated criterion is that the GPU path and the CPU build agree within 2
standard errors.  Asserting |z| < 2 on each of many observables would fail a
correct implementation about once in every 20 comparisons, so the tests assert
it in the pre-registered combined form: sum z^2 over all compared observables
must lie below the 95% quantile of chi^2 with that many degrees of freedom
(the multivariate statement of "within 2 sigma"), and every single |z| must be
below 4.
"""
import numpy as np


def per_chain(obs):
    """obs[chain, t] -> per-chain time averages."""
    return np.asarray(obs, np.float64).mean(axis=1)


def mean_err(x):
    x = np.asarray(x, np.float64)
    return float(x.mean()), float(x.std(ddof=1) / np.sqrt(len(x)))


def binder(m2, m4):
    """Binder cumulant from per-chain averages of m^2 and m^4, with its jackknife error."""
    m2 = np.asarray(m2, np.float64)
    m4 = np.asarray(m4, np.float64)
    n = len(m2)
    u = 1.0 - m4.mean() / (3.0 * m2.mean() ** 2)
    jk = np.array([1.0 - np.delete(m4, i).mean() / (3.0 * np.delete(m2, i).mean() ** 2) for i in range(n)])
    err = np.sqrt((n - 1) / n * ((jk - jk.mean()) ** 2).sum())
    return float(u), float(err)


def compare_chains(gpu, cpu, label, cols=(0, 1), names=("E/N", "|m|"), with_binder=True):
    """gpu[chain, t, k] and cpu[chain, t, k] with k = E/N, |m|, m^2, m^4 (any further columns
    are named in cols/names).  Returns rows (label, observable, gpu, gpu err, cpu, cpu err, z)."""
    rows = []
    for k, nm in zip(cols, names):
        g, sg = mean_err(per_chain(gpu[:, :, k]))
        c, sc = mean_err(per_chain(cpu[:, :, k]))
        rows.append((label, nm, g, sg, c, sc, (g - c) / np.hypot(sg, sc)))
    if with_binder:
        g, sg = binder(per_chain(gpu[:, :, 2]), per_chain(gpu[:, :, 3]))
        c, sc = binder(per_chain(cpu[:, :, 2]), per_chain(cpu[:, :, 3]))
        rows.append((label, "U", g, sg, c, sc, (g - c) / np.hypot(sg, sc)))
    return rows


def report(rows, tag):
    for lab, nm, g, sg, c, sc, z in rows:
        print(f"[{tag}] {lab} {nm}: gpu {g:.6f} +- {sg:.6f}  cpu {c:.6f} +- {sc:.6f}  z = {z:+.2f}", flush=True)


def chi2_quantile95(dof):
    from scipy.stats import chi2
    return float(chi2.ppf(0.95, dof))


def assert_within_2sigma_combined(rows, tag):
    """sum z^2 below the 95% chi^2 quantile, and every |z| < 4."""
    report(rows, tag)
    z = np.array([r[-1] for r in rows])
    q = chi2_quantile95(len(z))
    s = float((z ** 2).sum())
    print(f"[{tag}] chi2 = {s:.2f} over {len(z)} observables (95% quantile {q:.2f}); "
          f"{int((np.abs(z) < 2).sum())} of {len(z)} within 2 sigma, max |z| = {np.abs(z).max():.2f}", flush=True)
    assert np.all(np.abs(z) < 4.0), [(r[0], r[1], round(r[-1], 2)) for r in rows]
    assert s < q, (s, q, [(r[0], r[1], round(r[-1], 2)) for r in rows])
