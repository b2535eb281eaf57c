"""CPU restatement of the reference's MagFTComp observable (test infrastructure).

Only tests/, the smoke check and the bench's cpu_baseline leg may import this.

magft(spins, dim, L, axis) follows src/hamiltonian/util/observables.h:104-170
line for line: for every site i (natural order, x0 fastest) the coordinate
x = grid.crds(i)[axis] = (index_axis + 1/2) / L (regular_hypercubic.h:73-85),
the component sums sum_i s_i[j] * exp(2 pi i x) are divided by N, and the
result is the sum over components of their squared moduli.  Pinned in
tests/test_oracle.py against numpy's FFT (an independent algorithm: the
k = 1 coefficient of the axis profile, whose modulus the 1/2 shift does not
change) and against the closed form of a half-and-half stripe,
(2 / (L sin(pi / L)))^2.
"""
import numpy as np


def magft(spins, dim, L, axis):
    s = np.asarray(spins, np.float64).reshape(L ** dim, -1)
    N = s.shape[0]
    idx = np.arange(N)
    index = (idx // L ** axis) % L
    x = (index + 0.5) / L
    expi = np.cos(2.0 * np.pi * x) + 1j * np.sin(2.0 * np.pi * x)
    comp = (s * expi[:, None]).sum(axis=0) / N
    return float(np.sum(np.abs(comp) ** 2))
