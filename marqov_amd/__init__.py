"""MI355X-native MARQOV update path: Metropolis sweeps and Wolff clusters in HIP.

The compute path is libmarqov_amd.so (HIP kernels for gfx950 behind the C ABI
of include/marqov_amd.h); `marqov_amd.gpu` is its Python host binding.
"""
from . import _lib  # noqa: F401
