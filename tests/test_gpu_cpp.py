"""The C++ host API (include/marqov_amd.hpp) drives the 2D Ising L=64 EMCS schedule:
2D Ising L=64 at beta_c, 10 Wolff + 1 sweep per EMCS, energy vs Kaufman's exact value."""
import json
import os
import subprocess

import pytest

import exact_path  # noqa: F401
import exact

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "emcs_c1")


def test_cpp_emcs_c1_vs_kaufman():
    out = subprocess.run([EXE, "128", "200", "300"], capture_output=True, text=True, check=True, timeout=100)
    res = json.loads(out.stdout.strip().splitlines()[-1])
    ka = exact.kaufman_energy(0.4406868, 64, 64)
    assert abs(res["e"] - ka) < 4 * res["se"], (res, ka)
    assert 0.55 < res["binder"] < 0.66, res
    assert 1000 < res["mean_cluster"] < 2200, res
