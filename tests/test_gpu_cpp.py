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


DROP_IN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "drop_in_emcs")


def test_drop_in_emcs_specialisation():
    """The EMCS<Ising<int>, GpuRegular<dim>> specialisation of the integration notes, compiled
    and driven by a Core-like loop (50 warm-up EMCS, 20 game-loop EMCS with a host-side
    measurement each): temperature tables are uploaded once (beta never changes), the state
    comes back only at the 20 measurement points, and the final state equals the same EMCS
    run through Batch/GpuEMCS directly."""
    out = subprocess.run([DROP_IN, "50", "20"], capture_output=True, text=True, check=True, timeout=100)
    res = json.loads(out.stdout.strip().splitlines()[-1])
    assert res["equal"] is True, res
    assert res["downloads"] == 20 and res["beta_uploads"] == 1, res
    assert -1.6 < res["e_mean"] < -1.2, res
