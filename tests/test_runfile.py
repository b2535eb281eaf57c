"""HDF5 run files of the .ini front-end in MARQOV's layout (core.h:650-768 createstep,
core.h:841-896 dumprng / dumpstatespace, cachecontainer.h:166-205 scalars and strings,
cachecontainer.h:238-458 time series, doc/io.rst:33-47).  tests/cpp/runfile_check writes
two steps through the driver's writer (marqov_amd/driver/runfile.hpp) from synthetic
values and reads the last one back with the driver's reader; the file itself is checked
with the HDF5 project's h5ls and h5dump: every group and dataset the reference creates,
their types, extents, chunking and filters as the reference sets them, and the values."""
import json
import os
import subprocess

import numpy as np
import pytest

import h5tools

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHK = os.path.join(ROOT, "tests", "cpp", "runfile_check")

pytestmark = pytest.mark.skipif(not (os.path.exists(CHK) and h5tools.available()),
                                reason="runfile_check or the HDF5 tools absent (make -C tests/cpp)")

# step contents created by Core::createstep and its helpers, per group
STEP_DATASETS = {
    "environment": ["Code", "Version", "hostname", "startingdate"],  # marqov.cpp:30-52
    "config/marqovconfig": ["beta", "id", "repid", "seed", "nsteps", "warmupsteps", "gameloopsteps", "ncluster",
                            "nmetro"],  # core.h:756-757, 180-191
    "config/lattice": ["name", "dim", "npoints"],  # regular_hypercubic.h:151-156
    "config/hamiltonian": ["Model", "param0"],  # core.h:696-711
    "state": ["RNG", "rngstate", "hamiltonianstatespace"],  # core.h:841-896
}
OBSERVABLES = {"ising": ["m", "e", "dummy"], "heisenberg": ["m"]}
N, NMEAS = 16, 6


def _chunk(record_bytes, vector_width=None):
    """Chunk length of a series: 4 MiB of records (cachecontainer.h:262, core.h:887);
    vector-valued series 4 MiB / record + 1 (cachecontainer.h:403)."""
    if vector_width:
        return 4096 * 1024 // 8 // vector_width + 1
    return 4096 * 1024 // record_bytes


@pytest.fixture(params=["ising", "heisenberg"])
def run(request, tmp_path):
    path = tmp_path / f"{request.param}.h5"
    r = subprocess.run([CHK, str(path), request.param], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return request.param, path, json.loads(r.stdout)


def test_groups_and_datasets(run):
    kind, path, _ = run
    t = h5tools.tree(path)
    want = {}
    for s in (0, 1):
        st = f"/step{s}"
        want[st] = "Group"
        for g in ("environment", "config", "config/marqovconfig", "config/lattice", "config/hamiltonian", "state",
                  "observables"):
            want[f"{st}/{g}"] = "Group"
        for g, names in STEP_DATASETS.items():
            for n in names:
                want[f"{st}/{g}/{n}"] = "Dataset"
        for n in OBSERVABLES[kind]:
            want[f"{st}/observables/{n}"] = "Dataset"
    assert {k: v.split()[0] for k, v in t.items()} == want


def test_types_extents_chunks_filters(run):
    kind, path, _ = run
    h = h5tools.header
    assert "H5T_IEEE_F64LE" in h(path, "/step0/config/marqovconfig/beta")
    for n in ("id", "repid", "nsteps", "warmupsteps", "gameloopsteps", "ncluster", "nmetro"):
        hd = h(path, f"/step1/config/marqovconfig/{n}")
        assert "H5T_STD_I32LE" in hd and "DATASPACE  SCALAR" in hd, n
    assert "H5T_STD_I32LE" in h(path, "/step0/config/lattice/dim")
    assert "H5T_IEEE_F64LE" in h(path, "/step0/config/hamiltonian/param0")
    # strings: fixed length = the value's length, NUL-terminated C strings (dumpscalartoH5)
    for ds, val in (("/step1/config/hamiltonian/Model", "Ising" if kind == "ising" else "Heisenberg"),
                    ("/step1/config/lattice/name", "RegularHypercubic"), ("/step1/state/RNG", "Philox4x32-10")):
        hd = h(path, ds)
        assert f"STRSIZE {len(val)};" in hd and "H5T_STR_NULLTERM" in hd and "H5T_C_S1" in hd, ds
        assert h5tools.string(path, ds) == val
    # rngstate: a fixed-size uint64 vector (dumprng)
    hd = h(path, "/step1/state/rngstate")
    assert "H5T_STD_U64LE" in hd and "( 3 ) / ( 3 )" in hd
    # the state space: one fixed-size array per site, extendible, shuffle + deflate 9
    hd = h(path, "/step1/state/hamiltonianstatespace")
    rec = "H5T_ARRAY { [1] H5T_STD_I32LE }" if kind == "ising" else "H5T_ARRAY { [3] H5T_IEEE_F64LE }"
    assert rec in hd
    assert f"( {N} ) / ( H5S_UNLIMITED )" in hd
    assert f"CHUNKED ( {_chunk(4 if kind == 'ising' else 24)} )" in hd
    assert "PREPROCESSING SHUFFLE" in hd and "COMPRESSION DEFLATE { LEVEL 9 }" in hd
    # time series: extendible, 4 MiB chunks, deflate 9, no shuffle; the description as comment
    hd = h(path, "/step0/observables/m")
    assert "H5T_IEEE_F64LE" in hd and f"( {NMEAS} ) / ( H5S_UNLIMITED )" in hd
    assert f"CHUNKED ( {_chunk(8)} )" in hd and "COMPRESSION DEFLATE { LEVEL 9 }" in hd
    assert "SHUFFLE" not in hd and 'COMMENT "magnetization"' in hd
    if kind == "ising":
        hd = h(path, "/step0/observables/dummy")
        assert "H5T_ARRAY { [5] H5T_IEEE_F64LE }" in hd and f"CHUNKED ( {_chunk(8, 5)} )" in hd
        assert "COMMENT" not in h(path, "/step0/observables/e")


def test_values_and_last_step_reader(run):
    kind, path, got = run
    v = h5tools.values
    t = np.arange(NMEAS)
    for s in (0, 1):
        assert np.array_equal(v(path, f"/step{s}/observables/m"), 0.125 * (t + 1) + s + 1.0 / 3.0)
        if kind == "ising":
            assert np.array_equal(v(path, f"/step{s}/observables/e"), -1.5 - 0.01 * t - s)
            assert np.array_equal(v(path, f"/step{s}/observables/dummy"), np.tile(42 + 0.1 * np.arange(5), NMEAS))
        assert v(path, f"/step{s}/config/marqovconfig/warmupsteps")[0] == (0 if s else 20)
        assert v(path, f"/step{s}/config/marqovconfig/seed", integer=True) == [0xFEEDFACECAFE]
    i = np.arange(N)
    if kind == "ising":
        want = np.where((i * 7 + 1) % 3 != 0, 1, -1).astype(float)
    else:
        want = np.stack([np.full(N, 0.6), -0.8 + 1e-3 * i, np.full(N, 1.0 / 7.0)], axis=1).ravel()
    spins = v(path, "/step1/state/hamiltonianstatespace")
    assert np.array_equal(spins, want)
    # the driver's reader sees the last step as h5dump does
    assert got["step"] == 1 and got["rng"] == "Philox4x32-10"
    assert got["rngstate"] == [0xFEEDFACECAFE, 0x0123456789ABCDEF, 32]
    assert v(path, "/step1/state/rngstate", integer=True) == [0xFEEDFACECAFE, 0x0123456789ABCDEF, 32]
    assert got["model"] == ("Ising" if kind == "ising" else "Heisenberg")
    assert got["lattice"] == "RegularHypercubic" and got["npoints"] == N
    assert got["beta"] == 0.4406868 and got["repid"] == 2 and got["seed"] == 0xFEEDFACECAFE
    assert np.array_equal(np.array(got["spins"]), want)


def test_reference_cache_container_series(tmp_path):
    """The time series of the reference's own cache-container tests
    (testsuite/components/1-3: 2,000,000 doubles and ints; 500,000 records of a 7-vector;
    5 records of a 600,000-vector, longer than the 4 MiB cache) through the driver's series
    writer: extents, chunks (4 MiB of scalars; 4 MiB / 8 / width + 1 records of vectors),
    fill values (none for records of 8,192 values or more, cachecontainer.h:406-410),
    descriptions and values."""
    path = tmp_path / "series.h5"
    r = subprocess.run([CHK, str(path), "series"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    want = {
        "/obs1/testobs": ("H5T_IEEE_F64LE", 2000000, 4096 * 1024 // 8, "VALUE  0", "a test description"),
        "/obs2/testint": ("H5T_STD_I32LE", 2000000, 4096 * 1024 // 4, "VALUE  0", "a test int description"),
        "/obs3/name": ("H5T_ARRAY { [7] H5T_IEEE_F64LE }", 500000, 4096 * 1024 // 8 // 7 + 1,
                       "VALUE  [ 0, 0, 0, 0, 0, 0, 0 ]", "desc"),
        "/obs4/name": ("H5T_ARRAY { [600000] H5T_IEEE_F64LE }", 5, 4096 * 1024 // 8 // 600000 + 1,
                       "H5D_FILL_VALUE_DEFAULT", "desc"),
    }
    for ds, (ty, n, chunk, fill, desc) in want.items():
        hd = h5tools.header(path, ds)
        assert ty in hd and f"( {n} ) / ( H5S_UNLIMITED )" in hd and f"CHUNKED ( {chunk} )" in hd, (ds, hd)
        assert fill in hd and f'COMMENT "{desc}"' in hd and "COMPRESSION DEFLATE { LEVEL 9 }" in hd, (ds, hd)
    v = h5tools.values
    assert np.array_equal(v(path, "/obs1/testobs", start=1999990, count=10), np.full(10, 42.0))
    assert v(path, "/obs2/testint", integer=True, start=0, count=4) == [42] * 4
    assert np.array_equal(v(path, "/obs3/name", start=499999, count=1), np.arange(1, 8, dtype=float))
    assert np.array_equal(v(path, "/obs4/name", start=4, count=1), np.full(600000, 2.0))
