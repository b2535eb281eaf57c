"""The .ini front-end (marqov_amd/bin/marqov_gpu): configuration parsing and run
plans on the CPU (--dry-run touches no GPU), its graph builders against brute
force, and full runs on the GPU checked against exact enumeration."""
import json
import os
import subprocess

import numpy as np
import pytest

import h5tools

import exact_path  # noqa: F401
import exact

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "marqov_amd", "bin", "marqov_gpu")
CONF = os.path.join(ROOT, "tests", "data", "config")
LCHK = os.path.join(ROOT, "tests", "cpp", "lattice_check")


def plan(ham, conf=CONF):
    r = subprocess.run([BIN, ham, "--config", conf, "--dry-run"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return [json.loads(line) for line in r.stdout.splitlines() if line.strip()]


def test_plan_ising():
    p = plan("Ising")
    assert [x["L"] for x in p] == [4, 3]
    assert [x["lattice"] for x in p] == ["regular", "csr-hypercube"]  # odd L is not bipartite
    for x in p:
        # ncluster = int(2 * L^0); 10 * (205 // 10) warm-up and 10 * (4009 // 10) measured EMCS
        assert (x["ncluster"], x["nmetro"], x["warmup"], x["measure"]) == (2, 1, 200, 4000)
        assert x["out"] == f"out_ising/{x['L']}/"
        assert len(x["groups"]) == 1
    sims = p[0]["groups"][0] + p[1]["groups"][0]
    assert [s["id"] for s in sims] == list(range(48))
    # parameter sets first (beta fastest), replicas inner
    first = p[0]["groups"][0]
    assert [s["beta"] for s in first[:9]] == [0.3] * 8 + [0.4406868]
    assert [s["repid"] for s in first[:9]] == list(range(8)) + [0]
    assert first[8]["name"] == "beta0.440687_0"


def test_plan_potts_select_and_cart_prod():
    p = plan("Potts")  # q from PottsSelect.ini
    assert len(p) == 1 and p[0]["model"] == "Potts3" and p[0]["q"] == 3
    assert p[0]["ncluster"] == 2  # int(0.5 * 4^1)
    g0, g1 = p[0]["groups"]  # one state per J; beta varies fastest inside
    assert [s["J"] for s in g0] == [-1.0] * 8 and [s["J"] for s in g1] == [-0.5] * 8
    assert [s["beta"] for s in g0] == [0.8] * 4 + [1.1] * 4
    assert [s["id"] for s in g0 + g1] == list(range(16))
    assert g1[0]["name"] == "beta0.800000_0_g1"
    assert plan("Potts3")[0]["groups"] == p[0]["groups"]


def test_plan_phi4_and_rep_per_L():
    p = plan("Phi4")
    assert [x["L"] for x in p] == [6, 8]
    assert [x["ncluster"] for x in p] == [3, 4]
    # (beta, lambda, mass): beta fastest; one group per (lambda, mass); rep = 2 for L=6, 3 for L=8
    assert [len(x["groups"]) for x in p] == [2, 2]
    assert [len(g) for g in p[0]["groups"]] == [4, 4] and [len(g) for g in p[1]["groups"]] == [6, 6]
    assert {s["lambda"] for s in p[0]["groups"][0]} == {4.5} and {s["lambda"] for s in p[0]["groups"][1]} == {3.0}
    assert all(s["J"] == -1.0 for g in p[0]["groups"] for s in g)


def test_plan_ea_and_rgg():
    p = plan("EdwardsAnderson-Ising")
    assert [x["lattice"] for x in p] == ["csr-hypercube", "regular"]
    assert all(x["ncluster"] == 0 for x in p)
    p = plan("IsingRGG")
    assert p[0]["lattice"] == "rgg" and p[0]["ncluster"] == 8


def test_registry_errors():
    r = subprocess.run([BIN, "Ising", "--config", os.path.join(ROOT, "tests", "data", "config_noend"), "--dry-run"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "File not recognized as Config File" in r.stderr
    r = subprocess.run([BIN, "XXZAntiferro", "--config", CONF, "--dry-run"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "no GPU update path" in r.stderr
    r = subprocess.run([BIN, "Heisenberg", "--config", "/nonexistent", "--dry-run"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "Configuration Directory not found" in r.stderr


def _read_csr(lines, off):
    rp = np.array(lines[off].split(), np.int64)
    ci = np.array(lines[off + 1].split(), np.int64) if off + 1 < len(lines) else np.zeros(0, np.int64)
    return rp, ci


@pytest.mark.parametrize("dim,L", [(2, 3), (3, 5), (4, 3)])
def test_hypercube_csr_reference_order(dim, L):
    out = subprocess.run([LCHK, "cube", str(dim), str(L)], capture_output=True, text=True, check=True).stdout
    lines = out.splitlines()
    rp, ci = _read_csr(lines, 1)
    N = L ** dim
    assert list(rp) == list(range(0, 2 * dim * N + 1, 2 * dim))
    for i in range(N):
        x = [(i // L ** d) % L for d in range(dim)]
        want = []
        for d in range(dim):  # +x_d then -x_d, as RegularHypercubic
            for s in (1, -1):
                y = list(x)
                y[d] = (y[d] + s) % L
                want.append(sum(y[k] * L ** k for k in range(dim)))
        assert list(ci[rp[i]:rp[i + 1]]) == want, i


def test_csv_graph_reader_matches_python(tmp_path):
    """GraphFromCSV format: the C++ reader of the front-end and the Python reader agree,
    with 1-based neighbour indices shifted back by idxoffs."""
    from marqov_amd import graphs
    rp, ci, pts = graphs.rgg2d(500, graphs.rgg_radius(500), 9)
    path = str(tmp_path / "g.csv")
    graphs.write_csv_graph(path, rp, ci, pts)
    out = subprocess.run([LCHK, "csv", path], capture_output=True, text=True, check=True).stdout.splitlines()
    rp2, ci2 = _read_csr(out, 1)
    assert np.array_equal(rp2, rp) and np.array_equal(ci2, ci)
    path1 = str(tmp_path / "g1.csv")
    graphs.write_csv_graph(path1, rp, ci + 1, pts)
    out = subprocess.run([LCHK, "csv", path1, "1"], capture_output=True, text=True, check=True).stdout.splitlines()
    rp3, ci3 = _read_csr(out, 1)
    assert np.array_equal(rp3, rp) and np.array_equal(ci3, ci)
    rp4, ci4, _ = graphs.read_csv_graph(path1, idxoffs=1)
    assert np.array_equal(rp4, rp) and np.array_equal(ci4, ci)


def test_plan_csv_graph(tmp_path):
    from marqov_amd import graphs
    rp, ci = graphs.regular_csr(2, 3)
    path = str(tmp_path / "t33.csv")
    graphs.write_csv_graph(path, rp, ci)
    r = subprocess.run([BIN, "Ising", "--config", CONF, "--graph", path, "--dry-run"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    p = [json.loads(line) for line in r.stdout.splitlines()]
    assert len(p) == 1 and p[0]["lattice"] == "csv" and p[0]["nsites"] == 9 and p[0]["out"] == "out_ising/graph/"
    assert p[0]["ncluster"] == 2 and len(p[0]["groups"][0]) == 24
    r = subprocess.run([BIN, "IsingRGG", "--config", CONF, "--graph", path, "--dry-run"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "--graph replaces" in r.stderr


def _run(ham, out, seed=5):
    r = subprocess.run([BIN, ham, "--config", CONF, "--out", str(out), "--seed", str(seed)], capture_output=True,
                       text=True, timeout=110)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr


def _series(path):
    hdr = [line for line in open(path) if line.startswith("#")]
    data = np.loadtxt(path)
    return hdr, data


# observables of the reference tuples in the run files (Ising.h:82-94, Heisenberg.h:88-90,
# Potts.h:143-144, Phi4.h:133-140); the EA files hold e and m
RUNFILE_OBS = {"Ising": ["m", "e"], "IsingRGG": ["m", "e"], "Heisenberg": ["m"], "Phi4": ["m"], "Potts": ["m"],
               "Potts3": ["m"], "EdwardsAnderson-Ising": ["m", "e"]}


def _check_runfile(h5, data, ham, nft, steps=1):
    """The HDF5 run file of a simulation (read with h5ls / h5dump): steps /step0 ..
    /step<steps-1>, and the observable series of the last step equal, bit for bit, the
    text file's columns for that step's rows (data); returns the last step's spins."""
    if not h5tools.available():
        pytest.skip("HDF5 tools absent")
    t = h5tools.tree(h5)
    assert sorted(k for k in t if k.count("/") == 1) == [f"/step{s}" for s in range(steps)]
    st = f"/step{steps - 1}"
    names = RUNFILE_OBS[ham] + [f"magft{d}" for d in range(nft)] + (["dummy"] if ham in ("Ising", "IsingRGG") else [])
    assert sorted(k.rsplit("/", 1)[1] for k in t if k.startswith(st + "/observables/")) == sorted(names)
    cols = {"e": 1, "m": 2}
    for d in range(nft):
        cols[f"magft{d}"] = 6 + d
    for n in names:
        if n in cols:
            assert np.array_equal(h5tools.values(h5, f"{st}/observables/{n}"), data[:, cols[n]]), (h5, n)
    assert h5tools.string(h5, f"{st}/state/RNG") == "Philox4x32-10"
    return h5tools.values(h5, f"{st}/state/hamiltonianstatespace")


def _hypercube_graph(dim, L):
    out = subprocess.run([LCHK, "cube", str(dim), str(L)], capture_output=True, text=True, check=True).stdout
    rp, ci = _read_csr(out.splitlines(), 1)
    return rp.astype(np.int32), ci.astype(np.int32)


def _exact_magft_4x4(beta):
    """Boltzmann averages of MagFTComp along both axes of the 2D 4x4 Ising torus
    (J = -1) over all 2^16 states, with the reference's site coordinates (x + 1/2)/L."""
    L, N = 4, 16
    nb = exact.neighbour_table(2, L)
    S = (((np.arange(1 << N)[:, None] >> np.arange(N)) & 1) * 2 - 1).astype(np.float64)
    E = -((S * S[:, nb[:, 0]]).sum(1) + (S * S[:, nb[:, 2]]).sum(1))
    w = np.exp(-beta * (E - E.min()))
    i = np.arange(N)
    out = []
    for x in (i % L, i // L):
        ph = np.exp(2j * np.pi * (x + 0.5) / L)
        out.append(float((w * np.abs(S @ ph / N) ** 2).sum() / w.sum()))
    return out


@pytest.mark.gpu
def test_gpu_ising_ini_vs_exact(tmp_path):
    """The whole front-end on the GPU: Ising.ini runs 2D Ising on a 4x4 torus
    (checkerboard kernels) and a 3x3 torus (graph kernels), 8 replicas at each of
    3 temperatures, 2 Wolff clusters + 1 sweep per EMCS; <E/N> and <m^2> of each
    temperature agree with exact enumeration within 4 standard errors over the
    replicas."""
    _run("Ising", tmp_path)
    for L in (4, 3):
        rp, ci = _hypercube_graph(2, L)
        for beta in (0.3, 0.4406868, 0.7):
            e_rep, m2_rep, ft_rep = [], [], []
            for rep in range(8):
                hdr, d = _series(tmp_path / str(L) / f"beta{beta:.6f}_{rep}.txt")
                # the 4x4 torus runs on the checkerboard kernels and carries the MagFTComp columns
                assert d.shape == (4000, 8 if L == 4 else 6), d.shape
                if rep < 2:
                    spins = _check_runfile(tmp_path / str(L) / f"beta{beta:.6f}_{rep}.h5", d, "Ising", 2 if L == 4 else 0)
                    assert spins.shape == (L * L,) and set(np.unique(spins)) <= {-1.0, 1.0}
                assert np.allclose(d[:, 0], np.arange(200, 4200))
                assert np.all(d[:, 5] >= 1.0)
                e_rep.append(d[:, 1].mean())
                m2_rep.append(d[:, 3].mean())
                if L == 4:
                    ft_rep.append(d[:, 6:8].mean(axis=0))
            ex = exact.enumerate_ising(2, 4, beta, J=-1.0) if L == 4 else exact.enumerate_graph(rp, ci, beta, J=-1.0)
            ee, m2e = ex["e"], ex["m2"]
            checks = [(e_rep, ee, "E"), (m2_rep, m2e, "m2")]
            if L == 4:
                fte = _exact_magft_4x4(beta)
                checks += [([f[0] for f in ft_rep], fte[0], "magft0"), ([f[1] for f in ft_rep], fte[1], "magft1")]
            for got, want, what in checks:
                g = np.asarray(got)
                se = g.std(ddof=1) / np.sqrt(len(g))
                print(f"L={L} beta={beta} {what}: {g.mean():.5f} +- {se:.5f} exact {want:.5f}", flush=True)
                assert abs(g.mean() - want) < 4 * se + 1e-9, (L, beta, what, g.mean(), se, want)
    summ = (tmp_path / "4" / "summary.txt").read_text().splitlines()
    assert len(summ) == 1 + 24


@pytest.mark.gpu
@pytest.mark.parametrize("ham,dirs", [("Potts", ["4"]), ("Phi4", ["6", "8"]), ("Heisenberg", ["6"]),
                                      ("EdwardsAnderson-Ising", ["3", "4"]), ("IsingRGG", ["8"])])
def test_gpu_ini_models_run(tmp_path, ham, dirs):
    """Every supported Hamiltonian runs from its .ini file: one series file per
    simulation with the planned number of rows and finite observables, energies
    falling with beta."""
    _run(ham, tmp_path)
    p = plan(ham)
    for x, dname in zip(p, dirs):
        for g, grp in enumerate(x["groups"]):
            for s in grp:
                hdr, d = _series(tmp_path / dname / f"{s['name']}.txt")
                # MagFTComp columns of the reference's Ising / Phi4 observable tuples
                nft = {"Phi4": x["dim"]}.get(ham, 0) if x["lattice"] == "regular" else 0
                assert d.shape == (x["measure"], 6 + nft)
                h5 = tmp_path / dname / f"{s['name']}.h5"
                spins = _check_runfile(h5, d, ham, nft)
                assert spins.size == x["nsites"] * (3 if ham in ("Heisenberg", "Phi4") else 1)
                # lattice writers: RegularHypercubic (name, dim, npoints); the reference's default
                # writer for lattices without their own (name "unknown", npoints; core.h:63-68)
                lname = "unknown" if x["lattice"] == "rgg" else "RegularHypercubic"
                assert h5tools.string(h5, "/step0/config/lattice/name") == lname
                assert h5tools.values(h5, "/step0/config/lattice/npoints", integer=True) == [x["nsites"]]
                assert h5tools.string(h5, "/step0/config/hamiltonian/Model") == {
                    "Potts": "Potts", "Phi4": "Phi4", "Heisenberg": "Heisenberg",
                    "EdwardsAnderson-Ising": "EdwardsAndersonIsing", "IsingRGG": "Ising"}[ham]
                assert np.all(np.isfinite(d[:, 1:5]))
                assert np.all((d[:, 6:] >= 0) & (d[:, 6:] <= 1))
                assert hdr[-1].split()[7:] == [f"magft{k}" for k in range(nft)]
                if x["ncluster"] == 0:
                    assert np.all(np.isnan(d[:, 5]))
        lines = [line.split() for line in (tmp_path / dname / "summary.txt").read_text().splitlines()[1:]]
        by_beta = {}
        for f in lines:
            by_beta.setdefault(float(f[1]), []).append(float(f[6]))
        betas = sorted(by_beta)
        means = [np.mean(by_beta[b]) for b in betas]
        if ham != "Phi4":  # Phi4's on-site terms scale with 1/beta
            assert all(a >= b - 0.05 for a, b in zip(means, means[1:])), means


def _rows(path):
    """Data rows of a series file as text (bit-exact comparison), and the block headers."""
    rows, blocks = [], []
    for line in open(path):
        if line.startswith("# block"):
            blocks.append(line)
        elif not line.startswith("#"):
            rows.append(line)
    return rows, blocks


@pytest.mark.gpu
@pytest.mark.parametrize("ham,Ls", [("Ising", ["8", "5"]), ("Heisenberg", ["6"]), ("Phi4", ["6"]), ("Potts3", ["4"]),
                                    ("EdwardsAnderson-Ising", ["4"])])
def test_gpu_ini_resume_is_exact(tmp_path, ham, Ls):
    """Resume from the HDF5 run files: a run of 2 x 200 measured EMCS, stopped after the
    first 200 and continued from the last step of every simulation's run file (no second
    warm-up, the game loop recorded as /step1 and appended to the text file as the next
    block), writes exactly the rows of one uninterrupted run of 400 measured EMCS, and
    ends in the same spins: the Philox counters are keyed by the EMCS number, which the
    run file's state stores (core.h:543-607)."""
    once, twice = tmp_path / "once", tmp_path / "twice"
    for d, out in (("config_resume_once", once), ("config_resume_twice", twice)):
        conf = os.path.join(ROOT, "tests", "data", d)
        runs = 1 if d.endswith("once") else 2
        for _ in range(runs):
            r = subprocess.run([BIN, ham, "--config", conf, "--out", str(out), "--seed", "11"], capture_output=True,
                               text=True, timeout=110)
            print(r.stdout[-800:], r.stderr[-800:], flush=True)
            assert r.returncode == 0, r.stderr
    p = plan(ham, os.path.join(ROOT, "tests", "data", "config_resume_once"))
    nfiles = 0
    for x, L in zip(p, Ls):
        for grp in x["groups"]:
            for s in grp:
                a, ablocks = _rows(once / L / f"{s['name']}.txt")
                b, bblocks = _rows(twice / L / f"{s['name']}.txt")
                assert len(a) == 400 and not ablocks
                assert len(bblocks) == 1 and "resumed after" in bblocks[0]
                assert a == b, (L, s["name"])
                # run files: one step for the single run, two for the continued one; the
                # continued run's series and final spins are the single run's
                da = np.loadtxt(once / L / f"{s['name']}.txt", ndmin=2)
                db = np.loadtxt(twice / L / f"{s['name']}.txt", ndmin=2)
                nft = da.shape[1] - 6
                sa = _check_runfile(once / L / f"{s['name']}.h5", da, ham, nft, steps=1)
                _check_runfile(twice / L / f"{s['name']}.h5", db[200:], ham, nft, steps=2)
                h0 = h5tools.values(twice / L / f"{s['name']}.h5", "/step0/observables/m")
                assert np.array_equal(h0, db[:200, 2])
                sb = h5tools.values(twice / L / f"{s['name']}.h5", "/step1/state/hamiltonianstatespace")
                assert np.array_equal(sa, sb), (L, s["name"])
                wa = h5tools.values(once / L / f"{s['name']}.h5", "/step0/state/rngstate", integer=True)
                wb = h5tools.values(twice / L / f"{s['name']}.h5", "/step1/state/rngstate", integer=True)
                assert wa == wb and wa[0] == 11
                assert h5tools.values(twice / L / f"{s['name']}.h5", "/step1/config/marqovconfig/warmupsteps")[0] == 0
                nfiles += 1
    assert nfiles > 0
    # a different seed does not silently continue somebody else's chain
    conf = os.path.join(ROOT, "tests", "data", "config_resume_twice")
    r = subprocess.run([BIN, ham, "--config", conf, "--out", str(twice), "--seed", "12"], capture_output=True,
                       text=True, timeout=110)
    assert r.returncode == 1 and "--seed 11" in r.stderr
    # --fresh starts again: the run files are rewritten with one step, the other seed
    r = subprocess.run([BIN, ham, "--config", conf, "--out", str(twice), "--seed", "12", "--fresh"],
                       capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr
    s = p[0]["groups"][0][0]
    h5 = twice / Ls[0] / f"{s['name']}.h5"
    assert sorted(k for k in h5tools.tree(h5) if k.count("/") == 1) == ["/step0"]
    assert h5tools.values(h5, "/step0/state/rngstate", integer=True)[0] == 12
    # a run file whose last step has no state (copied without it) is refused, naming the file
    bad = tmp_path / "bad.h5"
    for grp in ("config", "environment", "observables"):
        subprocess.run([os.path.join(h5tools.H5BIN, "h5copy"), "-p", "-i", str(h5), "-o", str(bad), "-s",
                        f"/step0/{grp}", "-d", f"/step0/{grp}"], check=True, capture_output=True, timeout=60)
    os.replace(bad, h5)
    r = subprocess.run([BIN, ham, "--config", conf, "--out", str(twice), "--seed", "12"], capture_output=True,
                       text=True, timeout=110)
    assert r.returncode == 1 and str(h5) in r.stderr and "unreadable last step" in r.stderr, r.stderr


@pytest.mark.gpu
def test_gpu_ini_csv_graph_vs_exact(tmp_path):
    """--graph: Ising.ini's temperatures and replicas run on a GraphFromCSV file (the
    3x3 torus written with 1-based indices, read with idxoffs 1); <E/N> and <m^2>
    agree with exact enumeration of that graph within 4 standard errors."""
    from marqov_amd import graphs
    rp, ci = graphs.regular_csr(2, 3)
    path = str(tmp_path / "t33.csv")
    graphs.write_csv_graph(path, rp, ci + 1)
    out = tmp_path / "out"
    r = subprocess.run([BIN, "Ising", "--config", CONF, "--out", str(out), "--graph", path, "--idxoffs", "1"],
                       capture_output=True, text=True, timeout=110)
    print(r.stdout[-600:], r.stderr[-600:], flush=True)
    assert r.returncode == 0, r.stderr
    for beta in (0.3, 0.4406868, 0.7):
        e_rep, m2_rep = [], []
        for rep in range(8):
            hdr, d = _series(out / "graph" / f"beta{beta:.6f}_{rep}.txt")
            assert d.shape == (4000, 6) and "lattice=csv" in hdr[0]
            if rep == 0:  # the run file records the graph as GraphFromCSV does (graph_from_csv.h:65-69)
                h5 = out / "graph" / f"beta{beta:.6f}_{rep}.h5"
                _check_runfile(h5, d, "Ising", 0)
                assert h5tools.string(h5, "/step0/config/lattice/name") == "GraphFromCSV"
                assert h5tools.values(h5, "/step0/config/lattice/npoints", integer=True) == [9]
                assert "/step0/config/lattice/dim" not in h5tools.tree(h5)
            e_rep.append(d[:, 1].mean())
            m2_rep.append(d[:, 3].mean())
        ex = exact.enumerate_graph(rp, ci, beta, J=-1.0)
        for got, want in ((e_rep, ex["e"]), (m2_rep, ex["m2"])):
            g = np.asarray(got)
            se = g.std(ddof=1) / np.sqrt(len(g))
            assert abs(g.mean() - want) < 4 * se + 1e-9, (beta, g.mean(), se, want)


@pytest.mark.gpu
def test_gpu_ini_L2_vs_exact(tmp_path):
    """L = 2, where both x0 neighbours of a site are the same site and the reference's
    neighbour lists count that bond twice (regular_hypercubic.h:31-66), runs on the
    checkerboard kernels; <E/N> and <m^2> agree with exact enumeration of the 2x2 torus
    with doubled bonds within 4 standard errors over 8 replicas."""
    conf = tmp_path / "config"
    conf.mkdir()
    text = open(os.path.join(CONF, "Ising.ini")).read()
    text = text.replace("L   = 4, 3", "L   = 2").replace("beta = 0.3, 0.4406868, 0.7", "beta = 0.3, 0.7")
    (conf / "Ising.ini").write_text(text)
    out = tmp_path / "out"
    r = subprocess.run([BIN, "Ising", "--config", str(conf), "--out", str(out), "--seed", "3"], capture_output=True,
                       text=True, timeout=110)
    assert r.returncode == 0, r.stderr
    for beta in (0.3, 0.7):
        e_rep, m2_rep = [], []
        for rep in range(8):
            hdr, d = _series(out / "2" / f"beta{beta:.6f}_{rep}.txt")
            assert "lattice=regular" in hdr[0] and d.shape == (4000, 8), (hdr[0], d.shape)
            e_rep.append(d[:, 1].mean())
            m2_rep.append(d[:, 3].mean())
        ex = exact.enumerate_ising(2, 2, beta, J=-1.0)
        for got, want, what in ((e_rep, ex["e"], "E"), (m2_rep, ex["m2"], "m2")):
            g = np.asarray(got)
            se = g.std(ddof=1) / np.sqrt(len(g))
            print(f"L=2 beta={beta} {what}: {g.mean():.5f} +- {se:.5f} exact {want:.5f}", flush=True)
            assert abs(g.mean() - want) < 4 * se + 1e-9, (beta, what, g.mean(), se, want)


def test_csv_empty_lines_are_isolated_vertices(tmp_path):
    """import_geometry (lattice/util/io.h:99-123) reads every line after the two header lines
    as a vertex, an empty line as one with no coordinates and no neighbours; a file ending
    in an empty line therefore ends with an isolated vertex.  Both readers (C++ front-end,
    Python) read it so, also with CRLF line ends."""
    from marqov_amd import graphs
    lines = ["2", "coordinates and neighbours",
             "0.1\t0.2\t1\t3", "0.3\t0.4\t0", "", "0.5\t0.6\t0", ""]
    for eol in ("\n", "\r\n"):
        path = str(tmp_path / f"iso{len(eol)}.csv")
        with open(path, "w", newline="") as f:
            f.write(eol.join(lines) + eol)
        want_rp = np.array([0, 2, 3, 3, 4, 4], np.int32)
        want_ci = np.array([1, 3, 0, 0], np.int32)
        rp, ci, xy = graphs.read_csv_graph(path)
        assert np.array_equal(rp, want_rp) and np.array_equal(ci, want_ci), (rp, ci)
        assert xy.shape == (5, 2) and np.isnan(xy[2]).all() and xy[3, 1] == 0.6
        out = subprocess.run([LCHK, "csv", path], capture_output=True, text=True, check=True).stdout.splitlines()
        rp2, ci2 = _read_csr(out, 1)
        assert np.array_equal(rp2, want_rp) and np.array_equal(ci2, want_ci), (rp2, ci2)


LCHK_SAN = os.path.join(ROOT, "tests", "cpp", "lattice_check_san")
MQG_SAN = os.path.join(ROOT, "tests", "cpp", "marqov_gpu_san")
SAN_ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=97",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=97")


def _sanitizer_clean(r):
    bad = ("AddressSanitizer", "runtime error:", "LeakSanitizer")
    return r.returncode != 97 and not any(b in r.stderr for b in bad)


def test_host_parsers_under_address_and_ub_sanitizers(tmp_path):
    """The .ini registry reader and planner (registry.hpp, marqov_gpu.cpp) and the CSV graph
    reader (lattices.hpp), built with AddressSanitizer and UndefinedBehaviorSanitizer, on every
    configuration of the test data and on malformed inputs: each run either succeeds or ends
    with the parser's own error message, and no sanitizer reports anything."""
    if not (os.path.exists(LCHK_SAN) and os.path.exists(MQG_SAN)):
        pytest.skip("sanitizer builds absent (make -C tests/cpp)")
    from marqov_amd import graphs
    runs = []
    for conf in ("config", "config_noend", "config_resume_once", "config_resume_twice"):
        d = os.path.join(ROOT, "tests", "data", conf)
        for f in sorted(os.listdir(d)):
            runs.append([MQG_SAN, f[:-4], "--config", d, "--dry-run"])
    bad_ini = {
        "empty": "",
        "no_end": "[Ising]\nL = 8, 12\n",
        "no_value": "[Ising]\nL =\nbeta =\n[END]\n",
        "junk_numbers": "[Ising]\nL = 8x, -4, 99999999999999999999\nbeta = nan, 1e400, ,;\nrep = -3\n[MC]\nwarmupsteps=1e99\n[END]\n",
        "long_line": "[Ising]\nL = " + ", ".join(["16"] * 5000) + "\n[END]\n",
        "crlf": "[Ising]\r\nL = 8\r\nbeta = 0.4\r\n[MC]\r\nnmetro = 1\r\n[END]\r\n",
        "unbalanced": "[[Ising\n= 3\n]]\n[END\n",
    }
    for name, text in bad_ini.items():
        d = tmp_path / f"ini_{name}"
        d.mkdir()
        (d / "Ising.ini").write_text(text)
        runs.append([MQG_SAN, "Ising", "--config", str(d), "--dry-run"])
    rp, ci, pts = graphs.rgg2d(200, graphs.rgg_radius(200), 4)
    good = str(tmp_path / "good.csv")
    graphs.write_csv_graph(good, rp, ci, pts)
    runs.append([LCHK_SAN, "csv", good])
    bad_csv = {
        "empty": "",
        "header_only": "0\n",
        "no_number": "x\nheader\n1\t2\n",
        "bad_index": "0\nh\n1\tq\n0\n",
        "out_of_range": "0\nh\n5\n0\n",
        "huge_index": "0\nh\n99999999999999999999\n",
        "negative": "0\nh\n-1\n",
        "tabs": "1\nh\n\t\t\t\n0.5\t\t0\n",
        "isolated": "0\nh\n1\n0\n\n\n",
    }
    for name, text in bad_csv.items():
        pth = str(tmp_path / f"{name}.csv")
        with open(pth, "w") as f:
            f.write(text)
        runs.append([LCHK_SAN, "csv", pth])
        runs.append([LCHK_SAN, "csv", pth, "1"])
    for cmd in runs:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, env=SAN_ENV)
        assert _sanitizer_clean(r), (cmd, r.returncode, r.stderr[-2000:])
