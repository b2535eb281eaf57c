"""Launch-shape autotuner for the Metropolis sweep and the Wolff update.

For a (model, storage, lattice shape, replica count) key it times every candidate
launch shape with HIP events, keeps the fastest, caches the winner in a JSON file
and applies it to the caller's state through mq_state_set_tuning.

Sweep candidates (tune):
  kernels   0 (automatic: LDS multi-sweep kernel for lattices that fit in LDS,
            row-vectorised 3D kernel where it applies) or 1 (streaming only)
  block     threads per block of the streaming kernels: 64, 128, 256, 512
  xcd_remap 0 / 1 (contiguous slab of rows per XCD)
  lds_block threads per block of the LDS kernel: 128, 256, 512

Wolff candidates (tune_wolff):
  wolff_block  threads of the workgroup growing one replica's clusters on large
               lattices and graphs: 256, 512, 1024
  wolff_vset   LDS membership up to 4096 sites or none (0)
  wolff_ring   LDS frontier entries of the small-lattice kernel: 256 .. 2048

The chain a state produces does not depend on its launch shape (the kernels
draw the same Philox counters and grow the same clusters), so tuning never
changes results.  tune_wolff times every candidate on the same clusters from
the same state (saved before and restored after), so the caller's chain is
left as it was.
"""
import itertools
import json
import os

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "marqov_amd", "tuning.json")


def key_of(state):
    lat = state.lattices[0]
    shape = f"L{lat.L}d{lat.dim}" if getattr(lat, "dim", 0) else f"csr{lat.N}"
    sizes = sorted({(getattr(x, "L", 0), x.N) for x in state.lattices})
    return f"{state.model}/bits{state.params.spin_bits}/{shape}/nrep{state.nrep}/sizes{len(sizes)}"


def candidates(state):
    lat = state.lattices[0]
    small = all(x.N < 65536 for x in state.lattices) and state.model in ("ising", "ea") and \
        state.params.spin_bits != 1 and getattr(lat, "dim", 0) in (2, 3)
    out = []
    for kernels, block, xcd in itertools.product((0, 1), (64, 128, 256, 512), (0, 1)):
        if kernels == 0 and small:
            continue  # the LDS kernel ignores block / xcd: tuned below
        if xcd == 0 and not (getattr(lat, "dim", 0) == 3 and state.model in ("ising", "heisenberg", "phi4")):
            continue  # only the row-vectorised 3D Ising and the FP32 kernels remap blocks
        out.append({"kernels": kernels, "block": block, "xcd_remap": xcd, "lds_block": 256})
    if small:
        for lb in (128, 256, 512):
            out.append({"kernels": 0, "block": 256, "xcd_remap": 1, "lds_block": lb})
    return out


def time_shape(ctx, state, shape, sweeps=4, warm=1, seed=12345):
    state.set_tuning(**shape)
    state.metropolis(warm, seed, sweep0=1 << 28)
    ctx.sync()
    ctx.timer_begin()
    state.metropolis(sweeps, seed, sweep0=(1 << 28) + warm)
    return ctx.timer_end() / sweeps


def tune(ctx, state, cache_path=DEFAULT_CACHE, force=False, sweeps=4):
    """Pick the fastest launch shape for `state` and apply it; returns (shape, ms_per_sweep table).

    Sweeps run on the caller's state (they are Metropolis sweeps of the same chain
    with counters far from any production run); pass a scratch state if that matters."""
    key = key_of(state)
    cache = _read_cache(cache_path)
    if key in cache and not force:
        state.set_tuning(**cache[key]["shape"])
        return cache[key]["shape"], cache[key].get("timings", {})
    timings = {}
    best, best_ms = None, float("inf")
    for shape in candidates(state):
        ms = time_shape(ctx, state, shape, sweeps=sweeps)
        timings[json.dumps(shape, sort_keys=True)] = ms
        if ms < best_ms:
            best, best_ms = shape, ms
    state.set_tuning(**best)
    cache[key] = {"shape": best, "ms_per_sweep": best_ms, "timings": timings}
    _write_cache(cache_path, cache)
    return best, timings


def _read_cache(cache_path):
    if os.path.exists(cache_path):
        try:
            with open(cache_path) as f:
                return json.load(f)
        except (OSError, ValueError):
            return {}
    return {}


def _write_cache(cache_path, cache):
    os.makedirs(os.path.dirname(cache_path) or ".", exist_ok=True)
    tmp = cache_path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, cache_path)


def wolff_candidates(state):
    """Wolff launch shapes that matter for this state: the small-lattice kernel (int8
    Ising/Potts on regular lattices of at most 16,384 sites) takes only the ring size; the
    workgroup kernels take the block size and, on regular lattices, the LDS membership."""
    lat = state.lattices[0]
    regular = getattr(lat, "dim", 0) in (2, 3)
    small = regular and state.model in ("ising", "potts") and all(x.N <= 16384 for x in state.lattices)
    if small:
        return [{"wolff_ring": r} for r in (256, 512, 1024, 2048)]
    out = []
    for blk in (256, 512, 1024):
        for vset in ((4096, 0) if regular else (4096,)):
            out.append({"wolff_block": blk, "wolff_vset": vset})
    return out


def tune_wolff(ctx, state, nclusters=10, seed=12345, cluster0=1 << 30, cache_path=DEFAULT_CACHE, force=False):
    """Pick the fastest Wolff launch shape for `state` and apply it; returns (shape, ms table).

    Every candidate grows the same nclusters clusters per replica from the same state (saved
    first, uploaded again before each candidate and after the last), so the timings compare
    identical work and the caller's state is unchanged."""
    key = "wolff/" + key_of(state)
    cache = _read_cache(cache_path)
    if key in cache and not force:
        state.set_tuning(**cache[key]["shape"])
        return cache[key]["shape"], cache[key].get("timings", {})
    saved = [state.download(r) for r in range(state.nrep)]
    timings = {}
    best, best_ms = None, float("inf")
    for shape in wolff_candidates(state):
        for r, s in enumerate(saved):
            state.upload(r, s)
        state.set_tuning(**shape)
        ctx.sync()
        ctx.timer_begin()
        state.wolff(nclusters, seed, cluster0=cluster0, sizes=False)
        ms = ctx.timer_end()
        timings[json.dumps(shape, sort_keys=True)] = ms
        if ms < best_ms:
            best, best_ms = shape, ms
    for r, s in enumerate(saved):
        state.upload(r, s)
    state.set_tuning(**best)
    cache[key] = {"shape": best, "ms_per_call": best_ms, "timings": timings}
    _write_cache(cache_path, cache)
    return best, timings
