"""Reads HDF5 files through the HDF5 project's own command-line tools (h5ls,
h5dump from the image's HDF5 installation), so that the tests check the run
files with code independent of the writer (marqov_amd/driver/runfile.hpp)."""
import os
import re
import subprocess

import numpy as np

H5BIN = os.environ.get("MQ_H5BIN", "/opt/conda/bin")
H5LS = os.path.join(H5BIN, "h5ls")
H5DUMP = os.path.join(H5BIN, "h5dump")

_NUM = re.compile(r"[-+]?(?:\d+\.?\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|inf|nan)")


def available():
    return os.path.exists(H5LS) and os.path.exists(H5DUMP)


def _run(args):
    return subprocess.run(args, capture_output=True, text=True, check=True, timeout=120).stdout


def tree(path):
    """{name: kind} of every object below the root, kind "Group" or "Dataset {dims}"."""
    out = {}
    for line in _run([H5LS, "-r", str(path)]).splitlines():
        name, _, kind = line.partition(" ")
        if name != "/":
            out[name] = kind.strip()
    return out


def header(path, name):
    """h5dump's description of a dataset (type, space, layout, filters, fill value, comment)."""
    return _run([H5DUMP, "-p", "-H", "-d", name, str(path)])


def values(path, name, integer=False, start=None, count=None):
    """The values of a numeric dataset, flattened (array records concatenated), as
    printed by h5dump with 17 significant digits (exact for doubles); integer=True
    keeps integers exact (Python ints, for 64-bit words); start/count select records."""
    sel = ["-s", str(start), "-c", str(count)] if start is not None else []
    text = _run([H5DUMP, "-m", "%.17g", "-y", "-w", "0", "-d", name] + sel + [str(path)])
    body = text[text.index("DATA {") + len("DATA {"):]
    body = body[:body.index("}")]
    toks = _NUM.findall(body)
    if integer:
        return [int(t) for t in toks]
    return np.array([float(t) for t in toks])


def string(path, name):
    text = _run([H5DUMP, "-y", "-w", "0", "-d", name, str(path)])
    body = text[text.index("DATA {"):]
    return re.search(r'"((?:[^"\\]|\\.)*)"', body).group(1)
