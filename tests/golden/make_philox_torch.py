"""Writes tests/golden/philox_torch.json: Philox4x32-10 outputs computed by torch's
independent implementation (ATen/core/PhiloxRNGEngine.h, header-only) for random
counters and keys.  Run once; the fixture is committed.

torch's engine uses key = seed and counter = (offset_lo, offset_hi, subseq_lo, subseq_hi);
its first operator() call returns word 0 of philox(counter, key).
"""
import json
import os
import subprocess
import tempfile

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = r'''
#include <cstdio>
#include <cstdlib>
#include <ATen/core/PhiloxRNGEngine.h>
int main(int argc, char** argv) {
  unsigned c[6];
  while (std::scanf("%u %u %u %u %u %u", &c[0], &c[1], &c[2], &c[3], &c[4], &c[5]) == 6) {
    unsigned long long seed = (unsigned long long)c[4] | ((unsigned long long)c[5] << 32);
    unsigned long long sub = (unsigned long long)c[2] | ((unsigned long long)c[3] << 32);
    unsigned long long off = (unsigned long long)c[0] | ((unsigned long long)c[1] << 32);
    at::philox_engine e(seed, sub, off);
    unsigned a = e(), b = e(), d = e(), f = e();
    std::printf("%u %u %u %u\n", a, b, d, f);
  }
  return 0;
}
'''


def main():
    inc = os.path.join(os.path.dirname(torch.__file__), "include")
    rng = np.random.default_rng(2024)
    cases = rng.integers(0, 2 ** 32, size=(64, 6), dtype=np.uint64).tolist()
    cases[0] = [0, 0, 0, 0, 0, 0]
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "p.cpp")
        exe = os.path.join(td, "p")
        open(src, "w").write(SRC)
        subprocess.run(["g++", "-std=c++17", "-O1", "-I", inc, src, "-o", exe], check=True)
        inp = "\n".join(" ".join(str(v) for v in c) for c in cases) + "\n"
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    res = []
    for c, o in zip(cases, out):
        res.append({"ctr": c[:4], "key": c[4:], "out": [int(x) for x in o.split()]})
    with open(os.path.join(HERE, "philox_torch.json"), "w") as f:
        json.dump({"source": "torch ATen/core/PhiloxRNGEngine.h (torch " + torch.__version__ + ")",
                   "cases": res}, f, indent=0)


if __name__ == "__main__":
    main()
