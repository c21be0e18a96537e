"""Writes the golden fixtures of the GPU parity tests at the workload sizes (run here, on the CPU).

    python tests/golden/make_fixtures.py [name ...]

For each case the oracle (oracle/stage1 C restatement + oracle/stage2 numpy restatement, both pinned
by Fock-space brute force in tests/test_oracle_brute.py) evaluates the seeded synthetic problem, and
the fixture keeps the inputs' recipe (sizes, seed, electrons), the per-class table (norm_sc, e_sc,
norm_pc, e_pc) and, per Koopmans matrix, its sum, sum of squares and the values at fixed sampled
positions.  The GPU tests rebuild the same inputs from the recipe (synth is deterministic) and compare.
"""
from __future__ import annotations

import json
import os
import sys
import time
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from nevpt2_amd import synth            # noqa: E402
from oracle import stage1 as S1, stage2 as S2   # noqa: E402

CASES = {   # name: (nact, ncore, nvirt, seed, active electrons, description)
    "c1": (6, 10, 40, 1, 6, "CAS(6,6) nc=10 nv=40: the reference's CPU-runnable configuration"),
    "c2": (10, 20, 200, 1, 10, "CAS(10,10) nc=20 nv=200: the FP64 drop-in configuration (split-K GEMMs, "
                               "band form of K' at order 1000)"),
    "cas8": (8, 4, 20, 2, 8, "CAS(8,8) nc=4 nv=20: (+-1)' metric order 512, the smallest band-form case"),
    "cas9": (9, 3, 12, 3, 8, "CAS(9,9) with 8 electrons nc=3 nv=12: odd active size, order 729"),
}
KOOP = ["1-IPKO", "1-EAKO", "2-HRDM", "2-IPKO", "2-EAKO", "gk0pa", "gk0pd", "gk0pf", "amatIP", "amatEA",
        "bmatIP", "bmatEA", "cmatIP", "cmatEA", "dmatIP", "dmatEA"]
NSAMPLE = 64


def sample_positions(shape, seed):
    rng = np.random.default_rng(seed)
    return [[int(rng.integers(0, d)) for d in shape] for _ in range(NSAMPLE)]


def make(name):
    n, nc, nv, seed, nele, desc = CASES[name]
    t0 = time.time()
    P = synth.Problem(n, nc, nv, seed=seed, nele=nele)
    t1 = time.time()
    koop = S1.koopmans_all(P)
    t2 = time.time()
    tab, _ = S2.state_energies(P, koop, 0)
    t3 = time.time()
    fx = {"name": name, "description": desc, "nact": n, "ncore": nc, "nvirt": nv, "seed": seed, "nele": nele,
          "roots": 1, "classes": S2.CLASS_NAMES, "table": tab.tolist(), "koopmans": {},
          "seconds": {"rdms": t1 - t0, "stage1": t2 - t1, "stage2": t3 - t2}}
    for k in KOOP:
        a = np.asarray(koop[k][0])
        pos = sample_positions(a.shape, zlib.crc32(k.encode()) % 1000 + 17 * n)
        fx["koopmans"][k] = {"shape": list(a.shape), "sum": float(a.sum()), "sumsq": float((a * a).sum()),
                             "positions": pos, "values": [float(a[tuple(p)]) for p in pos]}
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(fx, f, indent=1)
    print(f"{name}: rdms {t1 - t0:.1f}s stage1 {t2 - t1:.1f}s stage2 {t3 - t2:.1f}s", flush=True)


if __name__ == "__main__":
    for name in (sys.argv[1:] or list(CASES)):
        make(name)
