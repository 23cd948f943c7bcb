"""Writes tests/golden/accd_stall_cases.txt: ACCD inputs whose iteration reaches the 10001-step
cap of accd.hpp:59-66 (result 0.001), as exact FP32 hex floats, one case per line:

    kind offset  x0[4][3]  x1[4][3]   (kind 0: point-triangle, 1: edge-edge)

The first point-triangle case is a pair taken from the cloth-tube workload (a pinned end-ring
vertex against a sheared first-ring triangle); the others are nearby perturbations of it and
random edge-edge configurations, kept when the CPU oracle needs more than 10000 iterations.
Inputs only: the test recomputes both results.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

TUBE_PAIR = (
    [["-0x1.0p-1", "-0x1.c08514p-4", "-0x1.56a0a6p-7"], ["-0x1.0p-1", "-0x1.bf8374p-4", "-0x1.a2745ep-7"],
     ["-0x1.fcccb4p-2", "-0x1.c6624cp-4", "-0x1.f6f32ap-7"], ["-0x1.0p-1", "-0x1.be4e8ep-4", "-0x1.ee1824p-7"]],
    [["-0x1.00001cp-1", "-0x1.b995c2p-4", "-0x1.65dcecp-6"], ["-0x1.fffffep-2", "-0x1.b79844p-4", "-0x1.8b2f18p-6"],
     ["-0x1.feb7e4p-2", "-0x1.bbea14p-4", "-0x1.41b78cp-6"], ["-0x1.000006p-1", "-0x1.b56864p-4", "-0x1.b03c8p-6"]],
)


def stalls(kind, a, b, off):
    _, it = oracle.accd(a.astype(np.float64), b.astype(np.float64), float(off), kind, 32)
    return it > 10000


def main():
    off = np.float32(1e-3)
    x0, x1 = (np.array([[float.fromhex(c) for c in row] for row in v], dtype=np.float32) for v in TUBE_PAIR)
    assert stalls(0, x0, x1, off)
    cases = [(0, x0, x1)]
    rng = np.random.default_rng(7)
    while sum(c[0] == 0 for c in cases) < 6:
        a = (x0 + rng.normal(0, 2e-5, x0.shape)).astype(np.float32)
        b = (x1 + rng.normal(0, 2e-4, x1.shape)).astype(np.float32)
        if stalls(0, a, b, off):
            cases.append((0, a, b))
    while sum(c[0] == 1 for c in cases) < 6:
        a = (rng.uniform(-1, 1, (4, 3)) * 0.01).astype(np.float32)
        b = (a + rng.uniform(-1, 1, (4, 3)) * 0.02).astype(np.float32)
        if stalls(1, a, b, off):
            cases.append((1, a, b))
    with open(os.path.join(HERE, "accd_stall_cases.txt"), "w") as f:
        for kind, a, b in cases:
            vals = [float(off)] + [float(v) for v in a.reshape(-1)] + [float(v) for v in b.reshape(-1)]
            f.write(f"{kind} " + " ".join(v.hex() for v in vals) + "\n")


if __name__ == "__main__":
    main()
