"""Writes tests/golden/pcg_vectors.json: golden vectors of the RNG of the PatchMatch path.

The hash is the PCG-RXS-M-XS output permutation published as `pcg_hash` in Jarzynski & Olano,
"Hash Functions for GPU Rendering", JCGT 9(3), 2020 (the function the reference uses at
frontend/src/app/vision/mvs/patchmatch.wgsl.ts:56-60).  This script evaluates it with numpy
uint64 arithmetic reduced mod 2^32 — independently of the C oracle — over a fixed set of inputs,
and records the per-pixel stream construction of PM:251-253 and PM:62-65 for a few seeds.
"""
import json
import os

import numpy as np

M = np.uint64(0xFFFFFFFF)


def pcg_hash(v):
    v = np.asarray(v, dtype=np.uint64) & M
    state = (v * np.uint64(747796405) + np.uint64(2891336453)) & M
    word = (((state >> ((state >> np.uint64(28)) + np.uint64(4))) ^ state) * np.uint64(277803737)) & M
    return ((word >> np.uint64(22)) ^ word) & M


def main():
    rng = np.random.default_rng(20201102)
    inputs = np.concatenate([np.array([0, 1, 2, 3, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF], dtype=np.uint64),
                             rng.integers(0, 2**32, 57, dtype=np.uint64)])
    outputs = pcg_hash(inputs)
    streams = []
    for seed, it, phase, index in [(0x9E3779B1, 0, 0, 0), (0x3C6EF362, 3, 1, 12345), (7, 7, 0, 921599)]:
        state = pcg_hash(np.uint64(index) ^ pcg_hash((np.uint64(seed) + np.uint64(it) * np.uint64(7919) +
                                                      np.uint64(phase) * np.uint64(104729)) & M))
        draws = []
        for _ in range(8):
            state = pcg_hash(state)
            draws.append(int(state >> np.uint64(8)))  # random() = draw / 2^24
        streams.append({"seed": seed, "iteration": it, "phase": phase, "index": index,
                        "draws_24bit": draws})
    out = {"inputs": [int(x) for x in inputs], "pcg": [int(x) for x in outputs], "streams": streams}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pcg_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
