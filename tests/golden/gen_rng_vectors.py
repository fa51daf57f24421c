"""Writes tests/golden/rng_vectors.json from the reference's own RNG sources.

The reference vendors the public-domain reference implementations
Source/Externals/xoshiro/{splitmix64.c,xoshiro128starstar.c} and tests its
shader RNG against them (Source/Tools/FalcorTest/Tests/Sampling/
PseudorandomTests.cpp:94-157: 256 instances x 64 outputs, seeds from
std::mt19937). oracle/Makefile compiles those two files, where they lie, into
oracle/_ref/; this script drives them to produce:

  * splitmix64: 64 instances x 64 outputs
  * xoshiro128**: 64 instances x 64 outputs
  * the pass's per-pixel sample generator (UT/Sampling/UniformSampleGenerator.slang:49-68):
    seed64 = (sampleNumber << 32) | Morton(pixel), two SplitMix64 outputs as the
    xoshiro128** state, then 16 outputs, for a set of (pixel, sampleNumber)
    including the pass's per-stage sample numbers.

The Morton code here is the bit-by-bit definition (bit i of x -> bit 2i,
bit i of y -> bit 2i+1), independent of the oracle's magic-number version.
Run:  python tests/golden/gen_rng_vectors.py   (needs oracle/_ref built)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle.binding import ref_splitmix64, ref_xoshiro128ss  # noqa: E402


def morton_bitwise(x, y):
    r = 0
    for i in range(16):
        r |= ((x >> i) & 1) << (2 * i)
        r |= ((y >> i) & 1) << (2 * i + 1)
    return r


def main():
    rng = np.random.Generator(np.random.MT19937(5489))
    sm_seeds = [int(s) for s in rng.integers(0, 2**64, size=64, dtype=np.uint64)]
    xo_seeds = [[int(v) for v in rng.integers(1, 2**32, size=4, dtype=np.uint64)] for _ in range(64)]
    out = {"splitmix64": [], "xoshiro128ss": [], "pixel_streams": []}
    for s in sm_seeds:
        out["splitmix64"].append({"seed": str(s), "outputs": [str(int(v)) for v in ref_splitmix64(s, 64)]})
    for s in xo_seeds:
        out["xoshiro128ss"].append({"seed": s, "outputs": [int(v) for v in ref_xoshiro128ss(np.array(s, np.uint32), 64)]})
    pixels = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (639, 359, 0), (1919, 1079, 0), (5, 7, 4), (5, 7, 5),
              (5, 7, 9), (100, 200, 6 * 3 + 5), (1234, 567, 6 * 59), (65535, 65535, 12345), (3, 3, 2**31 + 7)]
    for (px, py, sn) in pixels:
        seed = (sn << 32) | morton_bitwise(px, py)
        st = ref_splitmix64(seed, 2)
        s0, s1 = int(st[0]), int(st[1])
        state = np.array([s0 & 0xFFFFFFFF, s0 >> 32, s1 & 0xFFFFFFFF, s1 >> 32], np.uint32)
        stream = ref_xoshiro128ss(state, 16)
        out["pixel_streams"].append({"pixel": [px, py], "sampleNumber": sn, "morton": morton_bitwise(px, py),
                                     "outputs": [int(v) for v in stream]})
    path = os.path.join(HERE, "rng_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", path)


if __name__ == "__main__":
    main()
