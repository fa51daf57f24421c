"""RNG of the pass against the reference's own generator sources (golden vectors)
and the reference's statistical checks (FT/Sampling/SampleGeneratorTests.cpp:107-162)."""
import numpy as np

from oracle import binding as ob


def test_morton_known_answer():
    # FT/Utils/BitTricksTests.cpp:55
    assert ob.lib().orc_morton(0xE38E, 0xBE8B) == 0xDEADC0DE


def test_morton_matches_bitwise_definition(golden_rng):
    for rec in golden_rng["pixel_streams"]:
        x, y = rec["pixel"]
        assert ob.lib().orc_morton(x, y) == rec["morton"]


def test_splitmix64_golden(golden_rng):
    L = ob.lib()
    for rec in golden_rng["splitmix64"]:
        out = np.zeros(64, np.uint64)
        L.orc_splitmix64(int(rec["seed"]), 64, out.ctypes.data)
        assert [int(v) for v in out] == [int(s) for s in rec["outputs"]]


def test_xoshiro128ss_golden(golden_rng):
    L = ob.lib()
    for rec in golden_rng["xoshiro128ss"]:
        seed = np.array(rec["seed"], np.uint32)
        out = np.zeros(64, np.uint32)
        L.orc_xoshiro128ss(seed.ctypes.data, 64, out.ctypes.data)
        assert out.tolist() == rec["outputs"]


def test_pixel_sample_generator_golden(golden_rng):
    for rec in golden_rng["pixel_streams"]:
        x, y = rec["pixel"]
        out = ob.rng_stream(x, y, rec["sampleNumber"], 16)
        assert out.tolist() == rec["outputs"], rec["pixel"]


def test_uniform_statistics():
    # mean in [0.499, 0.501] and |corr| <= 0.002 across dimensions / neighbouring pixels
    W, H, D = 256, 64, 16
    u = np.zeros((H, W, D))
    for y in range(H):
        for x in range(W):
            u[y, x] = (ob.rng_stream(x, y, 3, D) >> 8) * 2.0**-24
    assert u.min() >= 0.0 and u.max() < 1.0
    assert 0.499 <= u.mean() <= 0.501
    flat = u.reshape(-1, D)
    bound = 5.0 / np.sqrt(flat.shape[0])  # 5 sigma of the sample correlation of independent uniforms
    for d in range(D - 1):
        assert abs(np.corrcoef(flat[:, d], flat[:, d + 1])[0, 1]) <= bound
    assert abs(np.corrcoef(u[:, :-1, 0].ravel(), u[:, 1:, 0].ravel())[0, 1]) <= bound
