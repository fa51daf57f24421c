"""Pins the oracle's learned free-flight proposal (oracle/oracle_mlp.hpp) by what
the mathematics fixes, independently of its own formulas:

- the frequency encoding equals sin/cos(pi 2^k x) evaluated by numpy;
- the exact network equals a numpy forward pass of the same parameter layout;
- backpropagation equals central finite differences of the loss (FP64 network);
- the 16-bit CDF is a valid distribution that rounds the mixture pdf, and the
  inverse-CDF sampler reproduces it exactly on a stratified grid of u;
- Adam on the cross-entropy drives the network to a target distribution that
  depends on its inputs;
- one training ray per block, and TraceRays' training records sum to the
  ray's weight;
- with a strongly non-uniform network as the source pdf, the pass still
  converges to the volume path tracer (the expectation check of
  tests/test_oracle_unbiased.py): a wrong pdf for an in-volume sample or for
  an escape changes that mean.

The learned proposal has no counterpart in the reference; it replaces the
distance pdf of the initial candidates (VR/ComputeInitialSample.slang:86,107).
"""
import numpy as np
import pytest

from oracle.binding import (MLP_OFFSETS, MLP_OUT, MLP_PARAMS, MLP_REC, Oracle, mlp_adam, mlp_backward, mlp_encode,
                            mlp_forward, mlp_loss, mlp_mass_above, mlp_random_weights, mlp_sample, mlp_train_slot,
                            mlp_unpack)
from volumetricrestirrelease_amd import abi, scenes

ALPHA = 0.25


def random_raw(rng, n):
    raw = np.zeros((n, 8), np.float32)
    raw[:, :3] = rng.uniform(0, 1, size=(n, 3))
    d = rng.normal(size=(n, 3))
    raw[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    raw[:, 6] = rng.uniform(0.05, 1.0, size=n)
    return raw


def random_records(rng, n, peaked=False):
    rec = np.zeros((n, MLP_REC), np.float32)
    rec[:, :8] = random_raw(rng, n)
    w = rng.gamma(0.5, size=(n, MLP_OUT)).astype(np.float32)
    if peaked:
        w[:, 5] += 3.0
    rec[:, 8:40] = w
    rec[:, 40:72] = w * w * rng.uniform(0.5, 2.0, size=w.shape).astype(np.float32)
    rec[:, 7] = w.sum(axis=1)
    return rec


def numpy_forward(w, raw, alpha=ALPHA):
    P = {k: v.astype(np.float64) for k, v in mlp_unpack(w).items()}
    x = np.stack([mlp_encode(r) for r in raw]).astype(np.float64)
    h = np.maximum(0, x @ P["W1"] + P["b1"])
    h = np.maximum(0, h @ P["W2"] + P["b2"])
    h = np.maximum(0, h @ P["W3"] + P["b3"])
    z = h @ P["W4"] + P["b4"]
    q = np.exp(z - z.max(axis=1, keepdims=True))
    q /= q.sum(axis=1, keepdims=True)
    return z, alpha / MLP_OUT + (1 - alpha) * q


def test_encoding_is_sin_cos_octaves():
    rng = np.random.default_rng(0)
    for raw in random_raw(rng, 20):
        enc = mlp_encode(raw)
        want = []
        for i in range(3):
            for k in range(3):
                want += [np.sin(np.pi * 2 ** k * raw[i]), np.cos(np.pi * 2 ** k * raw[i])]
        for i in range(3):
            for k in range(2):
                want += [np.sin(np.pi * 2 ** k * raw[3 + i]), np.cos(np.pi * 2 ** k * raw[3 + i])]
        for k in range(3):
            want += [np.sin(np.pi * 2 ** k * raw[6]), np.cos(np.pi * 2 ** k * raw[6])]
        np.testing.assert_allclose(enc, np.float32(want), rtol=0, atol=1e-6)


def test_exact_forward_equals_numpy_network():
    rng = np.random.default_rng(1)
    w = mlp_random_weights(3)
    raw = random_raw(rng, 64)
    z, cdf = mlp_forward(w, raw, ALPHA, emulate=False)
    zn, pn = numpy_forward(w, raw)
    np.testing.assert_allclose(z, zn, rtol=1e-5, atol=1e-5)
    # the CDF rounds the mixture pdf: every bin within one 2^-16 step (the last takes the rest)
    n = np.diff(np.concatenate([cdf.astype(np.int64), np.full((64, 1), 65536)], axis=1), axis=1)
    assert np.all(np.abs(n[:, :-1] - pn[:, :-1] * 65536) <= 0.5 + 1e-3)
    assert np.all(np.abs(n[:, -1] - pn[:, -1] * 65536) <= 16)


def test_emulated_forward_stays_close_to_exact():
    """bf16 operands (8 mantissa bits) with exact sums: logits within a few bf16 steps."""
    rng = np.random.default_rng(2)
    w = mlp_random_weights(4)
    raw = random_raw(rng, 256)
    ze, _ = mlp_forward(w, raw, ALPHA, emulate=True)
    zx, _ = mlp_forward(w, raw, ALPHA, emulate=False)
    scale = np.abs(zx).max()
    assert np.abs(ze - zx).max() < 0.03 * scale


def test_cdf_is_a_distribution_and_uniform_for_a_zero_output_layer():
    rng = np.random.default_rng(3)
    w = mlp_random_weights(5, last=3.0)
    _, cdf = mlp_forward(w, random_raw(rng, 500), ALPHA, emulate=True)
    assert np.all(cdf[:, 0] == 0)
    n = np.diff(np.concatenate([cdf.astype(np.int64), np.full((500, 1), 65536)], axis=1), axis=1)
    assert np.all(n >= 1)
    assert np.all(n.sum(axis=1) == 65536)
    # the mixture keeps at least alpha / 32 in every bin
    assert np.all(n >= np.floor(ALPHA / 32 * 65536) - 16)
    w0 = w.copy()
    o, _ = MLP_OFFSETS["W4"]
    w0[o:o + 64 * 32] = 0
    w0[MLP_OFFSETS["b4"][0]:MLP_OFFSETS["b4"][0] + 32] = 0
    _, cdf0 = mlp_forward(w0, random_raw(rng, 10), ALPHA, emulate=True)
    assert np.all(cdf0 == (2048 * np.arange(32)).astype(np.uint16))


def test_inverse_cdf_sampler_reproduces_the_bins():
    rng = np.random.default_rng(4)
    _, cdf = mlp_forward(mlp_random_weights(6, last=3.0), random_raw(rng, 1), ALPHA, emulate=True)
    cdf = cdf[0]
    n = np.diff(np.concatenate([cdf.astype(np.int64), [65536]]))
    counts = np.zeros(32, np.int64)
    cs = []
    K = 1 << 16
    for i in range(K):
        u = (i + 0.5) / K
        c, b, g = mlp_sample(cdf, u)
        assert b * (1 / 32) <= c < (b + 1) / 32 + 1e-7
        assert g == pytest.approx(32 * n[b] / 65536, rel=1e-6)
        counts[b] += 1
        cs.append(c)
    # a stratified grid of u lands exactly n_j times in bin j
    np.testing.assert_array_equal(counts, n)
    cs = np.array(cs)
    for x in [0.0, 0.1, 0.37, 0.5, 0.93, 0.999]:
        assert mlp_mass_above(cdf, x) == pytest.approx((cs >= x).mean(), abs=2e-4)
    assert mlp_mass_above(cdf, 0.0) == 1.0
    assert mlp_mass_above(cdf, 1.0) == 0.0


@pytest.mark.parametrize("alpha", [0.25, 1.0 / 64])
def test_backward_equals_finite_differences(alpha):
    rng = np.random.default_rng(5)
    w = mlp_random_weights(7, last=1.5)
    recs = random_records(rng, 6)
    g, st = mlp_backward(w, recs, alpha, 1.0, emulate=False)
    assert st[0] == pytest.approx(mlp_loss(w, recs, alpha, emulate=False), rel=1e-9)
    idx = np.concatenate([rng.choice(MLP_PARAMS, 240, replace=False),
                          np.arange(MLP_OFFSETS["b4"][0], MLP_PARAMS)])
    errs = []
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] = np.float32(w[i] + np.float32(1e-3))
        wm[i] = np.float32(w[i] - np.float32(1e-3))
        fd = (mlp_loss(wp, recs, alpha) - mlp_loss(wm, recs, alpha)) / (float(wp[i]) - float(wm[i]))
        errs.append(abs(fd - g[i]) / max(1e-4, abs(g[i]), abs(fd)))
    errs = np.array(errs)
    # ReLU kinks crossed by a perturbation can spoil an isolated difference
    assert np.median(errs) < 1e-4
    assert (errs < 1e-2).mean() >= 0.98


def test_emulated_backward_stays_close_to_exact():
    rng = np.random.default_rng(6)
    w = mlp_random_weights(8, last=1.5)
    recs = random_records(rng, 32)
    ge, se = mlp_backward(w, recs, ALPHA, 1.0 / 32, emulate=True)
    gx, sx = mlp_backward(w, recs, ALPHA, 1.0 / 32, emulate=False)
    for k, (o, sh) in MLP_OFFSETS.items():
        n = int(np.prod(sh))
        a, b = ge[o:o + n], gx[o:o + n]
        # bf16 rounding of dz and of the activations compounds over the layers below
        assert np.abs(a - b).max() < 0.15 * np.abs(b).max() + 1e-7, k
        assert np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.995, k
    np.testing.assert_allclose(se, sx, rtol=2e-2)


def test_adam_fits_an_input_dependent_target():
    """Records whose weight sits in bin 4 for entry x < 0.5 and in bin 27 otherwise:
    after Adam steps on the mean cross-entropy the network puts its mass there."""
    rng = np.random.default_rng(7)
    N = 128
    recs = np.zeros((N, MLP_REC), np.float32)
    recs[:, :8] = random_raw(rng, N)
    target = np.where(recs[:, 0] < 0.5, 4, 27)
    recs[np.arange(N), 8 + target] = 1.0
    recs[np.arange(N), 40 + target] = 1.0
    recs[:, 7] = 1.0
    w = mlp_random_weights(9)
    o, _ = MLP_OFFSETS["W4"]
    w[o:] = 0.0
    m, v = np.zeros_like(w), np.zeros_like(w)
    l0 = mlp_loss(w, recs, ALPHA, emulate=True) / N
    for t in range(1, 151):
        g, _ = mlp_backward(w, recs, ALPHA, 1.0 / N, emulate=True)
        mlp_adam(w, m, v, g, t, 1e-2)
    l1 = mlp_loss(w, recs, ALPHA, emulate=True) / N
    # the best mixture (alpha/32 + (1 - alpha) on the target) has loss -log(32 (alpha/32 + 1 - alpha))
    best = -np.log(32 * (ALPHA / 32 + 1 - ALPHA))
    assert l0 == pytest.approx(0.0, abs=1e-6)   # uniform g: log 1
    assert l1 < best + 0.1 * abs(best)
    z, cdf = mlp_forward(w, recs[:, :8], ALPHA, emulate=True)
    assert (z.argmax(axis=1) == target).mean() > 0.95


def test_one_training_ray_per_block():
    W, H, BX, BY = 37, 21, 16, 8
    for frame in (0, 1, 7):
        slots = {}
        for y in range(H):
            for x in range(W):
                s = mlp_train_slot(x, y, W, H, BX, BY, frame)
                if s >= 0:
                    assert s not in slots
                    slots[s] = (x, y)
                    assert s == (y // BY) * ((W + BX - 1) // BX) + x // BX
        assert sorted(slots) == list(range(((W + BX - 1) // BX) * ((H + BY - 1) // BY)))


def test_trace_rays_training_records():
    spec = scenes.config0(res=32, width=24, height=16)
    o = Oracle()
    o.set_scene(spec)
    o.set_learned(enable=True, train=False, block=(8, 4), judge=False)
    o.set_learned_weights(mlp_random_weights(11, last=2.0))
    W, H = spec.width, spec.height
    o.prepare(W, H)
    o.run_stage(abi.STAGE_GENERATE_FEATURES)
    o.run_stage(abi.STAGE_LEARNED_PROPOSAL)
    o.run_stage(abi.STAGE_TRACE_RAYS)
    recs = o.read(abi.BUF_TRAIN_RECORDS, 3 * 4 * MLP_REC * 4).view(np.float32).reshape(-1, MLP_REC)
    cdf = o.read(abi.BUF_LEARNED_CDF, W * H * 32 * 2).view(np.uint16).reshape(H, W, 32)
    frame = 0
    hit = 0
    for slot in range(len(recs)):
        by, bx = divmod(slot, 3)
        px = [(x, y) for y in range(by * 4, by * 4 + 4) for x in range(bx * 8, bx * 8 + 8)
              if mlp_train_slot(x, y, W, H, 8, 4, frame) == slot]
        assert len(px) == 1
        x, y = px[0]
        raw = o.learned_raw(x, y, W, H)
        r = recs[slot]
        if raw[6] > 0:
            np.testing.assert_array_equal(r[:7], raw[:7])
            assert r[7] == pytest.approx(r[8:40].sum(), rel=1e-5)
            assert np.all(r[8:40] >= 0) and np.all(r[40:72] >= 0)
            hit += r[7] > 0
        else:
            assert r[7] == 0
        # the proposal stage wrote this pixel's CDF from the network
        _, want = mlp_forward(o.learned_weights(), raw[None], ALPHA, emulate=True)
        np.testing.assert_array_equal(cdf[y, x], want[0])
    assert hit >= 6


def _frames(spec_params, n, reference, learned_weights=None):
    spec = scenes.config0(res=32, width=16, height=12)
    spec.params.update(finalVisibilityTrackingMethod=3, finalLightTrackingMethod=3)
    spec.params.update(spec_params)
    if reference:
        spec.params.update(useReference=1, baselineSamplePerPixel=4)
    o = Oracle()
    o.set_scene(spec)
    if learned_weights is not None:
        o.set_learned(enable=True, train=False, judge=False)
        o.set_learned_weights(learned_weights)
    out = np.stack([o.execute(16, 12, nthreads=8)[..., :3].astype(np.float64) for _ in range(n)])
    o.close()
    return out


@pytest.mark.parametrize("linear", [0, 1])
def test_learned_source_pdf_keeps_the_pass_unbiased(linear):
    """RIS + one spatial round with every initial distance drawn through a strongly
    non-uniform network (point and trilinear trackers): per-pixel means against
    the volume path tracer, as tests/test_oracle_unbiased.py does for the plain pass."""
    from test_oracle_unbiased import compare
    n = 2000
    params = dict(enableTemporalReuse=0, initialVisibilityUseLinearSampler=linear)
    pt = _frames({}, n // 4, reference=True)
    ps = _frames(params, n, reference=False, learned_weights=mlp_random_weights(12, last=3.0))
    z, zi, rel = compare(pt, ps, 1)
    chi2 = float((z ** 2).mean())
    print(f"learned linear={linear}: image z {zi:.2f} (rel {rel:+.4f}), chi2/dof {chi2:.3f}, max|z| {np.abs(z).max():.2f}")
    assert abs(zi) < 4.5
    assert chi2 < 1.4
    assert np.abs(z).max() < 6.0
