"""The learned free-flight proposal on the GPU against the CPU oracle, through
the C ABI:

- the BF16 MFMA network: logits within 1e-5 of their scale on most rays (bf16
  rounding of an activation that sits on a rounding boundary moves a few by
  up to 1e-2), the 16-bit CDF identical on >= 99 % of rays and never off by
  more than 2;
- the training kernels' gradient (summed over records, BF16 operands, FP32
  accumulation) against the oracle's gradient at the same rounding points:
  cosine >= 0.9999 and max deviation <= 1e-2 of each tensor's largest
  entry; the statistics within 1e-4;
- one Adam step against the oracle's Adam on the oracle's gradient;
- TraceRays with the learned proposal (the oracle given the GPU's CDF): the
  stage-parity bar of the plain pass, and the training records' raw inputs
  within 1e-5 and weights within 2e-3 where the reservoirs match;
- the whole pass with a strongly non-uniform network still converges to the
  volume path tracer (expectation check), and online training lowers the
  network's estimated second moment of the RIS weights below the analytic
  proposal's.
"""
import numpy as np
import pytest

from gpu_common import Pair, compare_reservoirs, rel, say, small_scene
from oracle.binding import (MLP_OFFSETS, MLP_PARAMS, MLP_REC, Oracle, mlp_adam, mlp_backward, mlp_forward,
                            mlp_random_weights)
from volumetricrestirrelease_amd import abi
from volumetricrestirrelease_amd.volumetric_restir import VolumetricReSTIR

pytestmark = pytest.mark.gpu

ALPHA = 0.25


def random_raw(rng, n):
    raw = np.zeros((n, 8), np.float32)
    raw[:, :3] = rng.uniform(0, 1, size=(n, 3))
    d = rng.normal(size=(n, 3))
    raw[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    raw[:, 6] = rng.uniform(0.05, 1.0, size=n)
    return raw


def random_records(rng, n):
    rec = np.zeros((n, MLP_REC), np.float32)
    rec[:, :8] = random_raw(rng, n)
    w = rng.gamma(0.5, size=(n, 32)).astype(np.float32)
    rec[:, 8:40] = w
    rec[:, 40:72] = w * w * rng.uniform(0.5, 2.0, size=w.shape).astype(np.float32)
    rec[:, 7] = w.sum(axis=1)
    rec[::7, 7] = 0.0  # records without weight contribute nothing
    return rec


def pass_with_network(w, spec=None):
    g = VolumetricReSTIR()
    if spec is not None:
        g.setScene(spec)
    g.setLearned(enable=1, judge=0)
    g.setLearnedWeights(w)
    return g


def test_network_forward_matches_oracle():
    import torch
    say("case: 1000 random rays through a random network (bf16 MFMA forward, 16-bit CDF)")
    rng = np.random.default_rng(0)
    w = mlp_random_weights(1, last=2.0)
    raw = random_raw(rng, 1000)
    g = pass_with_network(w)
    cdf = torch.zeros((1000, 32), dtype=torch.int16, device="cuda")
    logits = torch.zeros((1000, 32), dtype=torch.float32, device="cuda")
    g.learnedForward(torch.from_numpy(raw).cuda(), cdf, logits)
    torch.cuda.synchronize()
    zg = logits.cpu().numpy()
    cg = cdf.cpu().numpy().view(np.uint16)
    zo, co = mlp_forward(w, raw, ALPHA, emulate=True)
    scale = np.abs(zo).max()
    d = np.abs(zg - zo) / scale
    say(f"  logits: median |d| {np.median(d):.2e}, max {d.max():.2e} of scale {scale:.3f}")
    assert np.median(d) < 1e-5
    assert d.max() < 1e-2
    same = np.all(cg == co, axis=1)
    dc = np.abs(cg.astype(np.int64) - co.astype(np.int64)).max()
    say(f"  CDF identical on {same.mean():.4f} of rays, max deviation {dc}")
    assert same.mean() >= 0.99
    assert dc <= 2


def test_network_gradient_matches_oracle():
    import torch
    say("case: gradient of 200 random training records (BF16 MFMA backward) against the oracle")
    rng = np.random.default_rng(1)
    w = mlp_random_weights(2, last=1.5)
    recs = random_records(rng, 200)
    g = pass_with_network(w)
    out = torch.zeros(MLP_PARAMS + 3, dtype=torch.float32, device="cuda")
    g.learnedGradient(torch.from_numpy(recs).cuda(), out, scale=1.0 / 200)
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.float64)
    want, st = mlp_backward(w, recs, ALPHA, 1.0 / 200, emulate=True)
    for k, (o, sh) in MLP_OFFSETS.items():
        n = int(np.prod(sh))
        a, b = got[o:o + n], want[o:o + n]
        cos = np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b))
        dev = np.abs(a - b).max() / np.abs(b).max()
        say(f"  {k}: cosine {cos:.6f}, max deviation {dev:.2e}")
        assert cos >= 0.9999, k
        assert dev <= 1e-2, k
    np.testing.assert_allclose(got[MLP_PARAMS:], st, rtol=1e-4)


def test_adam_step_matches_oracle():
    import torch
    say("case: one Adam step on 96 records")
    rng = np.random.default_rng(2)
    w = mlp_random_weights(3, last=1.5)
    recs = random_records(rng, 96)
    g = pass_with_network(w)
    g.learnedTrain(torch.from_numpy(recs).cuda())
    got = g.learnedWeights()
    st = g.learnedStats()
    grad, sto = mlp_backward(w, recs, ALPHA, 1.0 / 96, emulate=True)
    want = w.copy()
    m, v = np.zeros_like(w), np.zeros_like(w)
    mlp_adam(want, m, v, grad.astype(np.float32), 1, 2e-3)
    # Adam's first step moves every parameter by about lr sign(g): compare where the
    # gradient is clearly non-zero, and bound the step everywhere
    big = np.abs(grad) > 1e-3 * np.abs(grad).max()
    say(f"  {big.mean():.3f} of the parameters with a clear gradient; max |step| {np.abs(got - w).max():.3e}")
    assert np.abs(got - want)[big].max() < 2e-5
    assert np.abs(got - w).max() <= 2e-3 * 1.001
    assert st[4] == 1
    np.testing.assert_allclose(st[:3], sto, rtol=1e-4)


def test_learned_trace_rays_matches_oracle():
    say("case: TraceRays with the learned proposal, 32^3, 48x32, b=2 and b=3, oracle given the GPU's CDF")
    for params in (dict(), dict(maxBounces=3, initialM=8, initialVisibilityUseLinearSampler=1)):
        spec = small_scene(res=32, width=48, height=32, **params)
        pair = Pair(spec)
        w = mlp_random_weights(4, last=2.0)
        pair.g.setLearned(enable=1, judge=0, trainBlockX=8, trainBlockY=4)
        pair.g.setLearnedWeights(w)
        pair.o.set_learned(enable=True, train=False, block=(8, 4), judge=False)
        pair.o.set_learned_weights(w)
        pair.begin(0)
        pair.stage(abi.STAGE_GENERATE_FEATURES)
        pair.inject(abi.BUF_FEATURES, pair.n * 8)
        pair.g.run_stage(abi.STAGE_LEARNED_PROPOSAL)
        pair.o.run_stage(abi.STAGE_LEARNED_PROPOSAL)
        cg = pair.g.read(abi.BUF_LEARNED_CDF, pair.n * 64).view(np.uint16).reshape(-1, 32)
        co = pair.o.read(abi.BUF_LEARNED_CDF, pair.n * 64).view(np.uint16).reshape(-1, 32)
        same = np.all(cg == co, axis=1).mean()
        say(f"  {params}: proposal CDF identical to the oracle's on {same:.4f} of pixels")
        assert same >= 0.99
        pair.o.write(abi.BUF_LEARNED_CDF, cg)
        pair.stage(abi.STAGE_TRACE_RAYS)
        m, match = compare_reservoirs(pair.o.read(abi.BUF_RESERVOIRS0, pair.nres),
                                      pair.g.read(abi.BUF_RESERVOIRS0, pair.nres))
        say(f"  reservoirs match on {m:.4f} of pixels")
        assert m >= 0.995
        nrec = 6 * 8
        rg = pair.g.read(abi.BUF_TRAIN_RECORDS, nrec * MLP_REC * 4).view(np.float32).reshape(nrec, MLP_REC)
        ro = pair.o.read(abi.BUF_TRAIN_RECORDS, nrec * MLP_REC * 4).view(np.float32).reshape(nrec, MLP_REC)
        assert np.abs(rg[:, :7] - ro[:, :7]).max() <= 1e-5
        wrel = rel(rg[:, 7:], ro[:, 7:]).max(axis=1)
        say(f"  training records: weights within 2e-3 on {(wrel <= 2e-3).mean():.3f} of rays")
        assert (wrel <= 2e-3).mean() >= 0.9
        assert (rg[:, 7] > 0).sum() >= 10


def test_learned_pass_expectation_equals_path_tracer():
    """tests/test_gpu_unbiased.py's check with every initial distance drawn
    through a strongly non-uniform network."""
    import torch
    from test_gpu_unbiased import compare, gpu_frames
    say("case: expectation of the pass with a random learned proposal vs the path tracer")
    w = mlp_random_weights(5, last=3.0)

    def setup(g):
        g.setLearned(enable=1, judge=0, train=0)
        g.setLearnedWeights(w)

    n = 3000
    pt = gpu_frames({}, n // 4, reference=True)
    ps = gpu_frames(dict(enableTemporalReuse=0), n, reference=False, setup=setup)
    z, zi, relm = compare(pt, ps, 1)
    chi2 = float((z ** 2).mean())
    say(f"  image z {zi:.2f} (rel {relm:+.4f}), chi2/dof {chi2:.3f}, max|z| {np.abs(z).max():.2f}")
    assert abs(zi) < 4.5
    assert chi2 < 1.4
    assert np.abs(z).max() < 6.5


def test_online_training_beats_the_analytic_proposal():
    """Trains on the pass's own candidates for 40 frames of a 160x90 config0 view:
    the loss falls and the network's estimated second moment of the normalised RIS
    weights ends below the analytic proposal's (the judge then keeps it in use)."""
    import torch
    from volumetricrestirrelease_amd import scenes
    say("case: online training, config0 scene at 160x90, 40 frames")
    spec = scenes.config0(res=128, width=160, height=90)
    g = VolumetricReSTIR()
    g.setScene(spec)
    g.setLearned(enable=1, trainBlockX=4, trainBlockY=4, learningRate=3e-3, trainStepsPerFrame=4)
    out = torch.zeros((90, 160, 4), device="cuda")
    hist = []
    for f in range(40):
        g.execute(out)
        hist.append(g.learnedStats())
    hist = np.array(hist)
    for f in (0, 1, 5, 10, 20, 39):
        say(f"  frame {f}: loss {hist[f, 0]:.4f}, second moments analytic {hist[f, 1]:.4f} network "
            f"{hist[f, 2]:.4f}, in use {hist[f, 3]:.0f}")
    assert hist[-5:, 0].mean() < hist[0, 0] - 0.01
    assert hist[-5:, 2].mean() < hist[-5:, 1].mean()
    assert hist[-1, 3] == 1
