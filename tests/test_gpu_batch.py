"""Batched persistent launches: passes of uneven size (views, volumes, light counts
drawn from a seed, as the batched workload draws them) rendered together -- one
persistent launch per step over every pass's tiles, the frame's launches captured
into a HIP graph and replayed -- give the same images as each pass rendered alone
(the first frame bit for bit, later ones to a few ulp), on every frame (the graph's
replays included), with a small persistent grid so that each workgroup drains many
queue items, and after the autotuner has picked the queue shape."""
import numpy as np
import pytest

from gpu_common import say
from volumetricrestirrelease_amd import abi, scenes
from volumetricrestirrelease_amd.volumetric_restir import (VolumetricReSTIR, VolumetricReSTIRBatch,
                                                           VolumetricReSTIRError)

pytestmark = pytest.mark.gpu

FRAMES = 3


def batch_specs():
    return scenes.config2(seed=5, count=5, min_res=24, max_res=112, min_grid=64, max_grid=128, max_lights=8)


def make_pass(spec, setup=None):
    """A pass on the spec. A batch runs TraceRays as one kernel, so the passes
    rendered alone do too (the split TraceRays agrees with it to float rounding
    only, which would hide whether frame 0 is bit for bit)."""
    g = VolumetricReSTIR()
    g.setScene(spec)
    assert abi.load_library().vrestir_set_tuning(g.h, abi.TUNE_SPLIT_TRACE, 0) == 0
    if setup:
        setup(g)
    return g


def singles(specs, setup=None, before_frame=None):
    import torch
    out = []
    for i, spec in enumerate(specs):
        g = make_pass(spec, setup(i) if setup else None)
        img = torch.zeros((spec.height, spec.width, 4), device="cuda")
        frames = []
        for f in range(FRAMES):
            g.setCamera(spec, f)
            if before_frame:
                before_frame(g, f)
            g.execute(img)
            torch.cuda.synchronize()
            frames.append(img.cpu().numpy().copy())
        out.append(frames)
    return out


def batched(specs, tuning=None, autotune=False, setup=None, before_frame=None):
    import torch
    passes = [make_pass(spec, setup(i) if setup else None) for i, spec in enumerate(specs)]
    b = VolumetricReSTIRBatch(passes)
    outs = [torch.zeros((s.height, s.width, 4), device="cuda") for s in specs]
    if tuning:
        b.setTuning(*tuning)
    if autotune:
        for g, spec in zip(passes, specs):
            g.setCamera(spec, 0)
        ms = b.autotune(outs, frames=1)
        say(f"   autotuner: {ms:.3f} ms per batched frame with (grid, tiles per item) = {b.tuning()}")
    res = [[] for _ in specs]
    for f in range(FRAMES):
        for g, spec in zip(passes, specs):
            g.setCamera(spec, f)
            if before_frame:
                before_frame(g, f)
        b.execute(outs)
        torch.cuda.synchronize()
        for i, o in enumerate(outs):
            res[i].append(o.cpu().numpy().copy())
    return res


def compare(a, b, specs):
    """Frame 0 bit for bit; later frames (temporal reuse on) within 1e-4 relative on
    >= 99.9 % of pixels: the batched kernels inline the same stage bodies into other
    kernels, and the compiler may fuse a multiply-add differently there (a few ulp)."""
    for i, spec in enumerate(specs):
        for f in range(FRAMES):
            x, y = a[i][f][..., :3].reshape(-1, 3), b[i][f][..., :3].reshape(-1, 3)
            r = np.abs(x - y) / np.maximum(np.maximum(np.abs(x), np.abs(y)), 1e-12)
            r[(x == y)] = 0.0
            r = r.max(axis=1)
            exact = float(np.mean(r == 0.0))
            close = float(np.mean(r <= 1e-4))
            say(f"   pass {i} ({spec.name}) frame {f}: identical {exact:.4f}, within 1e-4 {close:.4f}, "
                f"max rel {r.max():.2e}")
            if f == 0:
                assert np.array_equal(a[i][f], b[i][f]), (i, f)
            assert close >= 0.999, (i, f)
            assert abs(x.mean() - y.mean()) <= 1e-5 * x.mean(), (i, f)
        assert a[i][-1][..., :3].mean() > 0


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_passes_rendered_alone(use_graph):
    specs = batch_specs()
    say("case: batch of " + ", ".join(f"{s.width}x{s.height}/{s.density.shape[0]}^3/{len(s.lights)}L" for s in specs)
        + f", 64 persistent workgroups x 2 tiles per item, graph {use_graph}, {FRAMES} frames")
    compare(singles(specs), batched(specs, tuning=(64, 2, use_graph)), specs)


def test_batch_with_emissive_triangles_matches_passes_rendered_alone():
    specs = batch_specs()[:3]
    for s in specs:
        s.surfaces = scenes.emissive_panels()
        s.params.update(useEmissiveLights=1)
    say("case: batch of 3 passes lit by emissive triangles too, graph on, " + f"{FRAMES} frames")
    compare(singles(specs), batched(specs, tuning=(64, 2, True)), specs)


def test_batch_after_autotune_matches_passes_rendered_alone():
    specs = batch_specs()
    say("case: autotuned batch, then 3 frames from a reset history")
    compare(singles(specs), batched(specs, autotune=True), specs)


def test_batch_refuses_passes_with_other_kernels():
    import torch
    say("case: passes with different path lengths in one batch")
    specs = batch_specs()[:2]
    specs[1].params.update(maxBounces=2)
    passes = []
    for spec in specs:
        g = VolumetricReSTIR()
        g.setScene(spec)
        passes.append(g)
    b = VolumetricReSTIRBatch(passes)
    outs = [torch.zeros((s.height, s.width, 4), device="cuda") for s in specs]
    with pytest.raises(VolumetricReSTIRError, match="kernel options"):
        b.execute(outs)


def test_batch_of_learned_passes_matches_passes_rendered_alone():
    """Each pass proposes with and trains its own network (per-pass launches of the
    network stages between the batched steps, no graph)."""
    specs = batch_specs()
    say("case: batch of 5 passes with the learned proposal on (online training), 64 x 2, 3 frames")

    def setup(i):
        return lambda g: g.setLearned(enable=1, judge=0, trainBlockX=8, trainBlockY=4, seed=11 + i)

    compare(singles(specs, setup=setup), batched(specs, tuning=(64, 2, True), setup=setup), specs)


def test_batch_of_animated_passes_matches_passes_rendered_alone():
    """Moving volumes with previous-frame grids in temporal reuse: the batch runs the
    one-kernel temporal form for every pass."""
    seqs = [scenes.moving_sequence(res=32, width=40, height=32, seed=9),
            scenes.moving_sequence(res=32, shift=1, width=56, height=24, seed=4)]
    specs = [s0 for s0, _ in seqs]
    say("case: batch of 2 moving 32^3 sequences (40x32, 56x24) with usePrevVolumeForReproj, 3 frames")

    def setup(i):
        def add(g):
            for s in seqs[i][1][1:]:
                g.addVolumeFrame(s)
        return add

    def advance(g, f):
        g.setVolumeFrame(f, animating=f > 0)

    compare(singles(specs, setup=setup, before_frame=advance),
            batched(specs, tuning=(64, 1, True), setup=setup, before_frame=advance), specs)


def test_batch_refuses_mixed_learned_passes():
    import torch
    say("case: a learned and a plain pass in one batch")
    specs = batch_specs()[:2]
    passes = [make_pass(specs[0], lambda g: g.setLearned(enable=1)), make_pass(specs[1])]
    b = VolumetricReSTIRBatch(passes)
    outs = [torch.zeros((s.height, s.width, 4), device="cuda") for s in specs]
    with pytest.raises(VolumetricReSTIRError, match="learned proposal"):
        b.execute(outs)
