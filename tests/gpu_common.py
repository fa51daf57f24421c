"""Shared helpers of the GPU parity tests: run the HIP pass (through the C ABI)
and the CPU oracle on the same scene, stage by stage, and compare."""
import sys

import numpy as np

from oracle.binding import Oracle
from volumetricrestirrelease_amd import abi, scenes
from volumetricrestirrelease_amd.volumetric_restir import FEATURES_DTYPE, RESERVOIR_DTYPE, VolumetricReSTIR

FLT_MAX = 3.4e38


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_scene(res=64, width=64, height=48, fp16=False, **params):
    spec = scenes.config0(res=res, width=width, height=height)
    if fp16:
        spec.brick_format = abi.BRICK_FP16
    spec.params.update(params)
    return spec


def rel(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    both = (a >= FLT_MAX) & (b >= FLT_MAX)
    r = np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-12)
    r[both] = 0.0
    return r


def compare_reservoirs(ro, rg, rtol=2e-3):
    """Fraction of pixels whose reservoirs match: identical integer fields and M,
    depth / p_y / runningSum within rtol."""
    a = ro.view(RESERVOIR_DTYPE)
    b = rg.view(RESERVOIR_DTYPE)
    same = (a["lightID"] == b["lightID"]) & (a["sampledPixel"] == b["sampledPixel"]) & (a["M"] == b["M"])
    close = (rel(a["depth"], b["depth"]) <= rtol) & (rel(a["p_y"], b["p_y"]) <= rtol) & \
            (rel(a["runningSum"], b["runningSum"]) <= rtol)
    return float(np.mean(same & close)), same & close


class Pair:
    """Oracle and HIP pass on the same scene."""

    def __init__(self, spec):
        self.spec = spec
        self.o = Oracle()
        self.o.set_scene(spec)
        self.g = VolumetricReSTIR()
        self.g.setScene(spec)
        self.W, self.H = spec.width, spec.height
        self.n = self.W * self.H
        self.B = spec.params.get("maxBounces", 1)
        self.nres = self.n * 36
        self.nex = self.n * max(self.B - 1, 0) * 12

    def both(self, fn):
        fn(self.o)
        fn(self.g)

    def inject(self, buf, nbytes):
        if nbytes:
            self.g.write(buf, self.o.read(buf, nbytes))

    def begin(self, frame):
        self.o.set_camera(self.spec, frame)
        self.g.setCamera(self.spec, frame)
        self.o.prepare(self.W, self.H, 8)
        self.g.prepare_frame(self.W, self.H)

    def sync_history(self):
        self.inject(abi.BUF_TEMPORAL_RESERVOIRS, self.nres)
        self.inject(abi.BUF_TEMPORAL_FEATURES, self.n * 8)
        self.inject(abi.BUF_TEMPORAL_EXTRA, self.nex)

    def stage(self, st, rnd=0):
        import torch
        self.o.run_stage(st, rnd)
        self.g.run_stage(st, rnd)
        torch.cuda.synchronize()

    def end(self):
        self.o.end_frame()
        self.g.end_frame()

    def color(self):
        co = self.o.read(abi.BUF_COLOR, self.n * 16).view(np.float32).reshape(-1, 4)
        cg = self.g.read(abi.BUF_COLOR, self.n * 16).view(np.float32).reshape(-1, 4)
        return co, cg
