"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
always as the checker; the product package never imports this module.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
REF_SPLITMIX = os.path.join(HERE, "_ref", "libref_splitmix64.so")
REF_XOSHIRO = os.path.join(HERE, "_ref", "libref_xoshiro128ss.so")

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
    from volumetricrestirrelease_amd import abi
    L = ctypes.CDLL(LIB)
    sig = {
        "orc_create": (c_void_p, []),
        "orc_destroy": (None, [c_void_p]),
        "vrestir_default_params_oracle": (None, [POINTER(abi.Params)]),
        "orc_set_params": (c_int, [c_void_p, POINTER(abi.Params)]),
        "orc_set_volume": (c_int, [c_void_p, POINTER(abi.VolumeDesc)]),
        "orc_set_lights": (c_int, [c_void_p, POINTER(abi.PointLight), c_int, POINTER(c_float)]),
        "orc_set_surfaces": (c_int, [c_void_p, POINTER(abi.SurfacesDesc)]),
        "orc_sample_lights": (None, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
        "orc_emissive_eval": (None, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
        "orc_surface_hit": (None, [c_void_p, c_int, c_void_p, c_void_p]),
        "orc_surface_shade": (None, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
        "orc_set_camera": (c_int, [c_void_p, POINTER(abi.Camera), c_int, c_int]),
        "orc_execute": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_int, c_int]),
        "orc_prepare_frame": (c_int, [c_void_p, c_int, c_int, c_int]),
        "orc_run_stage": (c_int, [c_void_p, c_int, c_int]),
        "orc_set_rows": (c_int, [c_void_p, c_int, c_int]),
        "orc_set_slot": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_void_p, c_int]),
        "orc_end_frame": (c_int, [c_void_p]),
        "orc_read_buffer": (c_int, [c_void_p, c_int, c_void_p, c_size_t]),
        "orc_write_buffer": (c_int, [c_void_p, c_int, c_void_p, c_size_t]),
        "orc_frame_count": (c_int, [c_void_p]),
        "orc_total_round_id": (c_int, [c_void_p]),
        "orc_stage_seconds": (None, [c_void_p, POINTER(c_double)]),
        "orc_track_batch": (c_int, [c_void_p, c_int, c_int, c_int, c_float, c_int, c_int, c_void_p, c_void_p,
                                    c_void_p]),
        "orc_rng_stream": (None, [c_uint32, c_uint32, c_uint32, c_int, c_void_p]),
        "orc_morton": (c_uint32, [c_uint32, c_uint32]),
        "orc_splitmix64": (None, [c_uint64, c_int, c_void_p]),
        "orc_xoshiro128ss": (None, [c_void_p, c_int, c_void_p]),
        "orc_round_to_half": (c_float, [c_float]),
        "orc_neighbor_offsets": (None, [c_int, c_float, c_int, c_int, c_void_p]),
        "orc_grid_info": (c_int, [c_void_p, c_int, POINTER(c_int), POINTER(c_float), POINTER(c_int)]),
        "orc_brick_bounds": (c_int, [c_void_p, c_int, c_int, c_int, c_int, POINTER(c_float)]),
        "orc_voxel": (c_float, [c_void_p, c_int, c_int, c_int, c_int]),
        "orc_final_costs": (c_int, [c_void_p, c_void_p]),
        "orc_sample_phase": (c_float, [c_float, c_void_p, c_float, c_float, c_void_p]),
        "orc_phase_hg": (c_float, [c_float, c_float]),
        "orc_oct_to_dir": (None, [c_float, c_float, c_void_p]),
        "orc_light_uv_encode": (None, [c_void_p, c_void_p, POINTER(c_int)]),
        "orc_light_uv_decode": (None, [c_void_p, c_int, c_void_p]),
        "orc_wi_dist_encode": (None, [c_void_p, c_void_p]),
        "orc_wi_dist_decode": (None, [c_void_p, c_void_p]),
        "orc_pack_sampled_pixel": (c_int, [c_int, c_int, c_int]),
        "orc_unpack_sampled_pixel": (None, [c_int, POINTER(c_int), POINTER(c_int)]),
        "orc_wrs": (c_int, [c_int, c_void_p, c_void_p, c_float, c_uint32, c_uint32, c_uint32, POINTER(c_float),
                            POINTER(c_float)]),
        "orc_camera_ray": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p]),
        "orc_reproject": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p]),
        "orc_trace_bricks": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_void_p]),
        "orc_set_temperature": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_float, c_float,
                                        c_float]),
        "orc_set_blackbody_lut": (c_int, [c_void_p, c_void_p, c_int]),
        "orc_add_volume_frame": (c_int, [c_void_p, POINTER(abi.VolumeDesc)]),
        "orc_set_volume_frame": (c_int, [c_void_p, c_int, c_int]),
        "orc_set_freeze_frame": (c_int, [c_void_p, c_int]),
        "orc_set_env_map": (c_int, [c_void_p, c_int, c_int, c_void_p, c_float, c_void_p, c_void_p, c_int, c_int]),
        "orc_set_env_rotation": (c_int, [c_void_p, c_void_p]),
        "orc_set_env_intensity": (c_int, [c_void_p, c_float, c_void_p]),
        "orc_env_rotation": (None, [c_void_p, c_int, c_void_p, c_void_p]),
        "orc_env_eval": (None, [c_void_p, c_void_p, c_void_p]),
        "orc_env_sample": (None, [c_void_p, c_float, c_float, c_void_p]),
        "orc_env_pdf": (c_float, [c_void_p, c_void_p]),
        "orc_env_importance": (c_int, [c_void_p, c_int, c_void_p]),
        "orc_latlong": (None, [c_void_p, c_void_p]),
        "orc_dir_to_oct": (None, [c_void_p, c_void_p]),
        "orc_emission": (None, [c_void_p, c_void_p, c_void_p]),
        "orc_blackbody": (None, [c_void_p, c_float, c_void_p]),
        "orc_set_learned": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_float, c_float, c_int]),
        "orc_learned_set_weights": (c_int, [c_void_p, c_void_p, c_int]),
        "orc_learned_get_weights": (c_int, [c_void_p, c_void_p, c_int]),
        "orc_learned_stats": (None, [c_void_p, c_void_p]),
        "orc_learned_raw": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p]),
        "orc_mlp_encode": (None, [c_void_p, c_void_p]),
        "orc_mlp_forward": (None, [c_void_p, c_void_p, c_int, c_float, c_int, c_void_p, c_void_p]),
        "orc_mlp_loss": (c_double, [c_void_p, c_void_p, c_int, c_float, c_int]),
        "orc_mlp_backward": (None, [c_void_p, c_void_p, c_int, c_float, c_float, c_int, c_void_p, c_void_p]),
        "orc_mlp_adam": (None, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_float]),
        "orc_mlp_sample": (c_float, [c_void_p, c_float, POINTER(c_int), POINTER(c_float)]),
        "orc_mlp_mass_above": (c_float, [c_void_p, c_float]),
        "orc_mlp_train_slot": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class Oracle:
    """The pass on the CPU: same stage/buffer identities as the product's C ABI."""

    def __init__(self):
        self.L = lib()
        self.h = self.L.orc_create()

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        rc = self.L.orc_set_params(self.h, ctypes.byref(p))
        assert rc == 0, rc

    def set_scene(self, spec, frame=0):
        d = spec.volume_desc()
        assert self.L.orc_set_volume(self.h, ctypes.byref(d)) == 0
        if spec.temperature is not None:
            lut, rows = spec.blackbody_array()
            assert self.L.orc_set_blackbody_lut(self.h, lut.ctypes.data, rows) == 0
        arr, n, env = spec.light_array()
        assert self.L.orc_set_lights(self.h, arr, n, env) == 0
        sd, _keep = spec.surfaces_desc()
        assert self.L.orc_set_surfaces(self.h, ctypes.byref(sd) if sd is not None else None) == 0
        if spec.env_map is not None:
            self.set_env_map(spec.env_map, spec.env_intensity, spec.env_tint, spec.env_rotation_at(0),
                             spec.env_importance_dim, spec.env_importance_spp)
        else:
            self.set_env_map(None)
        self.set_camera(spec, frame)
        self.set_params(spec.make_params())

    def sample_lights(self, kinds, points: np.ndarray, seeds: np.ndarray) -> np.ndarray:
        """sampleSceneLights at points (kinds: 1 env, 2 point, 4 emissive): (n, 8)."""
        p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        sd = np.ascontiguousarray(seeds, dtype=np.uint32).reshape(-1, 3)
        out = np.zeros((p.shape[0], 8), np.float32)
        self.L.orc_sample_lights(self.h, kinds, p.shape[0], p.ctypes.data, sd.ctypes.data, out.ctypes.data)
        return out

    def surface_hit(self, rays: np.ndarray) -> np.ndarray:
        """Closest and any hit over [0, tMax] per ray (n, 8): (n, 8) {t, tri, bary, occluded, 0...}."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        out = np.zeros((r.shape[0], 8), np.float32)
        self.L.orc_surface_hit(self.h, r.shape[0], r.ctypes.data, out.ctypes.data)
        return out

    def surface_shade(self, variant, rays: np.ndarray, seeds: np.ndarray) -> np.ndarray:
        """Shading at each ray's closest hit: variant 0 a BSDF sample, 1 its evaluation: (n, 8)."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        sd = np.ascontiguousarray(seeds, dtype=np.uint32).reshape(-1, 3)
        out = np.zeros((r.shape[0], 8), np.float32)
        self.L.orc_surface_shade(self.h, variant, r.shape[0], r.ctypes.data, sd.ctypes.data, out.ctypes.data)
        return out

    def emissive_eval(self, points: np.ndarray, tris: np.ndarray, uvs: np.ndarray) -> np.ndarray:
        """The emissive-triangle light evaluation at points: (n, 8)."""
        p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1)
        u = np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2)
        out = np.zeros((p.shape[0], 8), np.float32)
        self.L.orc_emissive_eval(self.h, p.shape[0], p.ctypes.data, t.ctypes.data, u.ctypes.data, out.ctypes.data)
        return out

    def set_camera(self, spec, frame=0):
        c = spec.camera_struct(frame)
        assert self.L.orc_set_camera(self.h, ctypes.byref(c), spec.width, spec.height) == 0
        if spec.env_map is not None:
            self.set_env_rotation(spec.env_rotation_at(frame))

    # ---- animated volumes
    def add_volume_frame(self, desc):
        d = desc.volume_desc() if hasattr(desc, "volume_desc") else desc
        idx = self.L.orc_add_volume_frame(self.h, ctypes.byref(d))
        assert idx > 0, idx
        return idx

    def set_volume_frame(self, index, animating=True):
        assert self.L.orc_set_volume_frame(self.h, index, 1 if animating else 0) == 0

    def set_freeze_frame(self, freeze):
        assert self.L.orc_set_freeze_frame(self.h, 1 if freeze else 0) == 0

    # ---- environment map and emission
    def set_env_map(self, rgb, intensity=1.0, tint=(1.0, 1.0, 1.0), rotation=(0.0, 0.0, 0.0), dim=0, spp=0):
        if rgb is None:
            assert self.L.orc_set_env_map(self.h, 0, 0, None, 1.0, None, None, 0, 0) == 0
            return
        arr = np.ascontiguousarray(rgb, dtype=np.float32)
        t = (c_float * 3)(*tint)
        r = (c_float * 3)(*rotation)
        assert self.L.orc_set_env_map(self.h, arr.shape[1], arr.shape[0], arr.ctypes.data, intensity, t, r, dim,
                                      spp) == 0

    def set_env_rotation(self, deg):
        assert self.L.orc_set_env_rotation(self.h, (c_float * 3)(*deg)) == 0

    def env_rotation(self, prev=False):
        w = np.zeros(9, dtype=np.float32)
        l = np.zeros(9, dtype=np.float32)
        self.L.orc_env_rotation(self.h, 1 if prev else 0, w.ctypes.data, l.ctypes.data)
        return w.reshape(3, 3), l.reshape(3, 3)

    def env_eval(self, d):
        d = np.ascontiguousarray(d, dtype=np.float32)
        out = np.zeros(3, dtype=np.float32)
        self.L.orc_env_eval(self.h, d.ctypes.data, out.ctypes.data)
        return out

    def env_sample(self, u, v):
        out = np.zeros(8, dtype=np.float32)
        self.L.orc_env_sample(self.h, u, v, out.ctypes.data)
        return out

    def env_pdf(self, d):
        d = np.ascontiguousarray(d, dtype=np.float32)
        return float(self.L.orc_env_pdf(self.h, d.ctypes.data))

    def env_importance(self, mip):
        n = self.L.orc_env_importance(self.h, mip, None)
        assert n > 0
        out = np.zeros((n, n), dtype=np.float32)
        self.L.orc_env_importance(self.h, mip, out.ctypes.data)
        return out

    def set_temperature(self, t, index_to_world, fp16, le_scale, cutoff, scale):
        t = np.ascontiguousarray(t, dtype=np.float32)
        m = np.ascontiguousarray(index_to_world, dtype=np.float32)
        nz, ny, nx = t.shape
        assert self.L.orc_set_temperature(self.h, t.ctypes.data, nx, ny, nz, m.ctypes.data, 1 if fp16 else 0,
                                          le_scale, cutoff, scale) == 0

    def set_blackbody(self, lut):
        lut = np.ascontiguousarray(lut, dtype=np.float32)
        assert self.L.orc_set_blackbody_lut(self.h, lut.ctypes.data, lut.shape[0]) == 0

    def emission(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        out = np.zeros(4, dtype=np.float32)
        self.L.orc_emission(self.h, p.ctypes.data, out.ctypes.data)
        return out

    def blackbody(self, temperature):
        out = np.zeros(3, dtype=np.float32)
        self.L.orc_blackbody(self.h, temperature, out.ctypes.data)
        return out

    def execute(self, W, H, row_begin=0, row_end=0, nthreads=1):
        out = np.zeros((H, W, 4), dtype=np.float32)
        rc = self.L.orc_execute(self.h, W, H, out.ctypes.data, row_begin, row_end, nthreads)
        assert rc == 0, rc
        return out

    def prepare(self, W, H, nthreads=1):
        assert self.L.orc_prepare_frame(self.h, W, H, nthreads) == 0

    def set_slot(self, slot, density, index_to_world, fp16):
        """Replace grid slot `slot` with a dense (nz, ny, nx) grid and its own index -> world."""
        d = np.ascontiguousarray(density, dtype=np.float32)
        m = np.ascontiguousarray(index_to_world, dtype=np.float32)
        nz, ny, nx = d.shape
        assert self.L.orc_set_slot(self.h, slot, d.ctypes.data, nx, ny, nz, m.ctypes.data, int(fp16)) == 0

    def set_rows(self, row_begin, row_end):
        """Restrict the following stages to rows [row_begin, row_end)."""
        assert self.L.orc_set_rows(self.h, row_begin, row_end) == 0

    def run_stage(self, stage, rnd=0):
        assert self.L.orc_run_stage(self.h, stage, rnd) == 0

    def end_frame(self):
        self.L.orc_end_frame(self.h)

    def read(self, buf, nbytes):
        out = np.zeros(nbytes, dtype=np.uint8)
        rc = self.L.orc_read_buffer(self.h, buf, out.ctypes.data, nbytes)
        assert rc == 0, (buf, nbytes, rc)
        return out

    def write(self, buf, data: np.ndarray):
        data = np.ascontiguousarray(data)
        rc = self.L.orc_write_buffer(self.h, buf, data.ctypes.data, data.nbytes)
        assert rc == 0, (buf, data.nbytes, rc)

    def track(self, mode, mip, linear, tstep, nsamples, rays: np.ndarray, seeds: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
        n = rays.shape[0]
        out = np.zeros((n, 8), dtype=np.float32)
        rc = self.L.orc_track_batch(self.h, mode, mip, linear, tstep, nsamples, n, rays.ctypes.data,
                                    seeds.ctypes.data, out.ctypes.data)
        assert rc == 0, rc
        return out

    def stage_seconds(self):
        out = (c_double * 6)()
        self.L.orc_stage_seconds(self.h, out)
        return list(out)

    # ---- learned proposal
    def set_learned(self, enable=True, train=True, block=(16, 8), steps=1, lr=2e-3, alpha=0.25, judge=True):
        rc = self.L.orc_set_learned(self.h, int(enable), int(train), block[0], block[1], steps, lr, alpha, int(judge))
        assert rc == 0, rc

    def set_learned_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert self.L.orc_learned_set_weights(self.h, w.ctypes.data, w.size) == 0

    def learned_weights(self):
        w = np.zeros(MLP_PARAMS, dtype=np.float32)
        assert self.L.orc_learned_get_weights(self.h, w.ctypes.data, w.size) == 0
        return w

    def learned_stats(self):
        out = np.zeros(5, dtype=np.float32)
        self.L.orc_learned_stats(self.h, out.ctypes.data)
        return out

    def learned_raw(self, px, py, W, H):
        out = np.zeros(MLP_RAW, dtype=np.float32)
        assert self.L.orc_learned_raw(self.h, px, py, W, H, out.ctypes.data) == 0
        return out

    def grid_info(self, slot):
        dims = (c_int * 3)()
        mx = c_float()
        nb = c_int()
        rc = self.L.orc_grid_info(self.h, slot, dims, ctypes.byref(mx), ctypes.byref(nb))
        assert rc == 0
        return tuple(dims), mx.value, nb.value


def latlong(d):
    d = np.ascontiguousarray(d, dtype=np.float32)
    out = np.zeros(2, dtype=np.float32)
    lib().orc_latlong(d.ctypes.data, out.ctypes.data)
    return out


def dir_to_oct(d):
    d = np.ascontiguousarray(d, dtype=np.float32)
    out = np.zeros(2, dtype=np.float32)
    lib().orc_dir_to_oct(d.ctypes.data, out.ctypes.data)
    return out


def rng_stream(px, py, sample_number, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_rng_stream(px, py, sample_number, n, out.ctypes.data)
    return out


def ref_splitmix64(seed: int, n: int):
    """The reference's own splitmix64.c (compiled from the reference tree into oracle/_ref)."""
    L = ctypes.CDLL(REF_SPLITMIX)
    L.ref_splitmix64_seed.argtypes = [c_uint64]
    L.ref_splitmix64_next.restype = c_uint64
    L.ref_splitmix64_seed(seed)
    return np.array([L.ref_splitmix64_next() for _ in range(n)], dtype=np.uint64)


def ref_xoshiro128ss(seed, n: int):
    """The reference's own xoshiro128starstar.c (compiled into oracle/_ref)."""
    L = ctypes.CDLL(REF_XOSHIRO)
    L.ref_xoshiro128ss_seed.argtypes = [c_void_p]
    L.ref_xoshiro128ss_next.restype = c_uint32
    s = np.ascontiguousarray(seed, dtype=np.uint32)
    L.ref_xoshiro128ss_seed(s.ctypes.data)
    return np.array([L.ref_xoshiro128ss_next() for _ in range(n)], dtype=np.uint32)


# ---- the learned proposal's network (oracle/oracle_mlp.hpp)
MLP_IN, MLP_HID, MLP_OUT, MLP_RAW, MLP_REC = 36, 64, 32, 8, 72
ORC_BUF_MVEC = 100  # the oracle's motion-vector buffer (float2 per pixel)
MLP_PARAMS = 12768
MLP_OFFSETS = {"W1": (0, (36, 64)), "b1": (2304, (64,)), "W2": (2368, (64, 64)), "b2": (6464, (64,)),
               "W3": (6528, (64, 64)), "b3": (10624, (64,)), "W4": (10688, (64, 32)), "b4": (12736, (32,))}


def mlp_unpack(w):
    return {k: np.asarray(w[o:o + int(np.prod(sh))]).reshape(sh) for k, (o, sh) in MLP_OFFSETS.items()}


def mlp_encode(raw):
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    out = np.zeros(MLP_IN, dtype=np.float32)
    lib().orc_mlp_encode(raw.ctypes.data, out.ctypes.data)
    return out


def mlp_forward(w, raw, alpha=0.25, emulate=True):
    """Logits [n, 32] and 16-bit CDF [n, 32] of raw inputs [n, 8]."""
    w = np.ascontiguousarray(w, dtype=np.float32)
    raw = np.ascontiguousarray(raw, dtype=np.float32).reshape(-1, MLP_RAW)
    n = raw.shape[0]
    z = np.zeros((n, MLP_OUT), dtype=np.float32)
    cdf = np.zeros((n, MLP_OUT), dtype=np.uint16)
    lib().orc_mlp_forward(w.ctypes.data, raw.ctypes.data, n, alpha, int(emulate), z.ctypes.data, cdf.ctypes.data)
    return z, cdf


def mlp_loss(w, recs, alpha=0.25, emulate=False):
    w = np.ascontiguousarray(w, dtype=np.float32)
    recs = np.ascontiguousarray(recs, dtype=np.float32).reshape(-1, MLP_REC)
    return lib().orc_mlp_loss(w.ctypes.data, recs.ctypes.data, recs.shape[0], alpha, int(emulate))


def mlp_backward(w, recs, alpha=0.25, scale=1.0, emulate=False):
    """Gradient of scale * sum of record losses, and [scale * loss, s_plain, s_net]."""
    w = np.ascontiguousarray(w, dtype=np.float32)
    recs = np.ascontiguousarray(recs, dtype=np.float32).reshape(-1, MLP_REC)
    g = np.zeros(MLP_PARAMS, dtype=np.float64)
    st = np.zeros(3, dtype=np.float64)
    lib().orc_mlp_backward(w.ctypes.data, recs.ctypes.data, recs.shape[0], alpha, scale, int(emulate), g.ctypes.data,
                           st.ctypes.data)
    return g, st


def mlp_adam(w, m, v, g, t, lr):
    """One Adam step in place on float32 arrays."""
    g = np.ascontiguousarray(g, dtype=np.float32)
    lib().orc_mlp_adam(w.ctypes.data, m.ctypes.data, v.ctypes.data, g.ctypes.data, w.size, t, lr)


def mlp_sample(cdf, u):
    cdf = np.ascontiguousarray(cdf, dtype=np.uint16)
    b = c_int()
    g = c_float()
    c = lib().orc_mlp_sample(cdf.ctypes.data, u, ctypes.byref(b), ctypes.byref(g))
    return c, b.value, g.value


def mlp_mass_above(cdf, x):
    cdf = np.ascontiguousarray(cdf, dtype=np.uint16)
    return lib().orc_mlp_mass_above(cdf.ctypes.data, x)


def mlp_train_slot(x, y, W, H, bx, by, frame):
    return lib().orc_mlp_train_slot(x, y, W, H, bx, by, frame)


def mlp_random_weights(seed, scale=1.0, last=1.0):
    """He-uniform weights (output layer times `last`) and small random biases."""
    rng = np.random.default_rng(seed)
    w = np.zeros(MLP_PARAMS, dtype=np.float32)
    for k, (o, sh) in MLP_OFFSETS.items():
        n = int(np.prod(sh))
        if k.startswith("W"):
            lim = scale * np.sqrt(6.0 / sh[0]) * (last if k == "W4" else 1.0)
            w[o:o + n] = rng.uniform(-lim, lim, size=n)
        else:
            w[o:o + n] = rng.uniform(-0.1, 0.1, size=n) * scale
    return w
