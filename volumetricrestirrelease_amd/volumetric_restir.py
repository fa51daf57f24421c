"""Python mirror of the reference render pass (VR/VolumetricReSTIR.{h,cpp}) over the C ABI.

Names follow the reference: ``VolumetricReSTIRParams`` fields keep their
``m``-prefixed names (VR/VolumetricReSTIR.cpp:1356-1426), ``execute`` renders
one frame, ``getScriptingDictionary`` / ``updateDict`` serialise the options.
Device memory and streams come from PyTorch; all work runs in libvrestir.so.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import abi


class VolumetricReSTIRError(RuntimeError):
    pass


class VolumetricReSTIRParams:
    """Option surface with the reference's field names (``mMaxBounces`` ...)."""

    def __init__(self, **kwargs):
        object.__setattr__(self, "_p", abi.default_params())
        for k, v in kwargs.items():
            setattr(self, k, v)

    @staticmethod
    def _cname(name: str) -> str:
        if not name.startswith("m") or len(name) < 2:
            raise AttributeError(name)
        c = name[1].lower() + name[2:]
        if c not in dict(abi.PARAM_FIELDS):
            raise AttributeError(f"VolumetricReSTIRParams has no field {name}")
        return c

    def __getattr__(self, name):
        return getattr(object.__getattribute__(self, "_p"), self._cname(name))

    def __setattr__(self, name, value):
        c = self._cname(name)
        setattr(self._p, c, value)

    def to_dict(self) -> dict:
        return {"m" + k[0].upper() + k[1:]: getattr(self._p, k) for k, _ in abi.PARAM_FIELDS}

    @property
    def struct(self) -> abi.Params:
        return self._p


def _dev_ptr(x):
    """A device address: a torch tensor's data pointer, or an int address."""
    return int(x) if isinstance(x, int) else x.data_ptr()


def _stream_ptr(stream):
    return ctypes.c_void_p(stream.cuda_stream if stream is not None else 0)


class VolumetricReSTIR:
    """One instance of the pass, bound to the current HIP device."""

    def __init__(self, params: VolumetricReSTIRParams | abi.Params | None = None):
        self.lib = abi.load_library()
        p = params.struct if isinstance(params, VolumetricReSTIRParams) else params
        h = ctypes.c_void_p()
        rc = self.lib.vrestir_create(ctypes.byref(p) if p is not None else None, ctypes.byref(h))
        if rc != 0:
            raise VolumetricReSTIRError(f"vrestir_create: {self.lib.vrestir_last_error(None).decode()} ({rc})")
        self.h = h

    @classmethod
    def create(cls, params=None):
        return cls(params)

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.vrestir_last_error(self.h).decode()
            raise VolumetricReSTIRError(f"{what}: {msg} (status {rc})")

    def close(self):
        if getattr(self, "h", None):
            self.lib.vrestir_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- scene (setScene, VR/VolumetricReSTIR.cpp:1218-1256)
    def setScene(self, spec, frame: int = 0):
        d = spec.volume_desc()
        self._check(self.lib.vrestir_set_volume(self.h, ctypes.byref(d)), "set_volume")
        self._setSceneRest(spec, frame, spec.temperature is not None)

    def streamVolumeFrame(self, spec, density, temperature=None, stream=None):
        """vrestir_stream_volume_frame: a volume frame built on the GPU from float32
        device tensors shaped (nz, ny, nx); spec gives the dimensions, mips, format,
        transform and material (its host grids are not read). The first call starts a
        new volume, later ones advance its animation by one frame."""
        d = spec.volume_desc(grids=False)
        d.density = ctypes.cast(ctypes.c_void_p(_dev_ptr(density)), ctypes.POINTER(ctypes.c_float))
        d.temperature = (ctypes.cast(ctypes.c_void_p(_dev_ptr(temperature)), ctypes.POINTER(ctypes.c_float))
                         if temperature is not None else None)
        self._check(self.lib.vrestir_stream_volume_frame(self.h, _stream_ptr(stream), ctypes.byref(d)),
                    "stream_volume_frame")

    def shareVolume(self, owner):
        """vrestir_share_volume: bind owner's current volume frame (read-only)."""
        self._check(self.lib.vrestir_share_volume(self.h, owner.h), "share_volume")

    def setSceneShared(self, spec, owner, frame: int = 0, emissive: bool = False):
        """setScene with owner's volume instead of one of its own."""
        self.shareVolume(owner)
        self._setSceneRest(spec, frame, emissive)

    def setSceneStreamed(self, spec, density, temperature=None, frame: int = 0):
        """setScene with the volume built on the GPU from device tensors."""
        self.streamVolumeFrame(spec, density, temperature)
        self._setSceneRest(spec, frame, temperature is not None)

    def _setSceneRest(self, spec, frame, emissive):
        if emissive:
            lut, rows = spec.blackbody_array()
            self._check(self.lib.vrestir_set_blackbody_lut(self.h, lut.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                           rows), "set_blackbody_lut")
        arr, n, env = spec.light_array()
        self._check(self.lib.vrestir_set_lights(self.h, arr, n, env), "set_lights")
        self.setSurfaces(spec)
        em = spec.env_map_desc()
        self._check(self.lib.vrestir_set_env_map(self.h, ctypes.byref(em) if em is not None else None),
                    "set_env_map")
        self.setCamera(spec, frame)
        self.setParams(spec.make_params())

    def setSurfaces(self, spec):
        """The scene's triangles (their emissive ones are the emissive lights); none
        when spec.surfaces is None."""
        d, _keep = spec.surfaces_desc()
        self._check(self.lib.vrestir_set_surfaces(self.h, ctypes.byref(d) if d is not None else None),
                    "set_surfaces")

    def setCamera(self, spec, frame: int = 0):
        """Camera (and env rotation, for animated environments) of a frame."""
        c = spec.camera_struct(frame)
        self._check(self.lib.vrestir_set_camera(self.h, ctypes.byref(c)), "set_camera")
        if spec.env_map is not None:
            rot = (ctypes.c_float * 3)(*spec.env_rotation_at(frame))
            self._check(self.lib.vrestir_set_env_rotation(self.h, rot), "set_env_rotation")

    def setEnvMap(self, rgb, intensity=1.0, tint=(1.0, 1.0, 1.0), rotation=(0.0, 0.0, 0.0), importance_dim=0,
                  importance_spp=0):
        """Scene::setEnvMap + setEnvMapIntensity + setEnvMapRotation; rgb: (H, W, 3) or None."""
        if rgb is None:
            self._check(self.lib.vrestir_set_env_map(self.h, None), "set_env_map")
            return
        d = abi.EnvMapDesc()
        arr = np.ascontiguousarray(rgb, dtype=np.float32)
        d.height, d.width = int(arr.shape[0]), int(arr.shape[1])
        d.rgb = arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        d.intensity = intensity
        for k in range(3):
            d.tint[k] = tint[k]
            d.rotationDegrees[k] = rotation[k]
        d.importanceDim, d.importanceSpp = importance_dim, importance_spp
        self._check(self.lib.vrestir_set_env_map(self.h, ctypes.byref(d)), "set_env_map")

    # ---- animated volumes (addGVDBVolumeSequence, SC/Scene.cpp:3329-3355)
    def addVolumeFrame(self, desc) -> int:
        """Append one frame (an abi.VolumeDesc, or a SceneSpec) to the bound sequence."""
        d = desc.volume_desc() if hasattr(desc, "volume_desc") else desc
        idx = ctypes.c_int32()
        self._check(self.lib.vrestir_add_volume_frame(self.h, ctypes.byref(d), ctypes.byref(idx)), "add_volume_frame")
        return idx.value

    def setVolumeFrame(self, index: int, animating: bool = True):
        self._check(self.lib.vrestir_set_volume_frame(self.h, index, 1 if animating else 0), "set_volume_frame")

    def setFreezeFrame(self, freeze: bool):
        self._check(self.lib.vrestir_set_freeze_frame(self.h, 1 if freeze else 0), "set_freeze_frame")

    def setEnvMapRotation(self, degrees_xyz):
        self._check(self.lib.vrestir_set_env_rotation(self.h, (ctypes.c_float * 3)(*degrees_xyz)), "set_env_rotation")

    def setEnvMapIntensity(self, intensity, tint=(1.0, 1.0, 1.0)):
        self._check(self.lib.vrestir_set_env_intensity(self.h, intensity, (ctypes.c_float * 3)(*tint)),
                    "set_env_intensity")

    def setParams(self, p):
        st = p.struct if isinstance(p, VolumetricReSTIRParams) else p
        self._check(self.lib.vrestir_set_params(self.h, ctypes.byref(st)), "set_params")

    def getParams(self) -> abi.Params:
        p = abi.Params()
        self._check(self.lib.vrestir_get_params(self.h, ctypes.byref(p)), "get_params")
        return p

    # ---- serialisation (VR/VolumetricReSTIR.cpp:142-147, 1280-1347)
    def getScriptingDictionary(self) -> dict:
        p = self.getParams()
        return {"mParams": {"m" + k[0].upper() + k[1:]: getattr(p, k) for k, _ in abi.PARAM_FIELDS}}

    def updateDict(self, d: dict):
        p = self.getParams()
        for k, v in d.get("mParams", d).items():
            setattr(p, VolumetricReSTIRParams._cname(k), v)
        self.setParams(p)

    # ---- per frame (execute, VR/VolumetricReSTIR.cpp:303-906)
    def execute(self, out_rgba, out_mvec=None, stream=None):
        """Render one frame into ``out_rgba`` (a CUDA/HIP float32 tensor [H, W, 4])."""
        H, W = int(out_rgba.shape[0]), int(out_rgba.shape[1])
        if out_rgba.dtype.itemsize != 4 or out_rgba.shape[-1] != 4 or not out_rgba.is_contiguous():
            raise VolumetricReSTIRError("out_rgba must be a contiguous float32 [H, W, 4] device tensor")
        s = ctypes.c_void_p(stream.cuda_stream if stream is not None else 0)
        mv = ctypes.c_void_p(out_mvec.data_ptr()) if out_mvec is not None else None
        self._check(self.lib.vrestir_execute(self.h, s, W, H, ctypes.c_void_p(out_rgba.data_ptr()), mv), "execute")

    def resetHistory(self):
        """Drop the temporal history and restart the frame count (as an option change does)."""
        self._check(self.lib.vrestir_reset_history(self.h), "reset_history")

    def frameCount(self) -> int:
        v = ctypes.c_int32()
        self._check(self.lib.vrestir_frame_count(self.h, ctypes.byref(v)), "frame_count")
        return v.value

    # ---- stage-level access for parity tests
    def prepare_frame(self, W, H):
        self._check(self.lib.vrestir_prepare_frame(self.h, W, H), "prepare_frame")

    def run_stage(self, stage, rnd=0, out=None, stream=None):
        s = ctypes.c_void_p(stream.cuda_stream if stream is not None else 0)
        o = ctypes.c_void_p(out.data_ptr()) if out is not None else None
        self._check(self.lib.vrestir_run_stage(self.h, s, stage, rnd, o), f"run_stage({stage})")

    def end_frame(self):
        self._check(self.lib.vrestir_end_frame(self.h), "end_frame")

    def read(self, buf, nbytes) -> np.ndarray:
        out = np.zeros(nbytes, dtype=np.uint8)
        self._check(self.lib.vrestir_read_buffer(self.h, buf, ctypes.c_void_p(out.ctypes.data), nbytes), "read")
        return out

    def write(self, buf, data: np.ndarray):
        data = np.ascontiguousarray(data)
        self._check(self.lib.vrestir_write_buffer(self.h, buf, ctypes.c_void_p(data.ctypes.data), data.nbytes),
                    "write")

    # ---- ground truth (AccumulatePass double mode) and relMSE (ImageCompare "rmse")
    def accumulate(self, frame_rgba, reset=False, stream=None):
        H, W = int(frame_rgba.shape[0]), int(frame_rgba.shape[1])
        st = _stream_ptr(stream)
        self._check(self.lib.vrestir_accumulate(self.h, st, frame_rgba.data_ptr(), W, H, 1 if reset else 0),
                    "accumulate")

    def accumulated(self, out_rgba, stream=None) -> int:
        n = ctypes.c_int64()
        self._check(self.lib.vrestir_accumulated(self.h, _stream_ptr(stream), out_rgba.data_ptr(), ctypes.byref(n)),
                    "accumulated")
        return n.value

    def relmse(self, image_rgba, reference_rgba, stream=None) -> float:
        H, W = int(image_rgba.shape[0]), int(image_rgba.shape[1])
        out = ctypes.c_double()
        self._check(self.lib.vrestir_relmse(self.h, _stream_ptr(stream), image_rgba.data_ptr(),
                                            reference_rgba.data_ptr(), W, H, ctypes.byref(out)), "relmse")
        return out.value

    def referenceLattice(self, width, height, stride, spp, sample_number, out, stream=None):
        """The volume path tracer on the pixel lattice (stride i + stride/2, stride j + stride/2)
        of a width x height frame; out: device float32 [chunks, lh, lw, 4], image c from the
        streams of sample number sample_number + c."""
        chunks = int(out.shape[0])
        self._check(self.lib.vrestir_reference_lattice(self.h, _stream_ptr(stream), width, height, stride, spp,
                                                       sample_number, chunks, ctypes.c_void_p(out.data_ptr())),
                    "reference_lattice")

    def gridInfo(self, slot):
        """dims, occupied bricks, brick cells and device bytes of a bound grid slot."""
        dims = (ctypes.c_int32 * 3)()
        nb, cells, nbytes = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.vrestir_grid_info(self.h, slot, dims, ctypes.byref(nb), ctypes.byref(cells),
                                               ctypes.byref(nbytes)), "grid_info")
        return {"dims": tuple(dims), "bricks": nb.value, "brick_cells": cells.value, "bytes": nbytes.value}

    def volumeOccupancy(self, slot=0):
        """Fraction of a slot's 8^3 brick cells that hold a brick."""
        gi = self.gridInfo(slot)
        return gi["bricks"] / max(1, gi["brick_cells"])

    def brickPoolBytes(self, slot):
        return self.gridInfo(slot)["bytes"]

    @staticmethod
    def latticeShape(width, height, stride):
        """(lh, lw) of vrestir_reference_lattice's lattice."""
        return (height - stride // 2 + stride - 1) // stride, (width - stride // 2 + stride - 1) // stride

    # ---- learned free-flight proposal (an addition: no counterpart in the reference)
    def setLearned(self, **fields):
        """vrestir_set_learned with vrestir_default_learned's values overridden by `fields`;
        (re)initialises the network."""
        d = abi.default_learned()
        for k, v in fields.items():
            setattr(d, k, v)
        self._check(self.lib.vrestir_set_learned(self.h, ctypes.byref(d)), "set_learned")
        return d

    def learnedStats(self) -> np.ndarray:
        """[loss, analytic second moment, network second moment, network in use, Adam steps]."""
        out = np.zeros(5, dtype=np.float32)
        self._check(self.lib.vrestir_learned_stats(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                    "learned_stats")
        return out

    def learnedWeights(self) -> np.ndarray:
        out = np.zeros(abi.LEARNED_PARAMS, dtype=np.float32)
        self._check(self.lib.vrestir_learned_get_weights(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                         out.size), "learned_get_weights")
        return out

    def setLearnedWeights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(self.lib.vrestir_learned_set_weights(self.h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                         w.size), "learned_set_weights")

    def learnedForward(self, raw, cdf, logits=None, stream=None):
        """The network on raw inputs [n, 8] (device float32) -> cdf [n, 32] (device int16/uint16), logits [n, 32]."""
        n = int(raw.shape[0])
        lg = ctypes.c_void_p(logits.data_ptr()) if logits is not None else None
        self._check(self.lib.vrestir_learned_forward(self.h, _stream_ptr(stream), ctypes.c_void_p(raw.data_ptr()), n,
                                                     ctypes.c_void_p(cdf.data_ptr()), lg), "learned_forward")

    def learnedGradient(self, recs, out, scale=1.0, stream=None):
        """Summed gradient of scale x loss over records [n, 72] (device) into out [12768 + 3] (device)."""
        n = int(recs.shape[0])
        self._check(self.lib.vrestir_learned_gradient(self.h, _stream_ptr(stream), ctypes.c_void_p(recs.data_ptr()),
                                                      n, scale, ctypes.c_void_p(out.data_ptr())), "learned_gradient")

    def learnedTrain(self, recs, stream=None):
        n = int(recs.shape[0])
        self._check(self.lib.vrestir_learned_train(self.h, _stream_ptr(stream), ctypes.c_void_p(recs.data_ptr()), n),
                    "learned_train")

    def track(self, mode, mip, linear, tstep, nsamples, rays, seeds, out, stream=None):
        """Batched tracking primitive over device tensors rays [n,8] f32, seeds [n,3] u32/i32, out [n,8] f32."""
        n = int(rays.shape[0])
        s = ctypes.c_void_p(stream.cuda_stream if stream is not None else 0)
        self._check(self.lib.vrestir_track_batch(self.h, s, mode, mip, linear, tstep, nsamples, n,
                                                 ctypes.c_void_p(rays.data_ptr()), ctypes.c_void_p(seeds.data_ptr()),
                                                 ctypes.c_void_p(out.data_ptr())), "track_batch")


# Reservoir / features exchange layouts (include/vrestir.h)
RESERVOIR_DTYPE = np.dtype([("runningSum", "<f4"), ("M", "<f4"), ("depth", "<f4"), ("p_y", "<f4"),
                            ("lightUV", "<f4", 2), ("lightID", "<i4"), ("sampledPixel", "<i4"),
                            ("extraBounceStartId", "<i4")])
FEATURES_DTYPE = np.dtype([("noReflectiveSurface", "<i4"), ("transmittance", "<f4")])
VBUFFER_DTYPE = np.dtype([("tri", "<i4"), ("bary", "<f4", 2), ("t", "<f4")])


class VolumetricReSTIRBatch:
    """Many passes rendered together: each step of a frame is one persistent launch
    over every pass's tiles, the frame's launches one replayed HIP graph
    (vrestir_batch_*; no counterpart in the reference)."""

    def __init__(self, passes):
        self.lib = abi.load_library()
        self.passes = list(passes)
        arr = (ctypes.c_void_p * len(self.passes))(*[p.h.value for p in self.passes])
        b = ctypes.c_void_p()
        rc = self.lib.vrestir_batch_create(arr, len(self.passes), ctypes.byref(b))
        if rc != 0:
            raise VolumetricReSTIRError(f"vrestir_batch_create: {self.lib.vrestir_last_error(None).decode()} ({rc})")
        self.b = b

    def _check(self, rc, what):
        if rc != 0:
            raise VolumetricReSTIRError(f"{what}: {self.lib.vrestir_batch_last_error(self.b).decode()} (status {rc})")

    def close(self):
        if getattr(self, "b", None):
            self.lib.vrestir_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _views(self, outs):
        n = len(self.passes)
        W = (ctypes.c_int32 * n)(*[int(o.shape[1]) for o in outs])
        H = (ctypes.c_int32 * n)(*[int(o.shape[0]) for o in outs])
        O = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
        return W, H, O

    def execute(self, outs, stream=None):
        """One frame of every pass into outs[i] (float32 [H_i, W_i, 4] device tensors)."""
        W, H, O = self._views(outs)
        self._check(self.lib.vrestir_batch_execute(self.b, _stream_ptr(stream), W, H, O), "batch_execute")

    def setTuning(self, grid, tiles_per_item, use_graph=True):
        self._check(self.lib.vrestir_batch_set_tuning(self.b, grid, tiles_per_item, 1 if use_graph else 0),
                    "batch_set_tuning")

    def tuning(self):
        g, p = ctypes.c_int32(), ctypes.c_int32()
        self._check(self.lib.vrestir_batch_get_tuning(self.b, ctypes.byref(g), ctypes.byref(p)), "batch_get_tuning")
        return g.value, p.value

    def autotune(self, outs, frames=2, stream=None):
        """Times the candidate queue shapes, keeps the fastest and resets every pass's
        history; returns its milliseconds per frame."""
        W, H, O = self._views(outs)
        ms = ctypes.c_float()
        self._check(self.lib.vrestir_batch_autotune(self.b, _stream_ptr(stream), W, H, O, frames, ctypes.byref(ms)),
                    "batch_autotune")
        return ms.value


class Plume:
    """The seeded synthetic animated plume (vrestir_plume_*), simulated on the GPU:
    density and temperature on an n^3 grid over the unit cube."""

    def __init__(self, n: int, seed: int):
        self.lib = abi.load_library()
        p = ctypes.c_void_p()
        rc = self.lib.vrestir_plume_create(n, seed, ctypes.byref(p))
        if rc != 0:
            raise VolumetricReSTIRError(f"vrestir_plume_create: status {rc}")
        self.p, self.n = p, n

    def advance(self, steps: int = 1, stream=None):
        rc = self.lib.vrestir_plume_advance(self.p, _stream_ptr(stream), steps)
        if rc != 0:
            raise VolumetricReSTIRError(f"vrestir_plume_advance: status {rc}")

    def grids(self):
        """(density address, temperature address, steps simulated) of the current grids."""
        d, t, s = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int32()
        rc = self.lib.vrestir_plume_grids(self.p, ctypes.byref(d), ctypes.byref(t), ctypes.byref(s))
        if rc != 0:
            raise VolumetricReSTIRError(f"vrestir_plume_grids: status {rc}")
        return d.value, t.value, s.value

    def read(self):
        """Host copies (nz, ny, nx) of the current density and temperature."""
        d = np.zeros((self.n, self.n, self.n), np.float32)
        t = np.zeros_like(d)
        rc = self.lib.vrestir_plume_read(self.p, ctypes.c_void_p(d.ctypes.data), ctypes.c_void_p(t.ctypes.data))
        if rc != 0:
            raise VolumetricReSTIRError(f"vrestir_plume_read: status {rc}")
        return d, t

    def close(self):
        if getattr(self, "p", None):
            self.lib.vrestir_plume_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
