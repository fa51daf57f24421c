"""Animated volumes and camera animation (no GPU):
  * camera animation modes (VR/VolumetricReSTIR.cpp:801-905) keep the view
    direction, move mode 0/1 along the camera's right axis (|offset| <= moveScale,
    the sine of 2 pi moved / interval) and forward axis, and rotate mode 2 at a
    constant distance and height around the view ray's closest point to the volume
    centre, rotationSpeed degrees per update; mode 0 pauses for pauseInterval of
    every round;
  * the oracle's velocity lookup equals a numpy trilinear interpolation of the
    dense field, zero in bricks without velocity;
  * binding frame i of a sequence puts frame (i - 1) mod n in the previous-frame
    slots (19 + k);
  * a frozen frame re-shades the last reservoirs with the previous frame's random
    numbers, so with one spatial round and a still camera it repeats the previous
    frame's image exactly.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import binding as ob
from volumetricrestirrelease_amd import abi, scenes


def animate(base, desc, updates):
    out = abi.Camera()
    fr = ctypes.c_int32()
    rc = abi.load_library().vrestir_camera_animation(ctypes.byref(base), ctypes.byref(desc), updates,
                                                     ctypes.byref(out), ctypes.byref(fr))
    assert rc == 0
    return np.array(out.posW[:], float), np.array(out.target[:], float), fr.value


def base_camera():
    c = abi.Camera()
    for k, v in enumerate((0.4, 0.3, 3.0)):
        c.posW[k] = v
    for k, v in enumerate((0.0, 0.1, 0.0)):
        c.target[k] = v
    c.up[1] = 1.0
    c.fovY, c.nearZ, c.farZ = 0.7, 0.1, 100.0
    return c


def test_camera_shake_mode0():
    c = base_camera()
    d = abi.camera_animation_desc(0, frameInterval=20, pauseInterval=30, shakeRoundsBeforePause=2, shakeTotalRounds=2)
    P0 = np.array(c.posW[:], float)
    view = np.array(c.target[:], float) - P0
    view /= np.linalg.norm(view)
    right = np.cross(view, [0, 1, 0])
    right /= np.linalg.norm(right)
    round_frames = 2 * 20 + 30
    frozen = 0
    for u in range(1, 2 * round_frames - 1):
        pos, tgt, fr = animate(c, d, u)
        assert np.allclose(tgt - pos, view, atol=1e-5)
        off = pos - P0
        s = off @ right
        assert np.allclose(off, s * right, atol=1e-5)      # on the right axis through the start
        assert abs(s) <= 0.3 + 1e-6
        moved = u + 1                                     # the counter starts at 1 and is incremented first
        if not fr:
            assert abs(s - math.sin(2 * math.pi * moved / 20) * 0.3) < 1e-5
        frozen += fr
    assert abs(frozen - 2 * 30) <= 2                      # one pause per round


def test_camera_rotation_mode2():
    c = base_camera()
    center = (0.1, -0.2, 0.05)
    d = abi.camera_animation_desc(2, rotationSpeed=3.0, volumeCenter=center)
    P0 = np.array(c.posW[:], float)
    T0 = np.array(c.target[:], float)
    v = (T0 - P0) / np.linalg.norm(T0 - P0)
    rc = P0 + ((np.array(center) - P0) @ v) * v           # closest point of the view ray to the centre
    prev = None
    for u in range(1, 12):
        pos, tgt, fr = animate(c, d, u)
        assert not fr
        assert np.allclose(tgt, rc, atol=1e-5)
        assert abs(pos[1] - P0[1]) < 1e-5                  # the orbit keeps the camera's height
        r = np.hypot(pos[0] - rc[0], pos[2] - rc[2])
        assert abs(r - np.hypot(P0[0] - rc[0], P0[2] - rc[2])) < 1e-5
        ang = math.atan2(pos[2] - rc[2], pos[0] - rc[0])
        if prev is not None:
            dang = (prev - ang + math.pi) % (2 * math.pi) - math.pi
            assert abs(abs(dang) - math.radians(3.0)) < 1e-4
        prev = ang
    pos, _, _ = animate(c, abi.camera_animation_desc(2, rotationSpeed=3.0, rotationDistance=2.5,
                                                     volumeCenter=center), 5)
    assert abs(np.hypot(pos[0] - rc[0], pos[2] - rc[2]) - 2.5) < 1e-5


def test_camera_mode1_stays_in_the_view_plane():
    c = base_camera()
    d = abi.camera_animation_desc(1, frameInterval=16)
    P0 = np.array(c.posW[:], float)
    view = np.array(c.target[:], float) - P0
    view /= np.linalg.norm(view)
    right = np.cross(view, [0, 1, 0])
    right /= np.linalg.norm(right)
    n = np.cross(view, right)
    saw_forward = False
    for u in range(1, 6 * 16 + 2):
        pos, tgt, fr = animate(c, d, u)
        assert np.allclose((tgt - pos) / np.linalg.norm(tgt - pos), view, atol=1e-5)
        assert abs((pos - P0) @ n) < 1e-5
        saw_forward |= abs((pos - P0) @ view) > 1e-3
    assert saw_forward


def random_velocity_scene(n=16, seed=4):
    rng = np.random.default_rng(seed)
    dens = np.zeros((n, n, n), np.float32)
    dens[4:12, 4:12, 4:12] = 1.0
    vel = rng.normal(size=(n, n, n, 3)).astype(np.float32)
    vel[:8, :8, :8] = 0.0                                  # one empty brick
    spec = scenes.config0(res=n, width=8, height=8)
    spec.density = dens
    spec.num_mips = 1
    spec.velocity = vel
    return spec, vel


def numpy_trilinear(vel, p):
    n = vel.shape[0]
    x = p - 0.5
    i0 = np.floor(x).astype(int)
    f = x - i0
    out = np.zeros(3)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (f[0] if dx else 1 - f[0]) * (f[1] if dy else 1 - f[1]) * (f[2] if dz else 1 - f[2])
                ix, iy, iz = i0[0] + dx, i0[1] + dy, i0[2] + dz
                if 0 <= ix < n and 0 <= iy < n and 0 <= iz < n:
                    out += w * vel[iz, iy, ix]
    return out


def test_velocity_lookup_is_trilinear_with_brick_occupancy():
    spec, vel = random_velocity_scene()
    o = ob.Oracle()
    o.set_scene(spec)
    rng = np.random.default_rng(5)
    n = 16
    pts = rng.uniform(0.0, n, size=(400, 3))
    rays = np.zeros((400, 8), np.float32)
    rays[:, :3] = pts * 2.0 / n - 1.0
    out = o.track(abi.TRACK_VELOCITY, 0, 1, 1.0, 1, rays, np.zeros((400, 3), np.uint32))
    for p, got in zip(pts, out):
        q = np.asarray(p, np.float32).astype(np.float64)
        if (q < 8).all():
            assert np.allclose(got[:3], 0.0)                # the empty brick reads zero
        else:
            assert np.allclose(got[:3], numpy_trilinear(vel, q), atol=2e-5)


def test_sequence_binds_previous_frame_slots():
    _, specs = scenes.moving_sequence(res=32, frames=3)
    o = ob.Oracle()
    o.set_scene(specs[0])
    for s in specs[1:]:
        o.add_volume_frame(s)
    counts = []
    for i in range(3):
        o.set_volume_frame(i)
        counts.append(o.grid_info(0)[2])
    for i in range(3):
        o.set_volume_frame(i)
        assert o.grid_info(19)[2] == counts[(i - 1) % 3]
    assert len(set(counts)) > 1                             # the frames differ


def test_frozen_frame_repeats_previous_image():
    spec = scenes.config0(res=32, width=24, height=16)
    o = ob.Oracle()
    o.set_scene(spec)
    o.execute(24, 16)
    a = o.execute(24, 16)
    o.set_freeze_frame(True)
    b = o.execute(24, 16)
    c = o.execute(24, 16)
    assert np.array_equal(a, b) and np.array_equal(b, c)
    o.set_freeze_frame(False)
    d = o.execute(24, 16)
    assert not np.array_equal(c, d)
