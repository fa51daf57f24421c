#!/usr/bin/env python3
"""Frames/s of the VolumetricReSTIR pass on one MI355X.

Workload (default): 512^3 sparse fBm cloud in 8^3 fp16 bricks (~15% occupied, 4 mips),
1920x1080, path length 3, 32 point lights + constant environment, 8 initial candidates,
temporal reuse (cap 2.5 x M), 2 spatial rounds of 5 neighbours in 10 px, camera
orbiting 1 degree per frame. One step = one frame of the pass (all six kernels).

Prints one JSON line. `roofline` is for the kernel that takes the most time, with
the compulsory per-pixel buffer bytes of that kernel as the algorithmic traffic;
`cpu_baseline` times the CPU oracle on a band of rows of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

STAGE_NAMES = ["GenerateFeatures", "TraceRays", "TemporalReuse", "SpatialReuse", "CopyReservoirs", "FinalShading"]
HBM_PEAK_GBS = 8000.0
# HBM traffic per launch of each stage at config1 (SpatialReuse, TemporalReuse and
# FinalShading: their record, trace and replay kernels), from rocprofv3 --pmc FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00
# (factors calibrated on a 256 MiB copy in the same process); commands and figures
# in profiles/SpatialReuse_config1_1080p.md.
MEASURED_TRAFFIC_BYTES_CONFIG1 = {"SpatialReuse": 32.743e9, "TraceRays": 12.905e9, "FinalShading": 4.322e9,
                                  "TemporalReuse": 3.363e9, "GenerateFeatures": 0.183e9, "CopyReservoirs": 0.156e9}


def stage_bytes_per_pixel(stage, b, k):
    """Compulsory buffer bytes per pixel of each kernel in the reference's layout:
    reservoir R = 36 B, extra-bounce E = 12 (b - 1) B, features F = 8 B, colour C = 16 B."""
    R, E, F, C = 36, 12 * (b - 1), 8, 16
    return {0: F, 1: R + E, 2: 2 * R + 2 * F + R + 2 * E, 3: k * R + F + R + 2 * E, 4: 2 * (R + E),
            5: R + E + F + C}[stage]


def cpu_baseline(spec, rows, threads):
    from oracle.binding import Oracle
    o = Oracle()
    t0 = time.perf_counter()
    o.set_scene(spec)
    setup = time.perf_counter() - t0
    y0 = spec.height // 2 - rows // 2
    y1 = y0 + rows
    o.set_camera(spec, 0)
    o.execute(spec.width, spec.height, y0, y1, threads)      # frame 0 (no temporal history yet)
    o.set_camera(spec, 1)
    t0 = time.perf_counter()
    o.execute(spec.width, spec.height, y0, y1, threads)      # frame 1: all six kernels
    dt = time.perf_counter() - t0
    px = rows * spec.width
    frame_s = dt / px * spec.width * spec.height
    return {"value": 1.0 / frame_s, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"oracle frame 1 on rows {y0}-{y1} of {spec.width}x{spec.height} ({px} px, every kernel "
                      f"restricted to those rows), {dt:.2f} s on {threads} threads, per-pixel time scaled to the "
                      f"full frame; volume set-up {setup:.1f} s not counted"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="config1", choices=["config0", "config1"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=64)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (a known-size kernel that "
                         "calibrates FETCH_SIZE / WRITE_SIZE counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from volumetricrestirrelease_amd import abi, scenes
    from volumetricrestirrelease_amd.volumetric_restir import VolumetricReSTIR

    torch.cuda.set_device(0)
    if args.calibration_copy_mb > 0:
        n = args.calibration_copy_mb * (1 << 20) // 4
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    t0 = time.perf_counter()
    spec = scenes.config1() if args.config == "config1" else scenes.config0()
    if args.config == "config0":
        spec.orbit_degrees_per_frame = 1.0
    g = VolumetricReSTIR()
    g.setScene(spec)
    setup_s = time.perf_counter() - t0
    W, H = spec.width, spec.height
    P = spec.params
    b, k, rounds = P["maxBounces"], P["spatialSampleCount"], P["spatialReuseRounds"]
    stream = torch.cuda.current_stream()
    out = torch.zeros((H, W, 4), device="cuda")

    frame = 0
    for _ in range(args.warmup):
        g.setCamera(spec, frame)
        g.execute(out, stream=stream)
        frame += 1
    torch.cuda.synchronize()

    stage_ms = {s: 0.0 for s in range(6)}
    stage_calls = {s: 0 for s in range(6)}
    ev_pairs = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        g.setCamera(spec, frame)
        g.prepare_frame(W, H)
        seq = [(0, 0), (1, 0), (2, 0)] + [(3, r) for r in range(rounds)] + [(4, 0), (5, 0)]
        for st, r in seq:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            g.run_stage(st, r, out=out if st == 5 else None, stream=stream)
            e1.record(stream)
            ev_pairs.append((st, e0, e1))
        g.end_frame()
        frame += 1
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    for st, e0, e1 in ev_pairs:
        stage_ms[st] += e0.elapsed_time(e1)
        stage_calls[st] += 1

    fps = args.steps / elapsed
    dom = max(stage_ms, key=lambda s: stage_ms[s])
    avg_ms = stage_ms[dom] / stage_calls[dom]
    alg_bytes = stage_bytes_per_pixel(dom, b, k) * W * H
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    line = {
        "metric": "1080p frames/s at matched relMSE vs 64k-spp ref; learned vs plain proposal",
        "value": round(fps, 4),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (fp16 bricks)" if spec.brick_format == abi.BRICK_FP16 else "f32",
        "data": "synthetic (seeded fBm density, seeded point lights)",
        "config": {
            "workload": f"{spec.name}: {spec.density.shape[0]}^3 fBm, {W}x{H}, path length {b}, "
                        f"M={P['initialM']}, {rounds} spatial rounds x {k - 1} neighbours, {len(spec.lights)} "
                        f"point lights + env, orbiting camera",
            "proposal": "plain (analytic / ray-marching trackers; learned proposal not built)",
            "stage_ms_avg": {STAGE_NAMES[s]: round(stage_ms[s] / max(1, stage_calls[s]), 3) for s in range(6)},
            "volume_setup_s": round(setup_s, 2),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": STAGE_NAMES[dom],
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": MEASURED_TRAFFIC_BYTES_CONFIG1.get(STAGE_NAMES[dom]) if args.config == "config1" else None,
            "algorithmic_bytes_per_launch": alg_bytes,
            "avg_launch_ms": round(avg_ms, 3),
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(spec, args.cpu_rows, args.cpu_threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
