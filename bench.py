#!/usr/bin/env python3
"""1080p frames/s of the VolumetricReSTIR pass at matched relMSE, learned vs plain
proposal, on one MI355X.

Workload (config1): 512^3 sparse fBm cloud in 8^3 fp16 bricks (4 mips), 1920x1080,
path length 3, 32 point lights + constant environment, temporal reuse (cap 2.5 x M),
2 spatial rounds of 5 neighbours in 10 px, camera orbiting 1 degree per frame.
One step = one frame of the pass (every kernel, the learned proposal's inference and
training included when it is on).

How the headline number is made:
1. Quality runs: the plain pass (M = 8 initial candidates, the config's count) renders
   the 60-frame orbit; frames 56-59 are each compared with a 65,536-spp FP64
   reference of the volume path tracer (the same camera) on the uniform lattice of
   every 8th pixel in x and y (240 x 135 pixels: relMSE over a uniform pixel
   subsample is an unbiased estimate of the full frame's), and relMSE is the mean. The learned-proposal pass, trained online from
   frame 0, renders the same orbit at M = 8, 6, 4, 2 until its relMSE exceeds the
   plain pass's; the smallest M that still matches it is the learned configuration.
2. Timed run: W warmup and K timed frames of the learned configuration (the metric's
   configuration; the plain M = 8 pass when no learned M matches), bracketed by
   device synchronisation; `value` = K / elapsed. The plain M = 8 pass is timed the
   same way and reported beside it.

3. Batched: after the config1 timing, `batched` reports config2 (16 seeded views
   and volumes of uneven size, one persistent launch per step over all of them, the
   frame a replayed HIP graph) against the same passes rendered one after another
   (`--config config2` prints that measurement as its own line; `--no-batched`
   leaves it out).

`roofline` is for the kernel family that takes the most time per frame, with its
compulsory per-pixel buffer bytes plus the brick pools of the grids it traverses as
algorithmic bytes; `cpu_baseline` times the CPU oracle on bands of rows of the
same workload, single-threaded and on all cores of the host's share.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

STAGE_NAMES = {0: "GenerateFeatures", 1: "TraceRays", 2: "TemporalReuse", 3: "SpatialReuse", 4: "CopyReservoirs",
               5: "FinalShading", 6: "LearnedProposal", 7: "LearnedTrain"}
HBM_PEAK_GBS = 8000.0
# HBM traffic per launch of each stage at config1 (record, trace and replay kernels
# summed), from rocprofv3 --pmc FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00 (factors
# calibrated on a 256 MiB copy in the same process). A profile constant, not
# measured by this run: see the `traffic_source` field.
PROFILE_TRAFFIC = {
    "file": "profiles/SpatialReuse_config1_1080p.md",
    "bytes": {"SpatialReuse": 33.395e9, "TraceRays": 15.489e9, "FinalShading": 4.360e9, "TemporalReuse": 3.376e9,
              "GenerateFeatures": 0.183e9, "CopyReservoirs": 0.156e9},
}
LATTICE = 8
QUALITY_FRAMES = 60
EVAL_FRAMES = (56, 57, 58, 59)   # relMSE is the mean over these frames of the orbit
LEARNED_M = (8, 6, 4, 2)


def stage_bytes_per_pixel(stage, b, k):
    """Compulsory buffer bytes per pixel of each stage in the reference's layout:
    reservoir R = 36 B, extra-bounce E = 12 (b - 1) B, features F = 8 B, colour C = 16 B."""
    R, E, F, C = 36, 12 * (b - 1), 8, 16
    return {0: F, 1: R + E, 2: 2 * R + 2 * F + R + 2 * E, 3: k * R + F + R + 2 * E, 4: 2 * (R + E),
            5: R + E + F + C, 6: 64, 7: 0}[stage]


def stage_slots(stage, params):
    """Grid slots (mip k: k, conservative mip k: 8 + k) a stage's rays traverse."""
    P = params
    ini_vis = P.initialBaseMipLevel if P.initialVisibilityUseLinearSampler else P.initialBaseMipLevel + 8
    return {0: {0}, 1: {ini_vis, P.initialLightingMipLevel}, 2: {P.spatialVisibilityMipLevel, P.spatialLightingMipLevel},
            3: {P.spatialVisibilityMipLevel, P.spatialLightingMipLevel}, 4: set(), 5: {0}, 6: set(), 7: set()}[stage]


def set_candidates(g, M):
    p = g.getParams()
    p.initialM = M
    g.setParams(p)


def configure(g, M, learned):
    """M initial candidates, the learned proposal on (a freshly initialised network)
    or off, and an empty history: the state of a new pass on the same scene."""
    set_candidates(g, M)
    g.setLearned(enable=1 if learned else 0)
    g.resetHistory()


def lattice_of(img):
    return img[LATTICE // 2::LATTICE, LATTICE // 2::LATTICE].contiguous()


def quality_run(g, spec, M, learned, out, stream):
    """The 60-frame orbit from an empty history; returns the evaluation frames on the
    lattice and the learned statistics (None for the plain pass)."""
    import torch
    configure(g, M, learned)
    imgs = {}
    for f in range(QUALITY_FRAMES):
        g.setCamera(spec, f)
        g.execute(out, stream=stream)
        if f in EVAL_FRAMES:
            imgs[f] = lattice_of(out)
    torch.cuda.synchronize()
    return imgs, (g.learnedStats().tolist() if learned else None)


def relmse_of(g, imgs, refs):
    return sum(g.relmse(imgs[f], refs[f]) for f in EVAL_FRAMES) / len(EVAL_FRAMES)


def lattice_reference(g, spec, spp_total, stream):
    """spp_total path-traced samples per lattice pixel at the bound camera, summed in FP64."""
    import torch
    W, H = spec.width, spec.height
    lh, lw = g.latticeShape(W, H, LATTICE)
    chunks, spp = 16, 256
    while chunks * spp > spp_total:
        spp //= 2
    buf = torch.zeros((chunks, lh, lw, 4), device="cuda")
    ref = torch.zeros((lh, lw, 4), device="cuda")
    done, sn, first = 0, 1 << 24, True
    t0 = time.perf_counter()
    while done < spp_total:
        g.referenceLattice(W, H, LATTICE, spp, sn, buf, stream=stream)
        for c in range(chunks):
            g.accumulate(buf[c], reset=first, stream=stream)
            first = False
        sn += chunks
        done += chunks * spp
    g.accumulated(ref, stream=stream)
    torch.cuda.synchronize()
    return ref, done, time.perf_counter() - t0


def timed_run(g, spec, M, learned, warmup, steps, out, stream):
    """Warmup frames, then `steps` frames bracketed by device synchronisation, with HIP
    events around each stage on the stream the stages run on."""
    import torch
    configure(g, M, learned)
    W, H = spec.width, spec.height
    rounds = spec.params["spatialReuseRounds"]
    frame = 0
    for _ in range(warmup):
        g.setCamera(spec, frame)
        g.execute(out, stream=stream)
        frame += 1
    seq = [(0, 0)] + ([(6, 0)] if learned else []) + [(1, 0)] + ([(7, 0)] if learned else []) + [(2, 0)] + \
          [(3, r) for r in range(rounds)] + [(4, 0), (5, 0)]
    ev = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.setCamera(spec, frame)
        g.prepare_frame(W, H)
        for st, r in seq:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            g.run_stage(st, r, out=out if st == 5 else None, stream=stream)
            e1.record(stream)
            ev.append((st, e0, e1))
        g.end_frame()
        frame += 1
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms, calls = {}, {}
    for st, e0, e1 in ev:
        ms[st] = ms.get(st, 0.0) + e0.elapsed_time(e1)
        calls[st] = calls.get(st, 0) + 1
    return elapsed, {s: ms[s] / calls[s] for s in ms}, {s: ms[s] / steps for s in ms}


def cpu_baseline(spec, threads):
    """The oracle on bands of rows: 5 frames after 1 warm-up, median frame time, once
    single-threaded (8 rows) and once on `threads` threads (64 rows); per-pixel time
    scaled to the full frame."""
    from oracle.binding import Oracle
    o = Oracle()
    o.set_scene(spec)
    res = {}
    for nt, rows in ((1, 8), (threads, 64)):
        y0 = spec.height // 2 - rows // 2
        y1 = y0 + rows
        ts = []
        for f in range(6):
            o.set_camera(spec, f)
            t0 = time.perf_counter()
            o.execute(spec.width, spec.height, y0, y1, nt)
            if f > 0:
                ts.append(time.perf_counter() - t0)
        ts.sort()
        med = ts[len(ts) // 2]
        res[nt] = (med, rows, y0, y1, med / (rows * spec.width) * spec.width * spec.height)
    m1, mt = res[1], res[threads]
    nproc = os.cpu_count()
    return {"value": 1.0 / mt[4], "unit": "frames/s", "cores": threads, "kind": "port",
            "single_thread_value": 1.0 / m1[4],
            "sample": f"oracle, frames 1-5 after frame 0 (median), every kernel restricted to a band of full-width "
                      f"rows, per-pixel time scaled to {spec.width}x{spec.height}: {threads} threads on rows "
                      f"{mt[2]}-{mt[3]} ({mt[0]:.2f} s per band frame), 1 thread on rows {m1[2]}-{m1[3]} "
                      f"({m1[0]:.2f} s per band frame); host reports {nproc} CPUs, the GPU's share is 16"}


def bench_batched(args, steps=None, warmup=None, emit=True):
    """config2: 16 seeded views and volumes of uneven size. One step = one frame of
    every view. Timed twice over the same frames: the batch (one persistent launch per
    step, the frame a replayed HIP graph, queue shape from the autotuner) and the
    passes executed one after another on one stream. Prints its own line (emit) or
    returns the measurement."""
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    import torch

    from volumetricrestirrelease_amd import scenes
    from volumetricrestirrelease_amd.volumetric_restir import VolumetricReSTIR, VolumetricReSTIRBatch
    t0 = time.perf_counter()
    specs = scenes.config2()
    passes = []
    for spec in specs:
        g = VolumetricReSTIR()
        g.setScene(spec)
        passes.append(g)
    setup_s = time.perf_counter() - t0
    outs = [torch.zeros((s.height, s.width, 4), device="cuda") for s in specs]
    stream = torch.cuda.current_stream()
    px = sum(s.width * s.height for s in specs)
    b = VolumetricReSTIRBatch(passes)
    for g, spec in zip(passes, specs):
        g.setCamera(spec, 0)
    tune_ms = b.autotune(outs, frames=2, stream=stream)
    grid, per = b.tuning()

    def run(batched, first):
        frame = first
        for _ in range(warmup):
            for g, spec in zip(passes, specs):
                g.setCamera(spec, frame)
            if batched:
                b.execute(outs, stream=stream)
            else:
                for g, o in zip(passes, outs):
                    g.execute(o, stream=stream)
            frame += 1
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            for g, spec in zip(passes, specs):
                g.setCamera(spec, frame)
            if batched:
                b.execute(outs, stream=stream)
            else:
                for g, o in zip(passes, outs):
                    g.execute(o, stream=stream)
            frame += 1
        torch.cuda.synchronize()
        return time.perf_counter() - t

    el_b = run(True, 0)
    for g in passes:
        g.resetHistory()
    el_s = run(False, 0)
    b.close()
    del passes
    line = {
        "metric": "config2 batched frames/s (every view, one frame each)",
        "value": round(steps / el_b, 4),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(1000.0 * el_b / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (fp16 bricks above 128^3)",
        "data": "synthetic (16 seeded fBm volumes, seeded point lights, seeded views)",
        "config": {
            "workload": "config2: " + ", ".join(f"{s.width}x{s.height}/{s.density.shape[0]}^3/{len(s.lights)}L"
                                                for s in specs),
            "pixels_per_step": px,
            "autotuned": {"persistent_workgroups": grid, "tiles_per_item": per, "ms_per_frame": round(tune_ms, 3)},
            "one_after_another": {"value": round(steps / el_s, 4),
                                  "ms_per_step": round(1000.0 * el_s / steps, 3)},
            "speedup": round(el_s / el_b, 3),
            "volume_setup_s": round(setup_s, 2),
        },
    }
    if emit:
        print(json.dumps(line), flush=True)
    return line


def bench_config4(args):
    """config4: learned + batched at full scale. A res^3 animated emissive plume
    (density and temperature simulated on the GPU, 40 steps of warm-up, then one
    step per frame), seen by 4 views at 1080p that render as one batch; every frame
    the volume frame is rebuilt on the GPU from the simulation's grids. The
    60-frame animation runs twice from the same start: plain proposal, then learned
    proposal (trained online from frame 0). relMSE: the mean over the views and the
    last two frames against 65,536-spp FP64 lattice references of the same frames.
    The first view builds each frame's volume on the GPU; the other views bind it.
    `value` is the learned batched form's frames/s over rendering (the volume
    update -- simulation step and the GPU volume build -- is timed apart)."""
    import torch

    from volumetricrestirrelease_amd import scenes
    from volumetricrestirrelease_amd.volumetric_restir import Plume, VolumetricReSTIR, VolumetricReSTIRBatch
    say = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    res, frames, warm = args.c4_res, args.c4_frames, 40
    specs = scenes.config4_views(res=res, frames=frames)
    stream = torch.cuda.current_stream()
    outs = [torch.zeros((s.height, s.width, 4), device="cuda") for s in specs]
    evals = (frames - 2, frames - 1)
    refs = {}

    def run(learned):
        t0 = time.perf_counter()
        plume = Plume(res, 4)
        plume.advance(warm)
        d, t, _ = plume.grids()
        passes = []
        for v, s in enumerate(specs):
            g = VolumetricReSTIR()
            # the first view builds the volume, the others bind it
            if v == 0:
                g.setSceneStreamed(s, d, t)
            else:
                g.setSceneShared(s, passes[0], emissive=True)
            g.setLearned(enable=1 if learned else 0)
            passes.append(g)
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        b = VolumetricReSTIRBatch(passes)
        render, update, sim, ref_s, imgs = 0.0, 0.0, 0.0, 0.0, {}
        for f in range(frames):
            if f > 0:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                plume.advance(1)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                d, t, _ = plume.grids()
                passes[0].streamVolumeFrame(specs[0], d, t)
                for g in passes[1:]:
                    g.shareVolume(passes[0])
                torch.cuda.synchronize()
                update += time.perf_counter() - t0
                sim += t1 - t0
            for g, s in zip(passes, specs):
                g.setCamera(s, f)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            b.execute(outs, stream=stream)
            torch.cuda.synchronize()
            render += time.perf_counter() - t0
            if f in evals:
                for v, o in enumerate(outs):
                    imgs[(v, f)] = lattice_of(o)
                if (0, f) not in refs:  # the plain run makes the references, the learned run reuses them
                    for v, (g, s) in enumerate(zip(passes, specs)):
                        refs[(v, f)], spp, dt = lattice_reference(g, s, args.reference_spp, stream)
                        ref_s += dt
            if f % 10 == 0 or f == frames - 1:
                say(f"config4 {'learned' if learned else 'plain'} frame {f}: render {1000 * render / (f + 1):.1f} "
                    f"ms/frame, volume update {1000 * update / max(1, f):.1f} ms/frame")
        gi = passes[0].gridInfo(0)
        occ = gi["bricks"] / gi["brick_cells"]
        rel = sum(passes[v].relmse(imgs[(v, f)], refs[(v, f)]) for v in range(len(specs)) for f in evals)
        rel /= len(specs) * len(evals)
        stats = passes[0].learnedStats().tolist() if learned else None
        b.close()
        for g in passes:
            g.close()
        plume.close()
        torch.cuda.synchronize()
        return {"frames_per_s": round(frames / render, 4), "ms_per_frame": round(1000 * render / frames, 3),
                "volume_update_ms_per_frame": round(1000 * update / max(1, frames - 1), 3),
                "of_which_simulation_ms": round(1000 * sim / max(1, frames - 1), 3),
                "relmse": rel, "setup_s": round(setup, 2), "reference_s": round(ref_s, 2),
                "mip0_brick_occupancy": round(occ, 4), "learned_stats": stats}

    plain = run(False)
    say(f"config4 plain: {plain}")
    learned = run(True)
    say(f"config4 learned: {learned}")
    line = {
        "metric": "config4 learned + batched frames/s (4 views of a 1024^3 animated plume, 1080p each)",
        "value": learned["frames_per_s"],
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": frames,
        "warmup": 0,
        "ms_per_step": learned["ms_per_frame"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (fp16 bricks)",
        "data": "synthetic (seeded GPU plume simulation: buoyant updraft + curl noise, density + temperature)",
        "config": {
            "workload": f"config4: {res}^3 emissive plume, fp16 bricks, 4 mips with conservative twins, "
                        f"{frames} frames after {warm} warm-up steps, 4 views x 1920x1080 in one batch, "
                        "path length 3, M=8, 2 spatial rounds x 5 neighbours, 32 point lights + env, "
                        "volume rebuilt on the GPU every frame and shared by the views",
            "learned": learned,
            "plain": plain,
            "relmse_frames": list(evals),
            "reference_spp": args.reference_spp,
            "learned_matches_plain_relmse": learned["relmse"] <= plain["relmse"],
        },
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="config1", choices=["config0", "config1", "config2", "config4"],
                    help="config2: the batched workload (16 seeded views / volumes in one persistent launch per "
                         "step, replayed as a HIP graph) against the same passes rendered one after another; "
                         "config4: the learned + batched animated plume (4 views)")
    ap.add_argument("--c4-res", type=int, default=1024, help="config4 plume grid side")
    ap.add_argument("--c4-frames", type=int, default=60, help="config4 animation frames")
    ap.add_argument("--reference-spp", type=int, default=65536)
    ap.add_argument("--no-learned", action="store_true", help="time and report the plain pass only")
    ap.add_argument("--no-quality", action="store_true", help="skip the relMSE runs (profiling)")
    ap.add_argument("--time-learned", action="store_true",
                    help="time the learned-proposal pass at the config's M whatever the quality runs find "
                         "(profiling its network kernels)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="leave out the config2 batched measurement that follows the config1 line's timing")
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

    torch.cuda.set_device(0)
    if args.config == "config2":
        bench_batched(args)
        return
    if args.config == "config4":
        bench_config4(args)
        return
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
    W, H = spec.width, spec.height
    P = spec.params
    b, k, rounds, M0 = P["maxBounces"], P["spatialSampleCount"], P["spatialReuseRounds"], P["initialM"]
    stream = torch.cuda.current_stream()
    out = torch.zeros((H, W, 4), device="cuda")
    say = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731

    from volumetricrestirrelease_amd.volumetric_restir import VolumetricReSTIR
    g = VolumetricReSTIR()
    g.setScene(spec)
    setup_s = time.perf_counter() - t0
    quality = None
    learned_M = None
    if not args.no_quality:
        imgs, _ = quality_run(g, spec, M0, False, out, stream)
        refs, ref_s = {}, 0.0
        for f in EVAL_FRAMES:
            g.setCamera(spec, f)
            refs[f], spp, dt = lattice_reference(g, spec, args.reference_spp, stream)
            ref_s += dt
        plain_rel = relmse_of(g, imgs, refs)
        say(f"references: {spp} spp on the lattice, {len(EVAL_FRAMES)} frames in {ref_s:.1f} s; "
            f"plain M={M0} relMSE {plain_rel:.5f}")
        learned_runs = {}
        if not args.no_learned:
            for M in LEARNED_M:
                imgl, st = quality_run(g, spec, M, True, out, stream)
                r = relmse_of(g, imgl, refs)
                learned_runs[M] = {"relmse": r, "loss": st[0], "second_moment_analytic": st[1],
                                   "second_moment_network": st[2], "network_in_use": bool(st[3]),
                                   "adam_steps": int(st[4])}
                say(f"learned M={M}: relMSE {r:.5f}, stats {st}")
                if r <= plain_rel:
                    learned_M = M
                else:
                    break
        quality = {"reference_spp": spp, "reference_s": round(ref_s, 2),
                   "lattice": f"every {LATTICE}th pixel from ({LATTICE // 2}, {LATTICE // 2}): "
                              f"{imgs[EVAL_FRAMES[0]].shape[1]}x{imgs[EVAL_FRAMES[0]].shape[0]} pixels",
                   "frames": f"mean over frames {EVAL_FRAMES[0]}-{EVAL_FRAMES[-1]} of a {QUALITY_FRAMES}-frame orbit "
                             f"from an empty history, one reference per frame",
                   "plain": {"M": M0, "relmse": plain_rel},
                   "learned": {str(m): v for m, v in learned_runs.items()},
                   "learned_M_matched": learned_M}

    if args.time_learned:
        learned_M = M0
    headline_learned = learned_M is not None
    M_head = learned_M if headline_learned else M0
    elapsed, avg_ms, frame_ms = timed_run(g, spec, M_head, headline_learned, args.warmup, args.steps, out, stream)
    plain = None
    if headline_learned:
        e_p, avg_p, frame_p = timed_run(g, spec, M0, False, args.warmup, args.steps, out, stream)
        plain = {"value": round(args.steps / e_p, 4), "ms_per_step": round(1000.0 * e_p / args.steps, 3),
                 "M": M0, "stage_ms_avg": {STAGE_NAMES[s]: round(v, 3) for s, v in sorted(frame_p.items())}}
    fps = args.steps / elapsed
    dom = max(frame_ms, key=lambda s: frame_ms[s])
    prm = g.getParams()
    pools = {slot: g.brickPoolBytes(slot) for slot in stage_slots(dom, prm)}
    occupancy = g.volumeOccupancy(0)
    b_stream = stage_bytes_per_pixel(dom, b, k) * W * H
    b_vol = sum(pools.values())
    alg_bytes = b_stream + b_vol
    achieved = alg_bytes / (avg_ms[dom] * 1e-3) / 1e9
    traffic = PROFILE_TRAFFIC["bytes"].get(STAGE_NAMES[dom]) if args.config == "config1" else None
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
            "workload": f"{spec.name}: {spec.density.shape[0]}^3 fBm cloud, 8^3 "
                        f"{'fp16' if spec.brick_format == abi.BRICK_FP16 else 'fp32'} bricks, {spec.num_mips} "
                        f"average-reduction mips with conservative (dilated) twins, {W}x{H}, path length {b}, "
                        f"{rounds} spatial rounds x {k - 1} neighbours, {len(spec.lights)} point lights + env, "
                        f"orbiting camera; {occupancy:.1%} of mip-0 brick cells occupied",
            "proposal": (f"learned (online-trained MLP free-flight proposal), M={M_head}" if headline_learned else
                         f"plain (analytic trackers), M={M_head}"),
            "stage_ms_avg": {STAGE_NAMES[s]: round(v, 3) for s, v in sorted(frame_ms.items())},
            "volume_setup_s": round(setup_s, 2),
        },
        "quality": quality,
        "plain": plain,
        "roofline": {
            "bound": "hbm",
            "kernel": STAGE_NAMES[dom],
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": (f"profile constant ({PROFILE_TRAFFIC['file']}, plain M=8 pass), not measured in this "
                               f"run" if traffic else None),
            "algorithmic_bytes_per_launch": alg_bytes,
            "algorithmic_bytes": {"stream": b_stream, "volume_brick_pools": b_vol,
                                  "volume_slots": sorted(pools)},
            "avg_launch_ms": round(avg_ms[dom], 3),
        },
        "cpu_baseline": None,
    }
    if not args.no_batched and args.config == "config1":
        del g
        torch.cuda.synchronize()
        bl = bench_batched(args, steps=5, warmup=2, emit=False)
        line["batched"] = {"workload": bl["config"]["workload"], "frames_per_s": bl["value"],
                           "ms_per_step": bl["ms_per_step"], "steps": bl["steps"],
                           "pixels_per_step": bl["config"]["pixels_per_step"],
                           "autotuned": bl["config"]["autotuned"],
                           "one_after_another": bl["config"]["one_after_another"],
                           "speedup": bl["config"]["speedup"]}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(spec, args.cpu_threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
