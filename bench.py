"""Benchmark: plain-forward cone tracing at 1920x1080 on the depth-12 fBm heightfield.

One step = one primary frame (svo_trace_frame: ray generation, octree traversal
with the cone-footprint LOD stop, shading, usage ticks, child-request queue) of
the seeded 100-frame orbit.  Inputs (tree, usages, camera table) are resident
in HBM before the timed region; the timed region is bracketed by stream
synchronisation and measured with HIP events on the context's stream.

Prints one JSON line.  Its roofline prices the trace kernel's algorithmic bytes
(algorithmic_bytes below) against HBM, and its VALU instructions (from the committed
counter profile) against the VALU issue rate, the roof the kernel presses.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "svo-tracer_amd"))

HBM_PEAK_GBS = 8000.0
# VALU issue of MI355X: 256 CUs, each issuing one 64-lane VALU instruction per clock (four SIMDs,
# a wave64 instruction every fourth clock each) at the 2.4 GHz peak engine clock
VALU_PEAK_GINSTS = 256 * 2.4
TRAFFIC_PROFILE = "profiles/trace_frame_1080p_d12_traffic.txt"
COUNTER_PROFILE = "profiles/trace_frame_1080p_d12_counters.txt"
METRIC = "Mcones/s @1080p depth-12 (plain vs batched); fwd+adjoint Mcones/s; HBM % peak"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(stats, cones: int, groups_touched: int) -> int:
    """Cache-less byte model of one frame: 2 B per Bases chunk examined, 16 B per block
    record examined, 24 B of per-cone output (hit u32, visit count u32, RGBA f32x4) and
    8 B per usage group whose tick changed."""
    return 2 * stats.base_steps + 16 * stats.block_steps + 24 * cones + 8 * groups_touched


# the sources trace_frame_kernel is built and launched from: a counter profile describes this
# build only when it records the same digest
KERNEL_SOURCES = ("svo-tracer_amd/csrc/trace.hip", "svo-tracer_amd/csrc/svo_device.hpp",
                  "svo-tracer_amd/csrc/svo_kernels.hpp", "svo-tracer_amd/csrc/frame.cpp")


def kernel_sources_digest() -> str:
    h = hashlib.sha1()
    for f in KERNEL_SOURCES:
        h.update((ROOT / f).read_bytes())
    return h.hexdigest()[:16]


def read_profile_value(profile: Path, key: str, as_text: bool = False):
    """A `key: value` line of a committed rocprofv3 summary, if present."""
    if not profile.exists():
        return None
    for line in profile.read_text().splitlines():
        if line.startswith(key + ":"):
            v = line.split(":", 1)[1].strip()
            return v if as_text else float(v)
    return None


def profile_matches_build(profile: Path) -> bool:
    """True when the profile was taken of a kernel built from the current sources."""
    return read_profile_value(profile, "kernel_sources_sha1", as_text=True) == kernel_sources_digest()


def cpu_baseline(svo, tree, cams, seconds: float):
    """The oracle (OpenMP over cones, root-restart walk as in wavefront.cl) on whole
    frames of the same orbit, until about `seconds` of host time have been spent."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    bases, blocks = tree.bases, tree.blocks
    threads = max(1, min(16, os.cpu_count() or 1))
    n = cams[0].screen_size[0] * cams[0].screen_size[1]
    ids = np.arange(n, dtype=np.uint32)
    frames, spent = 0, 0.0
    while spent < seconds and frames < len(cams):
        t0 = time.perf_counter()
        oracle.trace_rays(cams[frames], bases, blocks, ids, threads=threads)
        spent += time.perf_counter() - t0
        frames += 1
    model = "unknown CPU"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": round(frames * n / spent / 1e6, 4),
        "unit": "Mcones/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{frames} whole frames of the same orbit ({frames * n} cones, {spent:.1f} s host time) "
                  f"on {threads} threads of {model} ({os.cpu_count()} logical CPUs visible)",
    }


def copy_bandwidth(svo, mib: int = 2048, reps: int = 5) -> float:
    """GB/s of a device-to-device copy of `mib` MiB (read + write bytes over time): the
    practical HBM ceiling on this GPU, reported beside the 8 TB/s peak."""
    B = svo.BufferName
    n = mib << 20
    with svo.ComputeManager(0) as cm:
        cm.InitDeviceBuffer(B.RootLocations, n)
        cm.InitDeviceBuffer(B.RootDirections, n)
        cm.CopyBuffer(B.RootDirections, B.RootLocations, n)
        cm.Wait()
        cm.TimerBegin()
        for _ in range(reps):
            cm.CopyBuffer(B.RootDirections, B.RootLocations, n)
        ms = cm.TimerEnd()
    return 2.0 * n * reps / (ms / 1e3) / 1e9


def graph_opts(opts: dict) -> dict:
    return {k: v for k, v in opts.items() if v}


def batched_leg(svo, cm, steps: int, warmup: int, usages0, opts) -> dict:
    """The 64-camera set of uneven resolution (log-uniform 256^2..1920x1080, apertures 1-4 px,
    seed 11) over the same scene: one svo_trace_batch launch per step, beside the same
    cameras traced with one launch each, both submitted directly and as replayed HIP graphs."""
    B = svo.BufferName
    cams, apertures = svo.scenes.config4_cameras(count=64, tick=1)
    n = sum(c.screen_size[0] * c.screen_size[1] for c in cams)
    cm.InitDeviceBuffer(B.OutputImage, n, 16)
    cm.InitDeviceBuffer(B.HitIndices, n, 4)
    cm.InitDeviceBuffer(B.VisitCounts, n, 4)

    def frame_cams(f):
        for c in cams:
            c.tick = f + 1
        return cams

    def timed(fn):
        cm.WriteBuffer(B.Usages, usages0)
        for f in range(warmup):
            fn(f)
        cm.Wait()
        cm.WriteBuffer(B.Usages, usages0)
        cm.Wait()
        cm.TimerBegin()
        for f in range(steps):
            fn(f)
        return cm.TimerEnd() / steps

    def batch(f):
        cm.TraceBatch(frame_cams(f), apertures, **opts)

    def loop(f):
        for c, a in zip(frame_cams(f), apertures):
            cm.TraceBatch([c], [a], **opts)

    ms_batch = timed(batch)
    ms_loop = timed(loop)
    # launch options for the batched frame chosen by the autotuner on the frame itself
    cm.WriteBuffer(B.Usages, usages0)
    tuned, tuned_ms = cm.Autotune(frame_cams(0), apertures, frames=3, **graph_opts(opts))
    tuned = {k: v for k, v in tuned.items() if v}
    log(f"autotuned batched options: {tuned} ({tuned_ms:.2f} ms/frame)")
    ms_batch_tuned = timed(lambda f: cm.TraceBatch(frame_cams(f), apertures, **tuned))
    # the same two submissions captured once into HIP graphs and replayed per step
    g_batch = cm.CaptureFrame(frame_cams(0), apertures, per_camera=False, **tuned)
    g_loop = cm.CaptureFrame(frame_cams(0), apertures, per_camera=True, **graph_opts(opts))
    try:
        ms_batch_graph = timed(lambda f: g_batch.Launch(frame_cams(f)))
        ms_loop_graph = timed(lambda f: g_loop.Launch(frame_cams(f)))
    finally:
        g_batch.Dispose()
        g_loop.Dispose()
    rate = lambda ms: round(n / (ms / 1e3) / 1e6, 2)
    return {
        "value": rate(ms_batch_graph),
        "unit": "Mcones/s",
        "ms_per_step": round(ms_batch_graph, 3),
        "cameras": len(cams),
        "cones_per_step": n,
        "submission": "hipGraph replay of one batched launch with autotuned options",
        "autotuned_options": tuned,
        "direct_launch": {"value": rate(ms_batch), "ms_per_step": round(ms_batch, 3)},
        "direct_launch_autotuned": {"value": rate(ms_batch_tuned), "ms_per_step": round(ms_batch_tuned, 3)},
        "per_camera_loop": {"value": rate(ms_loop), "ms_per_step": round(ms_loop, 3)},
        "per_camera_loop_graph": {"value": rate(ms_loop_graph), "ms_per_step": round(ms_loop_graph, 3)},
        "workload": "64 cameras, resolutions log-uniform 256x256..1920x1080, apertures 1-4 px, seed 11, "
                    "one launch per step vs one launch per camera",
    }


def adjoint_leg(svo, steps: int, fit_steps: int) -> dict:
    """Forward + adjoint on the depth-10 sphere+noise scene (seed 3), 32 views of 512x512:
    targets rendered from the true leaf colours and opacities, parameters perturbed by
    N(0, 0.1) (colour) and N(0, 0.05) (opacity), clipped to [0, 1]; L2 loss.  Times the
    plain forward of the same views, the compositing forward and the adjoint launch (which
    is forward + adjoint in one traversal), then runs the Adam fit."""
    B = svo.BufferName
    spec = svo.scenes.config5_scene()
    tree = spec.build()
    views = svo.scenes.config5_cameras()
    for v in views:
        v.tick = 1
    n = sum(v.screen_size[0] * v.screen_size[1] for v in views)
    blocks = tree.blocks
    truth = np.zeros((blocks.size, 4), np.float32)
    truth[:, 0] = blocks["colour_r"] / np.float32(255.0)
    truth[:, 1] = blocks["colour_g"] / np.float32(255.0)
    truth[:, 2] = blocks["colour_b"] / np.float32(255.0)
    truth[:, 3] = blocks["opacity"] / np.float32(255.0)
    rng = np.random.default_rng(3)
    params = truth.copy()
    params[:, :3] = np.clip(params[:, :3] + rng.normal(0, 0.1, (blocks.size, 3)), 0, 1)
    params[:, 3] = np.clip(params[:, 3] + rng.normal(0, 0.05, blocks.size), 0, 1)
    cm = svo.ComputeManager(0)
    try:
        cm.LoadTree(tree, with_usages=False)
        cm.InitDeviceBuffer(B.OutputImage, n, 16)
        cm.InitBuffer(B.LeafParams, truth)
        cm.Composite(views)
        target = cm.ReadBuffer(B.OutputImage, np.float32, 4 * n)
        cm.InitBuffer(B.TargetImage, target)
        cm.InitBuffer(B.LeafParams, params)
        cm.InitBuffer(B.LeafGradients, np.zeros_like(params))
        cm.InitBuffer(B.AdamM, np.zeros_like(params))
        cm.InitBuffer(B.AdamV, np.zeros_like(params))

        def timed(fn):
            for _ in range(2):
                fn()
            cm.Wait()
            cm.TimerBegin()
            for _ in range(steps):
                fn()
            return cm.TimerEnd() / steps

        ms_plain = timed(lambda: cm.TraceBatch(views))
        ms_fwd = timed(lambda: cm.Composite(views))
        ms_adj = timed(lambda: cm.Composite(views, adjoint=True))
        cm.InitBuffer(B.LeafParams, params)
        cm.InitBuffer(B.LeafGradients, np.zeros_like(params))
        losses = []
        t0 = time.perf_counter()
        for step in range(1, fit_steps + 1):
            losses.append(cm.Composite(views, adjoint=True))
            cm.AdamStep(step, learning_rate=0.01)
        cm.Wait()
        fit_s = time.perf_counter() - t0
        final = cm.Composite(views)
    finally:
        cm.Dispose()
    rate = lambda ms: round(n / (ms / 1e3) / 1e6, 2)
    # one adjoint launch is a whole forward + adjoint step: it composites the image, layers
    # and loss in the same traversal that back-propagates (the Adam loop runs exactly that)
    return {
        "value": rate(ms_adj),
        "unit": "Mcones/s",
        "ms_per_step": round(ms_adj, 3),
        "workload": "depth-10 sphere+noise (seed 3), 32 views of 512x512, 4 layers, L2 loss",
        "cones_per_step": n,
        "plain_forward_ms": round(ms_plain, 3),
        "composite_forward_ms": round(ms_fwd, 3),
        "forward_plus_adjoint_ms": round(ms_adj, 3),
        "cost_vs_plain_forward": round(ms_adj / ms_plain, 3),
        "separate_forward_then_adjoint_ms": round(ms_fwd + ms_adj, 3),
        "separate_cost_vs_plain_forward": round((ms_fwd + ms_adj) / ms_plain, 3),
        "adam_fit": {"steps": fit_steps, "loss_first": round(losses[0], 4), "loss_last": round(final, 4),
                     "seconds": round(fit_s, 2)},
    }


def wavefront_leg(svo, tree, cams, frames: int = 3) -> dict:
    """The reference's own frame structure on the same GPU: MainWindow.RunWavefront's loop
    (init, then runBaseTrace / runBlockTrace launches with a blocking queue-size read-back
    between them, evaluateBackground, resolveRemainders, drawTrace; MainWindow.cs:356-481)
    driven through svo_enqueue, as an unchanged caller would.  Times whole frames; checks
    the image against the fused kernel's."""
    K, B = svo.KernelName, svo.BufferName
    cm = svo.ComputeManager(0)
    try:
        cm.LoadTree(tree, with_usages=True, usage_tick=0)
        W, H = cams[0].screen_size[0], cams[0].screen_size[1]
        n = W * H
        cm.InitBuffer(B.ChildRequestID, np.zeros(1, np.uint32))
        cm.InitDeviceBuffer(B.ChildRequests, n, 40)
        for name, size in ((B.Origins, 16), (B.Directions, 16), (B.FoVs, 4), (B.Locations, 32), (B.Depths, 1),
                           (B.Weightings, 1), (B.ParentTraces, 4), (B.ColourRs, 1), (B.ColourGs, 1),
                           (B.ColourBs, 1), (B.BaseTraces, 4), (B.BaseTraceQueue, 4), (B.BlockTraces, 4),
                           (B.BlockTraceQueue, 4), (B.BackgroundQueue, 4), (B.MaterialQueue, 4),
                           (B.FinalColourRs, 4), (B.FinalColourGs, 4), (B.FinalColourBs, 4),
                           (B.FinalWeightings, 4), (B.OutputImage, 16)):
            cm.InitDeviceBuffer(name, n, size)
        for name in (B.BaseTraceQueueID, B.BlockTraceQueueID, B.BackgroundQueueID, B.MaterialQueueID,
                     B.AccumulatorID):
            cm.InitBuffer(name, np.zeros(1, np.uint32))

        def bind(kernel, names):
            for param in names:
                cm.SetArg(kernel, param, B[param])

        def frame(cam):
            for name in (B.BlockTraceQueueID, B.BaseTraceQueueID, B.BackgroundQueueID, B.MaterialQueueID,
                         B.AccumulatorID, B.ChildRequestID):
                cm.ZeroIDBuffer(name)
            cam.max_child_request_id = n
            cm.SetArg(K.Init, "input", cam)
            bind(K.Init, ["Origins", "Directions", "FoVs", "Locations", "Depths", "BaseTraceQueue",
                          "BaseTraceQueueID", "BackgroundQueue", "BackgroundQueueID", "ColourRs", "ColourGs",
                          "ColourBs", "Weightings", "ParentTraces", "FinalColourRs", "FinalColourGs",
                          "FinalColourBs", "FinalWeightings"])
            cm.Enqueue(K.Init, [W, H])
            cm.SetArg(K.RunBaseTrace, "input", cam)
            cm.SetArg(K.RunBlockTrace, "input", cam)
            iterations = 0
            while True:
                base_size = cm.ResetIDBuffer(B.BaseTraceQueueID)
                if base_size > 0:
                    cm.Swap(B.BaseTraces, B.BaseTraceQueue)
                    bind(K.RunBaseTrace, ["Bases", "Origins", "Directions", "FoVs", "Locations", "Weightings",
                                          "Depths", "BaseTraces", "BaseTraceQueue", "BaseTraceQueueID",
                                          "BlockTraceQueue", "BlockTraceQueueID", "BackgroundQueue",
                                          "BackgroundQueueID"])
                    cm.Enqueue(K.RunBaseTrace, [base_size])
                block_size = cm.ResetIDBuffer(B.BlockTraceQueueID)
                if block_size > 0:
                    cm.Swap(B.BlockTraces, B.BlockTraceQueue)
                    bind(K.RunBlockTrace, ["Blocks", "Usages", "ChildRequestID", "ChildRequests", "Origins",
                                           "Directions", "FoVs", "Locations", "Weightings", "Depths",
                                           "BlockTraces", "BlockTraceQueue", "BlockTraceQueueID", "BaseTraceQueue",
                                           "BaseTraceQueueID", "BackgroundQueue", "BackgroundQueueID",
                                           "MaterialQueue", "MaterialQueueID", "ParentTraces"])
                    cm.Enqueue(K.RunBlockTrace, [block_size])
                if base_size == 0 and block_size == 0:
                    break
                iterations += 1
            bg = cm.ResetIDBuffer(B.BackgroundQueueID)
            cm.SetArg(K.EvaluateBackground, "input", cam)
            if bg:
                bind(K.EvaluateBackground, ["BackgroundQueue", "Directions", "ParentTraces", "ColourRs",
                                            "ColourGs", "ColourBs", "Weightings", "FinalColourRs", "FinalColourGs",
                                            "FinalColourBs", "FinalWeightings"])
                cm.Enqueue(K.EvaluateBackground, [bg])
            cm.SetArg(K.ResolveRemainders, "input", cam)
            bind(K.ResolveRemainders, ["Blocks", "Usages", "Locations", "Depths", "MaterialQueue", "FinalColourRs",
                                       "FinalColourGs", "FinalColourBs", "FinalWeightings", "Weightings",
                                       "ParentTraces"])
            mat = cm.ResetIDBuffer(B.MaterialQueueID)
            if mat:
                cm.Enqueue(K.ResolveRemainders, [mat])
            bind(K.DrawTrace, ["FinalColourRs", "FinalColourGs", "FinalColourBs", "FinalWeightings"])
            cm.SetArg(K.DrawTrace, "outputImage", B.OutputImage)
            cm.SetArg(K.DrawTrace, "input", cam)
            cm.Enqueue(K.DrawTrace, [W, H])
            cm.Wait()
            return iterations

        frame(cams[0])  # warm-up
        ms, its = [], []
        for cam in cams[1:1 + frames]:
            t0 = time.perf_counter()
            its.append(frame(cam))
            ms.append((time.perf_counter() - t0) * 1e3)
        img = cm.ReadBuffer(B.OutputImage, np.float32, 4 * n)
        # the fused kernel on the same camera and tree, for the image check
        cm.InitDeviceBuffer(B.HitIndices, n, 4)
        cm.InitDeviceBuffer(B.VisitCounts, n, 4)
        cm.TraceFrame(cams[frames])
        cm.Wait()
        same = np.array_equal(img, cm.ReadBuffer(B.OutputImage, np.float32, 4 * n))
    finally:
        cm.Dispose()
    t = float(np.mean(ms))
    return {"ms_per_frame": round(t, 2), "value": round(n / (t / 1e3) / 1e6, 2), "unit": "Mcones/s",
            "queue_iterations_per_frame": float(np.mean(its)), "image_equals_fused_kernel": bool(same),
            "what": "the reference's RunWavefront loop through svo_enqueue (one launch per kernel per "
                    "iteration, blocking queue-size read-backs), same frames and tree"}


def stream_leg(svo, frames: int, width: int = 3840, height: int = 2160, log2_blocks: int = 26,
               keep_requests: bool = True) -> dict:
    """Streaming: a 3840x2160 camera flies the seeded 300-frame spline over the fBm terrain
    generated on the device (16384^2 grid) and refined on demand to depth 14 in a fixed pool of
    2^26 blocks; per frame the trace makes child requests and ticks usages, and the streaming
    step grafts up to 65,536 bricks, collects groups unused for 30 frames and compacts when
    fewer than 10 % of the groups are free.  keep_requests: the frames keep their requests for
    the streaming step (SVO_FRAME_KEEP_REQUESTS) instead of writing them to ChildRequests."""
    B = svo.BufferName
    Bd, D, n_groups = 4, 14, 1 << (log2_blocks - 3)
    n = width * height
    cm = svo.ComputeManager(0)
    try:
        t0 = time.perf_counter()
        cm.InitDeviceBuffer(B.Bases, sum(8 ** d for d in range(1, Bd + 1)), 2)
        cm.InitDeviceBuffer(B.Blocks, 8 * n_groups, 16)
        cm.InitDeviceBuffer(B.Usages, n_groups, 8)
        cm.InitDeviceBuffer(B.ChildRequests, n, 40)
        cm.InitBuffer(B.ChildRequestID, np.zeros(1, np.uint32))
        cm.InitDeviceBuffer(B.OutputImage, n, 16)
        cm.InitDeviceBuffer(B.HitIndices, n, 4)
        cm.InitDeviceBuffer(B.VisitCounts, n, 4)
        terrain = cm.CreateTerrain(svo.scenes.config3_geometry())
        log(f"stream: terrain 16384^2 generated in {time.perf_counter() - t0:.2f} s")
        cm.StreamSeed(terrain, Bd)
        cm.Wait()
        setup_s = time.perf_counter() - t0
        log(f"stream: seed tree ready after {setup_s:.2f} s")
        cams = [svo.scenes.config3_camera(f, width, height, frames=frames, base_depth=Bd) for f in range(frames)]
        frame_wall_ms, stats = [], []
        # the trace kernel is timed with events on the stream (no synchronisation between the
        # trace and the update, as in a streaming loop); the update ends with the stream idle
        cm.KernelTiming(True)
        cm.KernelTiming(None, reset=True)
        wall0 = time.perf_counter()
        for f, cam in enumerate(cams):
            cam.tick = f + 1
            cam.max_child_request_id = n
            t1 = time.perf_counter()
            cm.TraceFrame(cam, flags=svo.FRAME_KEEP_REQUESTS if keep_requests else 0)
            stats.append(cm.StreamUpdate(terrain, Bd, D, f + 1, max_bricks=65536, max_child_request_id=n,
                                         prune_age=30, gc_free_percent=10))
            frame_wall_ms.append((time.perf_counter() - t1) * 1e3)
            if f < 3 or f % 25 == 24:
                log(f"stream frame {f + 1}: {frame_wall_ms[-1]:.2f} ms, "
                    f"{stats[-1]['requests']} requests, {stats[-1]['bricks']} bricks, "
                    f"free groups {stats[-1]['free_groups']}")
        cm.Wait()
        wall_s = time.perf_counter() - wall0
        trace_total, launches = cm.KernelTiming(None, reset=True)
        cm.KernelTiming(False)
        trace_mean = trace_total / max(1, launches)
        # one garbage collection of the final pool, timed on its own: reclaim of groups unused
        # for 30 frames, then compaction
        ui = svo.UpdateInputData()
        ui.base_depth, ui.tick, ui.offset = Bd, frames, 30
        t1 = time.perf_counter()
        reclaimed = cm.Reclaim(ui)
        t2 = time.perf_counter()
        moved = cm.Compact(ui)
        t3 = time.perf_counter()
        log(f"stream: final garbage collection reclaimed {reclaimed} groups in {(t2 - t1) * 1e3:.2f} ms, "
            f"compaction moved {moved} in {(t3 - t2) * 1e3:.2f} ms")
        terrain.close()
    finally:
        cm.Dispose()
    frame_ms = wall_s * 1e3 / frames
    return {
        "value": round(n / (frame_ms / 1e3) / 1e6, 2),
        "unit": "Mcones/s",
        "ms_per_step": round(frame_ms, 3),
        "workload": f"{width}x{height}, {frames}-frame seeded spline (seed 3), fBm terrain 16384^2 generated on "
                    f"the device, refined on demand to depth {D}, pool 2^{log2_blocks} blocks, <= 65536 bricks/frame, "
                    "prune age 30 frames, GC below 10 % free, requests "
                    + ("kept by the frame for the update" if keep_requests else "through ChildRequests"),
        "frames": frames,
        "trace_ms_mean": round(trace_mean, 3),
        "update_ms_mean": round(float(np.mean(frame_wall_ms)) - trace_mean, 3),
        "frame_ms_max": round(float(np.max(frame_wall_ms)), 3),
        "timing": "trace_ms_mean: the trace kernel by HIP events; update_ms_mean: the mean frame wall time "
                  "less that (frame begin, compaction, the streaming update, host time)",
        "bricks_total": int(sum(s["bricks"] for s in stats)),
        "bricks_per_frame_max": int(max(s["bricks"] for s in stats)),
        "requests_per_frame_mean": round(float(np.mean([s["requests"] for s in stats])), 1),
        "collected_total": int(sum(s["collected"] for s in stats)),  # groups the grafts' collection gathered
        "gc_runs": int(sum(s["gc_runs"] for s in stats)),
        "reclaimed_total": int(sum(s["reclaimed"] for s in stats)),
        "moved_total": int(sum(s["moved"] for s in stats)),
        "free_groups_last": int(stats[-1]["free_groups"]),
        "final_gc": {"reclaimed": reclaimed, "reclaim_ms": round((t2 - t1) * 1e3, 3), "moved": moved,
                     "compact_ms": round((t3 - t2) * 1e3, 3)},
        "pool_groups": n_groups,
        "setup_s": round(setup_s, 2),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=12)
    ap.add_argument("--workgroup", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--waves", type=int, default=0, help="register budget of the trace kernel (0 = default 5, or 4..6)")
    ap.add_argument("--chunk", type=int, default=0, help="cones a wave claims per atomic")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-camera batched leg")
    ap.add_argument("--no-wavefront", action="store_true", help="skip the kernel-by-kernel reference-loop leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward + adjoint leg")
    ap.add_argument("--no-stream", action="store_true", help="skip the streaming leg")
    ap.add_argument("--stream-frames", type=int, default=300)
    ap.add_argument("--stream-requests", choices=("kept", "buffer"), default="kept",
                    help="streaming legs: frames keep their child requests for the update, or write ChildRequests")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import _loader

    svo = _loader.load()
    if svo.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1

    if args.calibrate_copy:
        B = svo.BufferName
        n = args.calibrate_copy << 20
        with svo.ComputeManager(0) as cm:
            cm.InitDeviceBuffer(B.RootLocations, n)
            cm.InitDeviceBuffer(B.RootDirections, n)
            for _ in range(5):
                cm.CopyBuffer(B.RootDirections, B.RootLocations, n)
            cm.Wait()
        print(json.dumps({"calibration_copy_bytes": n, "copies": 5}))
        return 0

    W, H, K, Wm = args.width, args.height, args.steps, args.warmup
    N = W * H
    t0 = time.perf_counter()
    spec = svo.scenes.config2_scene(max_depth=args.depth)
    tree = spec.build()
    host_build_s = time.perf_counter() - t0
    log(f"scene {spec.name}: {tree.block_count} blocks ({tree.block_count * 16 / 2**20:.0f} MiB), "
        f"built on the host in {host_build_s:.1f} s")

    B = svo.BufferName
    cm = svo.ComputeManager(0)
    # the same tree built on the device (svo_terrain_build_tree) and bound for the frames;
    # it must equal the host-built one block for block
    terrain = cm.CreateTerrain(spec.geometry)
    build_ms = []
    for rep in range(4):
        cm.Wait()
        t0 = time.perf_counter()
        count, _ = cm.BuildTreeOnDevice(terrain, spec.base_depth, spec.max_depth, spec.max_size, with_usages=True,
                                        usage_tick=0)
        cm.Wait()
        build_ms.append((time.perf_counter() - t0) * 1e3)
    terrain.close()
    identical = count == tree.block_count and np.array_equal(
        cm.ReadBuffer(B.Blocks, np.uint8), tree.blocks.view(np.uint8).reshape(-1)) and np.array_equal(
        cm.ReadBuffer(B.Bases, np.uint16), tree.bases)
    device_build_ms = float(np.median(build_ms[1:]))
    log(f"scene built on the device in {device_build_ms:.1f} ms (first call {build_ms[0]:.1f} ms), "
        f"identical to the host build: {identical}")
    if not identical:
        raise SystemExit("device-built tree differs from the host-built tree")
    usages0 = cm.ReadBuffer(B.Usages, svo.USAGE_DTYPE)
    cm.InitBuffer(B.ChildRequests, np.zeros(6000 * 40, np.uint8))
    cm.InitBuffer(B.ChildRequestID, np.zeros(1, np.uint32))
    cm.InitDeviceBuffer(B.OutputImage, N, 16)
    cm.InitDeviceBuffer(B.HitIndices, N, 4)
    cm.InitDeviceBuffer(B.VisitCounts, N, 4)
    opts = dict(workgroup_size=args.workgroup, refill_lanes=args.refill, waves_per_simd=args.waves,
                work_chunk=args.chunk)
    # the timed frames are spread over the whole 100-frame orbit (poses f * 100 / K)
    cams = [svo.scenes.config2_camera((f * 100) // K, W, H, frames=100, tick=f + 1) for f in range(K)]

    # untimed pass: exact per-frame work counts for the byte model
    alg = []
    requests = 0
    for f, cam in enumerate(cams):
        cm.TraceFrame(cam, **opts)
        st = cm.FrameStats()
        ticks = cm.ReadBuffer(B.Usages, svo.USAGE_DTYPE)["tick"]
        g = int(np.count_nonzero(ticks == cam.tick)) - int(np.count_nonzero(usages0["tick"] == cam.tick))
        alg.append(algorithmic_bytes(st, N, g))
        requests += st.child_requests
    log(f"work model: {np.mean(alg) / 1e6:.1f} MB/frame algorithmic, {requests / K:.0f} child requests/frame")

    def reset():
        cm.WriteBuffer(B.Usages, usages0)

    reset()
    for f in range(Wm):
        cm.TraceFrame(cams[f % K], **opts)
    cm.Wait()
    reset()

    cm.KernelTiming(True)
    cm.KernelTiming(None, reset=True)
    cm.Wait()
    cm.TimerBegin()
    for cam in cams:
        cm.TraceFrame(cam, **opts)
    total_ms = cm.TimerEnd()
    kernel_ms, launches = cm.KernelTiming(None, reset=True)
    cm.KernelTiming(False)
    log(f"{K} frames in {total_ms:.2f} ms; trace kernel {kernel_ms / launches:.3f} ms/launch")

    ms_per_step = total_ms / K
    value = N * K / (total_ms / 1e3) / 1e6
    kernel_s = kernel_ms / launches / 1e3
    bytes_per_launch = float(np.mean(alg))
    achieved = bytes_per_launch / kernel_s / 1e9
    # counter figures come from the committed profiles, and only when those were taken of this
    # kernel build (same source digest); otherwise they are reported as null
    traffic_fresh = profile_matches_build(ROOT / TRAFFIC_PROFILE)
    counters_fresh = profile_matches_build(ROOT / COUNTER_PROFILE)
    traffic = read_profile_value(ROOT / TRAFFIC_PROFILE, "traffic_bytes_per_dispatch") if traffic_fresh else None
    valu_insts = read_profile_value(ROOT / COUNTER_PROFILE, "valu_insts_per_dispatch") if counters_fresh else None
    valu_util = read_profile_value(ROOT / COUNTER_PROFILE, "valu_utilization_percent") if counters_fresh else None
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mcones/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 fixed point + f32",
        "data": "synthetic (seeded fBm heightfield scene, seeded orbit cameras)",
        "config": {
            "workload": f"plain forward, {W}x{H} primary cones, depth-{args.depth} fBm heightfield "
                        f"({tree.block_count} x 16-B blocks), {K} frames spread over the 100-frame seeded orbit",
            "cones_per_step": N,
            "tree_blocks": tree.block_count,
        },
        "scene_build": {
            "host_s": round(host_build_s, 2),
            "device_ms": round(device_build_ms, 2),
            "identical": bool(identical),
            "entry_point": "svo_terrain_build_tree (TreeBuilder.BuildTree on the device)",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": f"{TRAFFIC_PROFILE}: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE per dispatch, scaled by "
                              "the same counters over a device-to-device copy; kernel sources sha1 "
                              f"{kernel_sources_digest()}, profile of this build: {traffic_fresh}",
        },
    }
    if valu_insts:
        # the roof the kernel actually presses: VALU issue (SQ_INSTS_VALU of the committed
        # counter profile over this run's kernel time)
        valu_achieved = valu_insts / kernel_s / 1e9
        line["roofline"]["valu_issue"] = {
            "achieved": round(valu_achieved, 1),
            "peak": VALU_PEAK_GINSTS,
            "unit": "G wave-instructions/s",
            "frac": round(valu_achieved / VALU_PEAK_GINSTS, 4),
            "insts_per_dispatch": valu_insts,
            "lane_utilization_percent": valu_util,
            "source": f"{COUNTER_PROFILE}: rocprofv3 --pmc SQ_INSTS_VALU, VALUUtilization per dispatch "
                      "(profile of this build)",
        }
    copy_gbs = copy_bandwidth(svo)
    line["hbm_copy"] = {"value": round(copy_gbs, 1), "unit": "GB/s",
                        "what": "2 GiB device-to-device copy, read + write bytes",
                        "achieved_vs_copy": round(achieved / copy_gbs, 4)}
    log(f"device copy: {copy_gbs:.0f} GB/s; trace kernel at {achieved / copy_gbs:.3f} of it")
    if not args.no_wavefront:
        log("reference loop (kernel by kernel) leg...")
        line["kernel_by_kernel"] = wavefront_leg(svo, tree, cams)
        log(f"kernel-by-kernel: {line['kernel_by_kernel']['ms_per_frame']} ms/frame "
            f"({line['kernel_by_kernel']['queue_iterations_per_frame']:.0f} queue iterations), image equal: "
            f"{line['kernel_by_kernel']['image_equals_fused_kernel']}")
    if not args.no_batched:
        log("batched 64-camera leg...")
        line["batched"] = batched_leg(svo, cm, max(3, min(K, 10)), 2, usages0, opts)
        log(f"batched: {line['batched']['value']} Mcones/s, per-camera loop "
            f"{line['batched']['per_camera_loop']['value']} Mcones/s")
    if not args.no_adjoint:
        log("forward + adjoint leg...")
        line["adjoint"] = adjoint_leg(svo, max(3, min(K, 10)), 200)
        log(f"adjoint: {line['adjoint']['value']} Mcones/s, {line['adjoint']['cost_vs_plain_forward']}x plain forward")
    if not args.no_stream:
        log("streaming leg...")
        line["streaming"] = stream_leg(svo, args.stream_frames, keep_requests=args.stream_requests == "kept")
        log(f"streaming: {line['streaming']['value']} Mcones/s, update {line['streaming']['update_ms_mean']} ms/frame")
        log("streaming leg with a pool of 2^22 blocks (garbage collection under pressure)...")
        line["streaming_small_pool"] = stream_leg(svo, 150, log2_blocks=22,
                                                  keep_requests=args.stream_requests == "kept")
        sp = line["streaming_small_pool"]
        log(f"streaming, small pool: {sp['value']} Mcones/s, {sp['collected_total']} groups gathered by the grafts' "
            f"collection (free, or unused for 30 frames), {sp['gc_runs']} garbage collections "
            f"({sp['reclaimed_total']} reclaimed, {sp['moved_total']} moved), update {sp['update_ms_mean']} ms/frame")
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on a bounded sample)...")
        line["cpu_baseline"] = cpu_baseline(svo, tree, cams, args.cpu_seconds)
    cm.Dispose()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
