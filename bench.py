"""Benchmark: plain-forward cone tracing at 1920x1080 on the depth-12 fBm heightfield.

One step = one primary frame (svo_trace_frame: ray generation, octree traversal
with the cone-footprint LOD stop, shading, usage ticks, child-request queue) of
the seeded 100-frame orbit.  Inputs (tree, usages, camera table) are resident
in HBM before the timed region; the timed region is bracketed by stream
synchronisation and measured with HIP events on the context's stream.

Prints one JSON line; see DESIGN.md for the roofline model.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "svo-tracer_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "Mcones/s @1080p depth-12 (plain vs batched); fwd+adjoint Mcones/s; HBM % peak"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(stats, cones: int, groups_touched: int) -> int:
    """Cache-less byte model of one frame: 2 B per Bases chunk examined, 16 B per block
    record examined, 24 B of per-cone output (hit u32, visit count u32, RGBA f32x4) and
    8 B per usage group whose tick changed."""
    return 2 * stats.base_steps + 16 * stats.block_steps + 24 * cones + 8 * groups_touched


def read_traffic(profile: Path):
    """HBM bytes per trace-kernel dispatch from a committed rocprofv3 PMC summary, if present."""
    if not profile.exists():
        return None
    for line in profile.read_text().splitlines():
        if line.startswith("traffic_bytes_per_dispatch:"):
            return float(line.split(":", 1)[1])
    return None


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
    return {
        "value": round(frames * n / spent / 1e6, 4),
        "unit": "Mcones/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{frames} whole frames of the same orbit ({frames * n} cones, {spent:.1f} s host time)",
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
    ap.add_argument("--waves", type=int, default=0, help="register budget of the trace kernel (0 = default, or 6)")
    ap.add_argument("--chunk", type=int, default=0, help="cones a wave claims per atomic")
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
    log(f"scene {spec.name}: {tree.block_count} blocks ({tree.block_count * 16 / 2**20:.0f} MiB), "
        f"built in {time.perf_counter() - t0:.1f} s")

    B = svo.BufferName
    cm = svo.ComputeManager(0)
    cm.LoadTree(tree, with_usages=True, usage_tick=0)
    usages0 = cm.ReadBuffer(B.Usages, svo.USAGE_DTYPE)
    cm.InitBuffer(B.ChildRequests, np.zeros(6000 * 40, np.uint8))
    cm.InitBuffer(B.ChildRequestID, np.zeros(1, np.uint32))
    cm.InitDeviceBuffer(B.OutputImage, N, 16)
    cm.InitDeviceBuffer(B.HitIndices, N, 4)
    cm.InitDeviceBuffer(B.VisitCounts, N, 4)
    opts = dict(workgroup_size=args.workgroup, refill_lanes=args.refill, waves_per_simd=args.waves,
                work_chunk=args.chunk)
    cams = [svo.scenes.config2_camera(f, W, H, frames=100, tick=f + 1) for f in range(K)]

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
    traffic = read_traffic(ROOT / "profiles" / "trace_frame_1080p_d12_traffic.txt")
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
                        f"({tree.block_count} x 16-B blocks), {K}-frame seeded orbit",
            "cones_per_step": N,
            "tree_blocks": tree.block_count,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on a bounded sample)...")
        line["cpu_baseline"] = cpu_baseline(svo, tree, cams, args.cpu_seconds)
    cm.Dispose()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
