"""GPU parity of svo_trace_frame (the fused primary-frame kernel) against the oracle.

Bit-exact on every per-cone output (hit block, visit count, visit-sequence hash,
RGBA, status, final Depths/Locations), on the usage ticks left behind, on the
child-request queue (ascending cone id) and on the background/material queue
contents.
"""
import math

import numpy as np
import pytest

from frame_helpers import assert_frame_matches, gpu_frame, tree_arrays

pytestmark = pytest.mark.gpu


def run_case(svo, orc, gpu, spec, cam, request_capacity=6000, usages_tick=0, usages=None, **opts):
    tree, bases, blocks = tree_arrays(svo, spec)
    if usages is None:
        usages = tree.initial_usages(usages_tick)
    ref_usages = usages.copy()
    ref = orc.trace_frame(cam, bases, blocks, ref_usages, request_capacity=request_capacity, threads=8)
    out = gpu_frame(svo, gpu, bases, blocks, usages.copy(), cam, request_capacity=request_capacity, **opts)
    assert_frame_matches(out, ref, usages_ref=ref_usages)
    return out, ref


def test_f0_reference_camera(svo, orc, gpu):
    out, ref = run_case(svo, orc, gpu, svo.scenes.f0_scene(8), svo.scenes.f0_camera(64, 48))
    assert (ref["status"] & 1).sum() > 0  # some cones hit the spheres


def test_f0_close_camera_deep_lod(svo, orc, gpu):
    cam = svo.scenes.f0_camera(320, 240)
    eye = (0.05, 0.45, 0.42)
    facing = svo.scenes.look_at(eye, (0.3, 0.3, 0.3))
    for i in range(3):
        cam.origin[i] = eye[i]
    for i in range(9):
        cam.facing[i] = facing[i]
    out, ref = run_case(svo, orc, gpu, svo.scenes.f0_scene(8), cam)
    assert ref["request_count"] > 0  # cones reach childless leaves and request detail
    assert ref["block_steps"] > 0


@pytest.mark.parametrize("size", [(160, 120), (640, 480)])
def test_config1_scene(svo, orc, gpu, size):
    cam = svo.scenes.config1_camera(*size)
    out, ref = run_case(svo, orc, gpu, svo.scenes.config1_scene(), cam)
    assert (ref["status"] & 1).sum() > 0


@pytest.mark.parametrize("frame", [0, 37])
def test_heightfield_orbit(svo, orc, gpu, frame):
    spec = svo.scenes.config2_scene(max_depth=10, resolution=512)
    cam = svo.scenes.config2_camera(frame, 320, 180)
    run_case(svo, orc, gpu, spec, cam)


@pytest.mark.parametrize("wg", [64, 128, 512])
def test_workgroup_sizes(svo, orc, gpu, wg):
    run_case(svo, orc, gpu, svo.scenes.config1_scene(), svo.scenes.config1_camera(96, 72), workgroup_size=wg)


@pytest.mark.parametrize("levels", [1, 2])
def test_short_stack_overflow_rewalk(svo, orc, gpu, levels):
    """Fewer LDS stack levels than tree levels forces the re-walk below the stack."""
    cam = svo.scenes.f0_camera(160, 120)
    eye = (0.1, 0.5, 0.5)
    facing = svo.scenes.look_at(eye, (0.7, 0.7, 0.7))
    for i in range(3):
        cam.origin[i] = eye[i]
    for i in range(9):
        cam.facing[i] = facing[i]
    run_case(svo, orc, gpu, svo.scenes.f0_scene(8), cam, stack_levels=levels)


@pytest.mark.parametrize("refill", [1, 64])
def test_refill_policy(svo, orc, gpu, refill):
    run_case(svo, orc, gpu, svo.scenes.config1_scene(), svo.scenes.config1_camera(128, 96), refill_lanes=refill)


def test_base_depth_5_bases_in_global_memory(svo, orc, gpu):
    spec = svo.scenes.f0_scene(8)
    spec.base_depth = 5
    spec.name = "f0-b5"
    cam = svo.scenes.f0_camera(96, 72)
    cam.base_depth = 5
    eye = (0.1, 0.45, 0.45)
    facing = svo.scenes.look_at(eye, (0.3, 0.3, 0.3))
    for i in range(3):
        cam.origin[i] = eye[i]
    for i in range(9):
        cam.facing[i] = facing[i]
    run_case(svo, orc, gpu, spec, cam)


def test_child_request_capacity_overflow(svo, orc, gpu):
    cam = svo.scenes.f0_camera(160, 120)
    eye = (0.05, 0.45, 0.42)
    facing = svo.scenes.look_at(eye, (0.3, 0.3, 0.3))
    for i in range(3):
        cam.origin[i] = eye[i]
    for i in range(9):
        cam.facing[i] = facing[i]
    out, ref = run_case(svo, orc, gpu, svo.scenes.f0_scene(8), cam, request_capacity=7)
    assert ref["request_count"] > 7 and len(out["requests"]) == 7


@pytest.mark.parametrize("size", [(1, 1), (7, 5), (33, 17)])
def test_ragged_screen_sizes(svo, orc, gpu, size):
    run_case(svo, orc, gpu, svo.scenes.config1_scene(), svo.scenes.config1_camera(*size))


def test_all_background(svo, orc, gpu):
    cam = svo.scenes.f0_camera(64, 48)
    for i in range(9):
        cam.facing[i] = [-1.0, 0, 0, 0, -1.0, 0, 0, 0, 1.0][i]
    out, ref = run_case(svo, orc, gpu, svo.scenes.f0_scene(8), cam)
    assert (ref["status"] & 1).sum() == 0


def test_usage_tick_rules(svo, orc, gpu):
    """Multi-parent (65534), inviolate (65535) and already-current ticks are left alone."""
    spec = svo.scenes.config1_scene()
    tree, bases, blocks = tree_arrays(svo, spec)
    usages = tree.initial_usages(3)
    rng = np.random.default_rng(5)
    idx = rng.choice(usages.size, size=usages.size // 3, replace=False)
    usages["tick"][idx[: len(idx) // 3]] = 65534
    usages["tick"][idx[len(idx) // 3: 2 * len(idx) // 3]] = 7  # equals the frame tick below
    cam = svo.scenes.config1_camera(160, 120, tick=7)
    out, ref = run_case(svo, orc, gpu, spec, cam, usages=usages)
    assert ref["groups_touched"] > 0


def test_repeated_frames_identical(svo, orc, gpu):
    spec = svo.scenes.config2_scene(max_depth=10, resolution=512)
    tree, bases, blocks = tree_arrays(svo, spec)
    cam = svo.scenes.config2_camera(5, 256, 144)
    a = gpu_frame(svo, gpu, bases, blocks, tree.initial_usages(0), cam)
    b = gpu_frame(svo, gpu, bases, blocks, tree.initial_usages(0), cam)
    for k in ("hit", "visits", "hash", "rgba", "status"):
        assert np.array_equal(a[k], b[k]), k


def test_tile_order_from_previous_frame(svo, orc, gpu):
    """By default each band's tiles are taken most expensive first, by the cost the previous
    frame of the same tiling measured.  A frame traced in that order (its output buffers
    filled with a sentinel first, so a tile left out would show) and a frame in raster order
    are both bit-exact against the oracle."""
    B = svo.BufferName
    spec = svo.scenes.config2_scene(max_depth=10, resolution=512)
    cam = svo.scenes.config2_camera(12, 320, 200)
    n = 320 * 200
    _, ref = run_case(svo, orc, gpu, spec, cam, state=False, queues=False)  # measures the tile costs
    for flags in (0, svo.FRAME_RASTER_ORDER, 0):
        gpu.WriteBuffer(B.HitIndices, np.full(n, 0xDEADBEEF, np.uint32))
        gpu.WriteBuffer(B.VisitCounts, np.full(n, 0xDEADBEEF, np.uint32))
        gpu.WriteBuffer(B.ChildRequestID, np.zeros(1, np.uint32))
        gpu.TraceFrame(cam, flags=flags)
        gpu.Wait()
        assert np.array_equal(gpu.ReadBuffer(B.HitIndices, np.uint32, n), ref["hit"]), flags
        assert np.array_equal(gpu.ReadBuffer(B.VisitCounts, np.uint32, n), ref["visits"]), flags
