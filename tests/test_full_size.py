"""GPU parity at the measured workloads' full sizes: the 1920x1080 plain frame on the
depth-12 fBm heightfield (54.4M blocks) and the 64-camera batch on the same scene, bit-exact
against the oracle on every per-cone output, the child requests, the queues, the usage
ticks and the step totals.  The oracle runs on 16 threads (a few seconds per frame)."""
from __future__ import annotations

import numpy as np
import pytest

from frame_helpers import assert_frame_matches, gpu_frame
from test_batch_parity import _compare, gpu_batch

pytestmark = pytest.mark.gpu

THREADS = 16


@pytest.fixture(scope="module")
def config2_tree(svo):
    spec = svo.scenes.config2_scene(max_depth=12)
    tree = spec.build()
    return tree, tree.bases, tree.blocks


@pytest.mark.parametrize("frame", [0, 63])
def test_plain_frame_1080p_depth12(svo, orc, gpu, config2_tree, frame):
    tree, bases, blocks = config2_tree
    cam = svo.scenes.config2_camera(frame, 1920, 1080, frames=100, tick=frame + 1)
    usages = tree.initial_usages(0)
    ref_usages = usages.copy()
    cap = 1920 * 1080
    ref = orc.trace_frame(cam, bases, blocks, ref_usages, request_capacity=cap, threads=THREADS)
    out = gpu_frame(svo, gpu, bases, blocks, usages, cam, request_capacity=cap)
    assert_frame_matches(out, ref, usages_ref=ref_usages)
    assert ref["request_count"] > 100000 and (ref["status"] & 1).sum() > cap // 4


def test_batch_64_cameras_full_size(svo, orc, gpu, config2_tree):
    tree, bases, blocks = config2_tree
    cams, apertures = svo.scenes.config4_cameras(64, tick=5)
    usages = tree.initial_usages(0)
    ref_usages = usages.copy()
    ref = orc.trace_batch(cams, bases, blocks, ref_usages, apertures=apertures, request_capacity=6000,
                          threads=THREADS)
    got = gpu_batch(svo, gpu, bases, blocks, usages, cams, apertures)
    _compare(got, ref, ref_usages)
    assert got["hit"].size > 30_000_000


def test_streamed_pool_4k_frame(svo, orc, gpu):
    """The streaming workload at full size: a 3840x2160 camera over the 16384^2 terrain, the
    pool of 2^26 blocks seeded and streamed for three frames on the device (up to 65,536
    bricks each), then one more frame traced on the device and by the oracle on a copy of the
    pool: bit-exact hits, visit counts, status, RGBA, child requests and usage ticks."""
    N = svo.BufferName
    W, H, Bd, D, n_groups = 3840, 2160, 4, 14, 1 << 23
    n = W * H
    for name in (N.VisitHashes, N.RayStatus, N.Locations, N.Depths, N.BackgroundQueue, N.MaterialQueue,
                 N.Parents, N.ParentResidency):
        try:
            gpu.ReleaseBuffer(name)
        except svo.SvoError:
            pass
    gpu.InitDeviceBuffer(N.Bases, sum(8 ** d for d in range(1, Bd + 1)), 2)
    gpu.InitDeviceBuffer(N.Blocks, 8 * n_groups, 16)
    gpu.InitDeviceBuffer(N.Usages, n_groups, 8)
    gpu.InitDeviceBuffer(N.ChildRequests, n, 40)
    gpu.InitBuffer(N.ChildRequestID, np.zeros(1, np.uint32))
    for name, size in ((N.OutputImage, 16), (N.HitIndices, 4), (N.VisitCounts, 4), (N.RayStatus, 1)):
        gpu.InitDeviceBuffer(name, n, size)
    terrain = gpu.CreateTerrain(svo.scenes.config3_geometry())
    try:
        gpu.StreamSeed(terrain, Bd)
        for f in range(3):
            cam = svo.scenes.config3_camera(40 + f, W, H, frames=300, base_depth=Bd)
            cam.tick = f + 1
            cam.max_child_request_id = n
            gpu.TraceFrame(cam)
            st = gpu.StreamUpdate(terrain, Bd, D, f + 1, max_bricks=65536, max_child_request_id=n)
        assert st["bricks"] > 0
    finally:
        terrain.close()
    bases = gpu.ReadBuffer(N.Bases, np.uint16)
    blocks = gpu.ReadBuffer(N.Blocks, svo.BLOCK_DTYPE)
    usages = gpu.ReadBuffer(N.Usages, svo.USAGE_DTYPE)
    cam = svo.scenes.config3_camera(43, W, H, frames=300, base_depth=Bd)
    cam.tick = 4
    cam.max_child_request_id = n
    ref_usages = usages.copy()
    ref = orc.trace_frame(cam, bases, blocks, ref_usages, request_capacity=n, threads=THREADS)
    gpu.InitBuffer(N.ChildRequestID, np.zeros(1, np.uint32))
    gpu.TraceFrame(cam)
    gpu.Wait()
    got = {
        "hit": gpu.ReadBuffer(N.HitIndices, np.uint32, n),
        "visits": gpu.ReadBuffer(N.VisitCounts, np.uint32, n),
        "status": gpu.ReadBuffer(N.RayStatus, np.uint8, n),
        "rgba": gpu.ReadBuffer(N.OutputImage, np.float32, 4 * n).reshape(n, 4),
    }
    for key, value in got.items():
        assert np.array_equal(value, ref[key]), key
    count = int(gpu.ReadBuffer(N.ChildRequestID, np.uint32, 1)[0])
    assert count == ref["request_count"]
    reqs = gpu.ReadBuffer(N.ChildRequests, svo.CHILD_REQUEST_DTYPE, min(count, n))
    assert np.array_equal(reqs.view(np.uint8), ref["requests"].view(np.uint8))
    assert np.array_equal(gpu.ReadBuffer(N.Usages, svo.USAGE_DTYPE), ref_usages)
