"""Batched cameras (svo_trace_batch): parity with the oracle's batched frame, and equality
with one frame per camera."""
from __future__ import annotations

import numpy as np
import pytest

from frame_helpers import first_mismatch, tree_arrays

pytestmark = pytest.mark.gpu

CAP = 6000


def _cameras(svo, sizes, tick=3):
    base = svo.scenes.config1_camera(64, 48, tick=tick)
    out = []
    for k, (w, h) in enumerate(sizes):
        c = svo.scenes.config1_camera(w, h, tick=tick)
        # spread the poses a little so the cameras differ
        c.origin[0] = base.origin[0] + 0.02 * k
        c.origin[2] = base.origin[2] - 0.015 * k
        out.append(c)
    return out


def gpu_batch(svo, cm, bases, blocks, usages, cams, apertures, flags=0, **opts):
    B = svo.BufferName
    n = sum(c.screen_size[0] * c.screen_size[1] for c in cams)
    cm.InitReadBuffer(B.Bases, bases)
    cm.InitReadBuffer(B.Blocks, blocks)
    cm.InitBuffer(B.Usages, usages)
    cm.InitBuffer(B.ChildRequests, np.zeros(CAP * 40, np.uint8))
    cm.InitBuffer(B.ChildRequestID, np.zeros(1, np.uint32))
    for name, size in ((B.OutputImage, 16), (B.HitIndices, 4), (B.VisitCounts, 4), (B.VisitHashes, 8),
                       (B.RayStatus, 1), (B.BackgroundQueue, 4), (B.MaterialQueue, 4)):
        cm.InitDeviceBuffer(name, n, size)
    for name in (B.BackgroundQueueID, B.MaterialQueueID):
        cm.InitBuffer(name, np.zeros(1, np.uint32))
    for c in cams:
        c.max_child_request_id = CAP
    if len(cams) == 1 and apertures is None:
        cm.TraceFrame(cams[0], flags=flags | svo.FRAME_QUEUES, **opts)
    else:
        cm.TraceBatch(cams, apertures, flags=flags | svo.FRAME_QUEUES, **opts)
    cm.Wait()
    nbg = int(cm.ReadBuffer(B.BackgroundQueueID, np.uint32, 1)[0])
    nmat = int(cm.ReadBuffer(B.MaterialQueueID, np.uint32, 1)[0])
    cnt = int(cm.ReadBuffer(B.ChildRequestID, np.uint32, 1)[0])
    return {
        "rgba": cm.ReadBuffer(B.OutputImage, np.float32, 4 * n).reshape(n, 4),
        "hit": cm.ReadBuffer(B.HitIndices, np.uint32, n),
        "visits": cm.ReadBuffer(B.VisitCounts, np.uint32, n),
        "hash": cm.ReadBuffer(B.VisitHashes, np.uint64, n),
        "status": cm.ReadBuffer(B.RayStatus, np.uint8, n),
        "usages": cm.ReadBuffer(B.Usages, svo.USAGE_DTYPE, usages.size),
        "request_count": cnt,
        "requests": cm.ReadBuffer(B.ChildRequests, svo.CHILD_REQUEST_DTYPE, min(cnt, CAP)),
        "background_queue": cm.ReadBuffer(B.BackgroundQueue, np.uint32, nbg),
        "material_queue": cm.ReadBuffer(B.MaterialQueue, np.uint32, nmat),
        "stats": cm.FrameStats(),
    }


def _compare(got, ref, ref_usages):
    for key in ("hit", "visits", "status", "rgba", "hash"):
        m = first_mismatch(got[key], ref[key])
        assert m is None, f"{key}: {m}"
    assert got["request_count"] == ref["request_count"]
    m = first_mismatch(got["requests"], ref["requests"])
    assert m is None, f"requests: {m}"
    m = first_mismatch(got["usages"], ref_usages)
    assert m is None, f"usages: {m}"
    ids = np.arange(ref["status"].size, dtype=np.uint32)
    assert np.array_equal(got["background_queue"], ids[(ref["status"] & 1) == 0])
    assert np.array_equal(got["material_queue"], ids[(ref["status"] & 1) == 1])
    st = got["stats"]
    assert (st.base_steps, st.block_steps) == (ref["base_steps"], ref["block_steps"])


@pytest.fixture(scope="module")
def scene(svo):
    tree, bases, blocks = tree_arrays(svo, svo.scenes.config1_scene())
    return tree, bases, blocks, tree.initial_usages(tick=1)


@pytest.mark.parametrize("wg", [64, 256, 512, 1024])
def test_batch_matches_oracle(svo, orc, gpu, scene, wg):
    tree, bases, blocks, usages = scene
    cams = _cameras(svo, [(64, 48), (33, 17), (96, 72), (8, 8), (1, 1), (130, 9)])
    apertures = [1.0, 2.0, 3.0, 4.0, 1.0, 1.5]
    ref_usages = usages.copy()
    ref = orc.trace_batch(cams, bases, blocks, ref_usages, apertures=apertures, request_capacity=CAP, threads=8)
    got = gpu_batch(svo, gpu, bases, blocks, usages.copy(), cams, apertures, workgroup_size=wg)
    assert int(np.count_nonzero(got["hit"] != 0xFFFFFFFF)) > 0
    _compare(got, ref, ref_usages)


def test_batch_equals_one_frame_per_camera(svo, orc, gpu, scene):
    tree, bases, blocks, usages = scene
    cams = _cameras(svo, [(40, 30), (72, 40), (16, 64)])
    batch = gpu_batch(svo, gpu, bases, blocks, usages.copy(), cams, None)
    off = 0
    for c in cams:
        one = gpu_batch(svo, gpu, bases, blocks, usages.copy(), [c], None)
        n = c.screen_size[0] * c.screen_size[1]
        for key in ("hit", "visits", "status", "rgba", "hash"):
            assert first_mismatch(batch[key][off:off + n], one[key]) is None, key
        off += n


def test_camera_table_flag_single_camera(svo, orc, gpu, scene):
    tree, bases, blocks, usages = scene
    cam = svo.scenes.config1_camera(96, 72, tick=4)
    ref_usages = usages.copy()
    ref = orc.trace_frame(cam, bases, blocks, ref_usages, request_capacity=CAP, threads=8)
    got = gpu_batch(svo, gpu, bases, blocks, usages.copy(), [cam], None, flags=svo.FRAME_CAMERA_TABLE)
    _compare(got, ref, ref_usages)


def test_batch_rejects_mixed_ticks(svo, gpu, scene):
    tree, bases, blocks, usages = scene
    cams = _cameras(svo, [(8, 8), (8, 8)])
    cams[1].tick = 9
    with pytest.raises(svo.SvoError):
        gpu_batch(svo, gpu, bases, blocks, usages.copy(), cams, None)


def test_config4_cameras_scaled_down(svo, orc, gpu):
    """The config-4 camera set (poses, aspect ratios, apertures) at 1/8 of each dimension
    over a depth-9 version of the heightfield."""
    tree, bases, blocks = tree_arrays(svo, svo.scenes.config2_scene(max_depth=9, resolution=512))
    cams, apertures = svo.scenes.config4_cameras(count=16, tick=2)
    for c in cams:
        c.screen_size[0] = max(1, c.screen_size[0] // 8)
        c.screen_size[1] = max(1, c.screen_size[1] // 8)
    usages = tree.initial_usages(tick=1)
    ref_usages = usages.copy()
    ref = orc.trace_batch(cams, bases, blocks, ref_usages, apertures=apertures, request_capacity=CAP, threads=8)
    got = gpu_batch(svo, gpu, bases, blocks, usages.copy(), cams, apertures)
    _compare(got, ref, ref_usages)


def _read_outputs(svo, cm, n):
    B = svo.BufferName
    cnt = int(cm.ReadBuffer(B.ChildRequestID, np.uint32, 1)[0])
    return {
        "rgba": cm.ReadBuffer(B.OutputImage, np.float32, 4 * n).reshape(n, 4),
        "hit": cm.ReadBuffer(B.HitIndices, np.uint32, n),
        "visits": cm.ReadBuffer(B.VisitCounts, np.uint32, n),
        "hash": cm.ReadBuffer(B.VisitHashes, np.uint64, n),
        "status": cm.ReadBuffer(B.RayStatus, np.uint8, n),
        "request_count": cnt,
        "requests": cm.ReadBuffer(B.ChildRequests, svo.CHILD_REQUEST_DTYPE, min(cnt, CAP)),
    }


@pytest.mark.parametrize("per_camera", [False, True])
def test_captured_frame_replays_match_direct_frames(svo, gpu, scene, per_camera):
    """Replays of a captured frame with new poses and ticks, with direct frames and
    synchronisations in between, equal direct frames of the same cameras."""
    tree, bases, blocks, usages = scene
    B = svo.BufferName
    sizes = [(48, 40), (24, 16), (64, 24)]
    apertures = [1.0, 2.0, 3.0]
    cams0 = _cameras(svo, sizes, tick=2)
    gpu_batch(svo, gpu, bases, blocks, usages.copy(), cams0, apertures)  # binds every buffer
    n = sum(w * h for w, h in sizes)
    graph = gpu.CaptureFrame(cams0, apertures, per_camera=per_camera, flags=svo.FRAME_QUEUES)
    try:
        for frame in range(3):
            cams = _cameras(svo, sizes, tick=3 + frame)
            for c in cams:
                c.origin[1] += 0.01 * frame
                c.max_child_request_id = CAP
            gpu.WriteBuffer(B.Usages, usages)
            gpu.WriteBuffer(B.ChildRequestID, np.zeros(1, np.uint32))
            graph.Launch(cams)
            gpu.Wait()
            got = _read_outputs(svo, gpu, n)
            got_usages = gpu.ReadBuffer(B.Usages, svo.USAGE_DTYPE, usages.size)
            # the same frame submitted directly
            gpu.WriteBuffer(B.Usages, usages)
            if per_camera:
                for c, a in zip(cams, apertures):
                    gpu.TraceBatch([c], [a], flags=svo.FRAME_QUEUES)
            else:
                gpu.TraceBatch(cams, apertures, flags=svo.FRAME_QUEUES)
            gpu.Wait()
            want = _read_outputs(svo, gpu, n)
            want_usages = gpu.ReadBuffer(B.Usages, svo.USAGE_DTYPE, usages.size)
            for key in ("hit", "visits", "status", "rgba", "hash", "requests"):
                assert first_mismatch(got[key], want[key]) is None, f"frame {frame} {key}"
            assert got["request_count"] == want["request_count"]
            assert first_mismatch(got_usages, want_usages) is None
    finally:
        graph.Dispose()


def test_autotune_keeps_results(svo, orc, gpu, scene):
    """The tuner's frames leave the state one frame leaves, and the tuned options give the
    same outputs as the defaults."""
    tree, bases, blocks, usages = scene
    cams = _cameras(svo, [(64, 48), (40, 24)])
    apertures = [1.0, 2.0]
    want = gpu_batch(svo, gpu, bases, blocks, usages.copy(), cams, apertures)
    best, ms = gpu.Autotune(cams, apertures, frames=2, flags=svo.FRAME_QUEUES)
    assert ms > 0 and best["workgroup_size"] in (0, 64, 128, 256, 512, 1024)
    got_usages = gpu.ReadBuffer(svo.BufferName.Usages, svo.USAGE_DTYPE, usages.size)
    assert first_mismatch(got_usages, want["usages"]) is None
    opts = {k: v for k, v in best.items() if v and k != "flags"}
    got = gpu_batch(svo, gpu, bases, blocks, usages.copy(), cams, apertures, **opts)
    for key in ("hit", "visits", "status", "rgba", "hash", "requests", "background_queue", "material_queue"):
        assert first_mismatch(got[key], want[key]) is None, key
