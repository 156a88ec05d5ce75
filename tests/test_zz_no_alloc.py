"""svo_trace_frame_static (the frame without allocation) against svo_trace_frame, called
through the C ABI on the same bound buffers: every output, the child requests, the queues,
the usage ticks and the step totals equal, on two calls in a row -- a 1920x1080 frame with
child requests and queues, then a smaller one -- and a frame whose scratch exceeds the static
array fails with SVO_ERR_CAPACITY before anything runs."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from frame_helpers import tree_arrays

pytestmark = pytest.mark.gpu


def _bind(svo, cm, bases, blocks, usages, n, capacity):
    B = svo.BufferName
    cm.InitReadBuffer(B.Bases, bases)
    cm.InitReadBuffer(B.Blocks, blocks)
    cm.InitBuffer(B.Usages, usages)
    cm.InitBuffer(B.ChildRequests, np.zeros(capacity * 40, np.uint8))
    cm.InitBuffer(B.ChildRequestID, np.zeros(1, np.uint32))
    for name, size in [(B.OutputImage, 16), (B.HitIndices, 4), (B.VisitCounts, 4), (B.VisitHashes, 8),
                       (B.RayStatus, 1), (B.Locations, 32), (B.Depths, 1), (B.BackgroundQueue, 4),
                       (B.MaterialQueue, 4)]:
        cm.InitDeviceBuffer(name, n, size)
    for name in (B.BackgroundQueueID, B.MaterialQueueID):
        cm.InitBuffer(name, np.zeros(1, np.uint32))


def _outputs(svo, cm, n, usages_size):
    B = svo.BufferName
    out = {
        "rgba": cm.ReadBuffer(B.OutputImage, np.uint32, 4 * n),
        "hit": cm.ReadBuffer(B.HitIndices, np.uint32, n),
        "visits": cm.ReadBuffer(B.VisitCounts, np.uint32, n),
        "hash": cm.ReadBuffer(B.VisitHashes, np.uint64, n),
        "status": cm.ReadBuffer(B.RayStatus, np.uint8, n),
        "location": cm.ReadBuffer(B.Locations, np.uint64, 4 * n).reshape(n, 4)[:, :3].copy(),
        "depth": cm.ReadBuffer(B.Depths, np.uint8, n),
        "usages": cm.ReadBuffer(B.Usages, svo.USAGE_DTYPE, usages_size).view(np.uint8),
    }
    cnt = int(cm.ReadBuffer(B.ChildRequestID, np.uint32, 1)[0])
    out["request_count"] = np.array([cnt])
    out["requests"] = cm.ReadBuffer(B.ChildRequests, np.uint8, 40 * min(cnt, 6000))
    for q, qid in ((B.BackgroundQueue, B.BackgroundQueueID), (B.MaterialQueue, B.MaterialQueueID)):
        k = int(cm.ReadBuffer(qid, np.uint32, 1)[0])
        out[q.name] = cm.ReadBuffer(q, np.uint32, k)
    st = cm.FrameStats()
    out["stats"] = np.array([st.base_steps, st.block_steps], np.uint64)
    return out


def _frame(svo, cm, lib_call, cam, bases, blocks, usages0, n):
    _bind(svo, cm, bases, blocks, usages0.copy(), n, 6000)
    o = svo.FrameOptions()
    o.flags = svo.FRAME_STATE | svo.FRAME_QUEUES | svo.FRAME_RASTER_ORDER
    rc = lib_call(cm._h, C.byref(cam), C.byref(o))
    assert rc == 0, svo.last_error()
    cm.Wait()
    return _outputs(svo, cm, n, usages0.size)


def test_static_frame_equals_allocating_frame(svo, gpu):
    lib = svo.lib()
    tree, bases, blocks = tree_arrays(svo, svo.scenes.config2_scene(10, 512))
    usages0 = tree.initial_usages(0)
    for (w, h), frame in (((1920, 1080), 3), ((640, 360), 40)):
        cam = svo.scenes.config2_camera(frame, w, h, frames=100, tick=frame + 1)
        cam.max_child_request_id = 6000
        n = w * h
        print(f"no-alloc frame {w}x{h} (orbit frame {frame})", flush=True)
        want = _frame(svo, gpu, lib.svo_trace_frame, cam, bases, blocks, usages0, n)
        got = _frame(svo, gpu, lib.svo_trace_frame_static, cam, bases, blocks, usages0, n)
        assert want["request_count"][0] > 1000 and want["MaterialQueue"].size > n // 10
        for k in want:
            assert np.array_equal(got[k], want[k]), f"{w}x{h}: {k} differs"


def test_static_frame_rejects_oversized_scratch(svo, gpu):
    lib = svo.lib()
    tree, bases, blocks = tree_arrays(svo, svo.scenes.config2_scene(10, 512))
    # 3840x2160 with child requests: 8.3M cones x 41 bytes of staging, beyond the static array
    cam = svo.scenes.config2_camera(0, 3840, 2160, frames=100, tick=1)
    gpu.InitReadBuffer(svo.BufferName.Bases, bases)
    gpu.InitReadBuffer(svo.BufferName.Blocks, blocks)
    gpu.InitBuffer(svo.BufferName.ChildRequests, np.zeros(6000 * 40, np.uint8))
    gpu.InitBuffer(svo.BufferName.ChildRequestID, np.zeros(1, np.uint32))
    gpu.InitDeviceBuffer(svo.BufferName.OutputImage, 3840 * 2160, 16)
    for name in (svo.BufferName.HitIndices, svo.BufferName.VisitCounts, svo.BufferName.VisitHashes,
                 svo.BufferName.RayStatus, svo.BufferName.Locations, svo.BufferName.Depths):
        try:
            gpu.ReleaseBuffer(name)
        except svo.SvoError:
            pass
    o = svo.FrameOptions()
    rc = lib.svo_trace_frame_static(gpu._h, C.byref(cam), C.byref(o))
    assert rc == -10 and "static maximum" in svo.last_error()
