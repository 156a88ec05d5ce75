"""Shared helpers of the frame parity tests: run one frame on the GPU through the
C ABI (ComputeManager mirror) and compare it with the oracle."""
from __future__ import annotations

import numpy as np

_TREES = {}


def tree_arrays(svo, spec):
    """Builds (and caches) a scene: (tree, bases, blocks)."""
    if spec.name not in _TREES:
        t = spec.build()
        _TREES[spec.name] = (t, t.bases, t.blocks)
    return _TREES[spec.name]


def gpu_frame(svo, cm, bases, blocks, usages, cam, request_capacity=6000, state=True, queues=True,
              hashes=True, **opts):
    B = svo.BufferName
    n = cam.screen_size[0] * cam.screen_size[1]
    cm.InitReadBuffer(B.Bases, bases)
    cm.InitReadBuffer(B.Blocks, blocks)
    if usages is not None:
        cm.InitBuffer(B.Usages, usages)
    elif _bound(cm, svo, B.Usages):
        cm.ReleaseBuffer(B.Usages)
    cm.InitBuffer(B.ChildRequests, np.zeros(max(request_capacity, 1) * 40, np.uint8))
    cm.InitBuffer(B.ChildRequestID, np.zeros(1, np.uint32))
    cm.InitDeviceBuffer(B.OutputImage, n, 16)
    cm.InitDeviceBuffer(B.HitIndices, n, 4)
    cm.InitDeviceBuffer(B.VisitCounts, n, 4)
    if hashes:
        cm.InitDeviceBuffer(B.VisitHashes, n, 8)
    elif _bound(cm, svo, B.VisitHashes):
        cm.ReleaseBuffer(B.VisitHashes)
    cm.InitDeviceBuffer(B.RayStatus, n, 1)
    flags = 0
    if state:
        cm.InitDeviceBuffer(B.Locations, n, 32)
        cm.InitDeviceBuffer(B.Depths, n, 1)
        flags |= svo.FRAME_STATE
    if queues:
        for name in (B.BackgroundQueue, B.MaterialQueue):
            cm.InitDeviceBuffer(name, n, 4)
        for name in (B.BackgroundQueueID, B.MaterialQueueID):
            cm.InitBuffer(name, np.zeros(1, np.uint32))
        flags |= svo.FRAME_QUEUES
    cam.max_child_request_id = request_capacity
    cm.TraceFrame(cam, flags=flags, **opts)
    cm.Wait()
    out = {
        "rgba": cm.ReadBuffer(B.OutputImage, np.float32, 4 * n).reshape(n, 4),
        "hit": cm.ReadBuffer(B.HitIndices, np.uint32, n),
        "visits": cm.ReadBuffer(B.VisitCounts, np.uint32, n),
        "status": cm.ReadBuffer(B.RayStatus, np.uint8, n),
        "request_count": int(cm.ReadBuffer(B.ChildRequestID, np.uint32, 1)[0]),
    }
    if hashes:
        out["hash"] = cm.ReadBuffer(B.VisitHashes, np.uint64, n)
    cnt = min(out["request_count"], request_capacity)
    out["requests"] = cm.ReadBuffer(B.ChildRequests, svo.CHILD_REQUEST_DTYPE, cnt)
    if usages is not None:
        out["usages"] = cm.ReadBuffer(B.Usages, svo.USAGE_DTYPE, usages.size)
    if state:
        out["location"] = cm.ReadBuffer(B.Locations, np.uint64, 4 * n).reshape(n, 4)[:, :3]
        out["depth"] = cm.ReadBuffer(B.Depths, np.uint8, n)
    if queues:
        nbg = int(cm.ReadBuffer(B.BackgroundQueueID, np.uint32, 1)[0])
        nmat = int(cm.ReadBuffer(B.MaterialQueueID, np.uint32, 1)[0])
        out["background_queue"] = cm.ReadBuffer(B.BackgroundQueue, np.uint32, nbg)
        out["material_queue"] = cm.ReadBuffer(B.MaterialQueue, np.uint32, nmat)
    out["stats"] = cm.FrameStats()
    return out


def _bound(cm, svo, name):
    try:
        cm.BufferSize(name)
        return True
    except svo.SvoError:
        return False


def first_mismatch(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return f"shape {a.shape} vs {b.shape}"
    if a.size == 0:
        return None
    if a.dtype.names is None:
        bad = np.argwhere(a != b)
    else:
        bad = np.argwhere(a.view(np.uint8).reshape(len(a), -1) != b.view(np.uint8).reshape(len(b), -1))
    if bad.size == 0:
        return None
    i = tuple(bad[0])
    return f"{len(bad)} mismatches, first at {i}: gpu={a[i[0]]} oracle={b[i[0]]}"


def assert_frame_matches(gpu_out, ref, usages_ref=None, check_queues=True):
    """Bit-exact comparison of a GPU frame against the oracle's per-cone frame."""
    for key in ("hit", "visits", "status", "rgba", "hash", "depth", "location"):
        if key in gpu_out:
            m = first_mismatch(gpu_out[key], ref[key])
            assert m is None, f"{key}: {m}"
    assert gpu_out["request_count"] == ref["request_count"], (gpu_out["request_count"], ref["request_count"])
    m = first_mismatch(gpu_out["requests"], ref["requests"])
    assert m is None, f"child requests: {m}"
    if usages_ref is not None:
        m = first_mismatch(gpu_out["usages"], usages_ref)
        assert m is None, f"usages: {m}"
    if check_queues and "background_queue" in gpu_out:
        status = ref["status"]
        ids = np.arange(status.size, dtype=np.uint32)
        assert np.array_equal(gpu_out["background_queue"], ids[(status & 1) == 0])
        assert np.array_equal(gpu_out["material_queue"], ids[(status & 1) == 1])
    st = gpu_out["stats"]
    assert st.base_steps == ref["base_steps"] and st.block_steps == ref["block_steps"], (
        st.base_steps, st.block_steps, ref["base_steps"], ref["block_steps"])
