"""GPU parity of streaming (svo_terrain_create, svo_stream_seed, svo_stream_update, svo_reclaim,
svo_compact) with the oracle: the seed equals TreeBuilder's dense level, every brick equals the
statically built tree's children of its cell, and after every frame the traced outputs, the
requests, the graft records and the whole pool are bit-exact against the oracle's streaming
step run on its own copy of the pool.  Each flight also runs with the frames keeping their
requests for the update (SVO_FRAME_KEEP_REQUESTS) instead of writing ChildRequests."""
from __future__ import annotations

import numpy as np
import pytest

from frame_helpers import first_mismatch
from stream_helpers import U32MAX, pow_sum, seed_pool, static_children, static_tree, terrain_geometry
from tree_helpers import check_pool

pytestmark = pytest.mark.gpu


class GpuStream:
    """A streamed pool on the device: terrain, seed, per-frame trace and update."""

    def __init__(self, svo, cm, geom, B: int, n_groups: int, width: int, height: int):
        self.svo, self.cm, self.B = svo, cm, B
        self.n = width * height
        N = svo.BufferName
        for name in (N.Parents, N.ParentResidency, N.VisitHashes, N.BackgroundQueue, N.MaterialQueue):
            try:
                cm.ReleaseBuffer(name)
            except svo.SvoError:
                pass
        cm.InitDeviceBuffer(N.Bases, pow_sum(B), 2)
        cm.InitDeviceBuffer(N.Blocks, 8 * n_groups, 16)
        cm.InitDeviceBuffer(N.Usages, n_groups, 8)
        cm.InitDeviceBuffer(N.ChildRequests, self.n, 40)
        cm.InitBuffer(N.ChildRequestID, np.zeros(1, np.uint32))
        cm.InitDeviceBuffer(N.OutputImage, self.n, 16)
        cm.InitDeviceBuffer(N.HitIndices, self.n, 4)
        cm.InitDeviceBuffer(N.VisitCounts, self.n, 4)
        cm.InitDeviceBuffer(N.RayStatus, self.n, 1)
        cm.InitDeviceBuffer(N.Locations, self.n, 32)
        cm.InitDeviceBuffer(N.Depths, self.n, 1)
        self.terrain = cm.CreateTerrain(geom)
        cm.StreamSeed(self.terrain, B)
        self.requests_bound = True

    def pool(self):
        N, cm, svo = self.svo.BufferName, self.cm, self.svo
        return (cm.ReadBuffer(N.Bases, np.uint16), cm.ReadBuffer(N.Blocks, svo.BLOCK_DTYPE),
                cm.ReadBuffer(N.Usages, svo.USAGE_DTYPE))

    def trace(self, cam, tick: int, keep: bool = False, capacity: int = 0) -> dict:
        N, cm, svo, n = self.svo.BufferName, self.cm, self.svo, self.n
        cam.tick = tick
        cam.max_child_request_id = capacity or n
        cm.TraceFrame(cam, flags=svo.FRAME_STATE | (svo.FRAME_KEEP_REQUESTS if keep else 0))
        cm.Wait()
        count = int(cm.ReadBuffer(N.ChildRequestID, np.uint32, 1)[0])
        if keep or not self.requests_bound:
            requests = None
        else:
            requests = cm.ReadBuffer(N.ChildRequests, svo.CHILD_REQUEST_DTYPE, min(count, capacity or n))
        return {
            "rgba": cm.ReadBuffer(N.OutputImage, np.float32, 4 * n).reshape(n, 4),
            "hit": cm.ReadBuffer(N.HitIndices, np.uint32, n),
            "visits": cm.ReadBuffer(N.VisitCounts, np.uint32, n),
            "status": cm.ReadBuffer(N.RayStatus, np.uint8, n),
            "depth": cm.ReadBuffer(N.Depths, np.uint8, n),
            "location": cm.ReadBuffer(N.Locations, np.uint64, 4 * n).reshape(n, 4)[:, :3],
            "request_count": count,
            "requests": requests,
        }

    def update(self, tick: int, max_depth: int, capacity: int = 0, **kw) -> dict:
        N, cm, svo = self.svo.BufferName, self.cm, self.svo
        st = cm.StreamUpdate(self.terrain, self.B, max_depth, tick, max_child_request_id=capacity or self.n, **kw)
        nb = st["bricks"]
        st["records"] = cm.ReadBuffer(N.Grafting, svo.GRAFTING_DTYPE, nb) if nb else np.zeros(0, svo.GRAFTING_DTYPE)
        st["bricks_array"] = (cm.ReadBuffer(N.GraftingBlocks, svo.BLOCK_DTYPE, 8 * nb) if nb
                              else np.zeros(0, svo.BLOCK_DTYPE))
        return st

    def close(self):
        self.terrain.close()


def _eq(a, b, what):
    m = first_mismatch(a, b)
    assert m is None, f"{what}: {m}"


@pytest.mark.parametrize("B,resolution", [(2, 64), (3, 256)])
def test_seed_matches_the_builder(svo, gpu, B, resolution):
    geom = terrain_geometry(svo, resolution)
    n_groups = 8 ** (B + 1) + 100
    gs = GpuStream(svo, gpu, geom, B, n_groups, 8, 8)
    try:
        bases, blocks, usages = gs.pool()
        rb, rk, ru = seed_pool(svo, geom, B, n_groups)
        _eq(bases, rb, "bases")
        _eq(blocks, rk, "blocks")
        _eq(usages, ru, "usages")
        h = gs.terrain.heights()
        assert h.shape == (resolution, resolution) and np.all((h >= 0.36) & (h <= 0.44)) and h.std() > 1e-3
    finally:
        gs.close()


def _stream_and_compare(svo, orc, gpu, geom, B, D, n_groups, cams, hold, frames, keep=False, capacity=0,
                        unbind_requests=False, **kw):
    """keep: the frames keep their requests for the update; capacity: MaxChildRequestId of the
    frames and the update (0: one per cone); unbind_requests: ChildRequests is released (kept
    requests only), so that MaxChildRequestId alone limits them."""
    W, H = cams[0].screen_size[0], cams[0].screen_size[1]
    cap = capacity or W * H
    st = static_tree(svo, geom, B, D)
    gs = GpuStream(svo, gpu, geom, B, n_groups, W, H)
    N = svo.BufferName
    if unbind_requests:
        gpu.ReleaseBuffer(N.ChildRequests)
        gs.requests_bound = False
    elif keep:
        gpu.InitBuffer(N.ChildRequests, np.full(W * H * 40, 0xA5, np.uint8))
    try:
        bases, blocks, usages = gs.pool()
        pool = orc.TreePool(bases, blocks, usages)
        seq = list(cams) + [hold] * frames
        history = []
        for f, cam in enumerate(seq):
            tick = f + 1
            got = gs.trace(cam, tick, keep=keep, capacity=capacity)
            cam.tick = tick
            ref = orc.trace_frame(cam, pool.bases, pool.blocks, pool.usages, request_capacity=cap, threads=8)
            for key in ("hit", "visits", "status", "rgba", "depth", "location"):
                _eq(got[key], ref[key], f"frame {f}: {key}")
            assert got["request_count"] == ref["request_count"], f"frame {f}: request count"
            if got["requests"] is not None:
                _eq(got["requests"], ref["requests"], f"frame {f}: requests")
            upd = gs.update(tick, D, capacity=capacity, **kw)
            exp = orc.stream_update(pool, ref["requests"], ref["request_count"], cap, B, D, tick,
                                    lambda sel: static_children(st.blocks, B, sel), **kw)
            for key in ("requests", "unique", "bricks", "collected", "gc_runs", "reclaimed", "moved", "free_groups"):
                assert upd[key] == exp[key], f"frame {f}: {key} gpu={upd[key]} oracle={exp[key]}"
            _eq(upd["records"], exp["records"].view(upd["records"].dtype), f"frame {f}: graft records")
            _eq(upd["bricks_array"], exp["bricks_array"], f"frame {f}: bricks")
            _, gblocks, gusages = gs.pool()
            _eq(gusages, pool.usages, f"frame {f}: usages")
            _eq(gblocks, pool.blocks, f"frame {f}: blocks")
            history.append(upd)
            if f >= len(cams) and upd["unique"] == 0:
                break
        check_pool(pool.blocks, pool.usages, B)
        if keep and not unbind_requests:
            # the frames kept their requests: ChildRequests was never written
            assert np.all(gpu.ReadBuffer(N.ChildRequests, np.uint8) == 0xA5)
        return history, got, st
    finally:
        gs.close()


@pytest.mark.parametrize("keep", [False, True], ids=["child-requests", "kept-requests"])
def test_stream_matches_oracle_frame_by_frame(svo, orc, gpu, keep):
    """Fly 10 frames, then hold: bricks, grafts, reclaim and compaction all bit-exact."""
    geom = terrain_geometry(svo, 64)
    B, D = 2, 7
    cams = [svo.scenes.config3_camera(f, 48, 32, frames=10, base_depth=B) for f in range(10)]
    hold = svo.scenes.config3_camera(3, 48, 32, frames=10, base_depth=B)
    history, _, _ = _stream_and_compare(svo, orc, gpu, geom, B, D, 8 ** 3 + 560, cams, hold, 30, keep=keep,
                                        max_bricks=40, prune_age=2, gc_free_percent=50)
    assert sum(h["bricks"] for h in history) > 200
    assert sum(h["reclaimed"] for h in history) > 0 and sum(h["moved"] for h in history) > 0
    assert history[-1]["unique"] == 0


@pytest.mark.parametrize("keep", [False, True], ids=["child-requests", "kept-requests"])
def test_stream_brick_cap_and_forced_gc(svo, orc, gpu, keep):
    """Few bricks per frame, garbage collection forced every frame, a deeper tree."""
    geom = terrain_geometry(svo, 128)
    B, D = 2, 8
    cams = [svo.scenes.config3_camera(f, 40, 24, frames=6, base_depth=B) for f in range(6)]
    hold = svo.scenes.config3_camera(5, 40, 24, frames=6, base_depth=B)
    history, _, _ = _stream_and_compare(svo, orc, gpu, geom, B, D, 8 ** 3 + 2048, cams, hold, 4, keep=keep,
                                        max_bricks=7, prune_age=1, gc_free_percent=0,
                                        flags=svo.STREAM_FORCE_GC)
    assert all(h["gc_runs"] == 1 for h in history) and all(h["bricks"] <= 7 for h in history)


@pytest.mark.parametrize("keep,unbind", [(False, False), (True, False), (True, True)],
                         ids=["child-requests", "kept-requests", "kept-no-buffer"])
def test_stream_request_capacity(svo, orc, gpu, keep, unbind):
    """MaxChildRequestId far below the frame's requests: only the first 45 (in cone order) are
    stored or kept, and the update serves exactly those, with ChildRequests bound or not."""
    geom = terrain_geometry(svo, 64)
    B, D = 2, 7
    cams = [svo.scenes.config3_camera(f, 48, 32, frames=8, base_depth=B) for f in range(8)]
    hold = svo.scenes.config3_camera(2, 48, 32, frames=8, base_depth=B)
    history, _, _ = _stream_and_compare(svo, orc, gpu, geom, B, D, 8 ** 3 + 900, cams, hold, 6, keep=keep,
                                        capacity=45, unbind_requests=unbind, max_bricks=30, prune_age=3,
                                        gc_free_percent=20)
    assert any(h["requests"] > 45 for h in history) and sum(h["bricks"] for h in history) > 50


@pytest.mark.parametrize("dropper", ["direct", "static", "graph"])
def test_kept_requests_rules(svo, orc, gpu, dropper):
    """SVO_FRAME_KEEP_REQUESTS is refused by frame graphs and the allocation-free frame, and a
    later frame that does not keep its requests (a direct frame without the flag, the
    allocation-free frame, a graph replay) drops the kept ones: the update then reads
    ChildRequests."""
    geom = terrain_geometry(svo, 64)
    B, D = 2, 7
    W, H = 48, 32
    st = static_tree(svo, geom, B, D)
    gs = GpuStream(svo, gpu, geom, B, 8 ** 3 + 600, W, H)
    graph = None
    try:
        cam_a = svo.scenes.config3_camera(0, W, H, frames=10, base_depth=B)
        cam_b = svo.scenes.config3_camera(6, W, H, frames=10, base_depth=B)
        cam_a.max_child_request_id = W * H
        with pytest.raises(svo.SvoError, match="direct frames"):
            gpu.CaptureFrame([cam_a], flags=svo.FRAME_KEEP_REQUESTS)
        with pytest.raises(svo.SvoError, match="direct frames"):
            gpu.TraceFrameStatic(cam_a, flags=svo.FRAME_KEEP_REQUESTS)
        bases, blocks, usages = gs.pool()
        pool = orc.TreePool(bases, blocks, usages)
        cam_b.tick = 1
        cam_b.max_child_request_id = W * H
        if dropper == "graph":  # captured before the kept frame: capturing runs nothing
            graph = gpu.CaptureFrame([cam_b], flags=svo.FRAME_STATE)
        gs.trace(cam_a, 1, keep=True)
        # same tick, another view, requests to ChildRequests
        if dropper == "direct":
            got_requests = gs.trace(cam_b, 1)["requests"]
        else:
            if dropper == "static":
                gpu.TraceFrameStatic(cam_b, flags=svo.FRAME_STATE)
            else:
                graph.Launch([cam_b])
            gpu.Wait()
            N = svo.BufferName
            count = int(gpu.ReadBuffer(N.ChildRequestID, np.uint32, 1)[0])
            got_requests = gpu.ReadBuffer(N.ChildRequests, svo.CHILD_REQUEST_DTYPE, min(count, W * H))
        ref = orc.trace_frame(cam_b, pool.bases, pool.blocks, pool.usages, request_capacity=W * H, threads=8)
        _eq(got_requests, ref["requests"], "requests")
        upd = gs.update(1, D, max_bricks=50, prune_age=2, gc_free_percent=0)
        exp = orc.stream_update(pool, ref["requests"], ref["request_count"], W * H, B, D, 1,
                                lambda sel: static_children(st.blocks, B, sel), max_bricks=50, prune_age=2,
                                gc_free_percent=0)
        assert upd["bricks"] == exp["bricks"] > 0
        _eq(upd["records"], exp["records"].view(upd["records"].dtype), "graft records")
        _, gblocks, gusages = gs.pool()
        _eq(gblocks, pool.blocks, "blocks")
        _eq(gusages, pool.usages, "usages")
    finally:
        if graph is not None:
            graph.Dispose()
        gs.close()


def test_streamed_frame_converges_to_the_static_tree_gpu(svo, gpu):
    """Larger: a 640x360 camera over a 1024^2 terrain streamed to depth 10 until nothing more
    is served; the frame then equals the statically built tree's frame (addresses aside)."""
    geom = terrain_geometry(svo, 1024)
    B, D = 3, 10
    cam = svo.scenes.config3_camera(100, 640, 360, base_depth=B)
    n = 640 * 360
    gs = GpuStream(svo, gpu, geom, B, 1 << 18, 640, 360)
    try:
        history = []
        for f in range(2 * D):
            got = gs.trace(cam, f + 1)
            upd = gs.update(f + 1, D, max_bricks=1 << 20)
            history.append((upd["unique"], upd["bricks"]))
            if upd["unique"] == 0:
                break
        assert upd["unique"] == 0 and f > 2, history
        st = static_tree(svo, geom, B, D)
    finally:
        gs.close()
    N = svo.BufferName
    gpu.InitBuffer(N.Bases, st.bases)
    gpu.InitBuffer(N.Blocks, st.blocks)
    gpu.ReleaseBuffer(N.Usages)
    gpu.InitDeviceBuffer(N.ChildRequests, n, 40)
    cam.max_child_request_id = n
    gpu.TraceFrame(cam, flags=svo.FRAME_STATE)
    gpu.Wait()
    ref = {
        "rgba": gpu.ReadBuffer(N.OutputImage, np.float32, 4 * n).reshape(n, 4),
        "visits": gpu.ReadBuffer(N.VisitCounts, np.uint32, n),
        "status": gpu.ReadBuffer(N.RayStatus, np.uint8, n),
        "depth": gpu.ReadBuffer(N.Depths, np.uint8, n),
    }
    for key in ("rgba", "visits", "depth"):
        _eq(got[key], ref[key], key)
    _eq(got["status"] & 0xFD, ref["status"] & 0xFD, "status")
    assert np.count_nonzero(got["hit"] != U32MAX) > n // 4


def test_kept_requests_after_reseed(svo, orc, gpu):
    """The pool is seeded again, larger, between a kept frame and its update (the first-request
    table grows and starts over): the update serves the kept requests as the oracle's update of
    the new pool serves that frame's requests."""
    geom = terrain_geometry(svo, 64)
    B, D = 2, 7
    W, H = 48, 32
    st = static_tree(svo, geom, B, D)
    gs = GpuStream(svo, gpu, geom, B, 8 ** 3 + 300, W, H)
    try:
        cam = svo.scenes.config3_camera(4, W, H, frames=10, base_depth=B)
        bases, blocks, usages = gs.pool()
        got = gs.trace(cam, 1, keep=True)
        cam.tick = 1
        ref = orc.trace_frame(cam, bases, blocks, usages.copy(), request_capacity=W * H, threads=8)
        assert got["request_count"] == ref["request_count"] > 0
        N = svo.BufferName
        n_groups = 8 ** 3 + 5000
        gpu.InitDeviceBuffer(N.Blocks, 8 * n_groups, 16)
        gpu.InitDeviceBuffer(N.Usages, n_groups, 8)
        gpu.StreamSeed(gs.terrain, B)
        bases, blocks, usages = gs.pool()
        pool = orc.TreePool(bases, blocks, usages)
        kw = dict(max_bricks=60, prune_age=2, gc_free_percent=0)
        upd = gs.update(1, D, **kw)
        exp = orc.stream_update(pool, ref["requests"], ref["request_count"], W * H, B, D, 1,
                                lambda sel: static_children(st.blocks, B, sel), **kw)
        assert upd["unique"] == exp["unique"] and upd["bricks"] == exp["bricks"] > 0
        _eq(upd["records"], exp["records"].view(upd["records"].dtype), "graft records")
        _, gblocks, gusages = gs.pool()
        _eq(gblocks, pool.blocks, "blocks")
        _eq(gusages, pool.usages, "usages")
    finally:
        gs.close()


def test_second_update_of_a_kept_frame(svo, gpu):
    """Two updates after one frame give the same pools whether the frame kept its requests or
    wrote ChildRequests (the second update reads the same requests again)."""
    geom = terrain_geometry(svo, 64)
    B, D = 2, 7
    W, H = 48, 32
    cam = svo.scenes.config3_camera(5, W, H, frames=10, base_depth=B)
    results = []
    for keep in (False, True):
        gs = GpuStream(svo, gpu, geom, B, 8 ** 3 + 800, W, H)
        try:
            gs.trace(cam, 1, keep=keep)
            ups = [gs.update(1, D, max_bricks=25, prune_age=2, gc_free_percent=0) for _ in range(2)]
            results.append((ups, gs.pool()))
        finally:
            gs.close()
    (u0, p0), (u1, p1) = results
    for a, b in zip(u0, u1):
        for key in ("requests", "unique", "bricks", "collected", "free_groups"):
            assert a[key] == b[key], key
        _eq(a["records"], b["records"], "graft records")
    assert u0[0]["bricks"] > 0
    for a, b, what in zip(p0, p1, ("bases", "blocks", "usages")):
        _eq(a, b, what)


def test_kept_requests_of_a_batch(svo, gpu):
    """Two cameras traced in one batch (cone ids run over both): the update serves the batch's
    kept requests exactly as it serves them through ChildRequests."""
    geom = terrain_geometry(svo, 64)
    B, D = 2, 7
    W, H = 40, 24
    results = []
    for keep in (False, True):
        gs = GpuStream(svo, gpu, geom, B, 8 ** 3 + 800, 2 * W, H)
        try:
            cams = [svo.scenes.config3_camera(f, W, H, frames=10, base_depth=B) for f in (2, 7)]
            for c in cams:
                c.tick = 1
                c.max_child_request_id = 2 * W * H
            gpu.TraceBatch(cams, flags=svo.FRAME_KEEP_REQUESTS if keep else 0)
            count = int(gpu.ReadBuffer(svo.BufferName.ChildRequestID, np.uint32, 1)[0])
            ups = [gs.update(1, D, max_bricks=40, prune_age=2, gc_free_percent=0)]
            results.append((count, ups, gs.pool()))
        finally:
            gs.close()
    (c0, u0, p0), (c1, u1, p1) = results
    assert c0 == c1 > 0
    for a, b in zip(u0, u1):
        for key in ("requests", "unique", "bricks", "collected", "free_groups"):
            assert a[key] == b[key], key
        _eq(a["records"], b["records"], "graft records")
    assert u0[0]["bricks"] > 0
    for a, b, what in zip(p0, p1, ("bases", "blocks", "usages")):
        _eq(a, b, what)
