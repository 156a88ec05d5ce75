"""The streaming loop on the oracle alone (CPU): frames traced on a pool that starts as the
seed tree and grows by the bricks its own child requests ask for.

What fixes the answer: once no request can be served any more (every block a cone needs
detail below has children), each cone walks the same cells as in the statically built tree of
the same terrain, so the image, visit counts, stop depths and locations equal the static
tree's frame bit for bit.  Hit indices (pool addresses) may differ, and so may status bit 1:
a cone that stops on a childless block still records a request for it (TreeSize 0), where
the static tree's block has children.
"""
from __future__ import annotations

import numpy as np

from stream_helpers import seed_pool, static_children, static_tree, terrain_geometry


def _converge(svo, orc, geom, B, D, cam, n_groups, frames, path=(), **kw):
    """Streams the frames of `path` (cameras), then `cam` until nothing more is served."""
    st = static_tree(svo, geom, B, D)
    bases, blocks, usages = seed_pool(svo, geom, B, n_groups)
    pool = orc.TreePool(bases, blocks, usages)
    cap = cam.screen_size[0] * cam.screen_size[1]
    history = []
    for f, c in enumerate(path):
        c.tick = f + 1
        ref = orc.trace_frame(c, pool.bases, pool.blocks, pool.usages, request_capacity=cap, threads=8)
        history.append(orc.stream_update(pool, ref["requests"], ref["request_count"], cap, B, D, f + 1,
                                         lambda sel: static_children(st.blocks, B, sel), **kw))
    for f in range(len(path), len(path) + frames):
        cam.tick = f + 1
        ref = orc.trace_frame(cam, pool.bases, pool.blocks, pool.usages, request_capacity=cap, threads=8)
        upd = orc.stream_update(pool, ref["requests"], ref["request_count"], cap, B, D, f + 1,
                                lambda sel: static_children(st.blocks, B, sel), **kw)
        history.append(upd)
        if upd["unique"] == 0:
            return st, pool, ref, history
    raise AssertionError(f"no convergence in {frames} frames: {[h['unique'] for h in history]}")


def test_streamed_frame_converges_to_the_static_tree(svo, orc):
    geom = terrain_geometry(svo, 64)
    B, D = 2, 8
    cam = svo.scenes.config3_camera(10, 64, 48, frames=40, base_depth=B)
    st, pool, ref, history = _converge(svo, orc, geom, B, D, cam, n_groups=1 << 14, frames=12)
    assert history[0]["unique"] > 0 and sum(h["bricks"] for h in history) > 50
    cap = 64 * 48
    static = orc.trace_frame(cam, st.bases, st.blocks, None, request_capacity=cap, threads=8)
    for key in ("rgba", "visits", "depth", "location"):
        assert np.array_equal(ref[key], static[key]), key
    assert np.array_equal(ref["status"] & 0xFD, static["status"] & 0xFD)
    # what remains requested asks for no level, lies at the deepest level or cannot have children
    chunk = pool.blocks["chunk"][ref["requests"]["address"]].astype(np.int64)
    mixed = np.zeros(chunk.size, bool)
    for k in range(8):
        mixed |= ((chunk >> (2 * k)) & 3) == 3
    assert np.all((ref["requests"]["tree_size"] == 0) | (ref["requests"]["depth"] >= D) | ~mixed)


def test_streaming_under_pressure_still_converges(svo, orc):
    """A small pool, few bricks per frame and a short prune age while the camera flies 16
    frames along the path: garbage collection reclaims and compacts, and once the camera holds
    still the frame ends on the static tree's image."""
    geom = terrain_geometry(svo, 64)
    B, D = 2, 7
    path = [svo.scenes.config3_camera(f, 48, 32, frames=16, base_depth=B) for f in range(16)]
    cam = svo.scenes.config3_camera(3, 48, 32, frames=16, base_depth=B)
    st, pool, ref, history = _converge(svo, orc, geom, B, D, cam, n_groups=8 ** 3 + 560, frames=40, path=path,
                                       max_bricks=40, prune_age=2, gc_free_percent=50)
    assert sum(h["reclaimed"] for h in history) > 0 and sum(h["moved"] for h in history) > 0
    static = orc.trace_frame(cam, st.bases, st.blocks, None, request_capacity=48 * 32, threads=8)
    for key in ("rgba", "visits", "depth"):
        assert np.array_equal(ref[key], static[key]), key
    assert np.array_equal(ref["status"] & 0xFD, static["status"] & 0xFD)
