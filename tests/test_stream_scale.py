"""Streaming pool parity at the streaming workload's scale: the depth-14 fBm terrain on its
16384^2 grid, base depth 4, up to 65,536 bricks per frame, a 1920x1080 camera on the seeded
fly-through (reduced from 3840x2160 to bound the oracle's time; the pool, the depth and the
brick cap are the workload's, and the cap binds in the later frames).  After every frame the device's whole pool (Blocks and
Usages), the traced outputs, the child requests, the graft records and the bricks are
compared bit-exactly with the oracle's streaming step run on its own copy of the pool.  The
oracle's bricks come from the host builder's MakeBlock on the same terrain description
(svo_geometry_blocks: TreeBuilder's rules over the host's own grid and min/max pyramid), not
from the device's generator.  Two pools: 2^26 blocks with one forced garbage collection, and
2^22 blocks under real pressure (collections triggered by the 10 % rule)."""
from __future__ import annotations

import numpy as np
import pytest

from frame_helpers import first_mismatch
from stream_helpers import U32MAX, terrain_geometry
from test_stream_parity import GpuStream

pytestmark = pytest.mark.gpu

W, H, B, D = 1920, 1080, 4, 14


def _eq(a, b, what):
    m = first_mismatch(a, b)
    assert m is None, f"{what}: {m}"


def check_links(blocks: np.ndarray, usages: np.ndarray, B: int) -> int:
    """Vectorised pool invariant: every Child link out of a live group names a live group whose
    back-pointer (single-parent groups) is the linking block; dense groups stay inviolate.
    Returns the number of links."""
    ticks = usages["tick"]
    assert np.all(ticks[: 8 ** (B + 1)] == 65535)
    live_block = np.repeat(ticks != 0, 8)[: blocks.size]
    child = blocks["child"]
    src = np.nonzero(live_block & (child != U32MAX))[0]
    dst = (child[src] >> 3).astype(np.int64)
    assert np.all(dst < ticks.size), "a link leaves the pool"
    assert np.all(ticks[dst] != 0), "a linked group is marked free"
    single = ticks[dst] < 65534
    assert np.array_equal(usages["parent"][dst[single]], src[single].astype(np.uint32)), "a back-pointer differs"
    return int(src.size)


def host_bricks(geo, requests: np.ndarray) -> np.ndarray:
    """The 8 children (level d+1) of the cell at depth d holding each request's location."""
    n = requests.size
    d = requests["depth"].astype(np.uint64)
    s = np.uint64(64) - d
    xyz = [(requests["location"][:, a] >> s).astype(np.uint64) for a in range(3)]
    cells = np.zeros((n, 8, 4), np.uint32)
    for c in range(8):
        cells[:, c, 0] = (d + 1).astype(np.uint32)
        for a in range(3):
            cells[:, c, 1 + a] = (2 * xyz[a] + ((c >> a) & 1)).astype(np.uint32)
    return geo.blocks(cells.reshape(-1, 4))


@pytest.fixture(scope="module")
def terrain_host(svo):
    geom = terrain_geometry(svo, 16384)
    geo = svo.GeometryHost(geom)
    yield geom, geo
    geo.close()


@pytest.mark.timeout(1200)
@pytest.mark.parametrize("pool_log2,frames,prune_age,force_gc_frame", [(26, 8, 30, 5), (22, 24, 30, -1)])
def test_stream_pool_at_scale(svo, orc, gpu, terrain_host, pool_log2, frames, prune_age, force_gc_frame):
    geom, geo = terrain_host
    n_groups = 1 << (pool_log2 - 3)
    n = W * H
    gs = GpuStream(svo, gpu, geom, B, n_groups, W, H)
    try:
        bases, blocks, usages = gs.pool()
        pool = orc.TreePool(bases, blocks, usages)
        gc_total = bricks_total = max_bricks = 0
        for f in range(frames):
            tick = f + 1
            cam = svo.scenes.config3_camera(f, W, H, frames=60, base_depth=B)
            print(f"stream scale: pool 2^{pool_log2}, frame {f}", flush=True)
            got = gs.trace(cam, tick)
            cam.tick = tick
            ref = orc.trace_frame(cam, pool.bases, pool.blocks, pool.usages, request_capacity=n, threads=16)
            for key in ("hit", "visits", "status", "rgba", "depth", "location"):
                _eq(got[key], ref[key], f"frame {f}: {key}")
            assert got["request_count"] == ref["request_count"], f"frame {f}: request count"
            _eq(got["requests"], ref["requests"], f"frame {f}: requests")
            flags = svo.STREAM_FORCE_GC if f == force_gc_frame else 0
            upd = gs.update(tick, D, max_bricks=65536, prune_age=prune_age, gc_free_percent=10, flags=flags)
            exp = orc.stream_update(pool, ref["requests"], ref["request_count"], n, B, D, tick,
                                    lambda sel: host_bricks(geo, sel), max_bricks=65536, prune_age=prune_age,
                                    gc_free_percent=10, flags=flags)
            for key in ("requests", "unique", "bricks", "collected", "gc_runs", "reclaimed", "moved", "free_groups"):
                assert upd[key] == exp[key], f"frame {f}: {key} gpu={upd[key]} oracle={exp[key]}"
            _eq(upd["records"], exp["records"].view(upd["records"].dtype), f"frame {f}: graft records")
            _eq(upd["bricks_array"], exp["bricks_array"], f"frame {f}: bricks")
            _, gblocks, gusages = gs.pool()
            _eq(gusages, pool.usages, f"frame {f}: usages")
            _eq(gblocks, pool.blocks, f"frame {f}: blocks")
            gc_total += upd["gc_runs"]
            bricks_total += upd["bricks"]
            max_bricks = max(max_bricks, upd["bricks"])
            print(f"  requests {upd['requests']} unique {upd['unique']} bricks {upd['bricks']} "
                  f"gc {upd['gc_runs']} moved {upd['moved']} free {upd['free_groups']}", flush=True)
        assert check_links(pool.blocks, pool.usages, B) > 0
        if pool_log2 == 26:
            assert max_bricks == 65536, "the brick cap was meant to bind"
        if force_gc_frame >= 0:
            assert gc_total >= 1
        else:
            assert gc_total >= 1, "the small pool was meant to run out and collect"
    finally:
        gs.close()
