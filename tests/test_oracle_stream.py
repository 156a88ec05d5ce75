"""Pinning tests of the streaming oracle (request consumption, reclaim, compaction), CPU only.

Garbage collection has no reference counterpart (the reference's only eviction is the stale
scan of graft, wavefront.cl:1277-1305), so reclaim and compaction are pinned by hand-computed
known answers and by what a relabelling must preserve: the tree read from the dense level is
the same tree, block for block, before and after compaction.
"""
from __future__ import annotations

import numpy as np
import pytest

from frame_helpers import tree_arrays
from tree_helpers import TICK_INVIOLATE, TICK_MULTI, U32MAX, check_pool, dense_groups, pool_from_tree, tiny_pool


def _ui(orc, B, tick, offset=0):
    return orc.update_input(B, tick, offset=offset)


# ---------------------------------------------------------------------------
# reclaim
# ---------------------------------------------------------------------------

def test_reclaim_known_answer(svo, orc):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=6)
    t, par, ch = usages["tick"], usages["parent"], blocks["child"]
    # 64: stale (age 50) under dense block 200, holding 65 (fresh) below its block 3
    t[64], par[64], ch[200] = 50, 200, 8 * 64
    t[65], par[65], ch[8 * 64 + 3] = 99, 8 * 64 + 3, 8 * 65
    # 66: stale (age 40) under dense block 201; its block 1 shares multi-parent group 68
    t[66], par[66], ch[201] = 60, 201, 8 * 66
    # 67: fresh (age 10) under dense block 202
    t[67], par[67], ch[202] = 90, 202, 8 * 67
    # 68: shared by block 8*66+1 and dense block 205 (list head 0)
    t[68], par[68] = TICK_MULTI, 0
    ch[8 * 66 + 1] = 8 * 68
    ch[205] = 8 * 68
    parents = np.zeros(2, dtype=svo.PARENT_DTYPE)
    parents[0], parents[1] = (8 * 66 + 1, 1), (205, U32MAX)
    residency = np.array([1, 1], np.uint8)
    pool = orc.TreePool(bases, blocks, usages, parents=parents, residency=residency)
    n = pool.reclaim(_ui(orc, B, 100, offset=30))
    assert n == 2  # the stale groups themselves; their subtrees are freed besides
    assert pool.usages["tick"][64:70].tolist() == [0, 0, 0, 90, 100, 0]
    assert pool.usages["parent"][64] == U32MAX and pool.usages["parent"][66] == U32MAX
    assert pool.blocks["child"][200] == U32MAX and pool.blocks["child"][201] == U32MAX
    assert pool.blocks["child"][202] == 8 * 67
    # the shared group lost parent 8*66+1 and is single-parent under 205 again
    assert int(pool.usages["parent"][68]) == 205
    assert pool.residency.tolist() == [0, 0]
    assert np.all(pool.usages["tick"][: dense_groups(B)] == TICK_INVIOLATE)
    check_pool(pool.blocks, pool.usages, B)


def test_reclaim_without_stale_groups_changes_nothing(svo, orc):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=4)
    usages["tick"][64], usages["parent"][64], blocks["child"][100] = 98, 100, 8 * 64
    pool = orc.TreePool(bases, blocks, usages)
    assert pool.reclaim(_ui(orc, B, 100, offset=30)) == 0
    assert np.array_equal(pool.usages, usages) and np.array_equal(pool.blocks, blocks)


# ---------------------------------------------------------------------------
# compaction
# ---------------------------------------------------------------------------

def test_compact_known_answer(svo, orc):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=8)  # groups 64..71
    t, par, ch = usages["tick"], usages["parent"], blocks["child"]
    # live: 64 (stays), 69 under dense block 10 holding 71 below its block 2, 70 shared
    t[64], par[64], ch[7] = 5, 7, 8 * 64
    t[69], par[69], ch[10] = 6, 10, 8 * 69
    t[71], par[71], ch[8 * 69 + 2] = 7, 8 * 69 + 2, 8 * 71
    t[70], par[70] = TICK_MULTI, 0
    ch[8 * 71 + 5] = 8 * 70
    ch[300] = 8 * 70
    blocks["colour_r"][8 * 69: 8 * 70] = np.arange(8) + 10
    blocks["colour_r"][8 * 71: 8 * 72] = np.arange(8) + 20
    parents = np.zeros(2, dtype=svo.PARENT_DTYPE)
    parents[0], parents[1] = (8 * 71 + 5, 1), (300, U32MAX)
    pool = orc.TreePool(bases, blocks, usages, parents=parents, residency=np.array([1, 1], np.uint8))
    # live = 64 dense + 64, 69, 70, 71 = 68 groups: movable at >= 68 are 69 and 71 (70 is
    # shared and stays); free below 68: 65, 66, 67 -> 69 -> 65, 71 -> 66
    assert pool.compact(_ui(orc, B, 9)) == 2
    u, b = pool.usages, pool.blocks
    assert u["tick"][64:72].tolist() == [5, 6, 7, 0, 0, 0, TICK_MULTI, 0]
    assert u["parent"][65] == 10 and b["child"][10] == 8 * 65
    assert u["parent"][66] == 8 * 65 + 2 and b["child"][8 * 65 + 2] == 8 * 66
    assert b["colour_r"][8 * 65: 8 * 66].tolist() == list(range(10, 18))
    assert b["colour_r"][8 * 66: 8 * 67].tolist() == list(range(20, 28))
    # the shared group's parent-list entry follows its moved parent block
    assert pool.parents["parent_address"].tolist() == [8 * 66 + 5, 300]
    assert b["child"][8 * 66 + 5] == 8 * 70
    assert u["parent"][69] == U32MAX and u["parent"][71] == U32MAX
    assert b["child"][7] == 8 * 64 and u["parent"][64] == 7


def _tree_walk(blocks, B):
    """The tree as seen from the dense level: block records in depth-first order with Child
    replaced by whether it has children (Child 0 of solid dense cells counts as none)."""
    nd = dense_groups(B)
    fields = [f for f in blocks.dtype.names if f != "child"]
    out = []

    def rec(addr):
        blk = blocks[addr]
        c = int(blk["child"])
        has = c != U32MAX and (c >> 3) >= nd
        out.append(tuple(int(blk[f]) for f in fields) + (has,))
        if has:
            for k in range(8):
                rec(c + k)

    for a in range(8 ** (B + 2)):
        rec(a)
    return out


@pytest.fixture(scope="module")
def holey_pool(svo, orc):
    """The depth-7 sphere+noise pool with a third of its subtrees culled (holes everywhere)."""
    spec = svo.scenes.sphere_noise_scene(3, 7, base_depth=2)
    tree, bases, blocks = tree_arrays(svo, spec)
    bases, blocks, usages = pool_from_tree(svo, tree, bases, blocks, extra_groups=64, tick=1)
    B = tree.base_depth
    live = check_pool(blocks, usages, B)
    rng = np.random.default_rng(3)
    chosen = sorted(set(int(x) for x in rng.choice(list(live.values()), len(live) // 3, replace=False)))
    rec = np.zeros(len(chosen), dtype=svo.PRUNING_DTYPE)
    for i, a in enumerate(chosen):
        rec[i] = (1, 0xFF, 0, a, 0, 0, 0, U32MAX)
    pool = orc.TreePool(bases, blocks, usages)
    pool.prune(_ui(orc, B, 2), rec, np.zeros(0, dtype=svo.BLOCK_DTYPE), np.zeros((0, 3), np.uint64))
    check_pool(pool.blocks, pool.usages, B)
    return B, pool


def test_compact_is_a_relabelling(svo, orc, holey_pool):
    B, pool0 = holey_pool
    pool = orc.TreePool(pool0.bases, pool0.blocks, pool0.usages)
    before = _tree_walk(pool.blocks, B)
    live_before = int(np.count_nonzero(pool.usages["tick"] != 0))
    holes = int(np.count_nonzero(pool.usages["tick"][:live_before] == 0))
    assert holes > 10
    moved = pool.compact(_ui(orc, B, 3))
    assert moved == holes  # every live group above the live count is movable here
    ticks = pool.usages["tick"]
    assert int(np.count_nonzero(ticks != 0)) == live_before
    assert np.all(ticks[:live_before] != 0) and np.all(ticks[live_before:] == 0)
    assert _tree_walk(pool.blocks, B) == before
    check_pool(pool.blocks, pool.usages, B)
    # a compacted pool compacts to itself
    blocks, usages = pool.blocks.copy(), pool.usages.copy()
    assert pool.compact(_ui(orc, B, 3)) == 0
    assert np.array_equal(pool.blocks, blocks) and np.array_equal(pool.usages, usages)


# ---------------------------------------------------------------------------
# request consumption and the streaming step
# ---------------------------------------------------------------------------

def _requests(orc, rows):
    req = np.zeros(len(rows), dtype=orc._request_dtype())
    for i, row in enumerate(rows):
        req["address"][i], req["depth"][i] = row[0], row[1]
        req["tree_size"][i] = row[2] if len(row) > 2 else 1
    return req


def test_stream_requests_first_occurrence_known_answer(svo, orc):
    blocks = np.zeros(64, dtype=svo.BLOCK_DTYPE)
    blocks["child"] = U32MAX
    blocks["chunk"] = 0x0C00  # octant 5 holds air and geometry
    blocks["child"][9] = 16  # already has children
    blocks["chunk"][13] = 0xAAAA  # solid: every octant geometry only
    blocks["chunk"][14] = 0x5555  # empty: every octant air only
    req = _requests(orc, [(5, 6), (7, 6), (5, 6), (9, 6), (64, 6), (3, 12), (3, 11), (8, 0), (7, 6), (11, 6),
                          (13, 6), (14, 6), (20, 6, 0), (21, 6, 0), (20, 6, 1)])
    # 5 first at 0; 7 at 1; 9 has a child; 64 out of range; 3 at depth 12 = max_depth is
    # rejected and its later request at 11 is not a first occurrence; depth 0 rejected; 11
    # kept; 13 and 14 cannot have children; TreeSize-0 requests ask for nothing (21), and the
    # first request of 20 that asks for a level is the one at 14
    assert orc.stream_requests(req, 15, 15, blocks, 12).tolist() == [0, 1, 9, 14]
    # only the stored requests count
    assert orc.stream_requests(req, 10, 2, blocks, 12).tolist() == [0, 1]
    assert orc.stream_requests(req, 1, 10, blocks, 12).tolist() == [0]


def test_stream_update_grafts_in_request_order_and_collects_garbage(svo, orc):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=10)  # 74 groups, 10 free
    blocks["chunk"][[100, 200, 300]] = 0x0003  # octant 0 mixed: these can take children
    req = _requests(orc, [(300, 3), (100, 3), (300, 3), (200, 3)])

    def bricks(sel):
        out = np.zeros(8 * sel.size, dtype=svo.BLOCK_DTYPE)
        out["child"] = U32MAX
        for k, r in enumerate(sel):
            out["colour_r"][8 * k: 8 * k + 8] = int(r["address"]) % 251
        return out

    pool = orc.TreePool(bases, blocks, usages)
    st = orc.stream_update(pool, req, 4, 4, B, 6, tick=7, bricks_for=bricks, max_bricks=2, prune_age=30,
                           gc_free_percent=10)
    assert (st["unique"], st["bricks"], st["collected"], st["gc_runs"]) == (3, 2, 2, 0)
    assert st["records"]["graft_address"].tolist() == [300, 100]
    # groups in ascending order, records in request order
    assert pool.blocks["child"][300] == 8 * 64 and pool.blocks["child"][100] == 8 * 65
    assert pool.usages["tick"][64:66].tolist() == [7, 7]
    assert pool.usages["parent"][64:66].tolist() == [300, 100]
    assert pool.blocks["colour_r"][8 * 64] == 300 % 251 and pool.blocks["colour_r"][8 * 65] == 100
    assert st["free_groups"] == 8
    check_pool(pool.blocks, pool.usages, B)
    # 40 frames later the two bricks are stale; with 8 of 74 groups free (< 11 %) the step
    # collects garbage: both are reclaimed, nothing is left to move
    st = orc.stream_update(pool, req[:0], 0, 0, B, 6, tick=47, bricks_for=bricks, prune_age=30, gc_free_percent=11)
    assert (st["gc_runs"], st["reclaimed"], st["moved"], st["free_groups"]) == (1, 2, 0, 10)
    assert pool.blocks["child"][300] == U32MAX and pool.blocks["child"][100] == U32MAX
    check_pool(pool.blocks, pool.usages, B)
