"""Pinning tests of the tree-management oracle (prune, graft, collect), CPU only.

The reference kernels (wavefront.cl:1097-1350) cannot be executed (they never terminate
as written), so these tests pin the oracle with hand-computed known answers for the
allocation order, the collection passes and the parent-list unlink, with an independent
recursive restatement of the freed set, and with tree invariants over random sequences.
"""
from __future__ import annotations

import numpy as np
import pytest

from frame_helpers import tree_arrays
from tree_helpers import (TICK_INVIOLATE, TICK_MULTI, U32MAX, check_pool, pool_from_tree, random_graft,
                          random_prune, tiny_pool)


def _ui(orc, B, tick, offset=0, memory_size=0, cap=0):
    return orc.update_input(B, tick, offset=offset, memory_size=memory_size, max_child_request_id=cap)


# ---------------------------------------------------------------------------
# tick age on the 1..65533 cycle
# ---------------------------------------------------------------------------

def test_tick_age_follows_the_frame_clock(orc):
    """IncrementTick (StateManager.cs:113-122) counts 1..65533 and wraps to 1: the age of
    a tick is the number of increments since it was current."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        start = int(rng.integers(1, 65534))
        steps = int(rng.integers(0, 65533))
        t = start
        # closed-form clock: advance `steps` increments
        now = (start - 1 + steps) % 65533 + 1
        assert orc.tick_age(start, now) == steps
    # one explicit wrap: 65533 -> 1 is one frame
    assert orc.tick_age(65533, 1) == 1
    assert orc.tick_age(7, 7) == 0


# ---------------------------------------------------------------------------
# collection: ascending order, offset halving, stale detach
# ---------------------------------------------------------------------------

def test_collect_passes_known_answer(svo, orc):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=9)
    t = usages["tick"]
    t[64], t[65], t[66], t[67], t[68], t[69], t[70], t[71], t[72] = 0, 30, 60, 99, 65533, 65535, 65534, 100, 95
    # group 65 is live under dense block 200 and holds group 72 below its block 3
    usages["parent"][65] = 200
    blocks["child"][200] = 8 * 65
    blocks["child"][8 * 65 + 3] = 8 * 72
    usages["parent"][72] = 8 * 65 + 3
    pool = orc.TreePool(bases, blocks, usages)
    got = pool.collect(_ui(orc, B, 100, offset=64), 6)
    # off 64: free 64, ages 70 (65) and 100 (68); freeing below 65 frees 72;
    # off 32: age 40 (66) and the freed 72; off 16..1: nothing; off 0: age 1 (67)
    assert got.tolist() == [64, 65, 68, 66, 72, 67]
    assert all(pool.usages["tick"][g] == 100 for g in got)
    assert all(pool.usages["parent"][g] == U32MAX for g in got)
    assert pool.blocks["child"][200] == U32MAX  # stale group cut from its parent
    assert pool.usages["tick"][69] == TICK_INVIOLATE and pool.usages["tick"][70] == TICK_MULTI
    assert pool.usages["tick"][71] == 100  # age 0 is never stale
    # asking for more than exists returns what there is
    pool2 = orc.TreePool(bases, blocks, usages)
    assert pool2.collect(_ui(orc, B, 100, offset=64), 50).tolist() == [64, 65, 68, 66, 72, 67]


def test_collect_memory_size_limits_the_scan(svo, orc):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=8)
    pool = orc.TreePool(bases, blocks, usages)
    # memory_size counts blocks: 8 * 68 blocks -> groups 0..67 are scanned
    got = pool.collect(_ui(orc, B, 3, memory_size=8 * 68), 10)
    assert got.tolist() == [64, 65, 66, 67]


# ---------------------------------------------------------------------------
# graft: pre-order placement into collected groups
# ---------------------------------------------------------------------------

def _graft_three(svo):
    """Top group T (slot 0), X (slot 1), Y (slot 2): T.block2 -> Y, T.block5 -> X."""
    gb = np.zeros(24, dtype=svo.BLOCK_DTYPE)
    gb["child"] = U32MAX
    gb["colour_r"] = np.arange(24)
    gb["chunk"] = 0x5555 + np.arange(24)
    gb["child"][2] = 16
    gb["child"][5] = 8
    return gb


def test_graft_preorder_placement_known_answer(svo, orc):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=16)
    rec = np.zeros(1, dtype=svo.GRAFTING_DTYPE)
    rec[0] = (0, 24, 0xFF, (0, 0, 0), 100)
    pool = orc.TreePool(bases, blocks, usages)
    st = pool.graft(_ui(orc, B, 7), rec, _graft_three(svo), np.zeros((0, 3), np.uint64))
    assert st["placed"] == 1 and st["collected"] == 3
    b, u = pool.blocks, pool.usages
    # pre-order: T -> 64, Y (via T.block2) -> 65, X (via T.block5) -> 66
    assert b["child"][100] == 8 * 64
    assert b["child"][8 * 64 + 2] == 8 * 65 and b["child"][8 * 64 + 5] == 8 * 66
    assert [int(x) for x in b["colour_r"][8 * 64: 8 * 64 + 8]] == list(range(8))
    assert [int(x) for x in b["colour_r"][8 * 65: 8 * 65 + 8]] == list(range(16, 24))
    assert [int(x) for x in b["colour_r"][8 * 66: 8 * 66 + 8]] == list(range(8, 16))
    assert np.all(b["child"][8 * 65: 8 * 67] == U32MAX)
    assert [(int(u["tick"][g]), int(u["parent"][g])) for g in (64, 65, 66)] == [(7, 100), (7, 8 * 64 + 2), (7, 8 * 64 + 5)]
    assert np.all(u["tick"][67:] == 0)
    check_pool(b, u, B)


def test_graft_replaces_existing_child_and_releases_unused(svo, orc):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=16)
    # dense block 100 already holds group 70 (tick 5) which holds group 71
    blocks["child"][100] = 8 * 70
    usages["tick"][70], usages["parent"][70] = 5, 100
    blocks["child"][8 * 70 + 1] = 8 * 71
    usages["tick"][71], usages["parent"][71] = 5, 8 * 70 + 1
    gb = _graft_three(svo)
    rec = np.zeros(3, dtype=svo.GRAFTING_DTYPE)
    rec[0] = (0, 24, 0xFF, (0, 0, 0), 100)
    rec[1] = (0, 12, 0xFF, (0, 0, 0), 101)  # malformed size: not collected for
    rec[2] = (0, 24, 0xFF, (0, 0, 0), 100)  # same target: the later record wins
    pool = orc.TreePool(bases, blocks, usages)
    st = pool.graft(_ui(orc, B, 9), rec, gb, np.zeros((0, 3), np.uint64))
    assert st["collected"] == 6 and st["placed"] == 1 and st["invalid"] == 1 and st["skipped"] == 1
    u, b = pool.usages, pool.blocks
    # collected ascending: 64 65 66 (record 0, released) 67 68 69 (record 2, linked)
    assert b["child"][100] == 8 * 67
    assert [int(u["tick"][g]) for g in (64, 65, 66)] == [0, 0, 0]
    assert [int(u["tick"][g]) for g in (67, 68, 69)] == [9, 9, 9]
    assert u["tick"][70] == 0 and u["tick"][71] == 0  # the old subtree is freed
    check_pool(b, u, B)


def test_graft_too_few_groups_places_a_prefix(svo, orc):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=4)
    gb = _graft_three(svo)
    rec = np.zeros(2, dtype=svo.GRAFTING_DTYPE)
    rec[0] = (0, 24, 0xFF, (0, 0, 0), 100)
    rec[1] = (0, 24, 0xFF, (0, 0, 0), 101)
    pool = orc.TreePool(bases, blocks, usages)
    st = pool.graft(_ui(orc, B, 2), rec, gb, np.zeros((0, 3), np.uint64))
    assert st["collected"] == 4 and st["placed"] == 1
    assert pool.blocks["child"][100] == 8 * 64 and pool.blocks["child"][101] == U32MAX
    assert pool.usages["tick"][67] == 0  # the fourth group went back to the pool
    check_pool(pool.blocks, pool.usages, B)


# ---------------------------------------------------------------------------
# prune
# ---------------------------------------------------------------------------

def _descendants(blocks, usages, root_group):
    """Recursive restatement of the freed set below one detached group."""
    out = set()

    def rec(g):
        t = int(usages["tick"][g])
        if t in (TICK_INVIOLATE, TICK_MULTI) or g in out:
            return
        out.add(g)
        for k in range(8):
            c = int(blocks["child"][8 * g + k])
            if c != U32MAX:
                rec(c >> 3)

    rec(root_group)
    return out


@pytest.fixture(scope="module")
def noise_pool(svo):
    spec = svo.scenes.sphere_noise_scene(3, 7, base_depth=2)
    tree, bases, blocks = tree_arrays(svo, spec)
    return tree, pool_from_tree(svo, tree, bases, blocks, extra_groups=256, tick=1)


def test_cull_frees_exactly_the_subtrees(svo, orc, noise_pool):
    tree, (bases, blocks, usages) = noise_pool
    B = tree.base_depth
    live = check_pool(blocks, usages, B)
    rng = np.random.default_rng(11)
    # cull the children of 12 blocks that have one
    parents = [pb for pb in live.values()]
    chosen = sorted(set(int(x) for x in rng.choice(parents, 12, replace=False)))
    rec = np.zeros(len(chosen), dtype=svo.PRUNING_DTYPE)
    for i, a in enumerate(chosen):
        rec[i] = (1, 0xFF, 0, a, 0, 0, 0, U32MAX)
    expect = set()
    for a in chosen:
        expect |= _descendants(blocks, usages, int(blocks["child"][a]) >> 3)
    pool = orc.TreePool(bases, blocks, usages)
    st = pool.prune(_ui(orc, B, 4), rec, np.zeros(0, dtype=svo.BLOCK_DTYPE), np.zeros((0, 3), np.uint64))
    assert st["applied"] == len(chosen)
    freed = set(np.nonzero((pool.usages["tick"] == 0) & (usages["tick"] != 0))[0].tolist())
    assert freed == expect
    assert all(pool.blocks["child"][a] == U32MAX for a in chosen)
    check_pool(pool.blocks, pool.usages, B)


def test_multi_parent_unlink_known_answer(svo, orc):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=8)
    usages["tick"][64], usages["parent"][64] = TICK_MULTI, 0
    parents = np.zeros(4, dtype=svo.PARENT_DTYPE)
    parents[0], parents[1], parents[2], parents[3] = (300, 1), (301, 2), (302, U32MAX), (0, U32MAX)
    residency = np.array([1, 1, 1, 0], np.uint8)
    for pb in (300, 301, 302):
        blocks["child"][pb] = 8 * 64
    blocks["child"][8 * 64] = 8 * 65
    usages["tick"][65], usages["parent"][65] = 5, 8 * 64
    rec = np.zeros(2, dtype=svo.PRUNING_DTYPE)
    rec[0] = (1, 0xFF, 0, 302, 0, 0, 0, U32MAX)
    rec[1] = (1, 0xFF, 0, 300, 0, 0, 0, U32MAX)
    pool = orc.TreePool(bases, blocks, usages, parents=parents, residency=residency)
    st = pool.prune(_ui(orc, B, 9), rec, np.zeros(0, dtype=svo.BLOCK_DTYPE), np.zeros((0, 3), np.uint64))
    assert st["unlinks"] == 2
    # events in ascending (group, parent) order: 300 (the head) then 302 -> one parent left
    assert int(pool.usages["tick"][64]) == 9 and int(pool.usages["parent"][64]) == 301
    assert pool.residency.tolist() == [0, 0, 0, 0]
    assert int(pool.usages["tick"][65]) == 5  # a shared group's children stay
    assert pool.blocks["child"][300] == U32MAX and pool.blocks["child"][302] == U32MAX
    check_pool(pool.blocks, pool.usages, B)


def test_prune_fields_and_last_record_wins(svo, orc):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=8)
    blocks["child"][100] = 8 * 64
    usages["tick"][64], usages["parent"][64] = 3, 100
    surface = np.zeros(2, dtype=svo.BLOCK_DTYPE)
    surface["colour_r"] = [11, 22]
    surface["child"] = [7, 8]
    rec = np.zeros(6, dtype=svo.PRUNING_DTYPE)
    rec[0] = (0x10, 0, 0, 2, 0x1234, 0, 0, 0)              # BaseBlock
    rec[1] = (0x10, 0, 0, 2, 0x4321, 0, 0, 0)              # BaseBlock, same cell: wins
    rec[2] = (0x08, 0xFF, 0, 8 * 64 + 1, 0x00AA, 0, 0, 0)  # UpdateChunk from surface 0
    rec[3] = (0x08, 0xFF, 0, 8 * 64 + 1, 0x00BB, 0, 1, 0)  # UpdateChunk from surface 1: wins
    rec[4] = (0x06, 0xFF, 0, 8 * 64 + 2, 0, 0, 0, 0)       # MakeInviolate + AlterViolability
    rec[5] = (0x08, 0xFF, 0, 8 * 64 + 3, 0, 0, 9, 0)       # surface index out of range
    pool = orc.TreePool(bases, blocks, usages)
    st = pool.prune(_ui(orc, B, 6), rec, surface, np.zeros((0, 3), np.uint64))
    assert st["invalid"] == 1 and st["applied"] == 5
    assert int(pool.bases[2]) == 0x4321
    b = pool.blocks[8 * 64 + 1]
    assert int(b["colour_r"]) == 22 and int(b["chunk"]) == 0x00BB and int(b["child"]) == U32MAX
    assert int(pool.usages["tick"][64]) == TICK_INVIOLATE


def test_prune_location_miss_requests_child(svo, orc):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=8)
    loc = np.array([[3 << 60, 5 << 60, 9 << 59]], np.uint64)
    rec = np.zeros(2, dtype=svo.PRUNING_DTYPE)
    rec[0] = (0x02, 7, 0, 0, 0, 0, 0, 0)  # AlterViolability at depth 7 via location 0
    rec[1] = (0x02, 3, 0, 0, 0, 0, 0, 0)  # depth B+2: the dense block itself
    pool = orc.TreePool(bases, blocks, usages, request_capacity=4)
    st = pool.prune(_ui(orc, B, 12, cap=4), rec, np.zeros(0, dtype=svo.BLOCK_DTYPE), loc)
    assert st["requests"] == 1 and pool.request_count == 1
    q = pool.requests[0]
    dense = orc.base_location(B + 2, [int(v) for v in loc[0]])
    assert int(q["address"]) == dense and int(q["depth"]) == B + 2 and int(q["tree_size"]) == 7 - (B + 2)
    assert int(q["tick"]) == 12 and [int(v) for v in q["location"]] == [int(v) for v in loc[0]]


def test_request_capacity_drops_but_counts(svo, orc):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=8)
    rng = np.random.default_rng(2)
    loc = rng.integers(0, 2 ** 63, size=(5, 3), dtype=np.uint64)
    rec = np.zeros(5, dtype=svo.PRUNING_DTYPE)
    for i in range(5):
        rec[i] = (0, 6, 0, i, 0, 0, 0, 0)
    pool = orc.TreePool(bases, blocks, usages, request_capacity=3, request_count=1)
    pool.prune(_ui(orc, B, 2), rec, np.zeros(0, dtype=svo.BLOCK_DTYPE), loc)
    assert pool.request_count == 6
    assert [int(pool.requests[i]["location"][0]) for i in (1, 2)] == [int(loc[0][0]), int(loc[1][0])]


# ---------------------------------------------------------------------------
# random sequences: invariants hold after every dispatch
# ---------------------------------------------------------------------------

def test_random_sequences_keep_the_tree_consistent(svo, orc, noise_pool):
    tree, (bases, blocks, usages) = noise_pool
    B = tree.base_depth
    rng = np.random.default_rng(23)
    pool = orc.TreePool(bases, blocks, usages, request_capacity=64)
    for frame in range(2, 14):
        rec, surface, loc = random_prune(svo, rng, pool.blocks, pool.usages, B, 30, max_depth=tree.max_depth + 1)
        rec["child_address"] = U32MAX  # culls only cut here; linking arbitrary groups breaks the tree
        pool.prune(_ui(orc, B, frame, offset=4, cap=64), rec, surface, loc)
        check_pool(pool.blocks, pool.usages, B)
        grec, gb, gloc = random_graft(svo, rng, pool.blocks, pool.usages, B, 6, max_depth=tree.max_depth + 1)
        st = pool.graft(_ui(orc, B, frame, offset=4, cap=64), grec, gb, gloc)
        live = check_pool(pool.blocks, pool.usages, B)
        assert st["placed"] <= len(grec)
        # nothing reachable is free, and free + live never exceeds the pool
        assert len(live) + int(np.count_nonzero(pool.usages["tick"] == 0)) <= pool.usages.size
