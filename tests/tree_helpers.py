"""Pools, random management records and invariant checks shared by the tree-management tests."""
from __future__ import annotations

import numpy as np

U32MAX = 0xFFFFFFFF
TICK_MULTI = 65534
TICK_INVIOLATE = 65535


def dense_groups(B: int) -> int:
    return 8 ** (B + 1)


def pool_from_tree(svo, tree, bases, blocks, extra_groups: int, tick: int = 1):
    """Tree arrays padded to a pool: `extra_groups` free groups after the tree's groups."""
    n_groups = tree.group_count + extra_groups
    pool_blocks = np.zeros(8 * n_groups, dtype=svo.BLOCK_DTYPE)
    pool_blocks["child"] = U32MAX
    pool_blocks[: blocks.size] = blocks
    usages = tree.initial_usages(tick=tick, groups=n_groups)
    return bases.copy(), pool_blocks, usages


def tiny_pool(svo, B: int = 1, extra_groups: int = 16, tick: int = 1):
    """A dense level and nothing below it: every dense block is a childless full cell."""
    n_dense_blocks = 8 ** (B + 2)
    n_groups = dense_groups(B) + extra_groups
    bases = np.full(sum(8 ** d for d in range(1, B + 1)), 0xAAAA, np.uint16)  # every octant geometry
    blocks = np.zeros(8 * n_groups, dtype=svo.BLOCK_DTYPE)
    blocks["child"] = U32MAX
    blocks["chunk"][:n_dense_blocks] = 0xAAAA
    usages = np.zeros(n_groups, dtype=svo.USAGE_DTYPE)
    usages["parent"] = U32MAX
    usages["tick"][: dense_groups(B)] = TICK_INVIOLATE
    return bases, blocks, usages


def reachable(blocks, usages, B: int):
    """Groups reachable from the dense level through Child links: {group: parent block}.
    Raises if a group is reached twice (the pool is a tree) or a link leaves the pool.
    Links into the dense level are the zero blocks TreeBuilder leaves for solid cells
    (Child 0, treated as no child as checkTree does, TreeBuilder.cs:381)."""
    n_dense_blocks = 8 ** (B + 2)
    nd = dense_groups(B)
    seen = {}
    frontier = [(int(c) >> 3, b) for b, c in enumerate(blocks["child"][:n_dense_blocks])
                if c != U32MAX and (int(c) >> 3) >= nd]
    while frontier:
        nxt = []
        for g, pb in frontier:
            assert g < usages.size, f"link from block {pb} leaves the pool (group {g})"
            assert g not in seen, f"group {g} reached from blocks {seen[g]} and {pb}"
            seen[g] = pb
            for k in range(8):
                c = int(blocks["child"][8 * g + k])
                if c != U32MAX and (c >> 3) >= nd:
                    nxt.append((c >> 3, 8 * g + k))
        frontier = nxt
    return seen


def check_pool(blocks, usages, B: int):
    """Pool invariants: every reachable group is live (tick != 0) and, when single-parent,
    points back at the block that links it; inviolate dense groups stay inviolate."""
    live = reachable(blocks, usages, B)
    ticks = usages["tick"]
    assert np.all(ticks[: dense_groups(B)] == TICK_INVIOLATE)
    for g, pb in live.items():
        t = int(ticks[g])
        assert t != 0, f"reachable group {g} (parent block {pb}) is marked free"
        if t < TICK_MULTI:
            assert int(usages["parent"][g]) == pb, f"group {g}: parent {usages['parent'][g]} != {pb}"
    return live


def random_location(rng):
    return [int(v) for v in rng.integers(0, 2 ** 63, size=3, dtype=np.uint64) * 2 + rng.integers(0, 2, 3).astype(np.uint64)]


def random_prune(svo, rng, blocks, usages, B: int, n: int, n_surface: int = 8, max_depth: int = 10):
    """Random Pruning records over the current pool: direct and location targets, every
    property bit, duplicates and a few out-of-range records."""
    live_groups = np.nonzero((usages["tick"] != 0) & (usages["tick"] < TICK_MULTI))[0]
    rec = np.zeros(n, dtype=svo.PRUNING_DTYPE)
    locs = []
    for i in range(n):
        kind = rng.random()
        props = int(rng.integers(0, 16))
        if kind < 0.05:
            rec[i]["properties"] = 0x10
            rec[i]["address"] = int(rng.integers(0, 12))
            rec[i]["chunk"] = int(rng.integers(0, 65536))
            continue
        rec[i]["properties"] = props
        rec[i]["chunk"] = int(rng.integers(0, 65536))
        rec[i]["colour_address"] = int(rng.integers(0, n_surface + (1 if kind > 0.97 else 0)))
        rec[i]["child_address"] = U32MAX
        if kind < 0.55:
            rec[i]["depth"] = 0xFF
            if live_groups.size and rng.random() < 0.9:
                rec[i]["address"] = 8 * int(rng.choice(live_groups)) + int(rng.integers(0, 8))
            else:
                rec[i]["address"] = int(rng.integers(0, blocks.size + 4))
        else:
            rec[i]["depth"] = int(rng.integers(B + 2, max_depth + 1))
            rec[i]["address"] = len(locs)
            locs.append(random_location(rng))
        if rec[i]["depth"] != 0xFF or int(rec[i]["address"]) < 8 * dense_groups(B):
            # keep the dense level inviolate: AlterViolability only on directly addressed live groups
            rec[i]["properties"] = int(rec[i]["properties"]) & 0xFD
    if n > 3:
        rec[n - 1] = rec[int(rng.integers(0, n - 1))]  # a duplicate target
    surface = np.zeros(n_surface, dtype=svo.BLOCK_DTYPE)
    for f in ("chunk", "colour_r", "colour_g", "colour_b", "opacity", "gloss"):
        surface[f] = rng.integers(0, 256, n_surface)
    surface["child"] = rng.integers(0, 2 ** 32, n_surface, dtype=np.uint64).astype(np.uint32)
    loc = np.array(locs, dtype=np.uint64).reshape(-1, 3)
    return rec, surface, loc


def random_graft_blocks(svo, rng, groups: int, preorder: bool):
    """One graft subtree of `groups` groups, packed in pre-order or shuffled.  Child fields
    are graft-relative block indices (rebased by the caller)."""
    parent_of = [None] + [int(rng.integers(0, g)) for g in range(1, groups)]
    # children of each group in octant order: pick distinct octants per parent
    free_oct = {g: list(rng.permutation(8)) for g in range(groups)}
    link = {}
    for g in range(1, groups):
        p = parent_of[g]
        while not free_oct[p]:
            p = int(rng.integers(0, g))
        link[g] = (p, free_oct[p].pop())
    order = list(range(groups)) if preorder else [0] + list(rng.permutation(np.arange(1, groups)))
    slot = {g: i for i, g in enumerate(order)}
    gb = np.zeros(8 * groups, dtype=svo.BLOCK_DTYPE)
    gb["child"] = U32MAX
    for f in ("chunk", "colour_r", "colour_g", "colour_b", "opacity", "gloss", "normal_pitch"):
        gb[f] = rng.integers(0, 256, gb.size)
    for g, (p, k) in link.items():
        gb["child"][8 * slot[p] + k] = 8 * slot[g]
    return gb


def random_graft(svo, rng, blocks, usages, B: int, n: int, max_groups: int = 4, max_depth: int = 10):
    """Random Grafting records, their packed blocks (Child fields rebased as WorldManager
    does, WorldManager.cs:84-103) and target locations."""
    live_groups = np.nonzero((usages["tick"] != 0) & (usages["tick"] < TICK_MULTI))[0]
    rec = np.zeros(n, dtype=svo.GRAFTING_DTYPE)
    parts, locs, base = [], [], 0
    for i in range(n):
        groups = int(rng.integers(1, max_groups + 1))
        gb = random_graft_blocks(svo, rng, groups, preorder=bool(rng.random() < 0.5))
        linked = gb["child"] != U32MAX
        gb["child"][linked] += base
        rec[i]["graft_data_address"] = base
        rec[i]["graft_total_size"] = gb.size
        parts.append(gb)
        base += gb.size
        if rng.random() < 0.5:
            rec[i]["depth"] = 0xFF
            if live_groups.size:
                rec[i]["graft_address"] = 8 * int(rng.choice(live_groups)) + int(rng.integers(0, 8))
            else:
                rec[i]["graft_address"] = int(rng.integers(0, 8 ** (B + 2)))
        else:
            rec[i]["depth"] = int(rng.integers(B + 2, max_depth + 1))
            rec[i]["graft_address"] = len(locs)
            locs.append(random_location(rng))
    if n > 3 and rng.random() < 0.5:
        rec[n - 1]["graft_total_size"] = 12  # malformed: not whole groups
    gbl = np.concatenate(parts) if parts else np.zeros(0, dtype=svo.BLOCK_DTYPE)
    return rec, gbl, np.array(locs, dtype=np.uint64).reshape(-1, 3)
