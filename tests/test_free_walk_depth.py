"""Free walks deeper than the levels a dispatch launches before reading the frontier back:
pruning the children of dense-level blocks of a depth-10 tree (base depth 2) frees subtrees
six levels deep, so the walk continues level by level after its first three; collection of
stale groups with deep subtrees does the same inside a graft.  Pool and statistics bit-exact
against the oracle after every dispatch."""
from __future__ import annotations

import numpy as np
import pytest

from frame_helpers import tree_arrays
from test_tree_parity import Mirror, _stats_equal, _ui
from tree_helpers import U32MAX, check_pool, dense_groups, pool_from_tree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def deep_tree(svo):
    return tree_arrays(svo, svo.scenes.sphere_noise_scene(5, 10, base_depth=2))


def _walk_depth(blocks, start_groups):
    """Levels of the subtree below the given groups (the free walk's level count)."""
    level, frontier = 0, list(start_groups)
    while frontier:
        level += 1
        nxt = []
        for g in frontier:
            for k in range(8):
                c = int(blocks["child"][8 * g + k])
                if c != U32MAX:
                    nxt.append(c >> 3)
        frontier = nxt
    return level


def test_prune_deep_subtrees(svo, orc, gpu, deep_tree):
    tree, bases0, blocks0 = deep_tree
    B = tree.base_depth
    bases, blocks, usages = pool_from_tree(svo, tree, bases0, blocks0, extra_groups=64, tick=1)
    m = Mirror(svo, orc, gpu, bases, blocks, usages)
    n_dense = 8 ** (B + 2)
    child = m.ref.blocks["child"][:n_dense]
    parents = np.nonzero((child != U32MAX) & ((child >> 3) >= dense_groups(B)))[0][:48]
    assert parents.size >= 8
    depth = _walk_depth(m.ref.blocks, [int(child[a]) >> 3 for a in parents])
    assert depth > 3, f"the subtrees were meant to be deeper than three levels (got {depth})"
    rec = np.zeros(parents.size, dtype=svo.PRUNING_DTYPE)
    rec["properties"] = 1  # CullChild
    rec["depth"] = 0xFF
    rec["address"] = parents.astype(np.uint32)
    rec["child_address"] = U32MAX
    got, ref = m.prune(_ui(orc, B, 3), rec, np.zeros(1, dtype=svo.BLOCK_DTYPE), np.zeros((0, 3), np.uint64))
    _stats_equal(got, ref)
    m.compare(f"prune of {parents.size} subtrees {depth} levels deep")
    check_pool(m.ref.blocks, m.ref.usages, B)


def test_collect_stale_deep_subtrees(svo, orc, gpu, deep_tree):
    tree, bases0, blocks0 = deep_tree
    B = tree.base_depth
    bases, blocks, usages = pool_from_tree(svo, tree, bases0, blocks0, extra_groups=8, tick=1)
    m = Mirror(svo, orc, gpu, bases, blocks, usages)
    # every tree group is stale at tick 200 with offset 50; the graft needs more groups than the
    # pool has free, so the collection takes stale groups and frees what hangs below them
    gb = np.zeros(8 * 40, dtype=svo.BLOCK_DTYPE)
    gb["child"] = U32MAX
    gb["chunk"] = 0xAAAA
    rec = np.zeros(40, dtype=svo.GRAFTING_DTYPE)
    rec["graft_total_size"] = 8
    rec["graft_data_address"] = 8 * np.arange(40, dtype=np.uint32)
    rec["depth"] = 0xFF
    rec["graft_address"] = np.arange(40, dtype=np.uint32) * 3 + 7
    got, ref = m.graft(_ui(orc, B, 200, offset=50), rec, gb, np.zeros((0, 3), np.uint64))
    _stats_equal(got, ref)
    assert ref["collected"] >= 40
    m.compare("graft collecting stale groups with deep subtrees")
