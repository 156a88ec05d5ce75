"""svo_graft at the sizes where its scans span several work-groups (more than 2 x 16,384
records): 32,768, 32,769 and 3 x 16,384 + 5 records of one to three groups each, with
records that need nothing mixed in (duplicate targets, malformed sizes, targets outside
Blocks).  The scan over the records' sizes places every graft's blocks and the scan over
their needs places every collected group, so a wrong prefix sum moves blocks or links and
the pool (Blocks, Usages) and the statistics differ from the oracle's graft."""
from __future__ import annotations

import numpy as np
import pytest

from test_tree_parity import Mirror, _stats_equal, _ui
from tree_helpers import U32MAX, random_graft_blocks, tiny_pool

pytestmark = pytest.mark.gpu

B = 3  # 8^5 = 32,768 dense blocks, every one a childless full cell


def large_graft(svo, rng, n: int, n_blocks: int):
    """n Grafting records and their packed blocks: distinct dense targets first, then a mix of
    duplicate targets, malformed sizes (not whole groups) and targets beyond Blocks."""
    n_dense = 8 ** (B + 2)
    targets = rng.permutation(n_dense)
    rec = np.zeros(n, dtype=svo.GRAFTING_DTYPE)
    parts, base = [], 0
    kinds = rng.random(n)
    for i in range(n):
        groups = 1 if rng.random() < 0.7 else int(rng.integers(2, 4))
        gb = random_graft_blocks(svo, rng, groups, preorder=bool(rng.random() < 0.5))
        linked = gb["child"] != U32MAX
        gb["child"][linked] += base
        rec[i]["graft_data_address"] = base
        rec[i]["graft_total_size"] = gb.size
        rec[i]["depth"] = 0xFF
        if i < min(n, 30000) or kinds[i] < 0.4:
            rec[i]["graft_address"] = int(targets[i % n_dense])
        elif kinds[i] < 0.7:
            rec[i]["graft_address"] = int(targets[int(rng.integers(0, 30000))])  # a duplicate target
        elif kinds[i] < 0.85:
            rec[i]["graft_address"] = int(targets[i % n_dense])
            rec[i]["graft_total_size"] = gb.size - 4  # malformed: not whole groups
        else:
            rec[i]["graft_address"] = n_blocks + int(rng.integers(0, 1000))  # outside Blocks
        parts.append(gb)
        base += gb.size
    return rec, np.concatenate(parts)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("n", [32768, 32769, 3 * 16384 + 5])
def test_graft_multi_workgroup_scan(svo, orc, gpu, n):
    rng = np.random.default_rng(n)
    extra = 3 * n + 64
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=extra)
    rec, gb = large_graft(svo, rng, n, blocks.size)
    print(f"graft: {n} records, {gb.size} graft blocks, pool {usages.size} groups", flush=True)
    m = Mirror(svo, orc, gpu, bases, blocks, usages)
    got, ref = m.graft(_ui(orc, B, 7), rec, gb, np.zeros((0, 3), np.uint64))
    _stats_equal(got, ref)
    assert ref["applied"] > 30000, "most records were meant to graft"
    assert ref["invalid"] > 0 and ref["skipped"] > 0, "records that need nothing were meant to be mixed in"
    m.compare(f"graft of {n} records")
