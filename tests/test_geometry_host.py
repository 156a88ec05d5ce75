"""svo_geometry_blocks (the builder's MakeBlock for arbitrary cells) against the blocks the host
TreeBuilder wrote for the same cells (CPU): every block of the dense level and of the levels
below it, reached by walking the built tree, equals MakeBlock of its cell apart from the Child
link; for the fBm heightfield and the sphere + noise SDF."""
from __future__ import annotations

import numpy as np
import pytest

from stream_helpers import U32MAX, interleave3


def _cells_of_tree(tree, B: int, max_cells: int = 4000):
    """(level, x, y, z) and address of blocks of the built tree, breadth first from the dense
    level.  The dense level holds blocks only under level-(B+1) cells with both air and
    geometry (TreeBuilder recycles the other dense addresses as free groups)."""
    blocks, bases = tree.blocks, tree.bases
    n = 1 << (B + 2)
    off = sum(8 ** i for i in range(1, B))  # powSum(B - 1)
    out = []
    frontier = []
    for z in range(n):
        for y in range(n):
            for x in range(n):
                chunk = int(bases[off + interleave3(x >> 2, y >> 2, z >> 2)])
                octant = ((x >> 1) & 1) | (((y >> 1) & 1) << 1) | (((z >> 1) & 1) << 2)
                if (chunk >> (2 * octant)) & 3 == 3:
                    frontier.append((B + 2, x, y, z, interleave3(x, y, z)))
    rng = np.random.default_rng(5)
    while frontier and len(out) < max_cells:
        pick = rng.choice(len(frontier), size=min(len(frontier), 64), replace=False)
        nxt = []
        for i in pick:
            level, x, y, z, a = frontier[i]
            out.append((level, x, y, z, a))
            c = int(blocks["child"][a])
            if c != U32MAX:
                for k in range(8):
                    nxt.append((level + 1, 2 * x + (k & 1), 2 * y + ((k >> 1) & 1), 2 * z + ((k >> 2) & 1), c + k))
        frontier = [f for j, f in enumerate(frontier) if j not in set(pick.tolist())] + nxt
    return out


@pytest.mark.parametrize("kind", ["heightfield", "sphere_noise"])
def test_geometry_blocks_equal_the_builder(svo, kind):
    if kind == "heightfield":
        geom, B, D = svo.scenes.config3_geometry(resolution=256), 2, 8
    else:
        spec = svo.scenes.sphere_noise_scene(seed=3, max_depth=7, base_depth=2)
        geom, B, D = spec.geometry, 2, 7
    tree = svo.build_tree(geom, B, D, 0xFFFFFFFF // 8)
    blocks = tree.blocks
    cells = _cells_of_tree(tree, B)
    assert len({c[0] for c in cells}) >= 3
    geo = svo.GeometryHost(geom)
    got = geo.blocks(np.array([c[:4] for c in cells], np.uint32))
    geo.close()
    want = blocks[[c[4] for c in cells]].copy()
    assert np.all(got["child"] == U32MAX)
    want["child"] = U32MAX
    assert np.array_equal(got.view(np.uint8).reshape(-1, 16), want.view(np.uint8).reshape(-1, 16))
