"""Helpers of the streaming tests: seed pools, brick lookups in a statically built tree and
the oracle's side of a streamed frame."""
from __future__ import annotations

import numpy as np

U32MAX = 0xFFFFFFFF


def pow_sum(d: int) -> int:
    return sum(8 ** i for i in range(1, d + 1))


def interleave3(x: int, y: int, z: int) -> int:
    out = 0
    for i in range(21):
        out |= ((x >> i) & 1) << (3 * i) | ((y >> i) & 1) << (3 * i + 1) | ((z >> i) & 1) << (3 * i + 2)
    return out


def terrain_geometry(svo, resolution: int, seed: int = 7):
    return svo.scenes.config3_geometry(resolution=resolution, seed=seed)


def static_tree(svo, geometry, B: int, max_depth: int):
    """TreeBuilder's tree of the terrain (svo_tree_build), every block refined to max_depth."""
    return svo.build_tree(geometry, B, max_depth, 0xFFFFFFFF // 8)


def seed_pool(svo, geometry, B: int, n_groups: int):
    """The pool svo_stream_seed makes: TreeBuilder's tree cut at the dense level (max_depth
    B+2), zero blocks after it, dense groups inviolate, every other group free."""
    tree = svo.build_tree(geometry, B, B + 2, 0)
    blocks = np.zeros(8 * n_groups, dtype=svo.BLOCK_DTYPE)
    blocks[: tree.block_count] = tree.blocks
    usages = np.zeros(n_groups, dtype=svo.USAGE_DTYPE)
    usages["parent"] = U32MAX
    usages["tick"][: 8 ** (B + 1)] = 65535
    return tree.bases.copy(), blocks, usages


def static_children(static_blocks: np.ndarray, B: int, requests: np.ndarray) -> np.ndarray:
    """The 8 children, in the static tree, of the cell each request names (depth d and the
    cell holding its location), without their own children: what a brick for that request
    must hold."""
    out = np.zeros(8 * requests.size, dtype=static_blocks.dtype)
    child = static_blocks["child"]
    for k, r in enumerate(requests):
        d = int(r["depth"])
        x, y, z = (int(v) >> (64 - d) for v in r["location"])
        sh = d - (B + 2)
        a = interleave3(x >> sh, y >> sh, z >> sh)
        for level in range(B + 2, d):
            s = d - level - 1
            a = int(child[a]) + (((x >> s) & 1) | (((y >> s) & 1) << 1) | (((z >> s) & 1) << 2))
        c = int(child[a])
        assert c != U32MAX, f"request {k}: the static tree has no children at depth {d}"
        out[8 * k: 8 * k + 8] = static_blocks[c: c + 8]
    out["child"] = U32MAX  # a brick is one level: its blocks are childless
    return out
