"""Writes the frame and pool fixtures of tests/golden (run from the repo root; needs the
built library for the scene builder and the built oracle):

* f0_frame.npz: the reference's default scene (two spheres, BaseDepth 4; Program.cs:24-71,
  built to depth 8) seen by its default camera at 64x48 -- per-cone hit index, visit count,
  visit hash, RGBA and status, the child requests, and digests of the tree and of the usage
  table after the frame (frame tick 1);
* f1_frame.npz: the same for the config-1 scene (sphere + value noise, depth 8) at 64x48;
* pool_script.npz: a small pool (BaseDepth 1, 48 free groups) and a scripted sequence of
  prune and graft dispatches (seed 31), with the pool after each one.

The expected outputs come from the oracle; the tests check that the oracle still produces
them (a regression pin of the checker) and that the GPU produces them."""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT / "svo-tracer_amd"), str(ROOT / "oracle"), str(ROOT / "tests")]
OUT = Path(__file__).resolve().parent


def tree_digest(bases: np.ndarray, blocks: np.ndarray) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(bases).tobytes())
    h.update(np.ascontiguousarray(blocks).tobytes())
    return h.hexdigest()


def frame_fixture(svo, orc, spec, cam, path: Path):
    tree = spec.build()
    usages = tree.initial_usages(tick=0)
    ref = orc.trace_frame(cam, tree.bases, tree.blocks, usages, request_capacity=256, threads=8)
    np.savez_compressed(path, digest=np.array(tree_digest(tree.bases, tree.blocks)),
                        hit=ref["hit"], visits=ref["visits"], hash=ref["hash"], rgba=ref["rgba"],
                        status=ref["status"], request_count=np.array(ref["request_count"]),
                        requests=ref["requests"].view(np.uint8),
                        usages_digest=np.array(hashlib.sha256(usages.tobytes()).hexdigest()))
    print(f"wrote {path}: {tree.block_count} blocks, {int(np.count_nonzero(ref['hit'] != 0xFFFFFFFF))} hits")


def pool_fixture(svo, orc, path: Path):
    from tree_helpers import random_graft, random_prune, tiny_pool

    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=48)
    rng = np.random.default_rng(31)
    pool = orc.TreePool(bases, blocks, usages, request_capacity=32)
    arrays = {"bases0": bases, "blocks0": blocks.view(np.uint8), "usages0": usages.view(np.uint8)}
    steps = []
    for k, tick in enumerate((2, 3, 4, 5)):
        ui = orc.update_input(B, tick, offset=1, max_child_request_id=32)
        if k % 2 == 0:
            grec, gb, gloc = random_graft(svo, rng, pool.blocks, pool.usages, B, 5, max_depth=6)
            pool.graft(ui, grec, gb, gloc)
            arrays[f"s{k}_records"] = grec.view(np.uint8)
            arrays[f"s{k}_side"] = gb.view(np.uint8)
            arrays[f"s{k}_locations"] = np.ascontiguousarray(gloc, np.uint64).reshape(-1)
            steps.append("graft")
        else:
            rec, surface, loc = random_prune(svo, rng, pool.blocks, pool.usages, B, 12, max_depth=6)
            rec["child_address"] = 0xFFFFFFFF
            pool.prune(ui, rec, surface, loc)
            arrays[f"s{k}_records"] = rec.view(np.uint8)
            arrays[f"s{k}_side"] = surface.view(np.uint8)
            arrays[f"s{k}_locations"] = np.ascontiguousarray(loc, np.uint64).reshape(-1)
            steps.append("prune")
        arrays[f"s{k}_input"] = ui.view(np.uint8)
        arrays[f"s{k}_blocks"] = pool.blocks.view(np.uint8).copy()
        arrays[f"s{k}_usages"] = pool.usages.view(np.uint8).copy()
        arrays[f"s{k}_request_count"] = np.array(pool.request_count)
    arrays["steps"] = np.array(steps)
    np.savez_compressed(path, **arrays)
    print(f"wrote {path}: {steps}")


if __name__ == "__main__":
    import _loader
    import oracle as orc

    svo = _loader.load()
    frame_fixture(svo, orc, svo.scenes.f0_scene(8), svo.scenes.f0_camera(64, 48, tick=1), OUT / "f0_frame.npz")
    frame_fixture(svo, orc, svo.scenes.config1_scene(), svo.scenes.config1_camera(64, 48, tick=1), OUT / "f1_frame.npz")
    pool_fixture(svo, orc, OUT / "pool_script.npz")
