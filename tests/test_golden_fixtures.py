"""The committed fixtures of tests/golden (made by tests/golden/make_frame_fixtures.py):
the oracle must still produce them (CPU), and the GPU must produce them (gpu)."""
from __future__ import annotations

import hashlib
from pathlib import Path

import numpy as np
import pytest

from frame_helpers import first_mismatch, gpu_frame, tree_arrays

GOLDEN = Path(__file__).resolve().parent / "golden"
FRAMES = {"f0": ("f0_frame.npz", lambda s: s.scenes.f0_scene(8), lambda s: s.scenes.f0_camera(64, 48, tick=1)),
          "f1": ("f1_frame.npz", lambda s: s.scenes.config1_scene(), lambda s: s.scenes.config1_camera(64, 48, tick=1))}


def _digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _frame_case(svo, name):
    fname, scene, camera = FRAMES[name]
    fx = np.load(GOLDEN / fname)
    tree, bases, blocks = tree_arrays(svo, scene(svo))
    assert _digest(bases, blocks) == str(fx["digest"]), "the scene builder no longer makes the fixture's tree"
    return fx, tree, bases, blocks, camera(svo)


def _check_frame(svo, got, fx, usages):
    for key in ("hit", "visits", "hash", "rgba", "status"):
        m = first_mismatch(got[key], fx[key])
        assert m is None, f"{key}: {m}"
    assert int(got["request_count"]) == int(fx["request_count"])
    req = got["requests"].view(np.uint8).reshape(-1)
    assert np.array_equal(req, fx["requests"].reshape(-1))
    assert hashlib.sha256(usages.tobytes()).hexdigest() == str(fx["usages_digest"])


@pytest.mark.parametrize("name", ["f0", "f1"])
def test_oracle_reproduces_frame_fixture(svo, orc, name):
    fx, tree, bases, blocks, cam = _frame_case(svo, name)
    usages = tree.initial_usages(tick=0)
    ref = orc.trace_frame(cam, bases, blocks, usages, request_capacity=256, threads=8)
    _check_frame(svo, ref, fx, usages)


def _pool_steps(fx):
    steps = [str(s) for s in fx["steps"]]
    for k, kind in enumerate(steps):
        yield k, kind


def test_oracle_reproduces_pool_fixture(svo, orc):
    fx = np.load(GOLDEN / "pool_script.npz")
    pool = orc.TreePool(fx["bases0"], fx["blocks0"].view(svo.BLOCK_DTYPE), fx["usages0"].view(svo.USAGE_DTYPE),
                        request_capacity=32)
    for k, kind in _pool_steps(fx):
        ui = fx[f"s{k}_input"].view(orc.UPDATE_INPUT_DTYPE)
        loc = fx[f"s{k}_locations"].reshape(-1, 3)
        if kind == "graft":
            pool.graft(ui, fx[f"s{k}_records"].view(svo.GRAFTING_DTYPE), fx[f"s{k}_side"].view(svo.BLOCK_DTYPE), loc)
        else:
            pool.prune(ui, fx[f"s{k}_records"].view(svo.PRUNING_DTYPE), fx[f"s{k}_side"].view(svo.BLOCK_DTYPE), loc)
        assert np.array_equal(pool.blocks.view(np.uint8), fx[f"s{k}_blocks"]), f"step {k} ({kind}): blocks"
        assert np.array_equal(pool.usages.view(np.uint8), fx[f"s{k}_usages"]), f"step {k} ({kind}): usages"
        assert pool.request_count == int(fx[f"s{k}_request_count"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["f0", "f1"])
def test_gpu_matches_frame_fixture(svo, gpu, name):
    fx, tree, bases, blocks, cam = _frame_case(svo, name)
    usages = tree.initial_usages(tick=0)
    got = gpu_frame(svo, gpu, bases, blocks, usages, cam, request_capacity=256, state=False, queues=False)
    _check_frame(svo, got, fx, got["usages"])


@pytest.mark.gpu
def test_gpu_matches_pool_fixture(svo, gpu):
    fx = np.load(GOLDEN / "pool_script.npz")
    N = svo.BufferName
    for name in (N.Parents, N.ParentResidency):
        try:
            gpu.ReleaseBuffer(name)
        except svo.SvoError:
            pass
    gpu.InitBuffer(N.Bases, fx["bases0"])
    gpu.InitBuffer(N.Blocks, fx["blocks0"])
    gpu.InitBuffer(N.Usages, fx["usages0"])
    gpu.InitBuffer(N.ChildRequests, np.zeros(32, dtype=svo.CHILD_REQUEST_DTYPE))
    gpu.InitBuffer(N.ChildRequestID, np.zeros(1, np.uint32))
    for k, kind in _pool_steps(fx):
        ui = fx[f"s{k}_input"]
        loc = fx[f"s{k}_locations"].reshape(-1, 3)
        side = fx[f"s{k}_side"].view(svo.BLOCK_DTYPE)
        if kind == "graft":
            gpu.Graft(ui, fx[f"s{k}_records"].view(svo.GRAFTING_DTYPE), side, loc)
        else:
            gpu.Prune(ui, fx[f"s{k}_records"].view(svo.PRUNING_DTYPE), side, loc)
        gpu.Wait()
        assert np.array_equal(gpu.ReadBuffer(N.Blocks, np.uint8), fx[f"s{k}_blocks"]), f"step {k} ({kind}): blocks"
        assert np.array_equal(gpu.ReadBuffer(N.Usages, np.uint8), fx[f"s{k}_usages"]), f"step {k} ({kind}): usages"
        assert int(gpu.ReadBuffer(N.ChildRequestID, np.uint32, 1)[0]) == int(fx[f"s{k}_request_count"])
