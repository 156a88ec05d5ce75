"""GPU parity of tree management (svo_prune / svo_graft / svo_collect and the Prune / Graft
entries of svo_enqueue) with the oracle: every pool array bit-exact after every call."""
from __future__ import annotations

import numpy as np
import pytest

from frame_helpers import tree_arrays
from tree_helpers import (TICK_MULTI, U32MAX, check_pool, pool_from_tree, random_graft, random_prune,
                          tiny_pool)

pytestmark = pytest.mark.gpu

CAP = 64


def _ui(orc, B, tick, offset=0, memory_size=0, cap=CAP):
    return orc.update_input(B, tick, offset=offset, memory_size=memory_size, max_child_request_id=cap)


class Mirror:
    """The same pool on the device (through the C ABI) and in the oracle."""

    def __init__(self, svo, orc, cm, bases, blocks, usages, parents=None, residency=None, cap=CAP, count=0):
        self.svo, self.cm = svo, cm
        B = svo.BufferName
        self.ref = orc.TreePool(bases, blocks, usages, parents=parents, residency=residency, request_capacity=cap,
                                request_count=count)
        cm.InitBuffer(B.Bases, bases)
        cm.InitBuffer(B.Blocks, blocks)
        cm.InitBuffer(B.Usages, usages)
        if parents is not None:
            cm.InitBuffer(B.Parents, parents)
            cm.InitBuffer(B.ParentResidency, self.ref.residency)
        else:
            for name in (B.Parents, B.ParentResidency):
                try:
                    cm.ReleaseBuffer(name)
                except svo.SvoError:
                    pass
        cm.InitBuffer(B.ChildRequests, np.zeros(cap, dtype=svo.CHILD_REQUEST_DTYPE))
        cm.InitBuffer(B.ChildRequestID, np.array([count], np.uint32))

    def device(self):
        B, cm, svo = self.svo.BufferName, self.cm, self.svo
        out = {
            "bases": cm.ReadBuffer(B.Bases, np.uint16),
            "blocks": cm.ReadBuffer(B.Blocks, svo.BLOCK_DTYPE),
            "usages": cm.ReadBuffer(B.Usages, svo.USAGE_DTYPE),
            "request_count": int(cm.ReadBuffer(B.ChildRequestID, np.uint32, 1)[0]),
            "requests": cm.ReadBuffer(B.ChildRequests, svo.CHILD_REQUEST_DTYPE),
        }
        if self.ref.parents is not None:
            out["parents"] = cm.ReadBuffer(B.Parents, svo.PARENT_DTYPE)
            out["residency"] = cm.ReadBuffer(B.ParentResidency, np.uint8)
        return out

    def compare(self, what: str):
        d, r = self.device(), self.ref
        assert d["request_count"] == r.request_count, f"{what}: request count {d['request_count']} != {r.request_count}"
        n = min(r.request_count, r.request_capacity)
        assert d["requests"][:n].tobytes() == r.requests[:n].tobytes(), f"{what}: child requests differ"
        assert np.array_equal(d["bases"], r.bases), f"{what}: bases differ"
        for name in ("usages", "blocks"):
            a, b = d[name], getattr(r, name)
            if a.tobytes() != b.tobytes():
                bad = np.nonzero(a.view(np.uint8).reshape(a.size, -1).any(axis=1) !=
                                 b.view(np.uint8).reshape(b.size, -1).any(axis=1))[0]
                diff = np.nonzero((a.view(np.uint8).reshape(a.size, -1) != b.view(np.uint8).reshape(b.size, -1)).any(axis=1))[0]
                i = int(diff[0])
                raise AssertionError(f"{what}: {name} differ at {diff.size} entries, first {i}: gpu {a[i]} oracle {b[i]}")
        if r.parents is not None:
            assert d["parents"].tobytes() == r.parents.tobytes(), f"{what}: parents differ"
            assert np.array_equal(d["residency"], r.residency), f"{what}: residency differs"

    def prune(self, ui, rec, surface, loc):
        got = self.cm.Prune(ui, rec, surface, loc)
        ref = self.ref.prune(ui, rec, surface, loc)
        return got, ref

    def graft(self, ui, rec, gb, loc):
        got = self.cm.Graft(ui, rec, gb, loc)
        ref = self.ref.graft(ui, rec, gb, loc)
        return got, ref


def _stats_equal(got, ref):
    for k, v in ref.items():
        assert got[k] == v, f"stat {k}: gpu {got[k]} oracle {v}"


def test_collect_known_answer_gpu(svo, orc, gpu):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=9)
    t = usages["tick"]
    t[64], t[65], t[66], t[67], t[68], t[69], t[70], t[71], t[72] = 0, 30, 60, 99, 65533, 65535, 65534, 100, 95
    usages["parent"][65] = 200
    blocks["child"][200] = 8 * 65
    blocks["child"][8 * 65 + 3] = 8 * 72
    usages["parent"][72] = 8 * 65 + 3
    m = Mirror(svo, orc, gpu, bases, blocks, usages)
    got = gpu.Collect(_ui(orc, B, 100, offset=64), 6)
    ref = m.ref.collect(_ui(orc, B, 100, offset=64), 6)
    assert got.tolist() == ref.tolist() == [64, 65, 68, 66, 72, 67]
    m.compare("collect")


def test_graft_known_answer_gpu(svo, orc, gpu):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=16)
    blocks["child"][100] = 8 * 70
    usages["tick"][70], usages["parent"][70] = 5, 100
    blocks["child"][8 * 70 + 1] = 8 * 71
    usages["tick"][71], usages["parent"][71] = 5, 8 * 70 + 1
    gb = np.zeros(24, dtype=svo.BLOCK_DTYPE)
    gb["child"] = U32MAX
    gb["colour_r"] = np.arange(24)
    gb["child"][2], gb["child"][5] = 16, 8
    rec = np.zeros(3, dtype=svo.GRAFTING_DTYPE)
    rec[0] = (0, 24, 0xFF, (0, 0, 0), 100)
    rec[1] = (0, 12, 0xFF, (0, 0, 0), 101)
    rec[2] = (0, 24, 0xFF, (0, 0, 0), 100)
    m = Mirror(svo, orc, gpu, bases, blocks, usages)
    got, ref = m.graft(_ui(orc, B, 9), rec, gb, np.zeros((0, 3), np.uint64))
    _stats_equal(got, ref)
    m.compare("graft")
    assert m.device()["blocks"]["child"][100] == 8 * 67


def test_prune_fields_and_unlink_gpu(svo, orc, gpu):
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
    surface = np.zeros(2, dtype=svo.BLOCK_DTYPE)
    surface["colour_r"] = [11, 22]
    surface["child"] = [7, 8]
    rec = np.zeros(8, dtype=svo.PRUNING_DTYPE)
    rec[0] = (1, 0xFF, 0, 302, 0, 0, 0, U32MAX)
    rec[1] = (1, 0xFF, 0, 300, 0, 0, 0, U32MAX)
    rec[2] = (0x10, 0, 0, 2, 0x1234, 0, 0, 0)
    rec[3] = (0x10, 0, 0, 2, 0x4321, 0, 0, 0)
    rec[4] = (0x08, 0xFF, 0, 8 * 65 + 1, 0x00AA, 0, 0, 0)
    rec[5] = (0x08, 0xFF, 0, 8 * 65 + 1, 0x00BB, 0, 1, 0)
    rec[6] = (0x06, 0xFF, 0, 8 * 65 + 2, 0, 0, 0, 0)
    rec[7] = (0x08, 0xFF, 0, 8 * 65 + 3, 0, 0, 9, 0)
    m = Mirror(svo, orc, gpu, bases, blocks, usages, parents=parents, residency=residency)
    got, ref = m.prune(_ui(orc, B, 9), rec, surface, np.zeros((0, 3), np.uint64))
    _stats_equal(got, ref)
    assert ref["unlinks"] == 2 and ref["invalid"] == 1
    m.compare("prune")


def test_prune_requests_and_capacity_gpu(svo, orc, gpu):
    B = 1
    bases, blocks, usages = tiny_pool(svo, B, extra_groups=8)
    rng = np.random.default_rng(2)
    loc = rng.integers(0, 2 ** 63, size=(9, 3), dtype=np.uint64)
    rec = np.zeros(9, dtype=svo.PRUNING_DTYPE)
    for i in range(9):
        rec[i] = (0, 3 + (i % 5), 0, i, 0, 0, 0, 0)
    m = Mirror(svo, orc, gpu, bases, blocks, usages, cap=4, count=1)
    got, ref = m.prune(_ui(orc, B, 2, cap=4), rec, np.zeros(0, dtype=svo.BLOCK_DTYPE), loc)
    _stats_equal(got, ref)
    m.compare("prune requests")


@pytest.fixture(scope="module")
def noise_tree(svo):
    spec = svo.scenes.sphere_noise_scene(3, 7, base_depth=2)
    return tree_arrays(svo, spec)


@pytest.mark.parametrize("seed", [23, 24, 25])
def test_random_sequences_gpu(svo, orc, gpu, noise_tree, seed):
    tree, bases0, blocks0 = noise_tree
    bases, blocks, usages = pool_from_tree(svo, tree, bases0, blocks0, extra_groups=256, tick=1)
    B = tree.base_depth
    m = Mirror(svo, orc, gpu, bases, blocks, usages)
    rng = np.random.default_rng(seed)
    for frame in range(2, 12):
        rec, surface, loc = random_prune(svo, rng, m.ref.blocks, m.ref.usages, B, 40, max_depth=tree.max_depth + 1)
        rec["child_address"] = U32MAX
        got, ref = m.prune(_ui(orc, B, frame, offset=4), rec, surface, loc)
        _stats_equal(got, ref)
        m.compare(f"seed {seed} frame {frame} prune")
        grec, gb, gloc = random_graft(svo, rng, m.ref.blocks, m.ref.usages, B, 8, max_depth=tree.max_depth + 1)
        got, ref = m.graft(_ui(orc, B, frame, offset=4), grec, gb, gloc)
        _stats_equal(got, ref)
        m.compare(f"seed {seed} frame {frame} graft")
    check_pool(m.ref.blocks, m.ref.usages, B)


def test_enqueue_prune_and_graft(svo, orc, gpu, noise_tree):
    """KernelName.Prune / KernelName.Graft through SetArg + Enqueue (the reference's
    MainWindow.RunPrune / RunGraft call sequence, MainWindow.cs:281-334, with the
    argument bindings by role)."""
    tree, bases0, blocks0 = noise_tree
    bases, blocks, usages = pool_from_tree(svo, tree, bases0, blocks0, extra_groups=128, tick=1)
    B, K, N = tree.base_depth, svo.KernelName, svo.BufferName
    m = Mirror(svo, orc, gpu, bases, blocks, usages)
    rng = np.random.default_rng(7)
    rec, surface, loc = random_prune(svo, rng, m.ref.blocks, m.ref.usages, B, 24, max_depth=tree.max_depth + 1)
    rec["child_address"] = U32MAX
    ui = _ui(orc, B, 3, offset=1)
    gpu.InitBuffer(N.Pruning, rec)
    gpu.InitBuffer(N.PruningBlockData, surface)
    gpu.InitBuffer(N.PruningAddresses, loc.reshape(-1) if loc.size else np.zeros(3, np.uint64))
    for param, name in (("bases", N.Bases), ("blocks", N.Blocks), ("usage", N.Usages),
                        ("childRequestId", N.ChildRequestID), ("childRequests", N.ChildRequests),
                        ("pruning", N.Pruning), ("pruningBlockData", N.PruningBlockData),
                        ("pruningAddresses", N.PruningAddresses)):
        gpu.SetArg(K.Prune, param, name)
    gpu.SetArg(K.Prune, "inputData", ui)
    gpu.Enqueue(K.Prune, [rec.size])
    gpu.Wait()
    m.ref.prune(ui, rec, surface, loc)
    m.compare("enqueue prune")

    grec, gb, gloc = random_graft(svo, rng, m.ref.blocks, m.ref.usages, B, 6, max_depth=tree.max_depth + 1)
    gpu.InitBuffer(N.Grafting, grec)
    gpu.InitBuffer(N.GraftingBlocks, gb)
    gpu.InitBuffer(N.GraftingAddresses, gloc.reshape(-1) if gloc.size else np.zeros(3, np.uint64))
    for param, name in (("blocks", N.Blocks), ("usage", N.Usages), ("childRequestId", N.ChildRequestID),
                        ("childRequests", N.ChildRequests), ("grafting", N.Grafting),
                        ("graftingBlocks", N.GraftingBlocks), ("graftingAddresses", N.GraftingAddresses)):
        gpu.SetArg(K.Graft, param, name)
    gpu.SetArg(K.Graft, "inputData", ui)
    gpu.Enqueue(K.Graft, [grec.size])
    gpu.Wait()
    m.ref.graft(ui, grec, gb, gloc)
    m.compare("enqueue graft")


def test_trace_after_management_matches(svo, orc, gpu, noise_tree):
    """A frame traced on the managed pool equals the oracle's frame on the oracle's pool."""
    from frame_helpers import assert_frame_matches, gpu_frame

    tree, bases0, blocks0 = noise_tree
    bases, blocks, usages = pool_from_tree(svo, tree, bases0, blocks0, extra_groups=256, tick=1)
    B = tree.base_depth
    m = Mirror(svo, orc, gpu, bases, blocks, usages)
    rng = np.random.default_rng(31)
    for frame in range(2, 6):
        rec, surface, loc = random_prune(svo, rng, m.ref.blocks, m.ref.usages, B, 60, max_depth=tree.max_depth + 1)
        rec["child_address"] = U32MAX
        m.prune(_ui(orc, B, frame, offset=2), rec, surface, loc)
        grec, gb, gloc = random_graft(svo, rng, m.ref.blocks, m.ref.usages, B, 12, max_depth=tree.max_depth + 1)
        m.graft(_ui(orc, B, frame, offset=2), grec, gb, gloc)
    m.compare("managed pool")
    cam = svo.scenes.config1_camera(96, 72)
    cam.base_depth = B
    cam.tick = 9
    pool = m.ref
    ref_usages = pool.usages.copy()
    ref = orc.trace_frame(cam, pool.bases, pool.blocks, ref_usages, request_capacity=6000)
    got = gpu_frame(svo, gpu, pool.bases, pool.blocks, pool.usages.copy(), cam, state=False, queues=False)
    assert int(np.count_nonzero(got["hit"] != U32MAX)) > 0
    assert_frame_matches(got, ref, usages_ref=ref_usages)
