"""GPU parity of svo_terrain_build_tree (TreeBuilder.BuildTree on the device) with the host
builder svo_tree_build on the same heightfield: Bases, every block at every address, the
block count (with its BuildBlocks counting rule), the group parents (through the initial
usages), a free-group queue that runs out inside the tree or at the dense level, and a
frame traced on the device-built tree against the oracle on the host-built one."""
from __future__ import annotations

import numpy as np
import pytest

from frame_helpers import first_mismatch
from stream_helpers import terrain_geometry

pytestmark = pytest.mark.gpu

BIG = 0xFFFFFFFF // 8


def _compare(svo, gpu, geom, B, D, max_size):
    host = svo.build_tree(geom, B, D, max_size)
    terrain = gpu.CreateTerrain(geom)
    try:
        count, dev = gpu.BuildTreeOnDevice(terrain, B, D, max_size, with_usages=True, usage_tick=3, host_copy=True)
    finally:
        terrain.close()
    assert count == host.block_count == dev.block_count
    assert np.array_equal(dev.bases, host.bases)
    m = first_mismatch(dev.blocks, host.blocks)
    assert m is None, f"blocks: {m}"
    hu, du = host.initial_usages(3), dev.initial_usages(3)
    assert np.array_equal(du, hu)
    bound = gpu.ReadBuffer(svo.BufferName.Usages, svo.USAGE_DTYPE)
    assert np.array_equal(bound, hu)
    assert np.array_equal(gpu.ReadBuffer(svo.BufferName.Blocks, svo.BLOCK_DTYPE), host.blocks)
    return host


@pytest.mark.parametrize("res,B,D", [(64, 2, 7), (256, 3, 9), (128, 2, 4), (128, 2, 3), (512, 4, 10)])
def test_device_build_matches_host_builder(svo, gpu, res, B, D):
    host = _compare(svo, gpu, terrain_geometry(svo, res), B, D, BIG)
    assert host.block_count > 0 and (host.allocated_groups > 0) == (D > B + 2)


def _device_or_error(svo, gpu, geom, B, D, max_size):
    terrain = gpu.CreateTerrain(geom)
    try:
        return gpu.BuildTreeOnDevice(terrain, B, D, max_size, host_copy=True)[1]
    except svo.SvoError:
        return None
    finally:
        terrain.close()


def test_device_build_with_a_budget(svo, gpu):
    """Budgets that end inside the tree (the blocks left are childless, as TreeBuilder leaves
    them) and budgets that end at a dense-level block (an error in both builders)."""
    geom = terrain_geometry(svo, 128)
    B, D = 2, 8
    full = svo.build_tree(geom, B, D, BIG)
    # the free-group queue: the empty level-(B+1) cells, then groups 8^(B+1)..maxSize/8-1
    lvl = full.bases[sum(8 ** i for i in range(1, B)):].astype(np.uint32)
    mixed = sum(int(np.count_nonzero(((lvl >> (2 * k)) & 3) == 3)) for k in range(8))
    n_empties = 8 ** (B + 1) - mixed
    inside = errors = 0
    for short in (1, 3, 10, 40, 200, full.allocated_groups // 3, full.allocated_groups - 2):
        queue = full.allocated_groups - short
        max_size = 8 * (8 ** (B + 1) + max(queue - n_empties, 0))
        try:
            host = svo.build_tree(geom, B, D, max_size)
        except svo.SvoError:
            host = None
        dev = _device_or_error(svo, gpu, geom, B, D, max_size)
        if host is None:
            assert dev is None, f"queue {queue}: the host builder fails, the device one does not"
            errors += 1
            continue
        assert dev is not None, f"queue {queue}: the device builder fails, the host one does not"
        assert dev.block_count == host.block_count, queue
        assert np.array_equal(dev.bases, host.bases)
        m = first_mismatch(dev.blocks, host.blocks)
        assert m is None, f"queue {queue}: {m}"
        assert np.array_equal(dev.initial_usages(1), host.initial_usages(1))
        inside += host.allocated_groups < full.allocated_groups
    assert inside > 0 and errors > 0


def test_frame_on_device_built_tree(svo, orc, gpu):
    """Trace a frame on the tree bound by the device build; the oracle traces the host
    tree."""
    geom = terrain_geometry(svo, 1024)
    B, D = 3, 10
    host = svo.build_tree(geom, B, D, BIG)
    terrain = gpu.CreateTerrain(geom)
    try:
        count, _ = gpu.BuildTreeOnDevice(terrain, B, D, BIG, with_usages=True, usage_tick=0)
    finally:
        terrain.close()
    assert count == host.block_count
    N = svo.BufferName
    cam = svo.scenes.config3_camera(40, 320, 180, base_depth=B)
    n = 320 * 180
    for name in (N.VisitHashes, N.RayStatus, N.Locations, N.Depths, N.BackgroundQueue, N.MaterialQueue,
                 N.Parents, N.ParentResidency):
        try:  # per-cone outputs an earlier test left bound, at its own size
            gpu.ReleaseBuffer(name)
        except svo.SvoError:
            pass
    for name, size in ((N.OutputImage, 16), (N.HitIndices, 4), (N.VisitCounts, 4)):
        gpu.InitDeviceBuffer(name, n, size)
    gpu.InitBuffer(N.ChildRequests, np.zeros(n * 40, np.uint8))
    gpu.InitBuffer(N.ChildRequestID, np.zeros(1, np.uint32))
    cam.tick = 1
    cam.max_child_request_id = n
    gpu.TraceFrame(cam)
    gpu.Wait()
    usages = host.initial_usages(0)
    ref = orc.trace_frame(cam, host.bases, host.blocks, usages, request_capacity=n, threads=8)
    for key, name, dt in (("hit", N.HitIndices, np.uint32), ("visits", N.VisitCounts, np.uint32)):
        got = gpu.ReadBuffer(name, dt, n)
        m = first_mismatch(got, ref[key])
        assert m is None, f"{key}: {m}"
    assert np.array_equal(gpu.ReadBuffer(N.Usages, svo.USAGE_DTYPE), usages)
    assert int(gpu.ReadBuffer(N.ChildRequestID, np.uint32, 1)[0]) == ref["request_count"]


def _cubes_scene(svo):
    import ctypes as C
    shapes = (C.c_float * 12)(0.1, 0.2, 0.15, 0.45, 0.6, 0.4, 0.5, 0.05, 0.3, 0.93, 0.41, 0.77)
    g = svo.Geometry()
    g.kind = svo.GEOMETRY_CUBES
    g.count = 2
    g.shapes = C.cast(shapes, C.POINTER(C.c_float))
    return g, shapes


@pytest.mark.parametrize("case", ["f0_spheres", "config1_noise", "noise_d9_b3", "cubes", "heightfield"])
def test_device_build_every_geometry(svo, gpu, case):
    """svo_tree_build_device == svo_tree_build for spheres, the sphere + noise SDF, boxes and a
    heightfield given as a geometry."""
    keep = None
    if case == "f0_spheres":
        spec = svo.scenes.f0_scene(8)
        geom, B, D, size = spec.geometry, spec.base_depth, spec.max_depth, spec.max_size
    elif case == "config1_noise":
        spec = svo.scenes.config1_scene()
        geom, B, D, size = spec.geometry, spec.base_depth, spec.max_depth, spec.max_size
    elif case == "noise_d9_b3":
        spec = svo.scenes.sphere_noise_scene(3, 9, base_depth=3)
        geom, B, D, size = spec.geometry, spec.base_depth, spec.max_depth, spec.max_size
    elif case == "cubes":
        geom, keep = _cubes_scene(svo)
        B, D, size = 3, 8, BIG
    else:
        geom, B, D, size = terrain_geometry(svo, 256), 3, 9, BIG
    host = svo.build_tree(geom, B, D, size)
    count, dev = gpu.BuildTreeDevice(geom, B, D, size, with_usages=True, usage_tick=2, host_copy=True)
    assert count == host.block_count == dev.block_count
    assert np.array_equal(dev.bases, host.bases)
    m = first_mismatch(dev.blocks, host.blocks)
    assert m is None, f"blocks: {m}"
    assert np.array_equal(gpu.ReadBuffer(svo.BufferName.Usages, svo.USAGE_DTYPE), host.initial_usages(2))
    assert host.allocated_groups > 0
    del keep
