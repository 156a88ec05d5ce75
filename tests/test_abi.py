"""The C ABI (include/svo.h): library loads, every declared entry point is exported,
ordinals and struct layouts match the reference, and host-side errors behave as
HandleResultCode expects (no GPU needed)."""
import ctypes as C
import json
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(svo_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported(svo):
    lib = svo.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_enum_ordinals_match_reference(svo):
    ref = json.loads((ROOT / "tests" / "golden" / "enums.json").read_text())
    for name, value in ref["KernelName"].items():
        assert svo.KernelName[name] == value
    for name, value in ref["BufferName"].items():
        assert svo.BufferName[name] == value
    header = (ROOT / "include" / "svo.h").read_text()
    assert "SVO_KERNEL_COUNT = 11" in header
    assert len(ref["BufferName"]) == 51  # SVO_BUF_REFERENCE_COUNT


def test_struct_layouts(svo):
    assert svo.BLOCK_DTYPE.itemsize == 16
    assert svo.BLOCK_DTYPE.fields["chunk"][1] == 4 and svo.BLOCK_DTYPE.fields["colour_r"][1] == 10
    assert svo.USAGE_DTYPE.itemsize == 8 and svo.USAGE_DTYPE.fields["parent"][1] == 4
    assert svo.CHILD_REQUEST_DTYPE.itemsize == 40 and svo.CHILD_REQUEST_DTYPE.fields["tree_size"][1] == 32
    assert svo.PRUNING_DTYPE.itemsize == 20 and svo.GRAFTING_DTYPE.itemsize == 16
    assert C.sizeof(svo.TraceInput) == 104 and svo.TraceInput.screen_size.offset == 64
    assert svo.TraceInput.ambient_light_level.offset == 96
    assert C.sizeof(svo.UpdateInputData) == 20 and svo.UpdateInputData.graft_size.offset == 16


def test_abi_version(svo):
    assert svo.lib().svo_abi_version() == 1


def test_errors_raise_with_method_and_code(svo):
    with pytest.raises(svo.SvoError) as e:
        svo.build_tree(svo.Geometry(kind=99), 4, 6)
    assert e.value.code == -1 and "BuildTree" in str(e.value)
    lib = svo.lib()
    assert lib.svo_set_arg_buffer(None, 0, b"bases", 0) == -1
    assert "null context" in svo.last_error()
    assert lib.svo_set_local_arg(None, 0, b"Bases", 0) == -1
    assert lib.svo_buffer_init(None, 0, 16, None, 0) == -1
    assert lib.svo_trace_frame(None, None, None) == -1
    assert lib.svo_prune(None, None, None, 0, None, 0, None, 0, None) == -1
    assert lib.svo_graft(None, None, None, 0, None, 0, None, 0, None) == -1
    assert lib.svo_collect(None, None, 0, None, None) == -1
    assert lib.svo_reclaim(None, None, None) == -1
    assert lib.svo_compact(None, None, None) == -1
    assert lib.svo_terrain_create(None, None, None) == -1
    assert lib.svo_stream_seed(None, None, 4) == -1
    assert lib.svo_stream_update(None, None, None, None) == -1
    assert lib.svo_terrain_destroy(None) == 0
    assert lib.svo_terrain_build_tree(None, None, 4, 8, 0, 0, 0, None, None) == -1
    assert lib.svo_tree_build_device(None, None, 4, 8, 0, 0, 0, None, None) == -1


def test_create_without_device_fails_cleanly(svo):
    if svo.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(svo.SvoError) as e:
        svo.ComputeManager(0)
    assert e.value.code == -2


def test_tree_build_and_oct_round_trip(svo, tmp_path):
    tree = svo.scenes.f0_scene(7).build()
    path = tmp_path / "f0.oct"
    tree.save(path)
    size = path.stat().st_size
    assert size == 1 + 4 + 2 * tree.base_count + 16 * tree.block_count
    back = svo.Octree.load(path)
    assert back.base_depth == tree.base_depth and back.block_count == tree.block_count
    assert np.array_equal(back.bases, tree.bases)
    assert np.array_equal(back.blocks, tree.blocks)
    u1, u2 = tree.initial_usages(5), back.initial_usages(5)
    assert np.array_equal(u1, u2)


def test_load_oct_rejects_truncated_file(svo, tmp_path):
    tree = svo.scenes.f0_scene(6).build()
    path = tmp_path / "t.oct"
    tree.save(path)
    data = path.read_bytes()
    path.write_bytes(data[: len(data) // 2])
    with pytest.raises(svo.SvoError) as e:
        svo.Octree.load(path)
    assert e.value.code == -11


def test_camera_lut_values(svo):
    cam = svo.scenes.f0_camera(8, 6)
    cu, su, cv, sv = svo.camera_lut(cam)
    for x in range(8):
        u = np.float32(cam.fov[0]) * ((np.float32(4.0) - np.float32(x)) / np.float32(8.0))
        assert cu[x] == np.float32(np.cos(np.float64(u))) and su[x] == np.float32(np.sin(np.float64(u)))
    assert cu[4] == 1.0 and su[4] == 0.0


def test_struct_layouts_match_the_c_compiler(svo, tmp_path):
    """Sizes and field offsets of the ABI structs as gcc lays them out from include/svo.h,
    against the Python binding's ctypes mirrors."""
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    checks = {
        "svo_trace_input": (svo.TraceInput, ["tick", "max_child_request_id", "ambient_light_level"]),
        "svo_update_input_data": (svo.UpdateInputData, ["tick", "offset", "graft_size"]),
        "svo_frame_options": (svo.FrameOptions, ["stack_levels", "work_chunk", "waves_per_simd"]),
        "svo_stream_options": (svo.StreamOptions, ["tick", "max_bricks", "prune_age", "gc_free_percent", "flags"]),
        "svo_stream_stats": (svo.StreamStats, ["unique", "moved", "free_groups", "pool_groups"]),
        "svo_composite_options": (svo.CompositeOptions, ["max_layers", "flags"]),
        "svo_adam_options": (svo.AdamOptions, ["epsilon", "step"]),
        "svo_tree_stats": (svo.TreeStats, ["requests", "unlinks"]),
        "svo_geometry": (svo.Geometry, ["seed", "resolution", "frequency", "z_max"]),
    }
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "svo.h"', "int main(void) {"]
    for st, (_, fields) in checks.items():
        lines.append(f'  printf("{st} size %zu\\n", sizeof({st}));')
        for f in fields:
            lines.append(f'  printf("{st} {f} %zu\\n", offsetof({st}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        st, key, val = line.split()
        got[(st, key)] = int(val)
    for st, (cls, fields) in checks.items():
        assert got[(st, "size")] == C.sizeof(cls), st
        for f in fields:
            assert got[(st, f)] == getattr(cls, f).offset, (st, f)
