"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Optional

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
# SVO_ORACLE_LIB names another build of the oracle (the sanitized one of `make asan`)
LIB_PATH = Path(os.environ["SVO_ORACLE_LIB"]) if os.environ.get("SVO_ORACLE_LIB") else ORACLE_DIR / "liboracle.so"

U64MAX = (1 << 64) - 1
U32MAX = (1 << 32) - 1


class FrameOut(C.Structure):
    _fields_ = [
        ("hit", C.c_void_p),
        ("visits", C.c_void_p),
        ("hash", C.c_void_p),
        ("rgba", C.c_void_p),
        ("status", C.c_void_p),
        ("depth", C.c_void_p),
        ("location", C.c_void_p),
        ("visit_log", C.c_void_p),
        ("log_stride", C.c_uint32),
        ("base_steps", C.c_uint64),
        ("block_steps", C.c_uint64),
        ("groups_touched", C.c_uint64),
    ]


class Pool(C.Structure):
    _fields_ = [
        ("bases", C.c_void_p), ("n_bases", C.c_uint32),
        ("blocks", C.c_void_p), ("n_blocks", C.c_uint32),
        ("usages", C.c_void_p), ("n_usages", C.c_uint32),
        ("parents", C.c_void_p), ("n_parents", C.c_uint32),
        ("residency", C.c_void_p),
        ("requests", C.c_void_p), ("request_capacity", C.c_uint32),
        ("request_count", C.POINTER(C.c_uint32)),
    ]


class MaterialIO(C.Structure):
    _fields_ = [
        ("queue", C.c_void_p), ("n", C.c_uint32),
        ("root_directions", C.c_void_p), ("root_locations", C.c_void_p),
        ("root_depths", C.c_void_p), ("root_weightings", C.c_void_p), ("root_parents", C.c_void_p),
        ("n_root", C.c_uint32),
        ("blocks", C.c_void_p), ("n_blocks", C.c_uint32),
        ("usages", C.c_void_p), ("n_usages", C.c_uint32),
        ("acc_counter", C.c_void_p),
        ("origins", C.c_void_p), ("directions", C.c_void_p), ("fovs", C.c_void_p), ("locations", C.c_void_p),
        ("depths", C.c_void_p), ("weightings", C.c_void_p), ("colour_r", C.c_void_p), ("colour_g", C.c_void_p),
        ("colour_b", C.c_void_p), ("luminosities", C.c_void_p), ("parents", C.c_void_p),
        ("n_out", C.c_uint32),
        ("final_w", C.c_void_p), ("n_final", C.c_uint32),
        ("bg_queue", C.c_void_p), ("bg_counter", C.c_void_p), ("bg_capacity", C.c_uint32),
    ]


class ShadeIO(C.Structure):
    _fields_ = [
        ("queue", C.c_void_p), ("n", C.c_uint32), ("n_rays", C.c_uint32),
        ("directions", C.c_void_p), ("parents", C.c_void_p), ("weightings", C.c_void_p),
        ("colour_r", C.c_void_p), ("colour_g", C.c_void_p), ("colour_b", C.c_void_p),
        ("locations", C.c_void_p), ("depths", C.c_void_p),
        ("final_r", C.c_void_p), ("final_g", C.c_void_p), ("final_b", C.c_void_p), ("final_w", C.c_void_p),
        ("n_pixels", C.c_uint32),
    ]


class TreeStats(C.Structure):
    _fields_ = [(k, C.c_uint32) for k in ("applied", "skipped", "invalid", "requests", "collected", "placed", "unlinks")]


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        L.orc_float_to_ulong.argtypes = [C.c_float]
        L.orc_float_to_ulong.restype = C.c_uint64
        L.orc_ulong_to_float.argtypes = [C.c_uint64]
        L.orc_ulong_to_float.restype = C.c_float
        L.orc_float_to_ulong_n.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.orc_float_to_ulong_n.restype = None
        L.orc_ulong_to_float_n.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.orc_ulong_to_float_n.restype = None
        L.orc_round_ulong.argtypes = [C.c_uint64, C.c_uint32, C.c_int]
        L.orc_round_ulong.restype = C.c_uint64
        L.orc_pow_sum.argtypes = [C.c_uint32]
        L.orc_pow_sum.restype = C.c_uint32
        L.orc_chunk_position.argtypes = [C.c_uint32, C.POINTER(C.c_uint64)]
        L.orc_chunk_position.restype = C.c_uint32
        L.orc_base_location.argtypes = [C.c_uint32, C.POINTER(C.c_uint64)]
        L.orc_base_location.restype = C.c_uint32
        L.orc_compare_positions.argtypes = [C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint8)]
        L.orc_compare_positions.restype = C.c_uint32
        L.orc_traverse_chunk.argtypes = [C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_uint8)]
        L.orc_traverse_chunk.restype = None
        L.orc_init_ray.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_uint64), C.POINTER(C.c_float)]
        L.orc_init_ray.restype = C.c_int
        L.orc_camera_lut.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_camera_lut.restype = None
        L.orc_trace_frame.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                      C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(FrameOut), C.c_int]
        L.orc_trace_frame.restype = C.c_int
        L.orc_trace_rays.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                     C.c_void_p, C.c_void_p, C.c_int]
        L.orc_trace_rays.restype = C.c_int
        L.orc_wavefront_frame.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                          C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.c_void_p, C.POINTER(C.c_uint32),
                                          C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(FrameOut)]
        L.orc_wavefront_frame.restype = C.c_int
        L.orc_trace_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                      C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32),
                                      C.POINTER(FrameOut), C.c_int]
        L.orc_trace_batch.restype = C.c_int
        L.orc_composite_layers.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                           C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_composite_layers.restype = C.c_int
        L.orc_composite_loss.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                         C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_composite_loss.restype = C.c_double
        L.orc_prune.argtypes = [C.POINTER(Pool), C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                C.c_uint32, C.POINTER(TreeStats)]
        L.orc_prune.restype = C.c_int
        L.orc_collect.argtypes = [C.POINTER(Pool), C.c_void_p, C.c_uint32, C.c_void_p]
        L.orc_collect.restype = C.c_uint32
        L.orc_graft.argtypes = [C.POINTER(Pool), C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                C.c_uint32, C.POINTER(TreeStats)]
        L.orc_graft.restype = C.c_int
        L.orc_reclaim.argtypes = [C.POINTER(Pool), C.c_void_p]
        L.orc_reclaim.restype = C.c_uint32
        L.orc_compact.argtypes = [C.POINTER(Pool), C.c_void_p]
        L.orc_compact.restype = C.c_uint32
        L.orc_evaluate_material.argtypes = [C.c_void_p, C.POINTER(MaterialIO)]
        L.orc_evaluate_material.restype = C.c_int
        L.orc_sincos.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_sincos.restype = None
        L.orc_evaluate_background.argtypes = [C.c_uint8, C.POINTER(ShadeIO)]
        L.orc_evaluate_background.restype = C.c_int
        L.orc_resolve_remainders.argtypes = [C.c_uint32, C.c_void_p, C.c_uint32, C.POINTER(ShadeIO)]
        L.orc_resolve_remainders.restype = C.c_int
        L.orc_draw_trace.argtypes = [C.c_uint32] + [C.c_void_p] * 5
        L.orc_draw_trace.restype = None
        L.orc_tick_age.argtypes = [C.c_uint16, C.c_uint16]
        L.orc_tick_age.restype = C.c_uint32
        _lib = L
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


def _u64x3(v):
    return (C.c_uint64 * 3)(*[int(x) & U64MAX for x in v])


def float_to_ulong(x: float) -> int:
    return int(lib().orc_float_to_ulong(x))


def float_to_ulong_n(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.size, np.uint64)
    lib().orc_float_to_ulong_n(x.ctypes.data, out.ctypes.data, x.size)
    return out


def ulong_to_float_n(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.uint64)
    out = np.empty(x.size, np.float32)
    lib().orc_ulong_to_float_n(x.ctypes.data, out.ctypes.data, x.size)
    return out


def ulong_to_float(x: int) -> float:
    return float(lib().orc_ulong_to_float(x))


def round_ulong(value: int, depth: int, up: bool) -> int:
    return int(lib().orc_round_ulong(value, depth, 1 if up else 0))


def pow_sum(depth: int) -> int:
    return int(lib().orc_pow_sum(depth))


def chunk_position(depth: int, loc) -> int:
    return int(lib().orc_chunk_position(depth, _u64x3(loc)))


def base_location(depth: int, loc) -> int:
    return int(lib().orc_base_location(depth, _u64x3(loc)))


def compare_positions(depth: int, prev, loc, sign) -> int:
    return int(lib().orc_compare_positions(depth, _u64x3(prev), _u64x3(loc), (C.c_uint8 * 3)(*sign)))


def traverse_chunk(depth: int, loc, direction, inv, sign):
    l = _u64x3(loc)
    lib().orc_traverse_chunk(depth, l, (C.c_float * 3)(*direction), (C.c_float * 3)(*inv), (C.c_uint8 * 3)(*sign))
    return [int(v) for v in l]


def init_ray(trace_input, x: int, y: int):
    d = (C.c_float * 3)()
    inv = (C.c_float * 3)()
    loc = (C.c_uint64 * 3)()
    fov = C.c_float()
    hit = lib().orc_init_ray(C.byref(trace_input), x, y, d, inv, loc, C.byref(fov))
    return bool(hit), list(d), list(inv), [int(v) for v in loc], fov.value


def camera_lut(trace_input):
    w, h = trace_input.screen_size[0], trace_input.screen_size[1]
    cu, su = np.empty(w, np.float32), np.empty(w, np.float32)
    cv, sv = np.empty(h, np.float32), np.empty(h, np.float32)
    lib().orc_camera_lut(C.byref(trace_input), cu.ctypes.data, su.ctypes.data, cv.ctypes.data, sv.ctypes.data)
    return cu, su, cv, sv


def _outputs(n: int, log_stride: int):
    o = {
        "hit": np.zeros(n, np.uint32),
        "visits": np.zeros(n, np.uint32),
        "hash": np.zeros(n, np.uint64),
        "rgba": np.zeros((n, 4), np.float32),
        "status": np.zeros(n, np.uint8),
        "depth": np.zeros(n, np.uint8),
        "location": np.zeros((n, 3), np.uint64),
    }
    if log_stride:
        o["visit_log"] = np.zeros((n, log_stride), np.uint64)
    fo = FrameOut()
    for k in ("hit", "visits", "hash", "rgba", "status", "depth", "location"):
        setattr(fo, k, o[k].ctypes.data)
    fo.visit_log = o["visit_log"].ctypes.data if log_stride else None
    fo.log_stride = log_stride
    return o, fo


def trace_frame(trace_input, bases: np.ndarray, blocks: np.ndarray, usages: Optional[np.ndarray] = None,
                request_capacity: int = 0, threads: int = 1, log_stride: int = 0) -> dict:
    """Per-cone tracing of one primary frame.  `usages` (if given) is updated in place."""
    n = trace_input.screen_size[0] * trace_input.screen_size[1]
    o, fo = _outputs(n, log_stride)
    req = np.zeros(max(request_capacity, 1), dtype=_request_dtype())
    count = C.c_uint32(0)
    rc = lib().orc_trace_frame(C.byref(trace_input), bases.ctypes.data, bases.size, blocks.ctypes.data, blocks.size,
                               _ptr(usages), 0 if usages is None else usages.size, req.ctypes.data, request_capacity,
                               C.byref(count), C.byref(fo), threads)
    if rc != 0:
        raise RuntimeError(f"orc_trace_frame failed: {rc}")
    o["requests"] = req[: min(count.value, request_capacity)]
    o["request_count"] = count.value
    o["base_steps"] = fo.base_steps
    o["block_steps"] = fo.block_steps
    o["groups_touched"] = fo.groups_touched
    return o


def trace_batch(trace_inputs, bases: np.ndarray, blocks: np.ndarray, usages: Optional[np.ndarray] = None,
                apertures=None, request_capacity: int = 0, threads: int = 1) -> dict:
    """Several cameras in one frame: per-cone outputs concatenated in camera order."""
    n_cams = len(trace_inputs)
    n = sum(t.screen_size[0] * t.screen_size[1] for t in trace_inputs)
    o, fo = _outputs(n, 0)
    ins = (type(trace_inputs[0]) * n_cams)(*trace_inputs)
    ap = None if apertures is None else np.ascontiguousarray(apertures, dtype=np.float32)
    req = np.zeros(max(request_capacity, 1), dtype=_request_dtype())
    count = C.c_uint32(0)
    rc = lib().orc_trace_batch(C.cast(ins, C.c_void_p), None if ap is None else ap.ctypes.data, n_cams,
                               bases.ctypes.data, bases.size, blocks.ctypes.data, blocks.size, _ptr(usages),
                               0 if usages is None else usages.size, req.ctypes.data, request_capacity, C.byref(count),
                               C.byref(fo), threads)
    if rc != 0:
        raise RuntimeError(f"orc_trace_batch failed: {rc}")
    o["requests"] = req[: min(count.value, request_capacity)]
    o["request_count"] = count.value
    o["base_steps"] = fo.base_steps
    o["block_steps"] = fo.block_steps
    o["groups_touched"] = fo.groups_touched
    return o


def trace_rays(trace_input, bases: np.ndarray, blocks: np.ndarray, ray_ids: np.ndarray, threads: int = 1):
    ids = np.ascontiguousarray(ray_ids, dtype=np.uint32)
    hit = np.zeros(ids.size, np.uint32)
    visits = np.zeros(ids.size, np.uint32)
    rc = lib().orc_trace_rays(C.byref(trace_input), bases.ctypes.data, bases.size, blocks.ctypes.data, blocks.size,
                              ids.ctypes.data, ids.size, hit.ctypes.data, visits.ctypes.data, threads)
    if rc != 0:
        raise RuntimeError(f"orc_trace_rays failed: {rc}")
    return hit, visits


def wavefront_frame(trace_input, bases: np.ndarray, blocks: np.ndarray, usages: Optional[np.ndarray] = None,
                    request_capacity: int = 0, log_stride: int = 0) -> dict:
    """Exact emulation of MainWindow.RunWavefront with order-preserving queues."""
    n = trace_input.screen_size[0] * trace_input.screen_size[1]
    o, fo = _outputs(n, log_stride)
    req = np.zeros(max(request_capacity, 1), dtype=_request_dtype())
    bgq = np.zeros(max(n, 1), np.uint32)
    matq = np.zeros(max(n, 1), np.uint32)
    count, nbg, nmat, iters = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
    rc = lib().orc_wavefront_frame(C.byref(trace_input), bases.ctypes.data, bases.size, blocks.ctypes.data, blocks.size,
                                   _ptr(usages), 0 if usages is None else usages.size, req.ctypes.data, request_capacity,
                                   C.byref(count), bgq.ctypes.data, C.byref(nbg), matq.ctypes.data, C.byref(nmat),
                                   C.byref(iters), C.byref(fo))
    if rc != 0:
        raise RuntimeError(f"orc_wavefront_frame failed: {rc}")
    o["requests"] = req[: min(count.value, request_capacity)]
    o["request_count"] = count.value
    o["background_queue"] = bgq[: nbg.value]
    o["material_queue"] = matq[: nmat.value]
    o["iterations"] = iters.value
    o["base_steps"] = fo.base_steps
    o["block_steps"] = fo.block_steps
    o["groups_touched"] = fo.groups_touched
    return o


def _request_dtype():
    return np.dtype([("address", "<u4"), ("tick", "<u2"), ("depth", "u1"), ("pad0", "u1"), ("location", "<u8", (3,)),
                     ("tree_size", "u1"), ("pad1", "u1", (7,))])


# ---------------------------------------------------------------------------
# tree management (prune / graft / collect)
# ---------------------------------------------------------------------------

UPDATE_INPUT_DTYPE = np.dtype([("base_depth", "u1"), ("pad0", "u1"), ("tick", "<u2"), ("max_child_request_id", "<u4"),
                               ("memory_size", "<u4"), ("offset", "<u4"), ("graft_size", "<u4")])
assert UPDATE_INPUT_DTYPE.itemsize == 20


def update_input(base_depth: int, tick: int, offset: int = 0, memory_size: int = 0, max_child_request_id: int = 0,
                 graft_size: int = 0) -> np.ndarray:
    u = np.zeros(1, UPDATE_INPUT_DTYPE)
    u["base_depth"], u["tick"], u["offset"] = base_depth, tick, offset
    u["memory_size"], u["max_child_request_id"], u["graft_size"] = memory_size, max_child_request_id, graft_size
    return u


class TreePool:
    """Mutable copies of the arrays a management dispatch works on (all numpy, in place)."""

    def __init__(self, bases, blocks, usages, parents=None, residency=None, request_capacity: int = 0,
                 request_count: int = 0):
        self.bases = np.ascontiguousarray(bases).copy()
        self.blocks = np.ascontiguousarray(blocks).copy()
        self.usages = np.ascontiguousarray(usages).copy()
        self.parents = None if parents is None else np.ascontiguousarray(parents).copy()
        n_par = 0 if self.parents is None else self.parents.size
        self.residency = np.zeros(n_par, np.uint8) if residency is None else np.ascontiguousarray(residency).copy()
        self.requests = np.zeros(max(request_capacity, 1), dtype=_request_dtype())
        self.request_capacity = request_capacity
        self.count = C.c_uint32(request_count)

    def _pool(self) -> Pool:
        p = Pool()
        p.bases, p.n_bases = self.bases.ctypes.data, self.bases.size
        p.blocks, p.n_blocks = self.blocks.ctypes.data, self.blocks.size
        p.usages, p.n_usages = self.usages.ctypes.data, self.usages.size
        if self.parents is not None and self.parents.size:
            p.parents, p.n_parents = self.parents.ctypes.data, self.parents.size
            p.residency = self.residency.ctypes.data
        p.requests, p.request_capacity = self.requests.ctypes.data, self.request_capacity
        p.request_count = C.pointer(self.count)
        return p

    @property
    def request_count(self) -> int:
        return self.count.value

    def prune(self, inp: np.ndarray, records: np.ndarray, surface: np.ndarray, locations: np.ndarray) -> dict:
        st = TreeStats()
        loc = np.ascontiguousarray(locations, np.uint64).reshape(-1)
        rc = lib().orc_prune(C.byref(self._pool()), inp.ctypes.data, records.ctypes.data, records.size,
                             _ptr(surface) if surface.size else None, surface.size, loc.ctypes.data if loc.size else None,
                             loc.size // 3, C.byref(st))
        if rc != 0:
            raise RuntimeError(f"orc_prune failed: {rc}")
        return {k: getattr(st, k) for k, _ in TreeStats._fields_}

    def collect(self, inp: np.ndarray, need: int) -> np.ndarray:
        out = np.zeros(max(need, 1), np.uint32)
        n = lib().orc_collect(C.byref(self._pool()), inp.ctypes.data, need, out.ctypes.data)
        return out[:n]

    def graft(self, inp: np.ndarray, records: np.ndarray, graft_blocks: np.ndarray, locations: np.ndarray) -> dict:
        st = TreeStats()
        loc = np.ascontiguousarray(locations, np.uint64).reshape(-1)
        rc = lib().orc_graft(C.byref(self._pool()), inp.ctypes.data, records.ctypes.data, records.size,
                             _ptr(graft_blocks) if graft_blocks.size else None, graft_blocks.size,
                             loc.ctypes.data if loc.size else None, loc.size // 3, C.byref(st))
        if rc != 0:
            raise RuntimeError(f"orc_graft failed: {rc}")
        return {k: getattr(st, k) for k, _ in TreeStats._fields_}


    def reclaim(self, inp: np.ndarray) -> int:
        """Frees every stale live group (age > offset); returns how many."""
        return int(lib().orc_reclaim(C.byref(self._pool()), inp.ctypes.data))

    def compact(self, inp: np.ndarray) -> int:
        """Moves live groups at or above the live count into the holes below it; returns how many."""
        return int(lib().orc_compact(C.byref(self._pool()), inp.ctypes.data))


# ---------------------------------------------------------------------------
# streaming: child requests -> bricks -> graft, then garbage collection
# ---------------------------------------------------------------------------

GRAFTING_DTYPE = np.dtype([("graft_data_address", "<u4"), ("graft_total_size", "<u4"), ("depth", "u1"),
                           ("pad", "u1", (3,)), ("graft_address", "<u4")])
assert GRAFTING_DTYPE.itemsize == 16


def stream_requests(requests: np.ndarray, count: int, capacity: int, blocks: np.ndarray, max_depth: int) -> np.ndarray:
    """Indices of the requests a streaming step turns into bricks, in request order: among the
    requests asking for a level (TreeSize >= 1; requestChild, wavefront.cl:362-375, also
    records cones that need nothing below the block, with TreeSize 0), the first request of
    every block address whose block is in range, childless, at depth 1 <= d < max_depth and
    able to have children (an octant whose air and geometry bits are both set,
    CanHaveChildren, TreeBuilder.cs:291-302)."""
    n = min(int(count), int(capacity), requests.size)
    seen = set()
    keep = []
    child, chunk = blocks["child"], blocks["chunk"]
    for i in range(n):
        a = int(requests["address"][i])
        if int(requests["tree_size"][i]) == 0 or a >= blocks.size or a in seen:
            continue
        seen.add(a)
        d = int(requests["depth"][i])
        c = int(chunk[a])
        mixed = any((c >> (2 * k)) & 3 == 3 for k in range(8))
        if 1 <= d < max_depth and int(child[a]) == U32MAX and mixed:
            keep.append(i)
    return np.asarray(keep, np.int64)


def stream_update(pool: "TreePool", requests: np.ndarray, count: int, capacity: int, base_depth: int,
                  max_depth: int, tick: int, bricks_for, max_bricks: int = 65536, prune_age: int = 30,
                  gc_free_percent: int = 10, flags: int = 0) -> dict:
    """One streaming step on the oracle pool: the kept requests (stream_requests), at most
    max_bricks of them, become graft records k = (8k, 8, depth 0xFF, block address) whose
    bricks `bricks_for(kept_requests)` returns (8 blocks each, the block's children); they
    are grafted (collection age threshold prune_age); then, when free groups are fewer than
    gc_free_percent % of the pool (or flags & 2, and never with flags & 1), every stale group
    is reclaimed and the pool compacted.  The order of these steps is the product's."""
    keep = stream_requests(requests, count, capacity, pool.blocks, max_depth)
    nb = min(keep.size, max_bricks)
    out = {"requests": int(count), "unique": int(keep.size), "bricks": 0, "collected": 0, "gc_runs": 0,
           "reclaimed": 0, "moved": 0}
    records = np.zeros(nb, GRAFTING_DTYPE)
    bricks = np.zeros(0, pool.blocks.dtype)
    inp = update_input(base_depth, tick, offset=prune_age)
    inp["graft_size"] = 8 * nb
    if nb:
        sel = requests[keep[:nb]]
        records["graft_data_address"] = 8 * np.arange(nb, dtype=np.uint32)
        records["graft_total_size"] = 8
        records["depth"] = 0xFF
        records["graft_address"] = sel["address"]
        bricks = np.ascontiguousarray(bricks_for(sel)).astype(pool.blocks.dtype)
        assert bricks.size == 8 * nb
        st = pool.graft(inp, records, bricks, np.zeros(0, np.uint64))
        out["bricks"], out["collected"] = st["placed"], st["collected"]
    n_groups = pool.usages.size
    free = int(np.count_nonzero(pool.usages["tick"] == 0))
    gc = not (flags & 1) and ((flags & 2) or free * 100 < gc_free_percent * n_groups)
    if gc:
        out["gc_runs"] = 1
        out["reclaimed"] = pool.reclaim(inp)
        out["moved"] = pool.compact(inp)
        free = int(np.count_nonzero(pool.usages["tick"] == 0))
    out["free_groups"] = free
    out["pool_groups"] = n_groups
    out["records"] = records
    out["bricks_array"] = bricks
    return out


def tick_age(tick: int, now: int) -> int:
    return int(lib().orc_tick_age(tick, now))


# ---------------------------------------------------------------------------
# differentiable compositing
# ---------------------------------------------------------------------------

def composite_layers(trace_inputs, bases: np.ndarray, blocks: np.ndarray, max_layers: int, apertures=None,
                     threads: int = 1):
    """Layer lists (block addresses, UINT32_MAX padded) and background colours of every cone of
    the given views, concatenated view by view."""
    n_views = len(trace_inputs)
    n = sum(t.screen_size[0] * t.screen_size[1] for t in trace_inputs)
    ins = (type(trace_inputs[0]) * n_views)(*trace_inputs)
    ap = None if apertures is None else np.ascontiguousarray(apertures, dtype=np.float32)
    layers = np.zeros((n, max_layers), np.uint32)
    bg = np.zeros((n, 3), np.float32)
    rc = lib().orc_composite_layers(C.cast(ins, C.c_void_p), None if ap is None else ap.ctypes.data, n_views,
                                    bases.ctypes.data, bases.size, blocks.ctypes.data, blocks.size, max_layers,
                                    layers.ctypes.data, bg.ctypes.data, threads)
    if rc != 0:
        raise RuntimeError(f"orc_composite_layers failed: {rc}")
    return layers, bg


def composite_loss(layers: np.ndarray, bg: np.ndarray, params: np.ndarray, target: Optional[np.ndarray] = None,
                   rays: Optional[np.ndarray] = None, want_image: bool = False, want_grads: bool = False):
    """FP64 compositing of layer lists: (loss, image or None, grads or None)."""
    layers = np.ascontiguousarray(layers, np.uint32)
    bg = np.ascontiguousarray(bg, np.float32)
    params = np.ascontiguousarray(params, np.float64)
    n_rays = layers.shape[0] if rays is None else len(rays)
    r = None if rays is None else np.ascontiguousarray(rays, np.uint32)
    t = None if target is None else np.ascontiguousarray(target, np.float64)
    image = np.zeros((layers.shape[0], 3), np.float64) if want_image else None
    grads = np.zeros_like(params) if want_grads else None
    loss = lib().orc_composite_loss(layers.ctypes.data, bg.ctypes.data, None if r is None else r.ctypes.data, n_rays,
                                    layers.shape[1], params.ctypes.data, params.shape[0],
                                    None if t is None else t.ctypes.data, None if image is None else image.ctypes.data,
                                    None if grads is None else grads.ctypes.data)
    return float(loss), image, grads


# ---------------------------------------------------------------------------
# evaluateMaterial (secondary cones)
# ---------------------------------------------------------------------------

def sincos(x: float):
    """The defined sin and cos of evaluateMaterial (see svo_material_oracle.h)."""
    s, c = C.c_float(0.0), C.c_float(0.0)
    lib().orc_sincos(C.c_float(x), C.byref(s), C.byref(c))
    return s.value, c.value


MATERIAL_OUTPUTS = ("origins", "directions", "fovs", "locations", "depths", "weightings", "colour_r", "colour_g",
                    "colour_b", "luminosities", "parents")


def evaluate_material(trace_input, queue: np.ndarray, root: dict, blocks: np.ndarray, usages: Optional[np.ndarray],
                      out: dict, final_w: np.ndarray, bg_queue: np.ndarray, counters: np.ndarray) -> None:
    """evaluateMaterial over `queue` with the primary state in `root` (directions float4,
    locations u64x4, depths, weightings, parents) into the accumulator arrays of `out`
    (origins/directions float4, fovs, locations u64x4, depths, weightings, colour_r/g/b,
    luminosities, parents), FinalWeightings and the background queue; counters =
    [AccumulatorID, BackgroundQueueID] as uint32, advanced in place.  All arrays in place."""
    io = MaterialIO()
    q = np.ascontiguousarray(queue, np.uint32)
    io.queue, io.n = q.ctypes.data, q.size
    io.root_directions = root["directions"].ctypes.data
    io.root_locations = root["locations"].ctypes.data
    io.root_depths = root["depths"].ctypes.data
    io.root_weightings = root["weightings"].ctypes.data
    io.root_parents = root["parents"].ctypes.data
    io.n_root = min(root["directions"].size // 4, root["locations"].size // 4, root["depths"].size,
                    root["weightings"].size, root["parents"].size)
    io.blocks, io.n_blocks = blocks.ctypes.data, blocks.size
    if usages is not None:
        io.usages, io.n_usages = usages.ctypes.data, usages.size
    io.acc_counter = counters.ctypes.data
    io.bg_counter = counters.ctypes.data + 4
    for k in MATERIAL_OUTPUTS:
        setattr(io, k, out[k].ctypes.data)
    io.n_out = min(out["origins"].size // 4, out["directions"].size // 4, out["fovs"].size, out["locations"].size // 4,
                   *(out[k].size for k in ("depths", "weightings", "colour_r", "colour_g", "colour_b", "luminosities",
                                          "parents")))
    io.final_w, io.n_final = final_w.ctypes.data, final_w.size
    io.bg_queue, io.bg_capacity = bg_queue.ctypes.data, bg_queue.size
    assert counters.dtype == np.uint32 and counters.size == 2 and counters.flags["C_CONTIGUOUS"]
    rc = lib().orc_evaluate_material(C.byref(trace_input), C.byref(io))
    if rc != 0:
        raise RuntimeError(f"orc_evaluate_material failed: {rc}")


def _shade_io(queue: np.ndarray, traces: dict, final: dict) -> tuple:
    q = np.ascontiguousarray(queue, np.uint32)
    io = ShadeIO()
    io.queue, io.n = q.ctypes.data, q.size
    n_rays = min(traces["parents"].size, traces["weightings"].size)
    for k in ("directions", "locations"):
        if k in traces:
            setattr(io, k, traces[k].ctypes.data)
            n_rays = min(n_rays, traces[k].size // 4)
    for k in ("depths", "colour_r", "colour_g", "colour_b"):
        if traces.get(k) is not None:
            setattr(io, k, traces[k].ctypes.data)
            n_rays = min(n_rays, traces[k].size)
    io.parents, io.weightings = traces["parents"].ctypes.data, traces["weightings"].ctypes.data
    io.n_rays = n_rays
    for k in ("final_r", "final_g", "final_b", "final_w"):
        assert final[k].dtype == np.uint32 and final[k].flags["C_CONTIGUOUS"]
        setattr(io, k, final[k].ctypes.data)
    io.n_pixels = min(final[k].size for k in ("final_r", "final_g", "final_b", "final_w"))
    return io, q


def evaluate_background(ambient: int, queue: np.ndarray, traces: dict, final: dict) -> None:
    """evaluateBackground over `queue` (wavefront.cl:1741-1797, with the light mix of cones of
    weighting below 255): traces = directions (float4), parents, weightings, colour_r/g/b;
    final = final_r/g/b/w (uint32 per pixel), accumulated in place."""
    io, _q = _shade_io(queue, traces, final)
    lib().orc_evaluate_background(C.c_uint8(ambient), C.byref(io))


def resolve_remainders(base_depth: int, blocks: np.ndarray, queue: np.ndarray, traces: dict, final: dict) -> None:
    """resolveRemainders over `queue` (wavefront.cl:1799-1833): traces = locations (u64x4),
    depths, parents, weightings; final accumulated in place."""
    io, _q = _shade_io(queue, traces, final)
    lib().orc_resolve_remainders(base_depth, blocks.ctypes.data, blocks.size, C.byref(io))


def draw_trace(final: dict) -> np.ndarray:
    """drawTrace (wavefront.cl:1835-1850): the RGBA32F image of the accumulators."""
    n = min(final[k].size for k in ("final_r", "final_g", "final_b", "final_w"))
    rgba = np.zeros((n, 4), np.float32)
    lib().orc_draw_trace(n, *(np.ascontiguousarray(final[k], np.uint32).ctypes.data
                             for k in ("final_r", "final_g", "final_b", "final_w")), rgba.ctypes.data)
    return rgba
