"""svo-tracer_amd -- MI355X-native wavefront sparse-voxel-octree cone tracer.

Python host mirror of the reference operator surface, over the C ABI of
libsvo.so (include/svo.h):

* ``KernelName`` / ``BufferName``: the ordinals of KernelName.cs:11-21 and
  BufferName.cs:11-63;
* ``ComputeManager``: the methods of SvoTracer.Kernel.ComputeManager
  (ComputeManager.cs:30-367) with the same names and argument meaning; every
  failure raises ``SvoError("{method}: {code}")`` as HandleResultCode does
  (ComputeManager.cs:283-287);
* struct layouts (numpy dtypes / ctypes) of SvoTracer.Domain (Models/...), in
  the natural layouts of wavefront.cl:1-84;
* ``build_tree`` -- TreeBuilder.BuildTree (TreeBuilder.cs:102-150).

There is no CPU fallback: without a built libsvo.so or without a HIP device
the calls that need them raise.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from pathlib import Path
from typing import Optional, Union

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
# SVO_LIB names another build of the library (the host-sanitized one of build.py --sanitize)
LIB_PATH = Path(os.environ["SVO_LIB"]) if os.environ.get("SVO_LIB") else PKG_DIR / "libsvo.so"


class KernelName(enum.IntEnum):
    """KernelName.cs:11-21."""

    Prune = 0
    Graft = 1
    Trace = 2
    Test = 3
    Init = 4
    RunBaseTrace = 5
    RunBlockTrace = 6
    EvaluateBackground = 7
    EvaluateMaterial = 8
    ResolveRemainders = 9
    DrawTrace = 10


class BufferName(enum.IntEnum):
    """BufferName.cs:11-63, plus the per-cone output extensions (>= 64)."""

    Bases = 0
    Blocks = 1
    Usages = 2
    ChildRequestID = 3
    ChildRequests = 4
    Origins = 5
    FoVs = 6
    Directions = 7
    Locations = 8
    Depths = 9
    Weightings = 10
    ParentTraces = 11
    RootDirections = 12
    RootLocations = 13
    RootDepths = 14
    RootWeightings = 15
    RootParentTraces = 16
    BaseTraces = 17
    BlockTraces = 18
    BaseTraceQueue = 19
    BaseTraceQueueID = 20
    BlockTraceQueue = 21
    BlockTraceQueueID = 22
    BackgroundQueue = 23
    BackgroundQueueID = 24
    MaterialQueue = 25
    MaterialQueueID = 26
    ColourRs = 27
    ColourGs = 28
    ColourBs = 29
    Luminosities = 30
    RayLengths = 31
    FinalColourRs = 32
    FinalColourGs = 33
    FinalColourBs = 34
    FinalWeightings = 35
    AccumulatorID = 36
    ParentSize = 37
    ParentResidency = 38
    Parents = 39
    DereferenceQueue = 40
    DereferenceRemaining = 41
    Semaphor = 42
    Pruning = 43
    PruningBlockData = 44
    PruningAddresses = 45
    Grafting = 46
    GraftingBlocks = 47
    GraftingAddresses = 48
    HoldingAddresses = 49
    AddressPosition = 50
    # extensions: per-cone outputs of TraceFrame
    OutputImage = 64
    HitIndices = 65
    VisitCounts = 66
    VisitHashes = 67
    RayStatus = 68
    LeafParams = 69
    TargetImage = 70
    LeafGradients = 71
    Layers = 72
    AdamM = 73
    AdamV = 74


STATUS_NAMES = {
    0: "OK",
    -1: "INVALID_ARGUMENT",
    -2: "NO_DEVICE",
    -3: "HIP",
    -4: "UNKNOWN_BUFFER",
    -5: "UNKNOWN_KERNEL",
    -6: "UNKNOWN_PARAMETER",
    -7: "BUFFER_TOO_SMALL",
    -8: "OUT_OF_MEMORY",
    -9: "NOT_SUPPORTED",
    -10: "CAPACITY",
    -11: "IO",
    -12: "MISSING_ARGUMENT",
}

FRAME_STATE = 1
FRAME_QUEUES = 2
FRAME_NO_XCD_SPLIT = 4
FRAME_CAMERA_TABLE = 8
FRAME_RASTER_ORDER = 16
FRAME_KEEP_REQUESTS = 32

GEOMETRY_SPHERES = 0
GEOMETRY_CUBES = 1
GEOMETRY_SPHERE_NOISE = 2
GEOMETRY_FBM_HEIGHTFIELD = 3

# ---- struct layouts ----
BLOCK_DTYPE = np.dtype(
    [
        ("child", "<u4"),
        ("chunk", "<u2"),
        ("normal_pitch", "<i2"),
        ("normal_yaw", "<i2"),
        ("colour_r", "u1"),
        ("colour_g", "u1"),
        ("colour_b", "u1"),
        ("opacity", "u1"),
        ("spare", "u1"),
        ("gloss", "u1"),
    ]
)
USAGE_DTYPE = np.dtype([("tick", "<u2"), ("pad", "<u2"), ("parent", "<u4")])
PARENT_DTYPE = np.dtype([("parent_address", "<u4"), ("next_element", "<u4")])
CHILD_REQUEST_DTYPE = np.dtype(
    [
        ("address", "<u4"),
        ("tick", "<u2"),
        ("depth", "u1"),
        ("pad0", "u1"),
        ("location", "<u8", (3,)),
        ("tree_size", "u1"),
        ("pad1", "u1", (7,)),
    ]
)
PRUNING_DTYPE = np.dtype(
    [
        ("properties", "u1"),
        ("depth", "u1"),
        ("pad0", "<u2"),
        ("address", "<u4"),
        ("chunk", "<u2"),
        ("pad1", "<u2"),
        ("colour_address", "<u4"),
        ("child_address", "<u4"),
    ]
)
GRAFTING_DTYPE = np.dtype(
    [
        ("graft_data_address", "<u4"),
        ("graft_total_size", "<u4"),
        ("depth", "u1"),
        ("pad", "u1", (3,)),
        ("graft_address", "<u4"),
    ]
)
assert BLOCK_DTYPE.itemsize == 16 and USAGE_DTYPE.itemsize == 8 and CHILD_REQUEST_DTYPE.itemsize == 40
assert PRUNING_DTYPE.itemsize == 20 and GRAFTING_DTYPE.itemsize == 16


class TraceInput(C.Structure):
    """TraceInput, wavefront.cl:53-75 (natural 104-byte layout)."""

    _fields_ = [
        ("origin", C.c_float * 3),
        ("facing", C.c_float * 9),
        ("fov", C.c_float * 2),
        ("dof", C.c_float * 2),
        ("screen_size", C.c_uint32 * 2),
        ("max_opacity", C.c_uint8),
        ("base_depth", C.c_uint8),
        ("tick", C.c_uint16),
        ("max_child_request_id", C.c_uint32),
        ("fov_multiplier", C.c_float),
        ("fov_constant", C.c_float),
        ("weighting_multiplier", C.c_float),
        ("weighting_constant", C.c_float),
        ("ambient_light_level", C.c_uint8),
        ("pad", C.c_uint8 * 7),
    ]


class TreeStats(C.Structure):
    """svo_tree_stats: what one prune / graft / collect call did."""

    _fields_ = [(k, C.c_uint32) for k in ("applied", "skipped", "invalid", "requests", "collected", "placed",
                                           "unlinks", "reserved")]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class CompositeOptions(C.Structure):
    _fields_ = [("max_layers", C.c_uint32), ("flags", C.c_uint32), ("reserved", C.c_uint32 * 2)]


class AdamOptions(C.Structure):
    _fields_ = [("learning_rate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float),
                ("step", C.c_uint32)]


COMPOSITE_ADJOINT = 1
COMPOSITE_LAYERS = 2

STREAM_NO_GC = 1
STREAM_FORCE_GC = 2


class StreamOptions(C.Structure):
    """svo_stream_options: one streaming step after a traced frame."""

    _fields_ = [("base_depth", C.c_uint32), ("max_depth", C.c_uint32), ("tick", C.c_uint16), ("pad", C.c_uint16),
                ("max_bricks", C.c_uint32), ("max_child_request_id", C.c_uint32), ("prune_age", C.c_uint32),
                ("gc_free_percent", C.c_uint32), ("flags", C.c_uint32)]


class StreamStats(C.Structure):
    """svo_stream_stats: what one streaming step did."""

    _fields_ = [(k, C.c_uint32) for k in ("requests", "unique", "bricks", "collected", "gc_runs", "reclaimed",
                                           "moved", "free_groups", "pool_groups")] + [("reserved", C.c_uint32 * 3)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class UpdateInputData(C.Structure):
    """UpdateInputData, wavefront.cl:77-84 (natural 20-byte layout)."""

    _fields_ = [
        ("base_depth", C.c_uint8),
        ("pad0", C.c_uint8),
        ("tick", C.c_uint16),
        ("max_child_request_id", C.c_uint32),
        ("memory_size", C.c_uint32),
        ("offset", C.c_uint32),
        ("graft_size", C.c_uint32),
    ]


class FrameOptions(C.Structure):
    _fields_ = [
        ("flags", C.c_uint32),
        ("max_tree_depth", C.c_uint32),
        ("workgroup_size", C.c_uint32),
        ("refill_lanes", C.c_uint32),
        ("stack_levels", C.c_uint32),
        ("work_chunk", C.c_uint32),
        ("waves_per_simd", C.c_uint32),
        ("reserved", C.c_uint32 * 1),
    ]


class FrameStats(C.Structure):
    _fields_ = [
        ("base_steps", C.c_uint64),
        ("block_steps", C.c_uint64),
        ("rays", C.c_uint64),
        ("child_requests", C.c_uint64),
    ]


class Geometry(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("count", C.c_uint32),
        ("shapes", C.POINTER(C.c_float)),
        ("seed", C.c_uint64),
        ("octaves", C.c_uint32),
        ("resolution", C.c_uint32),
        ("lacunarity", C.c_float),
        ("gain", C.c_float),
        ("amplitude", C.c_float),
        ("frequency", C.c_float),
        ("half_extent", C.c_float),
        ("z_min", C.c_float),
        ("z_max", C.c_float),
    ]


class TreeInfo(C.Structure):
    _fields_ = [
        ("base_depth", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("block_count", C.c_uint32),
        ("base_count", C.c_uint32),
        ("group_count", C.c_uint32),
        ("allocated_groups", C.c_uint32),
    ]


assert C.sizeof(TraceInput) == 104 and C.sizeof(UpdateInputData) == 20


class SvoError(Exception):
    """Raised for every non-OK status, as HandleResultCode throws (ComputeManager.cs:283-287)."""

    def __init__(self, method: str, code: int, detail: str = ""):
        self.method = method
        self.code = code
        self.detail = detail
        super().__init__(f"{method}: {STATUS_NAMES.get(code, code)}" + (f" ({detail})" if detail else ""))


_lib: Optional[C.CDLL] = None

_PROTOS = {
    "svo_device_count": [C.POINTER(C.c_int)],
    "svo_create": [C.c_int, C.POINTER(C.c_void_p)],
    "svo_destroy": [C.c_void_p],
    "svo_abi_version": [],
    "svo_buffer_init": [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p, C.c_uint32],
    "svo_buffer_release": [C.c_void_p, C.c_int],
    "svo_buffer_size": [C.c_void_p, C.c_int, C.POINTER(C.c_size_t)],
    "svo_buffer_write": [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p, C.c_size_t],
    "svo_buffer_read": [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p, C.c_size_t],
    "svo_buffer_fill": [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t],
    "svo_reset_id": [C.c_void_p, C.c_int, C.POINTER(C.c_uint32)],
    "svo_buffer_copy": [C.c_void_p, C.c_int, C.c_size_t, C.c_int, C.c_size_t, C.c_size_t],
    "svo_zero_id": [C.c_void_p, C.c_int],
    "svo_swap": [C.c_void_p, C.c_int, C.c_int],
    "svo_buffer_device_pointer": [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)],
    "svo_set_arg_buffer": [C.c_void_p, C.c_int, C.c_char_p, C.c_int],
    "svo_set_local_arg": [C.c_void_p, C.c_int, C.c_char_p, C.c_int],
    "svo_set_arg_bytes": [C.c_void_p, C.c_int, C.c_char_p, C.c_void_p, C.c_size_t],
    "svo_enqueue": [C.c_void_p, C.c_int, C.c_uint32, C.POINTER(C.c_size_t)],
    "svo_flush": [C.c_void_p],
    "svo_finish": [C.c_void_p],
    "svo_trace_frame": [C.c_void_p, C.POINTER(TraceInput), C.POINTER(FrameOptions)],
    "svo_geometry_create": [C.c_void_p, C.c_void_p],
    "svo_geometry_destroy": [C.c_void_p],
    "svo_geometry_blocks": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p],
    "svo_trace_frame_static": [C.c_void_p, C.POINTER(TraceInput), C.POINTER(FrameOptions)],
    "svo_frame_stats_get": [C.c_void_p, C.POINTER(FrameStats)],
    "svo_trace_batch": [C.c_void_p, C.POINTER(TraceInput), C.c_void_p, C.c_uint32, C.POINTER(FrameOptions)],
    "svo_frame_graph_create": [C.c_void_p, C.POINTER(TraceInput), C.c_void_p, C.c_uint32, C.POINTER(FrameOptions),
                               C.c_uint32, C.POINTER(C.c_void_p)],
    "svo_frame_graph_launch": [C.c_void_p, C.c_void_p, C.POINTER(TraceInput)],
    "svo_frame_graph_destroy": [C.c_void_p, C.c_void_p],
    "svo_composite": [C.c_void_p, C.POINTER(TraceInput), C.c_void_p, C.c_uint32, C.POINTER(CompositeOptions),
                      C.POINTER(C.c_double)],
    "svo_adam_step": [C.c_void_p, C.POINTER(AdamOptions)],
    "svo_autotune": [C.c_void_p, C.POINTER(TraceInput), C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(FrameOptions),
                     C.POINTER(FrameOptions), C.POINTER(C.c_float)],
    "svo_timer_begin": [C.c_void_p],
    "svo_timer_end": [C.c_void_p, C.POINTER(C.c_float)],
    "svo_kernel_timing_enable": [C.c_void_p, C.c_int],
    "svo_kernel_timing_get": [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.c_int],
    "svo_tree_build": [C.POINTER(Geometry), C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)],
    "svo_tree_info_get": [C.c_void_p, C.POINTER(TreeInfo)],
    "svo_tree_copy_bases": [C.c_void_p, C.c_void_p, C.c_size_t],
    "svo_tree_copy_blocks": [C.c_void_p, C.c_void_p, C.c_size_t],
    "svo_tree_initial_usages": [C.c_void_p, C.c_uint16, C.c_void_p, C.c_size_t],
    "svo_tree_save_oct": [C.c_void_p, C.c_char_p],
    "svo_tree_load_oct": [C.c_char_p, C.POINTER(C.c_void_p)],
    "svo_tree_free": [C.c_void_p],
    "svo_context_load_tree": [C.c_void_p, C.c_void_p, C.c_int, C.c_uint16],
    "svo_camera_lut": [C.POINTER(TraceInput), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "svo_prune": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                  C.c_void_p],
    "svo_graft": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                  C.c_void_p],
    "svo_collect": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(C.c_uint32)],
    "svo_reclaim": [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)],
    "svo_compact": [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)],
    "svo_terrain_create": [C.c_void_p, C.POINTER(Geometry), C.POINTER(C.c_void_p)],
    "svo_terrain_destroy": [C.c_void_p],
    "svo_terrain_heights": [C.c_void_p, C.c_void_p, C.c_size_t],
    "svo_stream_seed": [C.c_void_p, C.c_void_p, C.c_uint32],
    "svo_terrain_build_tree": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_uint16,
                               C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)],
    "svo_convert_probe": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_size_t],
    "svo_tree_build_device": [C.c_void_p, C.POINTER(Geometry), C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_uint16,
                              C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)],
    "svo_stream_update": [C.c_void_p, C.c_void_p, C.POINTER(StreamOptions), C.POINTER(StreamStats)],
}


def lib() -> C.CDLL:
    """Loads svo-tracer_amd/libsvo.so (built by build.py); raises if it is missing."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `python svo-tracer_amd/build.py`")
        L = C.CDLL(str(LIB_PATH))
        for name, args in _PROTOS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.svo_last_error.argtypes = []
        L.svo_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def last_error() -> str:
    return (lib().svo_last_error() or b"").decode(errors="replace")


def _check(rc: int, method: str) -> None:
    if rc != 0:
        raise SvoError(method, rc, last_error())


def device_count() -> int:
    n = C.c_int(0)
    _check(lib().svo_device_count(C.byref(n)), "DeviceCount")
    return n.value


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _as_bytes(data) -> np.ndarray:
    if isinstance(data, (bytes, bytearray, memoryview)):
        return np.frombuffer(bytes(data), dtype=np.uint8)
    if isinstance(data, C.Structure):
        return np.frombuffer(bytes(data), dtype=np.uint8)
    arr = np.ascontiguousarray(data)
    return arr.reshape(-1).view(np.uint8)


class ComputeManager:
    """Mirror of SvoTracer.Kernel.ComputeManager over libsvo (one HIP stream per context)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().svo_create(device, C.byref(h)), "ComputeManagerFactory.Build")
        self._h = h
        self._sizes: dict[int, int] = {}

    # --- lifetime (Dispose, ComputeManager.cs:357-367) ---
    def Dispose(self) -> None:
        if self._h:
            lib().svo_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.Dispose()

    def __del__(self):
        try:
            self.Dispose()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    # --- buffers (ComputeManager.cs:126-273) ---
    def InitBuffer(self, name: BufferName, data) -> None:
        """InitBuffer<T>(BufferName, T[]): read-write buffer initialised from host data."""
        b = _as_bytes(data)
        _check(lib().svo_buffer_init(self._h, int(name), b.nbytes, b.ctypes.data if b.nbytes else None, 0), f"CreateBuffer:{BufferName(name).name}")

    def InitReadBuffer(self, name: BufferName, data) -> None:
        b = _as_bytes(data)
        _check(lib().svo_buffer_init(self._h, int(name), b.nbytes, b.ctypes.data if b.nbytes else None, 1), f"CreateBuffer:{BufferName(name).name}")

    def InitDeviceBuffer(self, name: BufferName, array_size: int, element_size: int = 1) -> None:
        """InitDeviceBuffer(BufferName, arraySize, elementSize): zero-filled device buffer (sizes in bytes = n * elem)."""
        _check(lib().svo_buffer_init(self._h, int(name), int(array_size) * int(element_size), None, 0), f"CreateBuffer:{BufferName(name).name}")

    def InitWriteBuffer(self, name: BufferName, array_size: int, element_size: int = 1) -> None:
        _check(lib().svo_buffer_init(self._h, int(name), int(array_size) * int(element_size), None, 2), f"CreateBuffer:{BufferName(name).name}")

    def ReleaseBuffer(self, name: BufferName) -> None:
        _check(lib().svo_buffer_release(self._h, int(name)), f"ReleaseBuffer:{BufferName(name).name}")

    def BufferSize(self, name: BufferName) -> int:
        n = C.c_size_t(0)
        _check(lib().svo_buffer_size(self._h, int(name), C.byref(n)), f"BufferSize:{BufferName(name).name}")
        return n.value

    def WriteBuffer(self, name: BufferName, data, offset: int = 0) -> None:
        b = _as_bytes(data)
        _check(lib().svo_buffer_write(self._h, int(name), offset, b.ctypes.data if b.nbytes else None, b.nbytes), f"EnqueueWriteBuffer:{BufferName(name).name}")

    def ReadBuffer(self, name: BufferName, dtype=np.uint8, count: Optional[int] = None, offset: int = 0) -> np.ndarray:
        dt = np.dtype(dtype)
        if count is None:
            count = (self.BufferSize(name) - offset) // dt.itemsize
        out = np.empty(count, dtype=dt)
        _check(lib().svo_buffer_read(self._h, int(name), offset, out.ctypes.data if out.nbytes else None, out.nbytes), f"EnqueueReadBuffer:{BufferName(name).name}")
        return out

    def FillBuffer(self, name: BufferName, pattern, size: Optional[int] = None, offset: int = 0) -> None:
        p = _as_bytes(pattern)
        if size is None:
            size = self.BufferSize(name) - offset
        _check(lib().svo_buffer_fill(self._h, int(name), p.ctypes.data, p.nbytes, offset, size), f"EnqueueFillBuffer:{BufferName(name).name}")

    def CopyBuffer(self, dst: BufferName, src: BufferName, size: int, dst_offset: int = 0, src_offset: int = 0) -> None:
        """Device-to-device copy (clEnqueueCopyBuffer role; no ComputeManager counterpart)."""
        _check(lib().svo_buffer_copy(self._h, int(dst), dst_offset, int(src), src_offset, size), "EnqueueCopyBuffer")

    def ResetIDBuffer(self, name: BufferName) -> int:
        v = C.c_uint32(0)
        _check(lib().svo_reset_id(self._h, int(name), C.byref(v)), f"EnqueueReadBuffer:{BufferName(name).name}")
        return v.value

    def ZeroIDBuffer(self, name: BufferName) -> None:
        _check(lib().svo_zero_id(self._h, int(name)), f"EnqueueWriteBuffer:{BufferName(name).name}")

    def Swap(self, buffer1: BufferName, buffer2: BufferName) -> None:
        _check(lib().svo_swap(self._h, int(buffer1), int(buffer2)), "Swap")

    def DevicePointer(self, name: BufferName) -> int:
        p = C.c_void_p()
        _check(lib().svo_buffer_device_pointer(self._h, int(name), C.byref(p)), "DevicePointer")
        return p.value or 0

    # --- kernels (ComputeManager.cs:60-120) ---
    def SetArg(self, kernel: KernelName, param: str, argument) -> None:
        """SetArg(kernel, param, BufferName) binds a buffer; any other argument is passed by value."""
        if isinstance(argument, BufferName):
            rc = lib().svo_set_arg_buffer(self._h, int(kernel), param.encode(), int(argument))
        else:
            b = _as_bytes(argument)
            rc = lib().svo_set_arg_bytes(self._h, int(kernel), param.encode(), b.ctypes.data, b.nbytes)
        _check(rc, f"SetKernelArg:{KernelName(kernel).name}:{param}")

    def SetLocalArg(self, kernel: KernelName, param: str, buffer: BufferName) -> None:
        """SetLocalArg(kernel, param, BufferName): a local-memory argument of the buffer's size."""
        _check(lib().svo_set_local_arg(self._h, int(kernel), param.encode(), int(buffer)),
               f"SetKernelArg:{KernelName(kernel).name}:{param}")

    def Enqueue(self, kernel: KernelName, work_size) -> None:
        ws = (C.c_size_t * len(work_size))(*[int(x) for x in work_size])
        _check(lib().svo_enqueue(self._h, int(kernel), len(work_size), ws), "EnqueueNDRangeKernel")

    def Flush(self) -> None:
        _check(lib().svo_flush(self._h), "Flush")

    def Wait(self) -> None:
        _check(lib().svo_finish(self._h), "WaitForEvents")

    # --- frame-level entry points (extensions) ---
    @staticmethod
    def _update_input(update_input) -> UpdateInputData:
        if isinstance(update_input, UpdateInputData):
            return update_input
        raw = np.ascontiguousarray(update_input).view(np.uint8)
        if raw.size != C.sizeof(UpdateInputData):
            raise ValueError("UpdateInputData must be 20 bytes")
        return UpdateInputData.from_buffer_copy(raw.tobytes())

    @staticmethod
    def _arr(a, dtype):
        if a is None:
            return None, 0
        a = np.ascontiguousarray(a, dtype=dtype)
        return a, a.size

    def Prune(self, update_input, records, surface=None, locations=None) -> dict:
        """svo_prune: Pruning records (PRUNING_DTYPE) with their surface blocks and target
        locations (three u64 per entry) applied to Bases/Blocks/Usages on the device."""
        ui = self._update_input(update_input)
        rec, n = self._arr(records, PRUNING_DTYPE)
        surf, ns = self._arr(surface, BLOCK_DTYPE)
        loc = None if locations is None else np.ascontiguousarray(locations, dtype=np.uint64).reshape(-1)
        st = TreeStats()
        _check(lib().svo_prune(self._h, C.byref(ui), _ptr(rec) if n else None, n, _ptr(surf) if ns else None, ns,
                               _ptr(loc) if loc is not None and loc.size else None,
                               0 if loc is None else loc.size // 3, C.byref(st)), "Prune")
        return st.as_dict()

    def Graft(self, update_input, records, graft_blocks=None, locations=None) -> dict:
        """svo_graft: Grafting records (GRAFTING_DTYPE) whose packed subtrees are linked
        below their targets after collecting the groups they need."""
        ui = self._update_input(update_input)
        rec, n = self._arr(records, GRAFTING_DTYPE)
        gb, ngb = self._arr(graft_blocks, BLOCK_DTYPE)
        loc = None if locations is None else np.ascontiguousarray(locations, dtype=np.uint64).reshape(-1)
        st = TreeStats()
        _check(lib().svo_graft(self._h, C.byref(ui), _ptr(rec) if n else None, n, _ptr(gb) if ngb else None, ngb,
                               _ptr(loc) if loc is not None and loc.size else None,
                               0 if loc is None else loc.size // 3, C.byref(st)), "Graft")
        return st.as_dict()

    def Collect(self, update_input, need: int) -> np.ndarray:
        """svo_collect: claims up to `need` free or stale groups; returns them in claim order."""
        ui = self._update_input(update_input)
        out = np.zeros(max(need, 1), np.uint32)
        count = C.c_uint32(0)
        _check(lib().svo_collect(self._h, C.byref(ui), need, out.ctypes.data, C.byref(count)), "Collect")
        return out[: count.value]

    def Reclaim(self, update_input) -> int:
        """svo_reclaim: frees every stale live group (age > UpdateInputData.Offset)."""
        ui = self._update_input(update_input)
        n = C.c_uint32(0)
        _check(lib().svo_reclaim(self._h, C.byref(ui), C.byref(n)), "Reclaim")
        return n.value

    def Compact(self, update_input) -> int:
        """svo_compact: moves live groups above the live count into the holes below it."""
        ui = self._update_input(update_input)
        n = C.c_uint32(0)
        _check(lib().svo_compact(self._h, C.byref(ui), C.byref(n)), "Compact")
        return n.value

    def CreateTerrain(self, geometry: "Geometry") -> "Terrain":
        """svo_terrain_create: the heightfield's grid and min/max pyramid on this device."""
        h = C.c_void_p()
        _check(lib().svo_terrain_create(self._h, C.byref(geometry), C.byref(h)), "CreateTerrain")
        return Terrain(h, geometry)

    def BuildTreeOnDevice(self, terrain: "Terrain", base_depth: int, max_depth: int, max_size: int = 0,
                          with_usages: bool = False, usage_tick: int = 0, host_copy: bool = False):
        """svo_terrain_build_tree: TreeBuilder.BuildTree of the terrain on the device, bound as
        this context's Bases / Blocks (/ Usages).  Returns (block_count, Octree or None)."""
        h = C.c_void_p()
        count = C.c_uint32(0)
        _check(lib().svo_terrain_build_tree(self._h, terrain.handle, base_depth, max_depth, max_size,
                                            1 if with_usages else 0, usage_tick,
                                            C.byref(h) if host_copy else None, C.byref(count)), "BuildTreeOnDevice")
        return int(count.value), (Octree(h) if host_copy else None)

    def BuildTreeDevice(self, geometry: "Geometry", base_depth: int, max_depth: int, max_size: int = 0,
                        with_usages: bool = False, usage_tick: int = 0, host_copy: bool = False):
        """svo_tree_build_device: TreeBuilder.BuildTree of any builder geometry on the device, bound
        as this context's Bases / Blocks (/ Usages).  Returns (block_count, Octree or None)."""
        h = C.c_void_p()
        count = C.c_uint32(0)
        _check(lib().svo_tree_build_device(self._h, C.byref(geometry), base_depth, max_depth, max_size,
                                           1 if with_usages else 0, usage_tick,
                                           C.byref(h) if host_copy else None, C.byref(count)), "BuildTreeDevice")
        return int(count.value), (Octree(h) if host_copy else None)

    def ConvertProbe(self, kind: int, values: np.ndarray) -> np.ndarray:
        """svo_convert_probe: the device's floatToUlong (kind 0, float32 in) or ulongToFloat
        (kind 1, uint64 in) over an array."""
        x = np.ascontiguousarray(values, dtype=np.float32 if kind == 0 else np.uint64)
        out = np.empty(x.size, np.uint64 if kind == 0 else np.float32)
        _check(lib().svo_convert_probe(self._h, kind, x.ctypes.data, out.ctypes.data, x.size), "ConvertProbe")
        return out

    def StreamSeed(self, terrain: "Terrain", base_depth: int) -> None:
        """svo_stream_seed: Bases, the dense block level and the usage table of a fresh pool."""
        _check(lib().svo_stream_seed(self._h, terrain.handle, base_depth), "StreamSeed")

    def StreamUpdate(self, terrain: "Terrain", base_depth: int, max_depth: int, tick: int, max_bricks: int = 65536,
                     max_child_request_id: int = 0, prune_age: int = 30, gc_free_percent: int = 10,
                     flags: int = 0) -> dict:
        """svo_stream_update: the frame's child requests become grafted bricks; garbage
        collection when the pool runs low."""
        o = StreamOptions()
        o.base_depth, o.max_depth, o.tick, o.max_bricks = base_depth, max_depth, tick, max_bricks
        o.max_child_request_id, o.prune_age, o.gc_free_percent, o.flags = (max_child_request_id, prune_age,
                                                                            gc_free_percent, flags)
        st = StreamStats()
        _check(lib().svo_stream_update(self._h, terrain.handle, C.byref(o), C.byref(st)), "StreamUpdate")
        return st.as_dict()

    def LoadTree(self, tree: "Octree", with_usages: bool = True, usage_tick: int = 0) -> None:
        _check(lib().svo_context_load_tree(self._h, tree.handle, 1 if with_usages else 0, usage_tick), "LoadTree")

    def TraceFrame(self, trace_input: TraceInput, flags: int = 0, max_tree_depth: int = 0, workgroup_size: int = 0,
                   refill_lanes: int = 0, stack_levels: int = 0, work_chunk: int = 0,
                   waves_per_simd: int = 0) -> None:
        o = FrameOptions()
        o.flags = flags
        o.max_tree_depth = max_tree_depth
        o.workgroup_size = workgroup_size
        o.refill_lanes = refill_lanes
        o.stack_levels = stack_levels
        o.work_chunk = work_chunk
        o.waves_per_simd = waves_per_simd
        _check(lib().svo_trace_frame(self._h, C.byref(trace_input), C.byref(o)), "TraceFrame")

    def TraceFrameStatic(self, trace_input: TraceInput, flags: int = 0, **opts) -> None:
        """svo_trace_frame_static: TraceFrame with its scratch in the library's static device
        array (no allocation; raster tile order)."""
        o = FrameOptions()
        o.flags = flags
        for k, v in opts.items():
            setattr(o, k, v)
        _check(lib().svo_trace_frame_static(self._h, C.byref(trace_input), C.byref(o)), "TraceFrameStatic")

    def TraceBatch(self, trace_inputs, apertures=None, flags: int = 0, max_tree_depth: int = 0,
                   workgroup_size: int = 0, refill_lanes: int = 0, stack_levels: int = 0, work_chunk: int = 0,
                   waves_per_simd: int = 0) -> None:
        """svo_trace_batch: several cameras in one persistent launch; per-cone buffers hold the
        cameras' cones concatenated in camera order."""
        n = len(trace_inputs)
        arr = (TraceInput * n)(*trace_inputs)
        ap = None if apertures is None else np.ascontiguousarray(apertures, dtype=np.float32)
        o = FrameOptions()
        o.flags = flags
        o.max_tree_depth = max_tree_depth
        o.workgroup_size = workgroup_size
        o.refill_lanes = refill_lanes
        o.stack_levels = stack_levels
        o.work_chunk = work_chunk
        o.waves_per_simd = waves_per_simd
        _check(lib().svo_trace_batch(self._h, arr, None if ap is None else ap.ctypes.data, n, C.byref(o)),
               "TraceBatch")

    def CaptureFrame(self, trace_inputs, apertures=None, per_camera: bool = False, **options) -> "FrameGraph":
        """svo_frame_graph_create: records the submission of a (batched) frame into a HIP graph."""
        n = len(trace_inputs)
        arr = (TraceInput * n)(*trace_inputs)
        ap = None if apertures is None else np.ascontiguousarray(apertures, dtype=np.float32)
        o = FrameOptions()
        for k, v in options.items():
            setattr(o, k, v)
        h = C.c_void_p()
        _check(lib().svo_frame_graph_create(self._h, arr, None if ap is None else ap.ctypes.data, n, C.byref(o),
                                            1 if per_camera else 0, C.byref(h)), "CaptureFrame")
        return FrameGraph(self, h, n)

    def Composite(self, views, apertures=None, max_layers: int = 0, adjoint: bool = False,
                  layers: bool = False) -> float:
        """svo_composite: differentiable compositing of the views into OutputImage; returns the
        L2 loss against TargetImage (0 when no target is bound)."""
        n = len(views)
        arr = (TraceInput * n)(*views)
        ap = None if apertures is None else np.ascontiguousarray(apertures, dtype=np.float32)
        o = CompositeOptions()
        o.max_layers = max_layers
        o.flags = (COMPOSITE_ADJOINT if adjoint else 0) | (COMPOSITE_LAYERS if layers else 0)
        loss = C.c_double(0.0)
        _check(lib().svo_composite(self._h, arr, None if ap is None else ap.ctypes.data, n, C.byref(o),
                                   C.byref(loss)), "Composite")
        return loss.value

    def Autotune(self, trace_inputs, apertures=None, frames: int = 3, **base) -> tuple:
        """svo_autotune: (best launch options as a dict, ms per frame with them)."""
        n = len(trace_inputs)
        arr = (TraceInput * n)(*trace_inputs)
        ap = None if apertures is None else np.ascontiguousarray(apertures, dtype=np.float32)
        b, best = FrameOptions(), FrameOptions()
        for k, v in base.items():
            setattr(b, k, v)
        ms = C.c_float(0.0)
        _check(lib().svo_autotune(self._h, arr, None if ap is None else ap.ctypes.data, n, frames, C.byref(b),
                                  C.byref(best), C.byref(ms)), "Autotune")
        keys = ("workgroup_size", "work_chunk", "refill_lanes", "stack_levels", "waves_per_simd", "flags")
        return {k: getattr(best, k) for k in keys}, ms.value

    def AdamStep(self, step: int, learning_rate: float = 0.01, beta1: float = 0.9, beta2: float = 0.999,
                 epsilon: float = 1e-8) -> None:
        o = AdamOptions(learning_rate, beta1, beta2, epsilon, step)
        _check(lib().svo_adam_step(self._h, C.byref(o)), "AdamStep")

    def FrameStats(self) -> FrameStats:
        s = FrameStats()
        _check(lib().svo_frame_stats_get(self._h, C.byref(s)), "FrameStats")
        return s

    def TimerBegin(self) -> None:
        _check(lib().svo_timer_begin(self._h), "TimerBegin")

    def TimerEnd(self) -> float:
        ms = C.c_float(0)
        _check(lib().svo_timer_end(self._h, C.byref(ms)), "TimerEnd")
        return ms.value

    def KernelTiming(self, enable: Optional[bool] = None, reset: bool = False):
        if enable is not None:
            _check(lib().svo_kernel_timing_enable(self._h, 1 if enable else 0), "KernelTiming")
            return None
        total = C.c_double(0)
        n = C.c_uint64(0)
        _check(lib().svo_kernel_timing_get(self._h, C.byref(total), C.byref(n), 1 if reset else 0), "KernelTiming")
        return total.value, n.value


class Terrain:
    """A device terrain (svo_terrain); keep it alive while streaming from it."""

    def __init__(self, handle: C.c_void_p, geometry: "Geometry"):
        self.handle = handle
        self.geometry = geometry
        res = geometry.resolution or 4096
        self.resolution = 1 << max(0, (res - 1).bit_length())

    def heights(self) -> np.ndarray:
        out = np.zeros((self.resolution, self.resolution), np.float32)
        _check(lib().svo_terrain_heights(self.handle, out.ctypes.data, out.size), "TerrainHeights")
        return out

    def close(self) -> None:
        if self.handle:
            lib().svo_terrain_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FrameGraph:
    """A captured frame submission (svo_frame_graph_*): Launch(new cameras) replays it."""

    def __init__(self, cm: "ComputeManager", handle: C.c_void_p, n: int):
        self.cm, self.handle, self.n = cm, handle, n

    def Launch(self, trace_inputs) -> None:
        if len(trace_inputs) != self.n:
            raise ValueError("a captured frame replays with the same number of cameras")
        arr = (TraceInput * self.n)(*trace_inputs)
        _check(lib().svo_frame_graph_launch(self.cm._h, self.handle, arr), "LaunchFrame")

    def Dispose(self) -> None:
        if self.handle:
            lib().svo_frame_graph_destroy(self.cm._h, self.handle)
            self.handle = None


class Octree:
    """SvoTracer.Domain.Models.Octree (Octree.cs) backed by a native tree handle."""

    def __init__(self, handle: C.c_void_p):
        self.handle = handle
        info = TreeInfo()
        _check(lib().svo_tree_info_get(handle, C.byref(info)), "TreeInfo")
        self.base_depth = info.base_depth
        self.max_depth = info.max_depth
        self.block_count = info.block_count
        self.group_count = info.group_count
        self.allocated_groups = info.allocated_groups
        self.base_count = info.base_count

    @property
    def bases(self) -> np.ndarray:
        out = np.empty(self.base_count, dtype=np.uint16)
        _check(lib().svo_tree_copy_bases(self.handle, out.ctypes.data, out.size), "CopyBases")
        return out

    @property
    def blocks(self) -> np.ndarray:
        out = np.empty(self.block_count, dtype=BLOCK_DTYPE)
        _check(lib().svo_tree_copy_blocks(self.handle, out.ctypes.data, out.size), "CopyBlocks")
        return out

    def initial_usages(self, tick: int = 0, groups: Optional[int] = None) -> np.ndarray:
        n = self.group_count if groups is None else groups
        out = np.zeros(n, dtype=USAGE_DTYPE)
        _check(lib().svo_tree_initial_usages(self.handle, tick, out.ctypes.data, n), "InitialUsages")
        return out

    def save(self, path: Union[str, os.PathLike]) -> None:
        _check(lib().svo_tree_save_oct(self.handle, str(path).encode()), "SaveTree")

    @classmethod
    def load(cls, path: Union[str, os.PathLike]) -> "Octree":
        h = C.c_void_p()
        _check(lib().svo_tree_load_oct(str(path).encode(), C.byref(h)), "LoadTree")
        return cls(h)

    def __del__(self):
        try:
            if self.handle:
                lib().svo_tree_free(self.handle)
                self.handle = None
        except Exception:
            pass


def build_tree(geometry: Geometry, base_depth: int, max_depth: int, max_size: int = 0) -> Octree:
    """TreeBuilder.BuildTree(baseDepth, maxDepth, maxSize) (TreeBuilder.cs:102-150)."""
    h = C.c_void_p()
    _check(lib().svo_tree_build(C.byref(geometry), base_depth, max_depth, max_size, C.byref(h)), "BuildTree")
    return Octree(h)


class GeometryHost:
    """The builder's geometry on the host (svo_geometry_create): MakeBlock of arbitrary cells."""

    def __init__(self, geometry: Geometry):
        self.handle = C.c_void_p()
        _check(lib().svo_geometry_create(C.byref(geometry), C.byref(self.handle)), "GeometryCreate")

    def blocks(self, cells: np.ndarray) -> np.ndarray:
        """cells: (n, 4) uint32 rows of (level, x, y, z) -> n blocks (BLOCK_DTYPE)."""
        c = np.ascontiguousarray(cells, np.uint32).reshape(-1, 4)
        out = np.zeros(c.shape[0], BLOCK_DTYPE)
        _check(lib().svo_geometry_blocks(self.handle, c.ctypes.data, c.shape[0], out.ctypes.data), "GeometryBlocks")
        return out

    def close(self) -> None:
        if self.handle:
            lib().svo_geometry_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def camera_lut(trace_input: TraceInput):
    w, h = trace_input.screen_size[0], trace_input.screen_size[1]
    cu = np.empty(w, np.float32)
    su = np.empty(w, np.float32)
    cv = np.empty(h, np.float32)
    sv = np.empty(h, np.float32)
    _check(lib().svo_camera_lut(C.byref(trace_input), cu.ctypes.data, su.ctypes.data, cv.ctypes.data, sv.ctypes.data), "CameraLut")
    return cu, su, cv, sv


from . import scenes  # noqa: E402,F401  (configuration scenes and cameras)
