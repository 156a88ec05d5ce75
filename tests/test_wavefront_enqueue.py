"""The kernel-by-kernel path behind svo_enqueue, driven exactly as
MainWindow.RunWavefront drives the reference (MainWindow.cs:356-622): counters
zeroed, init over (W,H), then {ResetIDBuffer + Swap + re-bind + runBaseTrace;
ResetIDBuffer + Swap + re-bind + runBlockTrace} until both queues are empty,
then evaluateBackground, resolveRemainders and drawTrace.

Checked bit-exactly against the oracle's emulation of the same loop, including
the order of every queue (both append in work-item order)."""
import numpy as np
import pytest

from frame_helpers import first_mismatch, tree_arrays

pytestmark = pytest.mark.gpu


def setup_buffers(svo, cm, bases, blocks, usages, n, request_capacity, trace_capacity=None):
    """Buffers of MainWindow (MainWindow.cs:121-178); per-trace arrays hold trace_capacity
    entries (default n, one per pixel)."""
    B = svo.BufferName
    m = trace_capacity or n
    cm.InitReadBuffer(B.Bases, bases)
    cm.InitReadBuffer(B.Blocks, blocks)
    cm.InitBuffer(B.Usages, usages)
    cm.InitBuffer(B.ChildRequestID, np.zeros(1, np.uint32))
    cm.InitBuffer(B.ChildRequests, np.zeros(max(request_capacity, 1) * 40, np.uint8))
    for name, size in [(B.Origins, 16), (B.Directions, 16), (B.FoVs, 4), (B.Locations, 32), (B.Depths, 1),
                       (B.Weightings, 1), (B.ParentTraces, 4), (B.ColourRs, 1), (B.ColourGs, 1), (B.ColourBs, 1),
                       (B.BaseTraces, 4), (B.BaseTraceQueue, 4), (B.BlockTraces, 4), (B.BlockTraceQueue, 4),
                       (B.BackgroundQueue, 4), (B.MaterialQueue, 4)]:
        cm.InitDeviceBuffer(name, m, size)
    for name, size in [(B.FinalColourRs, 4), (B.FinalColourGs, 4), (B.FinalColourBs, 4), (B.FinalWeightings, 4),
                       (B.OutputImage, 16)]:
        cm.InitDeviceBuffer(name, n, size)
    for name in (B.BaseTraceQueueID, B.BlockTraceQueueID, B.BackgroundQueueID, B.MaterialQueueID, B.AccumulatorID):
        cm.InitBuffer(name, np.zeros(1, np.uint32))


def bind(svo, cm, kernel, names):
    for param in names:
        cm.SetArg(kernel, param, svo.BufferName[param])


def run_wavefront(svo, cm, trace_input):
    """MainWindow.RunWavefront + RunTraces + per-kernel binders, through the ComputeManager mirror."""
    iterations, bg = run_primary(svo, cm, trace_input)
    mat = finish_frame(svo, cm, trace_input)
    return iterations, bg, mat


def run_primary(svo, cm, trace_input):
    """init, the RunTraces loop and evaluateBackground of MainWindow.RunWavefront."""
    K, B = svo.KernelName, svo.BufferName
    W, H = trace_input.screen_size[0], trace_input.screen_size[1]
    for name in (B.BlockTraceQueueID, B.BaseTraceQueueID, B.BackgroundQueueID, B.MaterialQueueID,
                 B.AccumulatorID, B.ChildRequestID):
        cm.ZeroIDBuffer(name)
    cm.SetArg(K.Init, "input", trace_input)
    bind(svo, cm, K.Init, ["Origins", "Directions", "FoVs", "Locations", "Depths", "BaseTraceQueue",
                           "BaseTraceQueueID", "BackgroundQueue", "BackgroundQueueID", "ColourRs", "ColourGs",
                           "ColourBs", "Weightings", "ParentTraces"])
    cm.SetArg(K.Init, "FinalColourRs", B.FinalColourRs)
    cm.SetArg(K.Init, "FinalColourGs", B.FinalColourGs)
    cm.SetArg(K.Init, "FinalColourBs", B.FinalColourBs)
    cm.SetArg(K.Init, "FinalWeightings", B.FinalWeightings)
    cm.Enqueue(K.Init, [W, H])
    cm.SetArg(K.RunBaseTrace, "input", trace_input)
    cm.SetArg(K.RunBlockTrace, "input", trace_input)
    iterations = 0
    while True:
        base_size = cm.ResetIDBuffer(B.BaseTraceQueueID)
        if base_size > 0:
            cm.Swap(B.BaseTraces, B.BaseTraceQueue)
            bind(svo, cm, K.RunBaseTrace, ["Bases", "Origins", "Directions", "FoVs", "Locations", "Weightings",
                                           "Depths", "BaseTraces", "BaseTraceQueue", "BaseTraceQueueID",
                                           "BlockTraceQueue", "BlockTraceQueueID", "BackgroundQueue",
                                           "BackgroundQueueID"])
            cm.Enqueue(K.RunBaseTrace, [base_size])
        block_size = cm.ResetIDBuffer(B.BlockTraceQueueID)
        if block_size > 0:
            cm.Swap(B.BlockTraces, B.BlockTraceQueue)
            bind(svo, cm, K.RunBlockTrace, ["Blocks", "Usages", "ChildRequestID", "ChildRequests", "Origins",
                                            "Directions", "FoVs", "Locations", "Weightings", "Depths",
                                            "BlockTraces", "BlockTraceQueue", "BlockTraceQueueID",
                                            "BaseTraceQueue", "BaseTraceQueueID", "BackgroundQueue",
                                            "BackgroundQueueID", "MaterialQueue", "MaterialQueueID",
                                            "ParentTraces"])
            cm.Enqueue(K.RunBlockTrace, [block_size])
        if base_size == 0 and block_size == 0:
            break
        iterations += 1
    trace_input.dof[0] = 0.0
    trace_input.dof[1] = 0.0
    bg = cm.ResetIDBuffer(B.BackgroundQueueID)
    cm.SetArg(K.EvaluateBackground, "input", trace_input)
    if bg:
        bind(svo, cm, K.EvaluateBackground, ["BackgroundQueue", "Directions", "ParentTraces", "ColourRs",
                                             "ColourGs", "ColourBs", "Weightings", "FinalColourRs",
                                             "FinalColourGs", "FinalColourBs", "FinalWeightings"])
        cm.Enqueue(K.EvaluateBackground, [bg])
    return iterations, bg


def finish_frame(svo, cm, trace_input):
    """resolveRemainders and drawTrace of MainWindow.RunWavefront; returns the material count."""
    K, B = svo.KernelName, svo.BufferName
    W, H = trace_input.screen_size[0], trace_input.screen_size[1]
    cm.SetArg(K.ResolveRemainders, "input", trace_input)
    bind(svo, cm, K.ResolveRemainders, ["Blocks", "Usages", "Locations", "Depths", "MaterialQueue",
                                        "FinalColourRs", "FinalColourGs", "FinalColourBs", "FinalWeightings",
                                        "Weightings", "ParentTraces"])
    mat = cm.ResetIDBuffer(B.MaterialQueueID)
    if mat:
        cm.Enqueue(K.ResolveRemainders, [mat])
    bind(svo, cm, K.DrawTrace, ["FinalColourRs", "FinalColourGs", "FinalColourBs", "FinalWeightings"])
    cm.SetArg(K.DrawTrace, "outputImage", B.OutputImage)
    cm.SetArg(K.DrawTrace, "input", trace_input)
    cm.Enqueue(K.DrawTrace, [W, H])
    cm.Wait()
    return mat


CASES = {
    "f0-reference": lambda s: (s.scenes.f0_scene(8), s.scenes.f0_camera(64, 48)),
    "config1": lambda s: (s.scenes.config1_scene(), s.scenes.config1_camera(96, 72)),
    "heightfield": lambda s: (s.scenes.config2_scene(10, 512), s.scenes.config2_camera(9, 160, 90)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_enqueue_loop_matches_oracle_emulation(svo, orc, gpu, case):
    spec, cam = CASES[case](svo)
    tree, bases, blocks = tree_arrays(svo, spec)
    n = cam.screen_size[0] * cam.screen_size[1]
    cap = 4000
    cam.max_child_request_id = cap
    usages = tree.initial_usages(0)
    ref_usages = usages.copy()
    ref = orc.wavefront_frame(cam, bases, blocks, ref_usages, request_capacity=cap)
    setup_buffers(svo, gpu, bases, blocks, usages, n, cap)
    iterations, nbg, nmat = run_wavefront(svo, gpu, cam)
    B = svo.BufferName
    assert iterations == ref["iterations"]
    assert nbg == len(ref["background_queue"]) and nmat == len(ref["material_queue"])
    assert np.array_equal(gpu.ReadBuffer(B.BackgroundQueue, np.uint32, nbg), ref["background_queue"])
    assert np.array_equal(gpu.ReadBuffer(B.MaterialQueue, np.uint32, nmat), ref["material_queue"])
    rgba = gpu.ReadBuffer(B.OutputImage, np.float32, 4 * n).reshape(n, 4)
    assert first_mismatch(rgba, ref["rgba"]) is None
    loc = gpu.ReadBuffer(B.Locations, np.uint64, 4 * n).reshape(n, 4)[:, :3]
    assert first_mismatch(loc, ref["location"]) is None
    assert first_mismatch(gpu.ReadBuffer(B.Depths, np.uint8, n), ref["depth"]) is None
    assert first_mismatch(gpu.ReadBuffer(B.Usages, svo.USAGE_DTYPE, usages.size), ref_usages) is None
    count = int(gpu.ReadBuffer(B.ChildRequestID, np.uint32, 1)[0])
    assert count == ref["request_count"]
    reqs = gpu.ReadBuffer(B.ChildRequests, svo.CHILD_REQUEST_DTYPE, min(count, cap))
    assert first_mismatch(reqs, ref["requests"]) is None
    fw = gpu.ReadBuffer(B.FinalWeightings, np.uint32, n)
    assert np.all(fw[ref["hit"] != 0xFFFFFFFF] == 255)


def test_init_append_is_a_permutation(svo, gpu):
    """AtomicTest's known answer (AtomicTest.cs:86-125): n appends give counter n and a permutation."""
    W, H = 333, 217
    spec, cam = svo.scenes.f0_scene(8), svo.scenes.f0_camera(W, H)
    tree, bases, blocks = tree_arrays(svo, spec)
    n = W * H
    setup_buffers(svo, gpu, bases, blocks, tree.initial_usages(0), n, 1)
    K, B = svo.KernelName, svo.BufferName
    for name in (B.BaseTraceQueueID, B.BackgroundQueueID):
        gpu.ZeroIDBuffer(name)
    gpu.SetArg(K.Init, "input", cam)
    bind(svo, gpu, K.Init, ["Origins", "Directions", "FoVs", "Locations", "Depths", "BaseTraceQueue",
                            "BaseTraceQueueID", "BackgroundQueue", "BackgroundQueueID", "ColourRs", "ColourGs",
                            "ColourBs", "Weightings", "ParentTraces", "FinalColourRs", "FinalColourGs",
                            "FinalColourBs", "FinalWeightings"])
    gpu.Enqueue(K.Init, [W, H])
    nb = gpu.ResetIDBuffer(B.BaseTraceQueueID)
    ng = gpu.ResetIDBuffer(B.BackgroundQueueID)
    assert nb + ng == n and nb > 0 and ng > 0
    a = gpu.ReadBuffer(B.BaseTraceQueue, np.uint32, nb)
    b = gpu.ReadBuffer(B.BackgroundQueue, np.uint32, ng)
    assert np.all(np.diff(a.astype(np.int64)) > 0) and np.all(np.diff(b.astype(np.int64)) > 0)  # work-item order
    assert np.array_equal(np.sort(np.concatenate([a, b])), np.arange(n, dtype=np.uint32))


def test_trace_kernel_name_runs_fused_frame(svo, orc, gpu):
    """KernelName.Trace with the megakernel's argument table (ComputeManager.cs:319-321)."""
    spec, cam = CASES["config1"](svo)
    tree, bases, blocks = tree_arrays(svo, spec)
    n = cam.screen_size[0] * cam.screen_size[1]
    usages = tree.initial_usages(0)
    ref_usages = usages.copy()
    ref = orc.trace_frame(cam, bases, blocks, ref_usages, request_capacity=6000)
    K, B = svo.KernelName, svo.BufferName
    gpu.InitReadBuffer(B.Bases, bases)
    gpu.InitReadBuffer(B.Blocks, blocks)
    gpu.InitBuffer(B.Usages, usages)
    gpu.InitBuffer(B.ChildRequestID, np.zeros(1, np.uint32))
    gpu.InitBuffer(B.ChildRequests, np.zeros(6000 * 40, np.uint8))
    gpu.InitDeviceBuffer(B.OutputImage, n, 16)
    for param, name in [("bases", B.Bases), ("blocks", B.Blocks), ("usage", B.Usages),
                        ("childRequestId", B.ChildRequestID), ("childRequests", B.ChildRequests),
                        ("outputImage", B.OutputImage)]:
        gpu.SetArg(K.Trace, param, name)
    gpu.SetArg(K.Trace, "input", cam)
    gpu.Enqueue(K.Trace, [cam.screen_size[0], cam.screen_size[1]])
    gpu.Wait()
    assert first_mismatch(gpu.ReadBuffer(B.OutputImage, np.float32, 4 * n).reshape(n, 4), ref["rgba"]) is None
    assert first_mismatch(gpu.ReadBuffer(B.Usages, svo.USAGE_DTYPE, usages.size), ref_usages) is None
    assert int(gpu.ReadBuffer(B.ChildRequestID, np.uint32, 1)[0]) == ref["request_count"]


def test_enqueue_errors(svo, gpu):
    K = svo.KernelName
    with pytest.raises(svo.SvoError) as e:
        gpu.Enqueue(K.Test, [4])
    assert e.value.code == -5
    gpu.InitDeviceBuffer(svo.BufferName.Origins, 4, 16)
    with pytest.raises(svo.SvoError) as e:
        gpu.SetArg(K.Init, "NoSuchParam", svo.BufferName.Origins)
    assert e.value.code == -6
    with pytest.raises(svo.SvoError) as e:
        gpu.SetArg(K.Init, "input", b"\x00" * 10)
    assert e.value.code == -1


def test_set_local_arg_checks_like_set_arg(svo, gpu):
    """SetLocalArg (ComputeManager.cs:101-107): unknown kernel, parameter or buffer fail as SetArg
    does; a valid call fails because no kernel of the surface has a local-memory parameter."""
    K, B = svo.KernelName, svo.BufferName
    gpu.InitDeviceBuffer(B.Origins, 64, 16)
    with pytest.raises(svo.SvoError) as e:
        gpu.SetLocalArg(K.Init, "NoSuchParameter", B.Origins)
    assert e.value.code == -6
    try:
        gpu.ReleaseBuffer(B.Pruning)
    except svo.SvoError:
        pass  # not initialised
    with pytest.raises(svo.SvoError) as e:
        gpu.SetLocalArg(K.Init, "Origins", B.Pruning)
    assert e.value.code == -4
    with pytest.raises(svo.SvoError) as e:
        gpu.SetLocalArg(K.Init, "Origins", B.Origins)
    assert e.value.code == -1 and "local-memory" in str(e.value) and "SetKernelArg:Init:Origins" in str(e.value)
