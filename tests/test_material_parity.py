"""evaluateMaterial through svo_enqueue (KernelName.EvaluateMaterial), driven as
MainWindow.EvaulateMaterial drives it (MainWindow.cs:533-571): after the primary frame's
traces and evaluateBackground, the five Root* swaps, the bindings by parameter name,
ResetIDBuffer(MaterialQueueID) and one launch over the material queue.  Every output array,
the usage table and both counters are compared bit-exactly with the oracle run on the same
inputs (read back from the device just before the launch)."""
import numpy as np
import pytest

from frame_helpers import first_mismatch, tree_arrays
from test_wavefront_enqueue import bind, run_primary, setup_buffers

pytestmark = pytest.mark.gpu

MATERIAL_ARGS = ["Blocks", "Usages", "Origins", "Directions", "FoVs", "Locations", "Depths", "Weightings", "ColourRs",
                 "ColourGs", "ColourBs", "RayLengths", "Luminosities", "MaterialQueue", "ParentTraces",
                 "RootDirections", "RootLocations", "RootDepths", "RootWeightings", "RootParentTraces",
                 "BaseTraceQueue", "BaseTraceQueueID", "FinalWeightings", "AccumulatorID", "BackgroundQueue",
                 "BackgroundQueueID"]
OUT = {"origins": ("Origins", np.float32, 4), "directions": ("Directions", np.float32, 4),
       "fovs": ("FoVs", np.float32, 1), "locations": ("Locations", np.uint64, 4), "depths": ("Depths", np.uint8, 1),
       "weightings": ("Weightings", np.uint8, 1), "colour_r": ("ColourRs", np.uint8, 1),
       "colour_g": ("ColourGs", np.uint8, 1), "colour_b": ("ColourBs", np.uint8, 1),
       "luminosities": ("Luminosities", np.uint8, 1), "parents": ("ParentTraces", np.uint32, 1)}
ROOT = {"directions": ("RootDirections", np.float32, 4), "locations": ("RootLocations", np.uint64, 4),
        "depths": ("RootDepths", np.uint8, 1), "weightings": ("RootWeightings", np.uint8, 1),
        "parents": ("RootParentTraces", np.uint32, 1)}

CASES = {
    "config1": lambda s: (s.scenes.config1_scene(), s.scenes.config1_camera(64, 48)),
    "heightfield": lambda s: (s.scenes.config2_scene(10, 512), s.scenes.config2_camera(9, 80, 45)),
}


def _read(cm, svo, spec, n):
    name, dtype, k = spec
    return cm.ReadBuffer(svo.BufferName[name], dtype, n * k)


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("params", [{}, {"fov_multiplier": 0.35, "fov_constant": 0.12, "weighting_multiplier": -0.1}])
def test_evaluate_material_matches_oracle(svo, orc, gpu, case, params):
    spec, cam = CASES[case](svo)
    for k, v in params.items():
        setattr(cam, k, v)
    tree, bases, blocks = tree_arrays(svo, spec)
    n = cam.screen_size[0] * cam.screen_size[1]
    m = 48 * n  # room for the secondary cones
    cam.max_child_request_id = 4000
    K, B = svo.KernelName, svo.BufferName
    setup_buffers(svo, gpu, bases, blocks, tree.initial_usages(0), n, 4000, trace_capacity=m)
    for name, size in [(B.RootDirections, 16), (B.RootLocations, 32), (B.RootDepths, 1), (B.RootWeightings, 1),
                       (B.RootParentTraces, 4), (B.RayLengths, 4), (B.Luminosities, 1)]:
        gpu.InitDeviceBuffer(name, m, size)
    run_primary(svo, gpu, cam)
    # MainWindow.EvaulateMaterial
    for a, b in [(B.RootDirections, B.Directions), (B.RootLocations, B.Locations), (B.RootDepths, B.Depths),
                 (B.RootWeightings, B.Weightings), (B.RootParentTraces, B.ParentTraces)]:
        gpu.Swap(a, b)
    bind(svo, gpu, K.EvaluateMaterial, MATERIAL_ARGS)
    gpu.SetArg(K.EvaluateMaterial, "input", cam)
    size = gpu.ResetIDBuffer(B.MaterialQueueID)
    assert size > 100
    # the oracle's inputs: the device state right before the launch
    queue = gpu.ReadBuffer(B.MaterialQueue, np.uint32, size)
    root = {k: _read(gpu, svo, v, m) for k, v in ROOT.items()}
    out = {k: _read(gpu, svo, v, m) for k, v in OUT.items()}
    usages = gpu.ReadBuffer(B.Usages, svo.USAGE_DTYPE)
    final_w = gpu.ReadBuffer(B.FinalWeightings, np.uint32, n)
    bg = gpu.ReadBuffer(B.BackgroundQueue, np.uint32, m)
    counters = np.array([gpu.ReadBuffer(B.AccumulatorID, np.uint32, 1)[0],
                         gpu.ReadBuffer(B.BackgroundQueueID, np.uint32, 1)[0]], np.uint32)
    gpu.Enqueue(K.EvaluateMaterial, [size])
    gpu.Wait()
    orc.evaluate_material(cam, queue, root, blocks, usages, out, final_w, bg, counters)
    got_counters = [int(gpu.ReadBuffer(B.AccumulatorID, np.uint32, 1)[0]),
                    int(gpu.ReadBuffer(B.BackgroundQueueID, np.uint32, 1)[0])]
    assert got_counters == counters.tolist()
    assert counters[0] > size  # several cones per material hit
    bits = lambda a: a.view(np.uint32) if a.dtype == np.float32 else a  # NaN directions compare by bits
    for k, spec in OUT.items():
        m_ = first_mismatch(bits(_read(gpu, svo, spec, m)), bits(out[k]))
        assert m_ is None, f"{k}: {m_}"
    assert first_mismatch(gpu.ReadBuffer(B.BackgroundQueue, np.uint32, m), bg) is None
    assert first_mismatch(gpu.ReadBuffer(B.FinalWeightings, np.uint32, n), final_w) is None
    assert first_mismatch(gpu.ReadBuffer(B.Usages, svo.USAGE_DTYPE), usages) is None
