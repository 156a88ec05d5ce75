"""The second bounce through svo_enqueue, in the order MainWindow.RunWavefront has it with the
material pass enabled (MainWindow.cs:391-412, 533-571): the primary frame (init, the trace loop,
evaluateBackground), evaluateMaterial, evaluateBackground over the secondary cones it queued
(their weightings are below 255, so the light mix of wavefront.cl:1775-1788 runs), then
resolveRemainders and drawTrace.

In that order resolveRemainders finds MaterialQueueID already read and cleared by the material
pass, so MainWindow launches nothing; the test checks that count and then also resolves the
primary material hits from their pre-material state (the Root* buffers after the swap), so
the remainders kernel runs over accumulators the light mix has written.  Each stage is
compared bit-exactly with the oracle run on the same inputs, and the oracle chain is carried
on its own outputs to the final image."""
import numpy as np
import pytest

from frame_helpers import first_mismatch, tree_arrays
from test_material_parity import CASES, MATERIAL_ARGS, OUT, ROOT, _read
from test_wavefront_enqueue import bind, run_primary, setup_buffers

pytestmark = pytest.mark.gpu

FINAL = {"final_r": "FinalColourRs", "final_g": "FinalColourGs", "final_b": "FinalColourBs",
         "final_w": "FinalWeightings"}
BACKGROUND_ARGS = ["BackgroundQueue", "Directions", "ParentTraces", "ColourRs", "ColourGs", "ColourBs", "Weightings",
                   "FinalColourRs", "FinalColourGs", "FinalColourBs", "FinalWeightings"]


def _finals(svo, gpu, n):
    return {k: gpu.ReadBuffer(svo.BufferName[v], np.uint32, n) for k, v in FINAL.items()}


def _same_finals(svo, gpu, n, want, stage):
    got = _finals(svo, gpu, n)
    for k in FINAL:
        m = first_mismatch(got[k], want[k])
        assert m is None, f"{stage}: {k} {m}"


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("ambient", [0, 3, 170])
def test_second_bounce_matches_oracle(svo, orc, gpu, case, ambient):
    spec, cam = CASES[case](svo)
    cam.fov_multiplier, cam.fov_constant, cam.weighting_multiplier = 0.35, 0.12, -0.1
    cam.ambient_light_level = ambient
    tree, bases, blocks = tree_arrays(svo, spec)
    W, H = cam.screen_size[0], cam.screen_size[1]
    n = W * H
    m = 48 * n
    cam.max_child_request_id = 4000
    K, B = svo.KernelName, svo.BufferName
    setup_buffers(svo, gpu, bases, blocks, tree.initial_usages(0), n, 4000, trace_capacity=m)
    for name, size in [(B.RootDirections, 16), (B.RootLocations, 32), (B.RootDepths, 1), (B.RootWeightings, 1),
                       (B.RootParentTraces, 4), (B.RayLengths, 4), (B.Luminosities, 1)]:
        gpu.InitDeviceBuffer(name, m, size)
    run_primary(svo, gpu, cam)
    final = _finals(svo, gpu, n)  # the accumulators after the primary background pass
    primary_w = final["final_w"].copy()

    # evaluateMaterial (MainWindow.EvaulateMaterial)
    for a, b in [(B.RootDirections, B.Directions), (B.RootLocations, B.Locations), (B.RootDepths, B.Depths),
                 (B.RootWeightings, B.Weightings), (B.RootParentTraces, B.ParentTraces)]:
        gpu.Swap(a, b)
    bind(svo, gpu, K.EvaluateMaterial, MATERIAL_ARGS)
    gpu.SetArg(K.EvaluateMaterial, "input", cam)
    n_mat = gpu.ResetIDBuffer(B.MaterialQueueID)
    assert n_mat > 100
    queue = gpu.ReadBuffer(B.MaterialQueue, np.uint32, n_mat)
    root = {k: _read(gpu, svo, v, m) for k, v in ROOT.items()}
    out = {k: _read(gpu, svo, v, m) for k, v in OUT.items()}
    usages = gpu.ReadBuffer(B.Usages, svo.USAGE_DTYPE)
    bg = gpu.ReadBuffer(B.BackgroundQueue, np.uint32, m)
    counters = np.array([gpu.ReadBuffer(B.AccumulatorID, np.uint32, 1)[0],
                         gpu.ReadBuffer(B.BackgroundQueueID, np.uint32, 1)[0]], np.uint32)
    assert counters.tolist() == [0, 0]
    gpu.Enqueue(K.EvaluateMaterial, [n_mat])
    gpu.Wait()
    orc.evaluate_material(cam, queue, root, blocks, usages, out, final["final_w"], bg, counters)
    _same_finals(svo, gpu, n, final, "evaluateMaterial")

    # evaluateBackground over the secondary cones (MainWindow.EvaulateBackground)
    n_bg = gpu.ResetIDBuffer(B.BackgroundQueueID)
    assert n_bg == int(counters[1]) and n_bg > 100
    queued = bg[:n_bg]
    assert first_mismatch(gpu.ReadBuffer(B.BackgroundQueue, np.uint32, n_bg), queued) is None
    w = out["weightings"][queued]
    # a cone along the face keeps the base weighting (255, no mix); the others take the mix
    assert np.all(w > 5) and np.count_nonzero(w < 255) > n_bg // 2
    gpu.SetArg(K.EvaluateBackground, "input", cam)
    bind(svo, gpu, K.EvaluateBackground, BACKGROUND_ARGS)
    gpu.Enqueue(K.EvaluateBackground, [n_bg])
    gpu.Wait()
    traces = {"directions": out["directions"], "parents": out["parents"], "weightings": out["weightings"],
              "colour_r": out["colour_r"], "colour_g": out["colour_g"], "colour_b": out["colour_b"]}
    orc.evaluate_background(ambient, queued, traces, final)
    _same_finals(svo, gpu, n, final, "evaluateBackground (secondary)")

    # resolveRemainders (MainWindow.ResolveRemainders): the material pass has read and cleared
    # the queue counter, so MainWindow's launch has no work
    assert gpu.ResetIDBuffer(B.MaterialQueueID) == 0
    # ... and the primary material hits resolved from their pre-material state
    gpu.SetArg(K.ResolveRemainders, "input", cam)
    bind(svo, gpu, K.ResolveRemainders, ["Blocks", "Usages", "MaterialQueue", "FinalColourRs", "FinalColourGs",
                                         "FinalColourBs", "FinalWeightings"])
    for param, name in [("Locations", B.RootLocations), ("Depths", B.RootDepths), ("Weightings", B.RootWeightings),
                        ("ParentTraces", B.RootParentTraces)]:
        gpu.SetArg(K.ResolveRemainders, param, name)
    gpu.Enqueue(K.ResolveRemainders, [n_mat])
    gpu.Wait()
    orc.resolve_remainders(cam.base_depth, blocks, queue,
                           {"locations": root["locations"], "depths": root["depths"], "parents": root["parents"],
                            "weightings": root["weightings"]}, final)
    _same_finals(svo, gpu, n, final, "resolveRemainders")

    # drawTrace
    bind(svo, gpu, K.DrawTrace, ["FinalColourRs", "FinalColourGs", "FinalColourBs", "FinalWeightings"])
    gpu.SetArg(K.DrawTrace, "outputImage", B.OutputImage)
    gpu.SetArg(K.DrawTrace, "input", cam)
    gpu.Enqueue(K.DrawTrace, [W, H])
    gpu.Wait()
    image = gpu.ReadBuffer(B.OutputImage, np.float32, 4 * n).reshape(n, 4)
    want = orc.draw_trace(final)
    assert first_mismatch(image.view(np.uint32), want.view(np.uint32)) is None
    # the secondary cones reached the pixels: their weightings were added to the primary ones
    assert np.count_nonzero(final["final_w"] != primary_w) > n // 20
