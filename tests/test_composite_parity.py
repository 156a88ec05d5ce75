"""Differentiable compositing on the GPU (svo_composite, svo_adam_step) against the oracle:
layer lists bit-exact, image within 1e-5, gradients within 1e-3 of the FP64 adjoint and of
central finite differences."""
from __future__ import annotations

import numpy as np
import pytest

from frame_helpers import tree_arrays

pytestmark = pytest.mark.gpu

U32MAX = 0xFFFFFFFF
K = 4


def _params_from_blocks(blocks):
    p = np.zeros((blocks.size, 4), np.float32)
    p[:, 0] = blocks["colour_r"] / np.float32(255.0)
    p[:, 1] = blocks["colour_g"] / np.float32(255.0)
    p[:, 2] = blocks["colour_b"] / np.float32(255.0)
    p[:, 3] = blocks["opacity"] / np.float32(255.0)
    return p


def _perturbed(params, seed):
    rng = np.random.default_rng(seed)
    q = params.copy()
    q[:, :3] = np.clip(q[:, :3] + rng.normal(0, 0.1, (q.shape[0], 3)), 0, 1)
    q[:, 3] = np.clip(q[:, 3] + rng.normal(0, 0.05, q.shape[0]), 0, 1)
    return q.astype(np.float32)


@pytest.fixture(scope="module")
def setup(svo):
    tree, bases, blocks = tree_arrays(svo, svo.scenes.config1_scene())
    views = [svo.scenes.config1_camera(64, 48, tick=1), svo.scenes.config1_camera(40, 56, tick=1)]
    views[1].origin[0] += 0.05
    return tree, bases, blocks, views


def _bind(svo, cm, bases, blocks, params, n, target=None):
    B = svo.BufferName
    cm.InitReadBuffer(B.Bases, bases)
    cm.InitReadBuffer(B.Blocks, blocks)
    cm.InitBuffer(B.LeafParams, params)
    cm.InitDeviceBuffer(B.OutputImage, n, 16)
    cm.InitDeviceBuffer(B.Layers, n * K, 4)
    cm.InitBuffer(B.LeafGradients, np.zeros_like(params))
    if target is not None:
        t4 = np.zeros((n, 4), np.float32)
        t4[:, :3] = target
        cm.InitBuffer(B.TargetImage, t4)
    else:
        try:
            cm.ReleaseBuffer(B.TargetImage)
        except svo.SvoError:
            pass


def test_layers_and_image_match_oracle(svo, orc, gpu, setup):
    tree, bases, blocks, views = setup
    n = sum(v.screen_size[0] * v.screen_size[1] for v in views)
    params = _perturbed(_params_from_blocks(blocks), 1)
    _bind(svo, gpu, bases, blocks, params, n)
    gpu.Composite(views, max_layers=K, layers=True)
    gpu.Wait()
    B = svo.BufferName
    got_layers = gpu.ReadBuffer(B.Layers, np.uint32, n * K).reshape(n, K)
    got_image = gpu.ReadBuffer(B.OutputImage, np.float32, 4 * n).reshape(n, 4)
    layers, bg = orc.composite_layers(views, bases, blocks, K, threads=8)
    assert np.array_equal(got_layers, layers)
    _, image, _ = orc.composite_loss(layers, bg, params.astype(np.float64), want_image=True)
    assert np.max(np.abs(got_image[:, :3] - image)) <= 1e-5
    assert np.all(got_image[:, 3] == 1.0)


def test_opaque_composite_equals_plain_frame(svo, orc, gpu, setup):
    tree, bases, blocks, views = setup
    v = views[0]
    n = v.screen_size[0] * v.screen_size[1]
    params = _params_from_blocks(blocks)
    params[:, 3] = 1.0
    _bind(svo, gpu, bases, blocks, params, n)
    gpu.Composite([v], max_layers=K)
    gpu.Wait()
    got = gpu.ReadBuffer(svo.BufferName.OutputImage, np.float32, 4 * n).reshape(n, 4)
    ref = orc.trace_frame(v, bases, blocks, None, threads=8)
    assert np.max(np.abs(got - ref["rgba"])) <= 1e-6


def test_adjoint_pass_is_also_the_forward(svo, orc, gpu, setup):
    """One adjoint launch computes the forward composite in the same traversal: its image and
    layers equal the forward launch's bit for bit and its loss to the rounding of a sum (an
    optimisation step is one launch)."""
    tree, bases, blocks, views = setup
    n = sum(v.screen_size[0] * v.screen_size[1] for v in views)
    truth = _params_from_blocks(blocks)
    params = _perturbed(truth, 3)
    layers, bg = orc.composite_layers(views, bases, blocks, K, threads=8)
    _, target, _ = orc.composite_loss(layers, bg, truth.astype(np.float64), want_image=True)
    _bind(svo, gpu, bases, blocks, params, n, target=target.astype(np.float32))
    B = svo.BufferName
    out = []
    for adjoint in (False, True):
        loss = gpu.Composite(views, max_layers=K, adjoint=adjoint, layers=True)
        gpu.Wait()
        out.append((loss, gpu.ReadBuffer(B.OutputImage, np.float32, 4 * n), gpu.ReadBuffer(B.Layers, np.uint32, n * K)))
    assert abs(out[0][0] - out[1][0]) <= 1e-12 * abs(out[0][0])  # a double sum; the order of its atomics varies
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])
    assert np.abs(gpu.ReadBuffer(B.LeafGradients, np.float32)).sum() > 0


def test_adjoint_matches_fp64_and_finite_differences(svo, orc, gpu, setup):
    tree, bases, blocks, views = setup
    n = sum(v.screen_size[0] * v.screen_size[1] for v in views)
    truth = _params_from_blocks(blocks)
    params = _perturbed(truth, 2)
    layers, bg = orc.composite_layers(views, bases, blocks, K, threads=8)
    _, target, _ = orc.composite_loss(layers, bg, truth.astype(np.float64), want_image=True)
    _bind(svo, gpu, bases, blocks, params, n, target=target.astype(np.float32))
    loss = gpu.Composite(views, max_layers=K, adjoint=True)
    gpu.Wait()
    got = gpu.ReadBuffer(svo.BufferName.LeafGradients, np.float32, 4 * blocks.size).reshape(-1, 4)
    # FP64 adjoint on the FP32 target the device saw
    ref_loss, _, ref = orc.composite_loss(layers, bg, params.astype(np.float64),
                                          target=target.astype(np.float32).astype(np.float64), want_grads=True)
    assert abs(loss - ref_loss) <= 1e-4 * abs(ref_loss)
    used = np.unique(layers[layers != U32MAX])
    assert np.linalg.norm(got[used] - ref[used]) <= 1e-3 * np.linalg.norm(ref[used])
    # central differences on random leaves (FP64, only the cones that composite each leaf)
    rng = np.random.default_rng(5)
    pick = rng.choice(used, size=min(200, used.size), replace=False)
    h = 1e-3
    p64 = params.astype(np.float64)
    t64 = target.astype(np.float32).astype(np.float64)
    for a in pick:
        rays = np.nonzero((layers == a).any(axis=1))[0].astype(np.uint32)
        for ch in range(4):
            p1, p0 = p64.copy(), p64.copy()
            p1[a, ch] += h
            p0[a, ch] -= h
            fd = (orc.composite_loss(layers, bg, p1, target=t64, rays=rays)[0] -
                  orc.composite_loss(layers, bg, p0, target=t64, rays=rays)[0]) / (2 * h)
            scale = max(abs(fd), 1e-2 * np.max(np.abs(ref[used])))
            assert abs(got[a, ch] - fd) <= 1e-3 * scale, (int(a), ch, float(got[a, ch]), fd)


def test_adam_step_matches_numpy(svo, gpu, setup):
    tree, bases, blocks, views = setup
    B = svo.BufferName
    rng = np.random.default_rng(7)
    p = rng.uniform(0, 1, (1000, 4)).astype(np.float32)
    g = rng.normal(0, 1, (1000, 4)).astype(np.float32)
    gpu.InitBuffer(B.LeafParams, p)
    gpu.InitBuffer(B.LeafGradients, g)
    gpu.InitBuffer(B.AdamM, np.zeros_like(p))
    gpu.InitBuffer(B.AdamV, np.zeros_like(p))
    gpu.AdamStep(1, learning_rate=0.01)
    gpu.Wait()
    got = gpu.ReadBuffer(B.LeafParams, np.float32, p.size).reshape(-1, 4)
    m = 0.1 * g
    v = 0.001 * g * g
    want = np.clip(p - 0.01 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8), 0, 1)
    assert np.max(np.abs(got - want)) <= 1e-6
    assert np.all(gpu.ReadBuffer(B.LeafGradients, np.float32, p.size) == 0)


def test_fitting_reduces_the_loss(svo, orc, gpu, setup):
    """A few Adam steps on the perturbed parameters bring the image towards the target."""
    tree, bases, blocks, views = setup
    n = sum(v.screen_size[0] * v.screen_size[1] for v in views)
    truth = _params_from_blocks(blocks)
    params = _perturbed(truth, 3)
    layers, bg = orc.composite_layers(views, bases, blocks, K, threads=8)
    _, target, _ = orc.composite_loss(layers, bg, truth.astype(np.float64), want_image=True)
    _bind(svo, gpu, bases, blocks, params, n, target=target.astype(np.float32))
    B = svo.BufferName
    gpu.InitBuffer(B.AdamM, np.zeros_like(params))
    gpu.InitBuffer(B.AdamV, np.zeros_like(params))
    losses = []
    for step in range(1, 21):
        losses.append(gpu.Composite(views, max_layers=K, adjoint=True))
        gpu.AdamStep(step, learning_rate=0.02)
    assert losses[-1] < 0.5 * losses[0], losses


@pytest.fixture(scope="module")
def config5(svo, orc):
    """The depth-10 sphere+noise scene (seed 3) and its 32 views of 512x512, with the FP64
    layer lists of every cone."""
    spec = svo.scenes.config5_scene()
    tree, bases, blocks = tree_arrays(svo, spec)
    views = svo.scenes.config5_cameras()
    for v in views:
        v.tick = 1
    layers, bg = orc.composite_layers(views, bases, blocks, K, threads=16)
    return tree, bases, blocks, views, layers, bg


def test_adjoint_at_full_size_on_1000_leaves(svo, orc, gpu, config5):
    """32 views of 512^2 over the depth-10 scene: targets rendered from the true leaves,
    leaves perturbed by N(0, 0.1) (colour) and N(0, 0.05) (opacity); the device gradient
    against the FP64 adjoint (relative error <= 1e-3) and, on 1000 random composited leaves,
    against central differences with h = 1e-3."""
    tree, bases, blocks, views, layers, bg = config5
    n = layers.shape[0]
    truth = _params_from_blocks(blocks)
    params = _perturbed(truth, 3)
    B = svo.BufferName
    _bind(svo, gpu, bases, blocks, truth, n)
    gpu.Composite(views, max_layers=K, layers=True)
    gpu.Wait()
    assert np.array_equal(gpu.ReadBuffer(B.Layers, np.uint32, n * K).reshape(n, K), layers)
    target = gpu.ReadBuffer(B.OutputImage, np.float32, 4 * n).reshape(n, 4)
    _, ref_image, _ = orc.composite_loss(layers, bg, truth.astype(np.float64), want_image=True)
    assert np.max(np.abs(target[:, :3] - ref_image)) <= 1e-5
    _bind(svo, gpu, bases, blocks, params, n, target=target[:, :3])
    loss = gpu.Composite(views, max_layers=K, adjoint=True)
    gpu.Wait()
    got = gpu.ReadBuffer(B.LeafGradients, np.float32, 4 * blocks.size).reshape(-1, 4).astype(np.float64)
    p64 = params.astype(np.float64)
    t64 = target[:, :3].astype(np.float64)
    ref_loss, _, ref = orc.composite_loss(layers, bg, p64, target=t64, want_grads=True)
    assert abs(loss - ref_loss) <= 1e-4 * abs(ref_loss)
    used = np.unique(layers[layers != U32MAX])
    assert np.linalg.norm(got[used] - ref[used]) <= 1e-3 * np.linalg.norm(ref[used])
    # the cones that composite each leaf, from one sort of the layer table
    flat = layers.reshape(-1)
    order = np.argsort(flat, kind="stable")
    sorted_leaves = flat[order]
    rng = np.random.default_rng(11)
    pick = rng.choice(used, size=1000, replace=False)
    floor = 1e-2 * np.max(np.abs(ref[used]))
    h = 1e-3
    work = p64.copy()
    for a in pick:
        lo, hi = np.searchsorted(sorted_leaves, [a, a + 1])
        rays = np.unique(order[lo:hi] // K).astype(np.uint32)
        for ch in range(4):
            x = work[a, ch]
            work[a, ch] = x + h
            l1 = orc.composite_loss(layers, bg, work, target=t64, rays=rays)[0]
            work[a, ch] = x - h
            l0 = orc.composite_loss(layers, bg, work, target=t64, rays=rays)[0]
            work[a, ch] = x
            fd = (l1 - l0) / (2 * h)
            scale = max(abs(ref[a, ch]), floor)
            assert abs(ref[a, ch] - fd) <= 1e-3 * scale, ("fp64 adjoint vs differences", int(a), ch, ref[a, ch], fd)
            assert abs(got[a, ch] - ref[a, ch]) <= 1e-3 * scale, ("device vs fp64", int(a), ch, got[a, ch], ref[a, ch])
