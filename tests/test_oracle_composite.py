"""Pinning tests of the differentiable compositing oracle (CPU only): the opaque case
reduces to the plain first-hit frame, the FP64 adjoint equals central finite differences,
and the layer rule (continue past a hit) follows the traversal."""
from __future__ import annotations

import numpy as np
import pytest

from frame_helpers import tree_arrays

U32MAX = 0xFFFFFFFF
K = 4


@pytest.fixture(scope="module")
def small(svo):
    tree, bases, blocks = tree_arrays(svo, svo.scenes.config1_scene())
    cams = [svo.scenes.config1_camera(48, 36, tick=1)]
    return tree, bases, blocks, cams


def _params_from_blocks(blocks):
    p = np.zeros((blocks.size, 4), np.float64)
    p[:, 0] = blocks["colour_r"] / 255.0
    p[:, 1] = blocks["colour_g"] / 255.0
    p[:, 2] = blocks["colour_b"] / 255.0
    p[:, 3] = blocks["opacity"] / 255.0
    return p


def test_first_layer_is_the_plain_hit(svo, orc, small):
    tree, bases, blocks, cams = small
    layers, bg = orc.composite_layers(cams, bases, blocks, K, threads=8)
    ref = orc.trace_frame(cams[0], bases, blocks, None, threads=8)
    hit = ref["hit"]
    assert np.array_equal(layers[:, 0], hit)  # material cones: their block; background: none
    assert np.count_nonzero(layers[:, 1] != U32MAX) > 0  # cones do continue past a hit


def test_opaque_composite_is_first_hit_colour(svo, orc, small):
    """With every opacity 1 the composite is the first layer's colour (or the sky)."""
    tree, bases, blocks, cams = small
    layers, bg = orc.composite_layers(cams, bases, blocks, K, threads=8)
    params = _params_from_blocks(blocks)
    params[:, 3] = 1.0
    _, image, _ = orc.composite_loss(layers, bg, params, want_image=True)
    ref = orc.trace_frame(cams[0], bases, blocks, None, threads=8)
    assert np.allclose(image, ref["rgba"][:, :3], atol=1e-6)


def test_adjoint_equals_central_differences(svo, orc, small):
    tree, bases, blocks, cams = small
    layers, bg = orc.composite_layers(cams, bases, blocks, K, threads=8)
    rng = np.random.default_rng(3)
    params = _params_from_blocks(blocks)
    truth = params.copy()
    params[:, :3] = np.clip(params[:, :3] + rng.normal(0, 0.1, (params.shape[0], 3)), 0, 1)
    params[:, 3] = np.clip(params[:, 3] + rng.normal(0, 0.05, params.shape[0]), 0, 1)
    _, target, _ = orc.composite_loss(layers, bg, truth, want_image=True)
    loss, _, grads = orc.composite_loss(layers, bg, params, target=target, want_grads=True)
    used = np.unique(layers[layers != U32MAX])
    pick = rng.choice(used, size=min(60, used.size), replace=False)
    h = 1e-3
    for a in pick:
        rays = np.nonzero((layers == a).any(axis=1))[0].astype(np.uint32)
        for ch in range(4):
            p1, p0 = params.copy(), params.copy()
            p1[a, ch] += h
            p0[a, ch] -= h
            fd = (orc.composite_loss(layers, bg, p1, target=target, rays=rays)[0] -
                  orc.composite_loss(layers, bg, p0, target=target, rays=rays)[0]) / (2 * h)
            assert abs(fd - grads[a, ch]) <= 1e-6 + 1e-5 * abs(fd), (int(a), ch, fd, grads[a, ch])
