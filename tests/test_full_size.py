"""GPU parity at the measured workloads' full sizes: the 1920x1080 plain frame on the
depth-12 fBm heightfield (54.4M blocks) and the 64-camera batch on the same scene, bit-exact
against the oracle on every per-cone output, the child requests, the queues, the usage
ticks and the step totals.  The oracle runs on 16 threads (a few seconds per frame)."""
from __future__ import annotations

import numpy as np
import pytest

from frame_helpers import assert_frame_matches, gpu_frame
from test_batch_parity import _compare, gpu_batch

pytestmark = pytest.mark.gpu

THREADS = 16


@pytest.fixture(scope="module")
def config2_tree(svo):
    spec = svo.scenes.config2_scene(max_depth=12)
    tree = spec.build()
    return tree, tree.bases, tree.blocks


@pytest.mark.parametrize("frame", [0, 63])
def test_plain_frame_1080p_depth12(svo, orc, gpu, config2_tree, frame):
    tree, bases, blocks = config2_tree
    cam = svo.scenes.config2_camera(frame, 1920, 1080, frames=100, tick=frame + 1)
    usages = tree.initial_usages(0)
    ref_usages = usages.copy()
    cap = 1920 * 1080
    ref = orc.trace_frame(cam, bases, blocks, ref_usages, request_capacity=cap, threads=THREADS)
    out = gpu_frame(svo, gpu, bases, blocks, usages, cam, request_capacity=cap)
    assert_frame_matches(out, ref, usages_ref=ref_usages)
    assert ref["request_count"] > 100000 and (ref["status"] & 1).sum() > cap // 4


def test_batch_64_cameras_full_size(svo, orc, gpu, config2_tree):
    tree, bases, blocks = config2_tree
    cams, apertures = svo.scenes.config4_cameras(64, tick=5)
    usages = tree.initial_usages(0)
    ref_usages = usages.copy()
    ref = orc.trace_batch(cams, bases, blocks, ref_usages, apertures=apertures, request_capacity=6000,
                          threads=THREADS)
    got = gpu_batch(svo, gpu, bases, blocks, usages, cams, apertures)
    _compare(got, ref, ref_usages)
    assert got["hit"].size > 30_000_000
