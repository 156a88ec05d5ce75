"""Pinning tests of the oracle's shading kernels (CPU only): evaluateBackground with the light
mix of non-primary cones (wavefront.cl:1741-1797), resolveRemainders (1799-1833) and drawTrace
(1835-1850).

Each branch is pinned by what its closed form fixes: a weighting of 255 takes the light's
colour unchanged; with ambient 0 the mix moves the base colour towards white by light/255 of
the distance (base + light (255 - base) / 255); an ambient above the luminosity makes the
boundary 0, and the reference's unsigned subtraction 0 - base wraps to 2^32 - base, so any
non-zero base saturates to 255 while a zero base stays 0; a luminosity equal to the ambient
leaves the base unchanged; the sun's luminosity is trunc(200 d).  The native_divide of the
reference is a correctly rounded division here (its device precision stays unpinned)."""
from __future__ import annotations

import math

import numpy as np
import pytest

SKY = (135, 206, 235)


def _final(n):
    return {k: np.zeros(n, np.uint32) for k in ("final_r", "final_g", "final_b", "final_w")}


def _traces(dirs, parents, weights, base):
    n = len(parents)
    d = np.zeros((n, 4), np.float32)
    d[:, :3] = np.asarray(dirs, np.float32)
    base = np.asarray(base, np.uint8).reshape(n, 3)
    return {"directions": d.reshape(-1), "parents": np.asarray(parents, np.uint32),
            "weightings": np.asarray(weights, np.uint8), "colour_r": base[:, 0].copy(),
            "colour_g": base[:, 1].copy(), "colour_b": base[:, 2].copy()}


def _mix(base, light, lum, amb):
    """The light mix in exact arithmetic, before the truncation to 0..255."""
    if lum == amb:
        return float(base)
    if lum > amb:
        return base + abs(lum - amb) / ((lum + amb) * 255.0) * light * (255 - base)
    return base + abs(lum - amb) / ((lum + amb) * 255.0) * light * ((0 - base) % 2**32)


def _one(orc, direction, weight, base, ambient):
    f = _final(1)
    orc.evaluate_background(ambient, np.array([0]), _traces([direction], [0], [weight], [base]), f)
    w = int(f["final_w"][0])
    return w, [int(f[k][0]) // max(w, 1) for k in ("final_r", "final_g", "final_b")], f


def test_primary_weighting_takes_the_light_unchanged(orc):
    w, c, f = _one(orc, (0.0, 0.6, 0.8), 255, (12, 34, 56), 0)  # d = -0.8: sky
    assert w == 255 and c == list(SKY)
    w, c, _ = _one(orc, (0.0, 0.3, -0.95), 255, (12, 34, 56), 0)  # d = 0.95 > 0.8: sun
    assert w == 255 and c == [255, 255, 255]
    rgba = orc.draw_trace(f)
    assert np.array_equal(rgba[0], np.float32([135 / 255.0, 206 / 255.0, 235 / 255.0, 1.0]))


@pytest.mark.parametrize("base", [(0, 0, 0), (10, 128, 250), (77, 3, 200), (255, 255, 255)])
def test_mix_towards_white_under_no_ambient(orc, base):
    # sky: luminosity 1 > ambient 0, boundary 255, quotient 1/255
    w, c, _ = _one(orc, (0.0, 0.0, 1.0), 100, base, 0)
    assert w == 100
    for k in range(3):
        exact = base[k] + SKY[k] * (255 - base[k]) / 255.0
        assert exact == pytest.approx(_mix(base[k], SKY[k], 1, 0))
        got = c[k]
        if abs(exact - round(exact)) > 1e-3:
            assert got == min(255, math.floor(exact)), (k, got, exact)
        else:
            assert abs(got - exact) <= 1


@pytest.mark.parametrize("ambient", [2, 9, 200])
def test_ambient_above_the_luminosity_wraps_and_saturates(orc, ambient):
    # boundary 0: (0 - base) is 2^32 - base as an unsigned int, so the mix overshoots to 255
    w, c, _ = _one(orc, (0.0, 0.0, 1.0), 60, (0, 1, 254), ambient)
    assert w == 60
    assert c == [0, 255, 255]
    assert _mix(1, 206, 1, ambient) > 255 and _mix(0, 135, 1, ambient) == 0


def test_luminosity_equal_to_ambient_keeps_the_base(orc):
    # sky luminosity 1 == ambient 1: quotient 0, boundary 0 (the wrapped distance times 0)
    w, c, _ = _one(orc, (0.0, 0.0, 1.0), 31, (17, 200, 3), 1)
    assert w == 31 and c == [17, 200, 3]


@pytest.mark.parametrize("dz,ambient", [(-0.9, 100), (-0.9, 180), (-0.9, 181), (-0.99, 198)])
def test_sun_luminosity_is_200_d(orc, dz, ambient):
    d = (0.0, math.sqrt(1.0 - dz * dz), dz)
    lum = int(np.float32(np.float32(-dz) * np.float32(200.0)))  # trunc(200 d): 180, 198
    w, c, _ = _one(orc, d, 50, (40, 0, 255), ambient)
    for k, b in enumerate((40, 0, 255)):
        exact = _mix(b, 255, lum, ambient)
        want = 255 if exact >= 255 else (0 if exact <= 0 else math.floor(exact))
        assert abs(c[k] - want) <= (1 if abs(exact - round(exact)) < 1e-3 else 0), (k, c[k], exact, lum)


def test_accumulation_skips_and_wraps(orc):
    dirs = [(0.0, 0.0, 1.0)] * 5
    t = _traces(dirs, [3, 3, 7, 3, 99], [255, 255, 255, 10, 255], [(0, 0, 0)] * 5)
    f = _final(8)
    # ids 0, 1, 3 to pixel 3; id 2 to pixel 7; id 4's parent is out of range; id 9 does not exist
    orc.evaluate_background(0, np.array([0, 1, 2, 3, 4, 9]), t, f)
    assert f["final_w"].tolist() == [0, 0, 0, 255 + 255 + 10, 0, 0, 0, 255, ]
    mixed = [math.floor(SKY[k] * 255 / 255.0) for k in range(3)]  # base 0 mixes to the sky colour
    assert f["final_r"][3] == 2 * 135 * 255 + mixed[0] * 10
    # u32 accumulators wrap
    f2 = _final(1)
    f2["final_w"][0] = 2**32 - 100
    orc.evaluate_background(0, np.array([0]), _traces([(0, 0, 1)], [0], [255], [(0, 0, 0)]), f2)
    assert int(f2["final_w"][0]) == 155


def test_draw_trace_closed_form(orc):
    f = _final(3)
    f["final_w"][:] = [0, 2, 1]
    f["final_r"][:] = [9, 2 * 255 * 255, 1]
    f["final_g"][:] = [9, 100, 0]
    f["final_b"][:] = [9, 2 * 255 * 300, 255 * 255]
    rgba = orc.draw_trace(f)
    assert np.array_equal(rgba[0], np.float32([0.75, 0, 0, 1]))  # no weighting: red
    assert rgba[1, 0] == 1.0 and rgba[1, 2] == 1.0  # clamped at 1
    assert rgba[1, 1] == np.float32(np.float32(100) / np.float32(2 * 255.0))
    assert rgba[2, 0] == np.float32(np.float32(1) / np.float32(255.0)) and rgba[2, 2] == 1.0
    assert np.all(rgba[:, 3] == 1.0)


def test_resolve_remainders_known_answer(svo, orc):
    # B = 1: dense level 3 (512 blocks); block m at level 3 -> child group G at level 4
    B = 1
    blocks = np.zeros(512 + 16, dtype=svo.BLOCK_DTYPE)
    blocks["child"] = 0xFFFFFFFF
    x, y, z = 0x6000_0000_0000_0000, 0x2000_0000_0000_0000, 0xA000_0000_0000_0000  # top bits 011, 001, 101
    lvl3 = lambda v: v >> 61
    m = 0
    for i in range(3):
        m |= ((lvl3(x) >> i) & 1) << (3 * i) | ((lvl3(y) >> i) & 1) << (3 * i + 1) | ((lvl3(z) >> i) & 1) << (3 * i + 2)
    blocks["child"][m] = 512
    oct3 = ((x >> 60) & 1) | (((y >> 60) & 1) << 1) | (((z >> 60) & 1) << 2)  # chunkPosition(3)
    a4 = 512 + oct3
    blocks["colour_r"][a4], blocks["colour_g"][a4], blocks["colour_b"][a4] = 11, 22, 33
    blocks["colour_r"][m], blocks["colour_g"][m], blocks["colour_b"][m] = 1, 2, 3
    loc = np.zeros(4 * 4, np.uint64)
    for i in range(4):
        loc[4 * i:4 * i + 3] = (x, y, z)
    t = {"locations": loc, "depths": np.array([4, 3, 2, 61], np.uint8), "parents": np.array([0, 1, 0, 1], np.uint32),
         "weightings": np.array([7, 255, 9, 9], np.uint8)}
    f = _final(2)
    orc.resolve_remainders(B, blocks, np.array([0, 1, 2, 3]), t, f)
    # depth 4: the child block; depth 3: the dense block; depth 2 < B + 2 and 61 > 60 skipped
    assert f["final_w"].tolist() == [7, 255]
    assert [int(f[k][0]) for k in ("final_r", "final_g", "final_b")] == [77, 154, 231]
    assert [int(f[k][1]) for k in ("final_r", "final_g", "final_b")] == [255, 510, 765]
