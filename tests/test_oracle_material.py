"""Pinning tests of the evaluateMaterial oracle (secondary cones, wavefront.cl:1568-1739), CPU only.

The reference's sin/cos/normalize are the OpenCL driver's, so the definitions are this
build's (svo_material_oracle.h); they are pinned against libm and by what the geometry
fixes: the first cone is the mirror reflection of the incoming direction, every direction is
a unit vector, the two cones of a ring pair make the same angle with the reflection, the new
location is the next cell along the normal's dominant axis, and the early exits of the
reference (low weighting, a direction parallel to the face, invalid FoV parameters)."""
from __future__ import annotations

import math

import numpy as np
import pytest

U32MAX = 0xFFFFFFFF


def test_sincos_against_libm(orc):
    xs = np.linspace(-8.0, 8.0, 40001).astype(np.float32)
    worst_s = worst_c = 0.0
    for x in xs[::3]:
        s, c = orc.sincos(float(x))
        worst_s = max(worst_s, abs(s - math.sin(float(x))))
        worst_c = max(worst_c, abs(c - math.cos(float(x))))
        assert abs(s * s + c * c - 1.0) < 5e-7
    assert worst_s < 2e-7 and worst_c < 2e-7, (worst_s, worst_c)
    assert orc.sincos(0.0) == (0.0, 1.0)


def _pitch_yaw(n):
    n = np.asarray(n, np.float64)
    n = n / np.linalg.norm(n)
    return int(32767 * math.asin(n[2]) / math.pi), int(32767 * math.atan2(n[1], n[0]) / math.pi)


class Case:
    """One material cone at depth 5 over a B=1 dense level, against one block whose normal
    and colour are given."""

    def __init__(self, svo, normal, direction, weighting=255, capacity=4096):
        self.B, self.depth = 1, 5
        self.blocks = np.zeros(8 ** 5, dtype=svo.BLOCK_DTYPE)
        self.blocks["child"] = U32MAX
        loc = np.array([0x5234_5678_9ABC_DEF0, 0x6A00_0000_0000_0000, 0x4F00_0000_0000_1234], np.uint64)
        # link the dense-level block down to depth 5 along `loc`
        x, y, z = (int(v) for v in loc)
        lvl = lambda v, d: (v >> (64 - d))
        morton = 0
        for i in range(3):
            morton |= ((lvl(x, 3) >> i) & 1) << (3 * i) | ((lvl(y, 3) >> i) & 1) << (3 * i + 1) | ((lvl(z, 3) >> i) & 1) << (3 * i + 2)
        a, nxt = morton, 8 ** 3
        for d in range(3, self.depth):
            self.blocks["child"][a] = nxt
            oct_ = ((x >> (63 - d)) & 1) | (((y >> (63 - d)) & 1) << 1) | (((z >> (63 - d)) & 1) << 2)
            a = nxt + oct_
            nxt += 8
        self.addr = a
        p, yw = _pitch_yaw(normal)
        self.blocks["normal_pitch"][a], self.blocks["normal_yaw"][a] = p, yw
        self.blocks["colour_r"][a], self.blocks["colour_g"][a], self.blocks["colour_b"][a] = 200, 100, 50
        self.usages = np.zeros(self.blocks.size // 8, dtype=svo.USAGE_DTYPE)
        self.usages["tick"] = 3
        d = np.asarray(direction, np.float32)
        self.root = {"directions": np.zeros(4 * 2, np.float32), "locations": np.zeros(4 * 2, np.uint64),
                     "depths": np.zeros(2, np.uint8), "weightings": np.zeros(2, np.uint8),
                     "parents": np.zeros(2, np.uint32)}
        self.root["directions"][4:7] = d
        self.root["locations"][4:7] = loc
        self.root["depths"][1] = self.depth
        self.root["weightings"][1] = weighting
        self.root["parents"][1] = 77
        self.loc = loc
        m = capacity
        self.out = {"origins": np.zeros(4 * m, np.float32), "directions": np.zeros(4 * m, np.float32),
                    "fovs": np.zeros(m, np.float32), "locations": np.zeros(4 * m, np.uint64),
                    "depths": np.zeros(m, np.uint8), "weightings": np.zeros(m, np.uint8),
                    "colour_r": np.zeros(m, np.uint8), "colour_g": np.zeros(m, np.uint8),
                    "colour_b": np.zeros(m, np.uint8), "luminosities": np.full(m, 9, np.uint8),
                    "parents": np.full(m, U32MAX, np.uint32)}
        self.final_w = np.zeros(2, np.uint32)
        self.bg = np.full(m, U32MAX, np.uint32)
        self.counters = np.zeros(2, np.uint32)

    def run(self, svo, orc, **params):
        cam = svo.scenes.trace_input((0, 0, 0), [1, 0, 0, 0, 1, 0, 0, 0, 1], 8, 8, 1.0, 1.0, self.B, tick=9)
        for k, v in params.items():
            setattr(cam, k, v)
        orc.evaluate_material(cam, np.array([1], np.uint32), self.root, self.blocks, self.usages, self.out,
                              self.final_w, self.bg, self.counters)
        return int(self.counters[0]), int(self.counters[1])

    def dirs(self, n):
        return self.out["directions"].reshape(-1, 4)[:n, :3].astype(np.float64)


def _normal_of(case):
    p = float(case.blocks["normal_pitch"][case.addr]) * math.pi / 32767
    yw = float(case.blocks["normal_yaw"][case.addr]) * math.pi / 32767
    return np.array([math.cos(yw) * math.cos(p), math.sin(yw) * math.cos(p), math.sin(p)])


def test_reflection_rings_and_outputs(svo, orc):
    c = Case(svo, normal=(0.3, -0.2, 0.93), direction=(0.6, 0.1, -0.79))
    n_cones, n_bg = c.run(svo, orc)
    assert n_cones > 10 and n_bg <= n_cones
    dirs = c.dirs(n_cones)
    normal = _normal_of(c)
    d = np.array([0.6, 0.1, -0.79], np.float64)
    refl = d - 2 * d.dot(normal) * normal
    assert np.allclose(dirs[0], refl, atol=2e-6)  # the first cone is the mirror reflection
    # rings (every cone after the reflection) are unit vectors
    assert np.allclose(np.linalg.norm(dirs[1:], axis=1), 1.0, atol=2e-6)
    # the two cones of each ring pair make the same angle with the reflection (a rotation
    # about it by +/- the ring angle); seeds have no partner, so find pairs as consecutive
    # cones with equal dot products
    r = refl / np.linalg.norm(refl)
    dots = dirs[1:] @ r
    pairs = sum(1 for i in range(len(dots) - 1) if abs(dots[i] - dots[i + 1]) < 2e-6)
    assert pairs >= (n_cones - 1) // 4
    # outputs of every cone: parent, colour, depth 1, origin = the surface point
    assert np.all(c.out["parents"][:n_cones] == 77)
    w = c.out["weightings"][:n_cones]
    assert np.all(w <= int(255 * 0.5)) and np.all(w > 0)
    stored = w > 5
    assert np.all(c.out["colour_r"][:n_cones][stored] == 200) and np.all(c.out["depths"][:n_cones][stored] == 1)
    assert np.all(c.out["luminosities"][:n_cones][stored] == 0) and np.all(c.out["luminosities"][:n_cones][~stored] == 9)
    assert c.bg[:n_bg].tolist() == np.nonzero(stored)[0].tolist()
    origin = c.out["origins"].reshape(-1, 4)[0, :3]
    assert np.allclose(origin, c.loc.astype(np.float64) / 2.0 ** 64, atol=1e-7)
    # the new location: along the normal to the farthest level-6 cell boundary -- the axis
    # that needs the longest move lands in the next cell, the others move along their sign
    new = c.out["locations"].reshape(-1, 4)[0, :3]
    cell = lambda v: int(v) >> (64 - 6)
    size = 1 << (64 - 6)

    def to_boundary(v, up):
        low = int(v) % size
        return size - low if up else low + 1

    t = [to_boundary(c.loc[i], normal[i] >= 0) / abs(normal[i]) for i in range(3)]
    m = int(np.argmax(t))
    assert cell(new[m]) - cell(c.loc[m]) == (1 if normal[m] >= 0 else -1)
    for i in range(3):
        assert (cell(new[i]) - cell(c.loc[i])) * (1 if normal[i] >= 0 else -1) >= 0
    assert np.all(c.out["locations"].reshape(-1, 4)[:n_cones][stored][:, :3] == new)
    # usage tick of the block's group
    assert int(c.usages["tick"][c.addr >> 3]) == 9


def test_parallel_low_weight_and_invalid_parameters(svo, orc):
    # direction along the face normal (dot >= 0): one cone, unchanged direction, base weighting
    c = Case(svo, normal=(0.0, 0.0, 1.0), direction=(0.1, 0.0, 0.99))
    assert c.run(svo, orc) == (1, 1)
    assert np.allclose(c.dirs(1)[0], [0.1, 0.0, 0.99], atol=1e-7)
    assert int(c.out["weightings"][0]) == 255
    # weighting_constant * base weighting < 6: no cone, FinalWeightings[trace] = base
    c = Case(svo, normal=(0.0, 0.0, 1.0), direction=(0.3, 0.0, -0.95), weighting=10)
    assert c.run(svo, orc) == (0, 0)
    assert int(c.final_w[1]) == 10
    # FovConstant <= 0 or |FovMultiplier| >= 1: the kernel does nothing (not even the tick)
    for params in ({"fov_constant": 0.0}, {"fov_multiplier": 1.0}, {"fov_multiplier": -1.0}):
        c = Case(svo, normal=(0.0, 0.0, 1.0), direction=(0.3, 0.0, -0.95))
        assert c.run(svo, orc, **params) == (0, 0)
        assert int(c.usages["tick"][c.addr >> 3]) == 3


def test_capacity_counts_but_drops(svo, orc):
    c = Case(svo, normal=(0.3, -0.2, 0.93), direction=(0.6, 0.1, -0.79), capacity=4096)
    full, full_bg = c.run(svo, orc)
    c2 = Case(svo, normal=(0.3, -0.2, 0.93), direction=(0.6, 0.1, -0.79), capacity=5)
    n, n_bg = c2.run(svo, orc)
    assert n == full  # the id counter counts every cone
    assert n_bg == int(np.count_nonzero(c.out["weightings"][:5] > 5))  # only stored cones are queued
    assert np.array_equal(c2.out["directions"], c.out["directions"][:20])


def _spine_closed_form(M: float, c: float, k: int):
    """Spine angle s_k and cone half-angle f_k of spine step k (k >= 1), wavefront.cl:1656-1660,
    from the closed form of the recurrence rather than by iterating it.  With t_k = s_{k-1} +
    f_{k-1} (t_1 = c): s_k = (1+M)/(1-M) t_k, f_k = (M t_k + c)/(1-M), and t_{k+1} = s_k + f_k =
    ((1+2M) t_k + c)/(1-M), a linear recurrence solved as t_k = a^(k-1) (c + c/(3M)) - c/(3M)
    with a = (1+2M)/(1-M); at M = 0 it is t_k = k c (so s_k = k c and f_k = c)."""
    if M == 0.0:
        t = k * c
    else:
        a = (1.0 + 2.0 * M) / (1.0 - M)
        t = a ** (k - 1) * (c + c / (3.0 * M)) - c / (3.0 * M)
    return (1.0 + M) / (1.0 - M) * t, (M * t + c) / (1.0 - M)


def _turn(v, axis, angle):
    """The rotation matrix of wavefront.cl:1663-1681 / 1697-1714 in closed form,
    cos(t) I + sin(t) [a]x + (1 - cos(t)) a a^T, applied to v and normalised (float64).  The
    reference builds it from the axis as given: the spine axis is a unit vector, the ring
    axis is the reflection, whose length is that of the incoming direction."""
    r = v * math.cos(angle) + np.cross(axis, v) * math.sin(angle) + axis * axis.dot(v) * (1.0 - math.cos(angle))
    return r / np.linalg.norm(r)


def _expected_cones(normal, d, M, c, WM, WC, base_w):
    """The cones evaluateMaterial emits, derived geometrically in float64: the mirror
    reflection (not normalised); then for spine step k the ring seed -- the reflection turned by the spine
    angle s_k about cross(d, n) in the negative sense -- and ring pairs j = 1, 2, ... (while
    j f_k < pi/2) turned about the reflection by -2 j f_k and +2 j f_k; a spine step or ring
    pair whose first vector points below the surface ends its loop.  Weights: the spine
    step's weighting s_k WM + WC, scaled by dot/f_k where a cone's dot with the normal is
    below f_k (the second vector of a pair takes the first one's), times the base weighting.
    Returns (direction, fov, weight, margin) per cone; margin is the closest any decision
    came to its boundary."""
    refl = d - 2.0 * d.dot(normal) * normal
    axis = np.cross(d, normal)
    axis /= np.linalg.norm(axis)
    out, margins = [], []

    def weight(w, dot, fov):
        return (w * dot / fov if dot < fov else w) * base_w

    dp = refl.dot(normal)
    out.append((refl, c, weight(WC, dp, c)))
    s_prev, k = 0.0, 0
    while s_prev < math.pi and k < 256:
        margins.append(math.pi - s_prev)
        k += 1
        s, f = _spine_closed_form(M, c, k)
        w = s * WM + WC
        seed = _turn(refl, axis, -s)
        dp = seed.dot(normal)
        margins.append(abs(dp))
        if dp < 0.0:
            break
        out.append((seed, f, weight(w, dp, f)))
        j = 1
        while j * f < math.pi / 2:
            margins.append(math.pi / 2 - j * f)
            r1 = _turn(seed, refl, -2.0 * j * f)
            dp = r1.dot(normal)
            margins.append(abs(dp))
            if dp < 0.0:
                break
            out.append((r1, f, weight(w, dp, f)))
            out.append((_turn(seed, refl, 2.0 * j * f), f, weight(w, dp, f)))
            j += 1
        s_prev = s
    return out, min(margins)


@pytest.mark.parametrize("params", [
    {"fov_multiplier": 0.0, "fov_constant": 0.3, "weighting_multiplier": 0.0, "weighting_constant": 0.5},
    {"fov_multiplier": 0.35, "fov_constant": 0.12, "weighting_multiplier": -0.1, "weighting_constant": 0.5},
    {"fov_multiplier": 0.1, "fov_constant": 0.17, "weighting_multiplier": 0.05, "weighting_constant": 0.4},
])
def test_spine_and_ring_schedule_closed_form(svo, orc, params):
    """Every cone of the oracle's schedule against the geometric derivation above: the same
    number of cones, each direction within 1e-4, each cone's FoV the closed-form f_k (a
    spine step that used the new FoV where the reference uses the previous one, or a ring
    angle of j f instead of 2 j f, moves cones by far more), and each weighting within one
    unit of its truncation."""
    normal_in, d = (0.3, -0.2, 0.93), np.array([0.6, 0.1, -0.79], np.float64)
    c = Case(svo, normal=normal_in, direction=d)
    n_cones, _ = c.run(svo, orc, **params)
    normal = _normal_of(c)
    d32 = d.astype(np.float32).astype(np.float64)
    exp, margin = _expected_cones(normal, d32, params["fov_multiplier"], params["fov_constant"],
                                  params["weighting_multiplier"], params["weighting_constant"], 255.0)
    assert margin > 1e-3, "the case sits too close to a loop or dot-product boundary to be decided in float32"
    assert n_cones == len(exp)
    dirs = c.dirs(n_cones)
    fovs = c.out["fovs"][:n_cones].astype(np.float64)
    w = c.out["weightings"][:n_cones].astype(np.int64)
    for i, (e_dir, e_fov, e_w) in enumerate(exp):
        assert np.allclose(dirs[i], e_dir, atol=1e-4), (i, dirs[i], e_dir)
        assert abs(fovs[i] - e_fov) <= 1e-5 * max(1.0, e_fov), (i, fovs[i], e_fov)
        assert abs(int(w[i]) - min(255, max(0, int(math.floor(e_w))))) <= 1, (i, w[i], e_w)
