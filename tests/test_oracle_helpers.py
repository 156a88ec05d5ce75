"""Pins the oracle's fixed-point helpers (wavefront.cl:197-437) against closed forms,
exact integer/rational arithmetic and exhaustive enumeration, independently of the
oracle's own code."""
import math
import random
from fractions import Fraction

import numpy as np
import pytest

U64 = (1 << 64) - 1
TWO64 = 1 << 64


def f32(x: float) -> float:
    return float(np.float32(x))


def rne_u64_to_f32(u: int) -> Fraction:
    """Exact round-to-nearest-even of a u64 to 24 significant bits."""
    if u == 0:
        return Fraction(0)
    e = u.bit_length() - 24
    if e <= 0:
        return Fraction(u)
    q, r = divmod(u, 1 << e)
    half = 1 << (e - 1)
    if r > half or (r == half and (q & 1)):
        q += 1
    return Fraction(q << e)


def test_pow_sum_closed_form(orc):
    for d in range(0, 11):
        assert orc.pow_sum(d) == (8 ** (d + 1) - 8) // 7
    # the reference's constant-shift form, wavefront.cl:224, valid for d <= 11
    for d in range(0, 11):
        assert orc.pow_sum(d) == (0b1001001001001001001001001001001000 >> ((11 - d) * 3)) - 1


def interleave(x: int, y: int, z: int) -> int:
    """TreeBuilder.Interleave (TreeBuilder.cs:316-326) restated."""
    out, i = 0, 0
    while x >= (1 << i) or y >= (1 << i) or z >= (1 << i):
        out += ((x >> i) & 1) << (3 * i)
        out += ((y >> i) & 1) << (3 * i + 1)
        out += ((z >> i) & 1) << (3 * i + 2)
        i += 1
    return out


@pytest.mark.parametrize("depth", [1, 2, 3, 4])
def test_base_location_is_morton_exhaustive(orc, depth):
    rng = random.Random(depth)
    n = 1 << depth
    for x in range(n):
        for y in range(n):
            for z in range(n):
                low = [rng.getrandbits(64 - depth) for _ in range(3)]
                loc = [(x << (64 - depth)) | low[0], (y << (64 - depth)) | low[1], (z << (64 - depth)) | low[2]]
                assert orc.base_location(depth, loc) == interleave(x, y, z)


def test_base_location_depth6_sampled(orc):
    rng = random.Random(6)
    for _ in range(4000):
        x, y, z = (rng.getrandbits(6) for _ in range(3))
        loc = [(v << 58) | rng.getrandbits(58) for v in (x, y, z)]
        assert orc.base_location(6, loc) == interleave(x, y, z)


def test_chunk_position_bits(orc):
    rng = random.Random(1)
    for _ in range(2000):
        loc = [rng.getrandbits(64) for _ in range(3)]
        d = rng.randrange(0, 63)
        want = sum(((loc[a] >> (63 - d)) & 1) << a for a in range(3))
        assert orc.chunk_position(d, loc) == want


def test_round_ulong_lands_on_boundaries(orc):
    rng = random.Random(2)
    values = [0, 1, U64, U64 - 1, 1 << 63, (1 << 63) - 1] + [rng.getrandbits(64) for _ in range(3000)]
    for v in values:
        for d in (0, 1, 3, 7, 12, 20, 40):
            cell = 1 << (63 - d)  # size of a level-(d+1) cell
            up = orc.round_ulong(v, d, True)
            assert 0 < up <= cell
            assert (v + up) % cell == 0  # lands exactly on the next boundary (mod 2^64)
            down = orc.round_ulong(v, d, False)
            start = v - v % cell
            if start == 0:
                assert down == v  # first cell: distance to 0
            else:
                assert v - down == start - 1  # last unit of the previous cell
                assert 0 < down <= cell


def test_float_to_ulong_exact_scaling(orc):
    rng = random.Random(3)
    thr = f32(0.999999911)
    samples = [f32(v) for v in (0.0, -0.0, 1e-30, 0.5, 0.25, thr, np.nextafter(np.float32(thr), np.float32(0)).item(),
                                1.0, 2.0, -0.5, -0.999)]
    samples += [f32(rng.random()) for _ in range(4000)]
    samples += [f32(rng.random() ** 8) for _ in range(1000)]
    for x in samples:
        got = orc.float_to_ulong(x)
        if x >= thr:
            assert got == U64
        else:
            want = math.floor(Fraction(abs(x)) * TWO64)
            assert got == min(want, U64), x


def test_ulong_to_float_exact_scaling(orc):
    rng = random.Random(4)
    values = [0, 1, 2, 3, (1 << 24) + 1, (1 << 25) + 3, U64, 1 << 63, (1 << 40) + (1 << 16)]
    values += [rng.getrandbits(rng.randrange(1, 65)) for _ in range(5000)]
    for u in values:
        want = rne_u64_to_f32(u) / TWO64
        assert Fraction(orc.ulong_to_float(u)) == want, u


def test_round_trip_float_ulong(orc):
    """ulongToFloat(floatToUlong(x)) == x for x in [0, threshold): exact scalings."""
    rng = random.Random(5)
    for _ in range(3000):
        x = f32(rng.random() * 0.99)
        assert orc.ulong_to_float(orc.float_to_ulong(x)) == x


def test_traverse_then_pop_invariants(orc):
    """A traverseChunk step ends exactly on a level-(d+1) boundary along at least one axis,
    moves every axis in its sign direction, keeps the direction ratios, and
    comparePositions returns the deepest level whose cell the step did not leave."""
    rng = random.Random(6)
    checked = 0
    for _ in range(3000):
        d = rng.randrange(1, 16)
        direction = [rng.uniform(-1, 1) for _ in range(3)]
        norm = math.sqrt(sum(v * v for v in direction))
        direction = [f32(v / norm) for v in direction]
        if min(abs(v) for v in direction) < 1e-3:
            continue
        inv = [f32(1.0 / v) for v in direction]
        sign = [1 if v >= 0 else 0 for v in direction]
        loc = [rng.getrandbits(64) >> 2 | (1 << 62) for _ in range(3)]  # keep away from the cube faces
        new = orc.traverse_chunk(d, loc, direction, inv, sign)
        cell = 1 << (63 - d)
        on_boundary = any((new[a] % cell == 0) if sign[a] else (new[a] % cell == cell - 1) for a in range(3))
        assert on_boundary
        delta = [(new[a] - loc[a]) if sign[a] else (loc[a] - new[a]) for a in range(3)]
        assert all(0 <= v <= cell for v in delta)
        dom = max(range(3), key=lambda a: delta[a] / abs(direction[a]))
        for a in range(3):
            if delta[dom] > (1 << 30):
                ratio = delta[a] / delta[dom]
                expect = abs(direction[a] / direction[dom])
                assert abs(ratio - expect) <= 1e-5 * expect + 1e-9
        dd = orc.compare_positions(d, loc, new, sign)
        assert 1 <= dd <= d
        if dd > 1:
            assert all((loc[a] >> (64 - dd)) == (new[a] >> (64 - dd)) for a in range(3))
        if dd < d:
            assert any((loc[a] >> (63 - dd)) != (new[a] >> (63 - dd)) for a in range(3))
        checked += 1
    assert checked > 2000


def test_camera_centre_pixel_is_forward(svo, orc):
    cam = svo.scenes.config1_camera(64, 48)
    hit, d, inv, loc, fov = orc.init_ray(cam, 32, 24)
    assert d == [cam.facing[0], cam.facing[1], cam.facing[2]]
    assert fov == f32(cam.fov[0] / 64)


def test_camera_directions_unit_length(svo, orc):
    cam = svo.scenes.config1_camera(40, 30)
    for y in range(0, 30, 7):
        for x in range(0, 40, 7):
            _, d, inv, _, _ = orc.init_ray(cam, x, y)
            assert abs(math.sqrt(sum(v * v for v in d)) - 1.0) < 2e-6
            assert all(inv[a] == f32(1.0 / d[a]) for a in range(3))


def test_camera_lut_matches_product_host_table(svo, orc):
    cam = svo.scenes.config2_camera(3, 320, 180)
    a = orc.camera_lut(cam)
    b = svo.camera_lut(cam)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_trace_into_volume_against_exact_slab(svo, orc):
    """traceIntoVolume vs an exact double-precision slab test, away from the cone margin."""
    rng = random.Random(7)
    agree = total = 0
    for _ in range(8000):
        eye = [rng.uniform(-1.5, 2.5) for _ in range(3)]
        if all(0 <= v <= 1 for v in eye):
            continue
        target = [rng.uniform(-0.2, 1.2) for _ in range(3)]
        facing = svo.scenes.look_at(eye, target)
        cam = svo.scenes.trace_input(eye, facing, 1, 1, 1e-4, 1e-4, 4)
        hit, d, inv, loc, fov = orc.init_ray(cam, 0, 0)
        o = [f32(v) for v in eye]
        t0, t1 = -math.inf, math.inf
        for a in range(3):
            if d[a] == 0:
                if not (0 <= o[a] <= 1):
                    t0, t1 = 1, 0
                continue
            ta, tb = (0 - o[a]) / d[a], (1 - o[a]) / d[a]
            t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
        exact = t0 <= t1 and t1 >= 0
        # distance of the entry point from the cube's edges decides whether the margin matters
        if exact:
            p = [o[a] + d[a] * t0 for a in range(3)]
            margin = min(min(v, 1 - v) for v in p)
        else:
            margin = 1.0
        c = fov * max(t0, 0) if exact else 0
        if exact and margin < c + 1e-5:
            continue
        if not exact:
            # a miss can still pass the reference's acceptance test within the cone margin;
            # only count clear misses
            tmid = max(t0, 0)
            p = [o[a] + d[a] * tmid for a in range(3)]
            if all(-1e-3 <= v <= 1 + 1e-3 for v in p):
                continue
        total += 1
        if hit == exact:
            agree += 1
            if exact:
                for a in range(3):
                    assert abs(loc[a] / TWO64 - min(max(p[a], 0.0), 1.0)) < 1e-5
    assert total > 1500
    assert agree == total
