"""floatToUlong / ulongToFloat (wavefront.cl:200-209), the fixed-point conversions every step
uses.  CPU: the oracle's floatToUlong over every float bit pattern in [0, 1] against its closed
form (floor(x * 2^64), saturating at the 0.999999911 threshold).  GPU: the device's
conversions (svo_convert_probe) equal the oracle's on a stride sample of all floats in
[-2, 2], on the boundary cases, and on u64 values of every bit length."""
from __future__ import annotations

import numpy as np
import pytest

U64 = np.uint64(0xFFFFFFFFFFFFFFFF)
THR_BITS = np.float32(0.999999911).view(np.uint32)


def closed_form_float_to_ulong(bits: np.ndarray) -> np.ndarray:
    """floor(|x| * 2^64) from the bit pattern of a float in [0, 1] (x = m 2^(e-150)), U64MAX
    from the threshold up."""
    b = bits.astype(np.uint64)
    e = (b >> np.uint64(23)).astype(np.int64)
    m = (b & np.uint64(0x7FFFFF)) | np.uint64(0x800000)
    s = e - 86
    left = np.where(s >= 0, s, 0).astype(np.uint64)
    right = np.where(s < 0, np.minimum(-s, 63), 0).astype(np.uint64)
    v = np.where(s >= 0, m << left, m >> right)
    v = np.where(e == 0, np.uint64(0), v)  # zero and the denormals: below 2^-64
    return np.where(bits >= THR_BITS, U64, v)


def test_oracle_float_to_ulong_every_float_in_unit_interval(orc):
    top = int(np.float32(1.0).view(np.uint32))
    step = 1 << 24
    for start in range(0, top + 1, step):
        bits = np.arange(start, min(start + step, top + 1), dtype=np.uint32)
        got = orc.float_to_ulong_n(bits.view(np.float32))
        want = closed_form_float_to_ulong(bits)
        bad = np.nonzero(got != want)[0]
        assert bad.size == 0, (hex(int(bits[bad[0]])), int(got[bad[0]]), int(want[bad[0]]))


def _u64_samples(rng, n):
    lens = rng.integers(1, 65, n)
    hi = rng.integers(0, 1 << 32, n, dtype=np.uint64)
    lo = rng.integers(0, 1 << 32, n, dtype=np.uint64)
    v = (hi << np.uint64(32)) | lo
    mask = np.where(lens >= 64, U64, (np.uint64(1) << lens.astype(np.uint64)) - np.uint64(1))
    edges = [0, 1, 2, 3, (1 << 24) - 1, 1 << 24, (1 << 24) + 1, (1 << 25) + 1, (1 << 25) + 3, (1 << 40) + (1 << 16),
             (1 << 63) - 1, 1 << 63, (1 << 64) - 1, (1 << 64) - (1 << 39), (1 << 64) - (1 << 40) - 1]
    return np.concatenate([v & mask, np.array(edges, dtype=np.uint64)])


@pytest.mark.gpu
def test_device_conversions_equal_the_oracle(orc, gpu):
    # every 97th float bit pattern in [-2, 2] (both signs), plus the boundary cases
    pos = np.arange(0, int(np.float32(2.0).view(np.uint32)) + 1, 97, dtype=np.uint32)
    edges = np.array([0.0, -0.0, 0.999999911, np.nextafter(np.float32(0.999999911), np.float32(0)), 1.0, 2.0,
                      np.inf, -np.inf, np.nan, 1e-30, 5.421010862e-20, 2.0 ** -64, 2.0 ** -63, -0.5],
                     dtype=np.float32)
    x = np.concatenate([pos.view(np.float32), (pos | np.uint32(0x80000000)).view(np.float32), edges])
    got = gpu.ConvertProbe(0, x)
    want = orc.float_to_ulong_n(x)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, (float(x[bad[0]]), int(got[bad[0]]), int(want[bad[0]]))
    u = _u64_samples(np.random.default_rng(7), 4_000_000)
    got_f = gpu.ConvertProbe(1, u)
    want_f = orc.ulong_to_float_n(u)
    assert np.array_equal(got_f.view(np.uint32), want_f.view(np.uint32))
