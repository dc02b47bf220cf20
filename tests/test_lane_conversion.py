"""The colour-lane conversion the kernels use (sw_device.h lane_to_u8 / lane_to_u16): the low byte
(16 bits) of cvtps2dq(minps(x, 255 or 65535)) (fast_math.rs:131-160; out-of-range and NaN give
0x80000000) equals that of the in-range conversion of max(min(x, M), -2^31), rounded to nearest
even. Checked on a dense stride through all 2^32 float bit patterns plus the edges (zeros,
infinities, quiet NaNs, the rounding ties around 0 and M, and the values around -2^31)."""
import numpy as np
import pytest


def _x86(x, m):
    y = np.where(np.isnan(x) | ~(x < m), np.float32(m), x)  # minps(x, m): the second operand unless x < m
    r = np.rint(y.astype(np.float64))
    ok = (r >= -2.0**31) & (r < 2.0**31)
    return np.where(ok, r, -2.0**31).astype(np.int64).astype(np.uint32)


def _clamped(x, m):
    y = np.fmin(x, np.float32(m))  # fminf: a quiet NaN gives m
    y = np.fmax(y, np.float32(-2.0**31))
    return np.rint(y).astype(np.int32).view(np.uint32)


def _samples():
    bits = np.arange(0, 1 << 32, 4093, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[~np.isnan(x) | ((bits & 0x00400000) != 0)]  # signalling NaNs never come out of arithmetic
    edges = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 0.5, 1.5, 2.5, -0.5, -1.5, 254.5, 255.5, 255.0, 256.0,
                      65534.5, 65535.5, 65536.0, -2.0**31, np.nextafter(np.float32(-2.0**31), np.float32(-np.inf)),
                      np.nextafter(np.float32(-2.0**31), np.float32(0)), -3e38, 3e38], dtype=np.float32)
    return np.concatenate([x, edges])


@pytest.mark.parametrize("m,mask", [(255.0, 0xFF), (65535.0, 0xFFFF)])
def test_clamped_conversion_keeps_low_bits(m, mask):
    x = _samples()
    with np.errstate(invalid="ignore"):
        a = _x86(x, m) & mask
        b = _clamped(x, m) & mask
    bad = np.nonzero(a != b)[0]
    assert bad.size == 0, (x[bad[:5]], a[bad[:5]], b[bad[:5]])
    # the cases the clamp exists for: below -2^31 and -inf give 0 in the low bits
    assert (_clamped(np.array([-np.inf, -1e30], np.float32), m) & mask == 0).all()
