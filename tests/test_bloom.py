"""Bloom postprocessing and auto-exposure (postprocessor.rs:68-223, 258-1051): the oracle's
restatement pinned by closed forms on the CPU, and the exposure update of the C ABI's host side.
GPU parity of sw_postprocess_device is in test_gpu_bloom.py."""
import math

import numpy as np
import pytest

import oracle
import sw_abi as A


def _pack(b, g, r, a):
    return (np.uint64(b) | (np.uint64(g) << np.uint64(16)) | (np.uint64(r) << np.uint64(32)) |
            (np.uint64(a) << np.uint64(48)))


def _lanes(img):
    return [((img >> np.uint64(16 * k)) & np.uint64(0xFFFF)).astype(np.int64) for k in range(4)]


def _f64_kernel(sigma, radius):
    out = []
    for x in range(-radius, radius + 1):
        val = 0.0
        for s in (-0.375, -0.125, 0.125, 0.375):
            c = x + s
            val += math.exp(-0.5 * (c / sigma) ** 2) / (sigma * math.sqrt(2 * math.pi))
        out.append(val / 4)
    return np.array(out)


@pytest.mark.parametrize("sigma,scale_log2", [(1.0, 1), (3.0, 1), (3.0, 2), (12.0, 2), (40.0, 1), (40.0, 3), (0.4, 1)])
def test_kernel_weights_and_radius(sigma, scale_log2):
    """Integer weights = trunc(gaussian(4 samples per tap) * 2^14) (postprocessor.rs:1026-1045,
    :855-862), radius = min(max(ceil(3 sigma / scale - 0.5), 0), 16) (:748-749): against a float64
    evaluation, +-1 for the float32 rounding before the truncation."""
    hdr = np.full((64, 64), _pack(100, 200, 300, 0), dtype=np.uint64)
    S = 1 << scale_log2
    _, _, radius, (k1, k2) = oracle.bloom(hdr, 1.0, (1, 1, 1), sigma, scale_log2, 0.25, False)
    assert radius == min(max(math.ceil(3 * sigma / S - 0.5), 0), 16)
    ref = _f64_kernel(sigma / S, radius)
    assert np.all(np.abs(k1[:2 * radius + 1] - np.floor(ref * 16384)) <= 1)
    assert np.all(np.abs(k2[:2 * radius + 1] - np.floor(ref * 0.25 * 16384)) <= 1)
    assert np.all(k1[2 * radius + 1:] == 0)
    if radius:
        # the taps cover the Gaussian's mass within the radius: their sum is that mass times 2^14
        # (less one per truncated tap), never above 2^14
        mass = float(ref.sum() * 16384)
        assert mass - 2 * radius - 2 <= int(k1[:2 * radius + 1].sum()) <= min(mass + 1, 16384)


def _blur_constant(v, k, radius):
    """One blur pass over a constant image: every tap (clamped edges included) sees v."""
    return (v * (int(k[radius]) + 2 * int(k[radius + 1:2 * radius + 1].sum()))) >> 14


@pytest.mark.parametrize("scale_log2", [1, 2, 3])
@pytest.mark.parametrize("size", [(48, 40), (53, 37)])
def test_constant_image_closed_form(scale_log2, size):
    """A constant frame: the box average is the colour itself, each blur pass scales it by its integer
    weight sum (>> 14), and every output pixel (nearest upsampling; leftover rows and columns
    included) is the plain tonemap of colour + bloom; the average colour is the colour."""
    w, h = size
    c = (1000, 2000, 3000, 50)
    hdr = np.full((h, w), _pack(*c), dtype=np.uint64)
    out, avg, radius, (k1, k2) = oracle.bloom(hdr, 1.5, (1.0, 0.5, 2.0), 6.0, scale_log2, 0.3, False)
    bloom = [_blur_constant(_blur_constant(v, k1, radius), k2, radius) for v in c]
    summed = np.full((h, w), _pack(*[v + b for v, b in zip(c, bloom)]), dtype=np.uint64)
    ref, _ = oracle.tonemap(summed, 1.5, (1.0, 0.5, 2.0))
    assert np.array_equal(out, ref)
    # the average: the exact sum of the (equal) box averages times the float 1 / texel count
    S = 1 << scale_log2
    n = (w // S) * (h // S)
    expect = [np.float32(v * n) * (np.float32(1.0) / np.float32(n)) for v in (c[2], c[1], c[0])]
    assert np.array_equal(avg, np.array(expect, dtype=np.float32))


def test_point_light_is_the_separable_kernel():
    """One bright bloom texel in a dark frame: after the two passes the bloom buffer is
    ((v * k1[dx]) >> 14) * k2[dy] >> 14 around it, so each output pixel (nearest) is the plain tonemap
    of its colour plus that value."""
    S, w, h = 2, 96, 80
    hdr = np.zeros((h, w), dtype=np.uint64)
    bx, by, v = 20, 17, 40000
    hdr[by * S:(by + 1) * S, bx * S:(bx + 1) * S] = _pack(v, v // 2, v // 4, 0)
    out, _, radius, (k1, k2) = oracle.bloom(hdr, 1.0, (1, 1, 1), 10.0, 1, 0.5, False)
    assert radius > 0
    bw, bh = w // S, h // S
    lanes = [v, v // 2, v // 4, 0]
    bloom = np.zeros((4, bh, bw), dtype=np.int64)
    for k, lv in enumerate(lanes):
        row = np.zeros(bw, dtype=np.int64)
        for dx in range(-radius, radius + 1):
            row[bx + dx] = (lv * int(k1[radius + abs(dx)])) >> 14
        for dy in range(-radius, radius + 1):
            bloom[k, by + dy, :] = (row * int(k2[radius + abs(dy)])) >> 14
    up = bloom.repeat(S, axis=1).repeat(S, axis=2)
    base = _lanes(hdr)
    summed = _pack(*[np.clip(base[k] + up[k], 0, 65535).astype(np.uint64) for k in range(4)]).astype(np.uint64)
    ref, _ = oracle.tonemap(summed, 1.0, (1, 1, 1))
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("scale_log2", [1, 2, 3])
def test_linear_upsampling_corners_and_smooth_field(scale_log2):
    """Bilinear upsampling: the four corner regions add the corner texel alone (as nearest does);
    on a constant bloom buffer every pixel is within one step of the nearest result."""
    w, h = 77, 61
    rng = np.random.default_rng(scale_log2)
    hdr = _pack(*[rng.integers(0, 3000, size=(h, w)).astype(np.uint64) for _ in range(3)], np.zeros((h, w), np.uint64))
    near, _, _, _ = oracle.bloom(hdr, 1.0, (1, 1, 1), 4.0, scale_log2, 0.25, False)
    lin, _, _, _ = oracle.bloom(hdr, 1.0, (1, 1, 1), 4.0, scale_log2, 0.25, True)
    S = 1 << scale_log2
    half = S // 2
    bw, bh = w // S, h // S
    rx, uy = (bw - 1) * S + half, (bh - 1) * S + half
    assert np.array_equal(lin[:half, :half], near[:half, :half])
    assert np.array_equal(lin[uy:, rx:], near[uy:, rx:])
    assert np.array_equal(lin[:half, rx:], near[:half, rx:])
    assert np.array_equal(lin[uy:, :half], near[uy:, :half])
    flat = np.full((h, w), _pack(800, 900, 1000, 0), dtype=np.uint64)
    a, _, _, _ = oracle.bloom(flat, 1.0, (1, 1, 1), 4.0, scale_log2, 0.25, False)
    b, _, _, _ = oracle.bloom(flat, 1.0, (1, 1, 1), 4.0, scale_log2, 0.25, True)
    d = np.abs(a.view(np.uint8).astype(int) - b.view(np.uint8).astype(int))
    assert d.max() <= 1


def _exposure_params(**kw):
    p = dict(exposure_update_speed=1.0, base_brightness=1.0, zero_level_brightness=0.5, brightness_scale_power=0.25,
             min_exposure=1.0 / 16.0, max_exposure=8.0)
    p.update(kw)
    return A.ExposureParams(**p)


def test_exposure_update_invariants():
    """update_exposure (postprocessor.rs:206-223): no time -> unchanged (up to exp(log x)); a long
    frame -> the clamped target; brighter frames never get a higher exposure; the target respects
    [min, max]."""
    p = _exposure_params()
    for cur in (0.1, 1.0, 3.0):
        e = A.update_exposure(p, cur, (100.0, 100.0, 100.0), 0.0)
        assert abs(e - cur) <= 2 * np.spacing(np.float32(cur))
    # target, from the formula's definition in float64
    avg = (60.0, 80.0, 100.0)
    b = min(max((avg[0] + avg[1] + avg[2]) * 0.33, 1 / 1024), 65536) / 255
    target = (b / 0.5) ** 0.25 * 0.5 * 1.0 / b
    e = A.update_exposure(p, 1.0, avg, 100.0)
    assert e == pytest.approx(min(max(target, 1 / 16), 8.0), rel=1e-5)
    prev = None
    for level in (1.0, 10.0, 50.0, 200.0, 1000.0, 30000.0):
        e = A.update_exposure(p, 1.0, (level,) * 3, 100.0)
        assert 1 / 16 - 1e-6 <= e <= 8.0 + 1e-6
        if prev is not None:
            assert e <= prev + 1e-6
        prev = e
    # half-way in log space after ln(2) / speed seconds
    e = A.update_exposure(_exposure_params(exposure_update_speed=2.0), 1.0, avg, math.log(2) / 2.0)
    assert e == pytest.approx(math.sqrt(min(max(target, 1 / 16), 8.0)), rel=1e-4)
