"""GPU parity of sw_postprocess_device (postprocessor.rs:68-204, 258-956): the tonemapped image
with bloom is bit-exact to the oracle, and so is the average colour that feeds the exposure
update (integer line sums, then a sequential float sum of the lines)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _hdr(rng, h, w, bright=0.02):
    """Colour64 frame: mostly dim lanes, a few bright spots (the pixels bloom spreads)."""
    lanes = rng.integers(0, 1200, size=(h, w, 4), dtype=np.uint64)
    hot = rng.random((h, w)) < bright
    lanes[hot] = rng.integers(20000, 65536, size=(int(hot.sum()), 4), dtype=np.uint64)
    return (lanes[..., 0] | (lanes[..., 1] << np.uint64(16)) | (lanes[..., 2] << np.uint64(32)) |
            (lanes[..., 3] << np.uint64(48)))


CASES = [
    # (w, h, sigma, scale_log2, bloom_scale, linear)
    (257, 131, 3.0, 1, 0.25, False),
    (257, 131, 3.0, 1, 0.25, True),
    (640, 360, 0.3, 2, 0.5, True),       # radius 0: no blur passes
    (640, 360, 12.0, 2, 0.125, False),
    (643, 365, 12.0, 3, 0.3, True),      # leftover rows and columns at scale 8
    (1000, 9, 40.0, 3, 0.0625, True),    # one bloom row, radius clamped to 16
    (2560, 1440, 8.0, 1, 0.2, True),
    (2560, 1440, 8.0, 1, 0.2, False),
    (17, 300, 5.0, 2, 0.4, True),        # narrow frame: more radius than bloom columns
]


@pytest.mark.parametrize("w,h,sigma,scale_log2,bloom_scale,linear", CASES)
def test_bloom_matches_oracle(gpu_ctx, w, h, sigma, scale_log2, bloom_scale, linear):
    rng = np.random.default_rng(w * 7919 + h * 31 + scale_log2)
    hdr = _hdr(rng, h, w)
    exposure, modulate = 1.7, (1.0, 0.8, 1.2)
    print(f"bloom {w}x{h} sigma {sigma} scale_log2 {scale_log2} linear {linear}", flush=True)
    out, avg = gpu_ctx.postprocess(hdr, exposure, modulate, bloom=(sigma, scale_log2, bloom_scale, linear))
    ref, ravg, radius, _ = oracle.bloom(hdr, exposure, modulate, sigma, scale_log2, bloom_scale, linear)
    bad = np.argwhere(out != ref)
    assert bad.size == 0, f"{len(bad)} mismatches (radius {radius}), first at {bad[:4].tolist()}"
    assert np.array_equal(avg, ravg), (avg, ravg)


def test_bloom_repeated_calls_reuse_buffers(gpu_ctx):
    """A large frame, then a small one: the grown bloom buffers hold no state between calls."""
    rng = np.random.default_rng(5)
    for w, h in ((1920, 1080), (300, 200), (1920, 1080)):
        hdr = _hdr(rng, h, w)
        out, avg = gpu_ctx.postprocess(hdr, 2.0, (1, 1, 1), bloom=(6.0, 1, 0.25, True))
        ref, ravg, _, _ = oracle.bloom(hdr, 2.0, (1, 1, 1), 6.0, 1, 0.25, True)
        assert np.array_equal(out, ref)
        assert np.array_equal(avg, ravg)


def test_no_bloom_is_plain_tonemap(gpu_ctx):
    """sigma 0 (postprocessor.rs:181-204): the plain tonemap, and the average is the pixel sums
    over the pixel count."""
    rng = np.random.default_rng(9)
    hdr = _hdr(rng, 240, 320)
    out, avg = gpu_ctx.postprocess(hdr, 0.9, (1.0, 1.0, 1.0), bloom=None)
    ref, sums = oracle.tonemap(hdr, 0.9, (1.0, 1.0, 1.0))
    assert np.array_equal(out, ref)
    # the GPU sums in a different order: float tolerance on the average only
    np.testing.assert_allclose(avg, sums * np.float32(1.0 / (320 * 240)), rtol=1e-5)
    out0, _ = gpu_ctx.postprocess(hdr, 0.9, (1.0, 1.0, 1.0), bloom=(0.0, 1, 0.25, True))
    assert np.array_equal(out0, ref)


def test_frame_smaller_than_a_bloom_texel_is_an_error(gpu_ctx):
    hdr = np.zeros((1, 64), dtype=np.uint64)
    with pytest.raises(Exception):
        gpu_ctx.postprocess(hdr, 1.0, (1, 1, 1), bloom=(3.0, 1, 0.25, False))
