"""sw_tonemap_device_noalloc: the tonemap that allocates nothing (its partial sums in a static device
array) gives the same image and colour sums as sw_tonemap_device, on two calls in a row, the second
with a smaller frame (the array keeps the first call's partials, which must not leak into the sums)."""
import ctypes as C

import numpy as np
import pytest

import sw_abi

pytestmark = pytest.mark.gpu


def _hdr(rng, h, w):
    lanes = rng.integers(0, 65536, size=(h, w, 4), dtype=np.uint64)
    return (lanes[..., 0] | (lanes[..., 1] << np.uint64(16)) | (lanes[..., 2] << np.uint64(32)) |
            (lanes[..., 3] << np.uint64(48)))


def test_noalloc_tonemap_equals_tonemap_device():
    ctx = sw_abi.Context(0)
    lib = ctx.lib
    try:
        rng = np.random.default_rng(2024)
        mod = (C.c_float * 3)(1.0, 0.75, 1.25)
        # the second call is smaller: fewer blocks, fewer rows and a narrower pitch
        for (w, h, pitch, exposure) in ((2560, 1440, 2560 + 32, 1.3), (333, 47, 340, 0.6)):
            print(f"noalloc tonemap {w}x{h} pitch {pitch}", flush=True)
            hdr = _hdr(rng, h, w)
            src = ctx.to_device(hdr)
            out_a = ctx.alloc(pitch * h * 4)
            out_b = ctx.alloc(pitch * h * 4)
            sums_b = ctx.alloc(16)
            try:
                sums_a = (C.c_float * 3)()
                ctx.check(lib.sw_tonemap_device(ctx.h, src.ptr, w, h, exposure, mod, out_a.ptr, pitch * 4, sums_a),
                          "sw_tonemap_device")
                ctx.check(lib.sw_tonemap_device_noalloc(ctx.h, src.ptr, w, h, exposure, mod, out_b.ptr, pitch * 4, sums_b.ptr),
                          "sw_tonemap_device_noalloc")
                ctx.check(lib.sw_sync(ctx.h), "sw_sync")
                a = out_a.read(np.uint32, pitch * h).reshape(h, pitch)[:, :w]
                b = out_b.read(np.uint32, pitch * h).reshape(h, pitch)[:, :w]
                assert np.array_equal(a, b)
                sb = sums_b.read(np.float32, 3)
                assert np.array_equal(np.array(list(sums_a), np.float32), sb), (list(sums_a), sb)
                assert np.all(sb > 0)
            finally:
                for buf in (src, out_a, out_b, sums_b):
                    buf.free()
        # sums are optional
        hdr = _hdr(rng, 16, 16)
        src = ctx.to_device(hdr)
        out = ctx.alloc(16 * 16 * 4)
        try:
            ctx.check(lib.sw_tonemap_device_noalloc(ctx.h, src.ptr, 16, 16, 1.0, mod, out.ptr, 64, None), "noalloc no sums")
            assert lib.sw_tonemap_device_noalloc(ctx.h, src.ptr, 16, 16, 1.0, mod, out.ptr, 60, None) == sw_abi.SW_ERR_INVALID_ARGUMENT
            ctx.check(lib.sw_sync(ctx.h), "sw_sync")
        finally:
            src.free()
            out.free()
    finally:
        ctx.close()
