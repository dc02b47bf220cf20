"""Per-frame texture animation (generative_texture_effect_turb.rs:89-138, generative_texture_effect_layered.rs:
162-381, surfaces.rs:903-971 for emissive layers; material_function.rs:90-156): the oracle's restatement
pinned on the CPU by closed forms (rolls, per-channel integer formulas), and the host-side parameter
helpers. GPU parity of sw_apply_texture_ops is in test_gpu_textures.py."""
import math

import numpy as np
import pytest

import oracle
import sw_abi as A
import sw_host


def _levels(*sizes, lite=False):
    lv = np.zeros(len(sizes), dtype=A.LEVEL_DTYPE)
    off = 0
    for i, (w, h) in enumerate(sizes):
        lv[i]["offset"], lv[i]["width"], lv[i]["height"] = off, w, h
        lv[i]["flags"] = A.TEX_LITE_LEVEL if lite else 0
        off += w * h
    return lv, off


def _data(rng, sizes, lite=False):
    lv, n = _levels(*sizes, lite=lite)
    tex = np.zeros(0 if lite else n, dtype=A.TEXEL_DTYPE)
    lit = np.zeros(n if lite else 0, dtype=np.uint32)
    if lite:
        lit[:] = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    else:
        tex["diffuse"] = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
        tex["packed_normal_roughness"] = rng.integers(-2 ** 31, 2 ** 31, size=n, dtype=np.int64).astype(np.int32)
    return A.MapData(tex, lit, np.zeros((1, 3), np.float32), np.zeros(1, A.DIRLM_DTYPE), lv)


def _op(kind, dst, src, blend=0, shift=(0, 0), light=(1.0, 1.0, 1.0), turb=0):
    o = np.zeros(1, dtype=A.TEXTURE_OP_DTYPE)
    o[0]["kind"], o[0]["dst_level"], o[0]["src_level"], o[0]["blend"] = kind, dst, src, blend
    o[0]["shift"], o[0]["light"], o[0]["turb_shifts"] = shift, light, turb
    return o


def _level(arr, lv, i):
    l = lv[i]
    return arr[int(l["offset"]):int(l["offset"]) + int(l["width"]) * int(l["height"])].reshape(int(l["height"]), int(l["width"]))


def _channels(c):
    c = np.asarray(c, dtype=np.uint32)
    return [((c >> np.uint32(8 * k)) & np.uint32(0xFF)).astype(np.int64) for k in range(4)]


def _pack(ch):
    return (ch[0].astype(np.uint32) | (ch[1].astype(np.uint32) << 8) | (ch[2].astype(np.uint32) << 16) |
            (ch[3].astype(np.uint32) << 24))


@pytest.mark.parametrize("lite", [False, True])
def test_turb_zero_and_constant_shifts(lite):
    """Zero shifts copy the source; a constant shift c everywhere rolls it by c along both axes."""
    rng = np.random.default_rng(1)
    d = _data(rng, [(32, 16), (32, 16)], lite=lite)
    for c in (0, 5, -7):
        shifts = np.full(32, c, dtype=np.int32)
        tex, lit = oracle.apply_texture_ops(d, _op(A.TEXOP_TURB, 1, 0), shifts)
        src = _level(d.lite if lite else d.texels, d.levels, 0)
        out = _level(lit if lite else tex, d.levels, 1)
        assert np.array_equal(out, np.roll(src, (-c, -c), axis=(0, 1)))


def test_turb_per_axis_shifts():
    """The row's shift moves along x, the column's along y: dst[y, x] = src[(y + s[x]) & 15, (x + s[y]) & 31]."""
    rng = np.random.default_rng(2)
    d = _data(rng, [(32, 16), (32, 16)], lite=True)
    s = rng.integers(-40, 40, size=32).astype(np.int32)
    _, lit = oracle.apply_texture_ops(d, _op(A.TEXOP_TURB, 1, 0), s)
    src = _level(d.lite, d.levels, 0)
    ys, xs = np.mgrid[0:16, 0:32]
    assert np.array_equal(_level(lit, d.levels, 1), src[(ys + s[xs]) & 15, (xs + s[ys]) & 31])


def test_turb_shift_table():
    """round(sin(fma(x, 2pi / (wave_length / 2^mip), t * frequency * 2pi)) * amplitude / 2^mip) for x <
    width against float64 (half-way cases excluded); the buffer grows with zeros and keeps what lies past
    the width, like the reference's temp buffer."""
    t = sw_host.TurbShifts()
    amp, wl, freq, time_s = 3.7, 24.0, 0.5, 1.25
    a = t.compute(amp, wl, freq, time_s, 64, 64, 0)
    for mip in range(3):
        scale = 1.0 / (1 << mip)
        w = 64 >> mip
        v = t.compute(amp, wl, freq, time_s, w, 128 >> mip, mip)
        assert len(v) == 128 >> mip
        x = np.arange(w, dtype=np.float64)
        f = np.sin(x * (2 * math.pi / (wl * scale)) + time_s * freq * 2 * math.pi) * amp * scale
        sure = np.abs(np.abs(f) - np.floor(np.abs(f)) - 0.5) > 1e-3
        ref = np.sign(f) * np.floor(np.abs(f) + 0.5)
        assert np.array_equal(v[:w][sure], ref[sure].astype(np.int32))
        if mip == 0:
            assert np.all(v[64:] == 0)  # grown past the previous 64 entries: zeros
        else:
            # entries past the width are the previous mip's values (computed over the wider row)
            assert np.array_equal(v[w:], prev[w:len(v)])
        prev = v
    assert np.array_equal(a, t.compute(amp, wl, freq, time_s, 64, 64, 0))


def test_layer_fast_path_and_roll():
    """Blend none, light 1: a same-size unshifted layer copies texels (normals too); a shifted one of
    another size tiles the source (rem_euclid of shift + coordinate)."""
    rng = np.random.default_rng(3)
    d = _data(rng, [(16, 8), (16, 8), (5, 3)])
    tex, _ = oracle.apply_texture_ops(d, _op(A.TEXOP_LAYER, 0, 1))
    assert np.array_equal(_level(tex, d.levels, 0), _level(d.texels, d.levels, 1))
    tex, _ = oracle.apply_texture_ops(d, _op(A.TEXOP_LAYER, 0, 2, shift=(-7, 4)))
    src = _level(d.texels, d.levels, 2)
    ys, xs = np.mgrid[0:8, 0:16]
    assert np.array_equal(_level(tex, d.levels, 0), src[(ys + 4) % 3, (xs - 7) % 5])


@pytest.mark.parametrize("blend", range(5))
@pytest.mark.parametrize("light", [(1.0, 1.0, 1.0), (1.01, 0.99, 1.0), (0.5, 2.0, 0.25)])
def test_layer_blend_closed_forms(blend, light):
    """Per channel: modulated = (c * round(l * 256)) >> 8 (alpha lane 0) when a light component lies
    outside [0.98, 1.02], else the texel itself; then none / average ((a + b) >> 1 modulated, pavg
    (a + b + 1) >> 1 not) / additive (saturating) / alpha test (alpha > 127) / alpha blend
    ((s a + d (255 - a)) >> 8); normals follow the texel for none and alpha test."""
    rng = np.random.default_rng(4 + blend)
    d = _data(rng, [(16, 8), (16, 8)])
    tex, _ = oracle.apply_texture_ops(d, _op(A.TEXOP_LAYER, 0, 1, blend=blend, shift=(0, 0), light=light))
    dst0, src = _level(d.texels, d.levels, 0).ravel(), _level(d.texels, d.levels, 1).ravel()
    out = _level(tex, d.levels, 0).ravel()
    modulate = any(c < 0.98 or c > 1.02 for c in light)
    if blend == 0 and not modulate:
        assert np.array_equal(out, src)
        return
    s, dd = _channels(src["diffuse"]), _channels(dst0["diffuse"])
    if modulate:
        lv = [np.int64(np.round(np.float32(light[2 - k]) * np.float32(256))) for k in range(3)]  # lanes B, G, R
        m = [(s[k] * lv[k]) >> 8 for k in range(3)] + [np.zeros_like(s[3])]
    else:
        m = s
    alpha = s[3]
    test = s[3] > 127
    sat = lambda v: np.clip(v, 0, 255)
    if blend == 0:
        exp, normals = _pack([sat(c) for c in m]), src["packed_normal_roughness"]
    elif blend == 1:
        exp = _pack([sat((a + b) >> 1) for a, b in zip(m, dd)]) if modulate else _pack([(a + b + 1) >> 1 for a, b in zip(s, dd)])
        normals = dst0["packed_normal_roughness"]
    elif blend == 2:
        exp, normals = _pack([sat(a + b) for a, b in zip(m, dd)]), dst0["packed_normal_roughness"]
    elif blend == 3:
        exp = np.where(test, _pack([sat(c) for c in m]), dst0["diffuse"])
        normals = np.where(test, src["packed_normal_roughness"], dst0["packed_normal_roughness"])
    else:
        exp = _pack([sat((a * alpha + b * (255 - alpha)) >> 8) for a, b in zip(m, dd)])
        normals = dst0["packed_normal_roughness"]
    assert np.array_equal(out["diffuse"], exp)
    assert np.array_equal(out["packed_normal_roughness"], normals)


@pytest.mark.parametrize("blend", range(5))
def test_lite_layer_closed_forms(blend):
    """Emissive layers are always modulated (mix_surface_with_texture)."""
    rng = np.random.default_rng(9 + blend)
    d = _data(rng, [(8, 8), (8, 8)], lite=True)
    light = (0.75, 1.0, 1.5)
    _, lit = oracle.apply_texture_ops(d, _op(A.TEXOP_LAYER_LITE, 0, 1, blend=blend, light=light))
    src, dst0 = _level(d.lite, d.levels, 1).ravel(), _level(d.lite, d.levels, 0).ravel()
    s, dd = _channels(src), _channels(dst0)
    lv = [np.int64(np.round(np.float32(light[2 - k]) * np.float32(256))) for k in range(3)]
    m = [(s[k] * lv[k]) >> 8 for k in range(3)] + [np.zeros_like(s[3])]
    sat = lambda v: np.clip(v, 0, 255)
    exp = {0: _pack([sat(c) for c in m]), 1: _pack([sat((a + b) >> 1) for a, b in zip(m, dd)]),
           2: _pack([sat(a + b) for a, b in zip(m, dd)]), 3: np.where(s[3] > 127, _pack([sat(c) for c in m]), dst0),
           4: _pack([sat((a * s[3] + b * (255 - s[3])) >> 8) for a, b in zip(m, dd)])}[blend]
    assert np.array_equal(_level(lit, d.levels, 0).ravel(), exp)


def test_layers_apply_in_order():
    """A list of layers on one destination equals applying them one call at a time."""
    rng = np.random.default_rng(20)
    d = _data(rng, [(16, 16), (16, 16), (8, 4), (4, 8)])
    ops = np.concatenate([_op(A.TEXOP_LAYER, 0, 1), _op(A.TEXOP_LAYER, 0, 2, blend=2, shift=(3, -1), light=(0.5, 0.5, 0.5)),
                          _op(A.TEXOP_LAYER, 0, 3, blend=4, shift=(-2, 9))])
    tex, _ = oracle.apply_texture_ops(d, ops)
    step = d
    for o in ops:
        t, _ = oracle.apply_texture_ops(step, o[None])
        step = A.MapData(t, step.lite, step.lm_simple, step.lm_dir, step.levels)
    assert np.array_equal(tex, step.texels)


def test_single_argument_functions():
    """Closed forms of the six function kinds (material_function.rs:90-156)."""
    f = sw_host.evaluate_function
    assert f(0, [3.5], 10.0) == 3.5
    assert f(1, [2.0, 1.0], 3.0) == 7.0
    assert f(2, [1.0, 0.0, 2.0, 1.0], 0.25) == pytest.approx(3.0, abs=1e-6)
    assert f(3, [1.0, 0.0, 1.0, 0.0], 0.25) == 1.0 and f(3, [1.0, 0.0, 1.0, 0.0], 0.75) == -1.0
    assert f(3, [1.0, 0.0, 1.0, 0.0], 0.5) == 0.0
    assert f(4, [2.0, 0.0, 1.0, 0.0], 0.25) == 0.0 and f(4, [1.0, 0.0, 3.0, 1.0], 0.0) == -2.0
    assert f(5, [1.0, 0.0, 2.0, 0.0], 0.3) == 2.0 and f(5, [1.0, 0.0, 2.0, 0.0], 0.7) == -2.0
    assert sw_host.layer_shift(-37.9, 2) == -10 and sw_host.layer_shift(1e10, 0) == 2 ** 31 - 1


@pytest.mark.parametrize("bits", [32, 64])
def test_skybox_side_closed_forms(bits):
    """make_skybox_side_texture{,64}: the top-left square of the image, (c * trunc(b * 256)) >> 8 per
    channel, alpha 0, unsigned saturation; four mips average 2x2 blocks ((sum) >> 2)."""
    rng = np.random.default_rng(30 + bits)
    lane_max = 255 if bits == 32 else 65535
    dt, sh = (np.uint32, 8) if bits == 32 else (np.uint64, 16)
    lanes = rng.integers(0, lane_max + 1, size=(4, 40, 35)).astype(np.uint64)  # odd: mips of 35, 17, 8, 4
    img = (lanes[0] | (lanes[1] << np.uint64(sh)) | (lanes[2] << np.uint64(2 * sh)) | (lanes[3] << np.uint64(3 * sh))).astype(dt)
    for brightness in (1.0, 0.5, 1.7, 3.999):
        out = oracle.make_skybox_side(img, brightness)
        assert len(out) == 35 * 35 + 17 * 17 + 8 * 8 + 4 * 4 + 2 * 2
        bi = int(np.float32(brightness) * np.float32(256))
        mip0 = out[:35 * 35].reshape(35, 35).astype(np.uint64)
        for k in range(3):
            ref = np.minimum((lanes[k][:35, :35].astype(np.int64) * bi) >> 8, lane_max)
            assert np.array_equal((mip0 >> np.uint64(k * sh)) & np.uint64(lane_max), ref.astype(np.uint64))
        assert np.all(mip0 >> np.uint64(3 * sh) == 0)
        prev, size, off = mip0, 35, 35 * 35
        for _ in range(4):  # five levels in all; an odd level's last row and column are dropped
            size //= 2
            mip = out[off:off + size * size].reshape(size, size).astype(np.uint64)
            for k in range(3):
                p = ((prev >> np.uint64(k * sh)) & np.uint64(lane_max)).astype(np.int64)[:2 * size, :2 * size]
                ref = (p[0::2, 0::2] + p[1::2, 0::2] + p[0::2, 1::2] + p[1::2, 1::2]) >> 2
                assert np.array_equal((mip >> np.uint64(k * sh)) & np.uint64(lane_max), ref.astype(np.uint64))
            prev, off = mip, off + size * size
