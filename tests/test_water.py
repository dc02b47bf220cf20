"""The water effect (generative_texture_effect_water.rs) on the CPU: the oracle's field step
(update_wave_field, :307-418, after the step's points, :68-76) and wavy texture (:420-682, with
build_texture_mip, textures.rs:309-362) pinned by closed forms (constant fields, a single impulse
spreading to its wrapped neighbours, flat fields giving the +z normal, equal texels giving equal mips)
and by an independent numpy formulation with rolled arrays (the oracle follows the reference's row
cases). GPU parity is in test_gpu_water.py."""
import numpy as np
import pytest

import oracle
import sw_abi as A
import sw_host

F32 = np.float32


def _pts(rows):
    p = np.zeros(len(rows), dtype=A.WAVE_POINT_DTYPE)
    for i, (x, y, v) in enumerate(rows):
        p[i] = (x, y, v)
    return p


def _np_step(f, fo, w, h, att, pts):
    f, fo = f.copy().reshape(h, w), fo.copy().reshape(h, w)
    for x, y, v in pts:
        f[y & (h - 1), x & (w - 1)] += F32(v)
        fo[y & (h - 1), x & (w - 1)] -= F32(v)
    s = ((np.roll(f, 1, axis=1) + np.roll(f, -1, axis=1)) + np.roll(f, 1, axis=0)) + np.roll(f, -1, axis=0)
    new = (s * F32(0.5) - fo) * F32(att)
    return new.reshape(-1), f.reshape(-1)


def test_constant_field_and_impulse():
    w, h, att = 32, 16, F32(0.75)
    f = np.full(w * h, 0.25, F32)
    fo = np.full(w * h, 0.125, F32)
    new, cur = oracle.water_step(f, fo, 5, 4, att, None)
    assert np.all(new == F32((2 * 0.25 - 0.125) * 0.75)) and np.array_equal(cur, f)
    # an impulse at the corner reaches its four wrapped neighbours with v / 2 * attenuation
    new, _ = oracle.water_step(np.zeros(w * h, F32), np.zeros(w * h, F32), 5, 4, att, _pts([(0, 0, 1.0)]))
    img = new.reshape(h, w)
    want = np.zeros((h, w), F32)
    for y, x in ((0, 1), (0, w - 1), (1, 0), (h - 1, 0)):
        want[y, x] = F32(0.5 * 0.75)
    want[0, 0] = F32(1.0 * 0.75)  # the old field's -1 at the point: (0 - (-1)) * attenuation
    assert np.array_equal(img, want)


def test_points_wrap_and_accumulate_in_order():
    w, h = 16, 16
    pts = _pts([(3, 4, 1e8), (3 + 16, (4 - 16) & 0xFFFFFFFF, 1.0), (3, 4, -1e8)])  # same texel: (1e8 + 1) - 1e8 = 0 in float32
    new, cur = oracle.water_step(np.zeros(w * h, F32), np.zeros(w * h, F32), 4, 4, F32(1.0), pts)
    assert cur[4 * w + 3] == F32(0.0)


@pytest.mark.parametrize("log2", [(4, 4), (6, 5), (7, 8)])
def test_steps_match_rolled_arrays(log2):
    lw, lh = log2
    w, h = 1 << lw, 1 << lh
    rng = np.random.default_rng(lw * 10 + lh)
    f = rng.normal(size=w * h).astype(F32)
    fo = rng.normal(size=w * h).astype(F32)
    att = F32(1.0 - 1.0 / 70.0)
    for _ in range(5):
        rows = [(int(rng.integers(0, 3 * w)), int(rng.integers(0, 3 * h)), float(rng.normal())) for _ in range(int(rng.integers(0, 9)))]
        pts = _pts(rows)
        a, b = oracle.water_step(f, fo, lw, lh, att, pts)
        ra, rb = _np_step(f, fo, w, h, att, list(zip(pts["x"], pts["y"], pts["value"])))
        assert np.array_equal(a.view(np.uint32), ra.view(np.uint32))
        assert np.array_equal(b.view(np.uint32), rb.view(np.uint32))
        f, fo = a, b


# --- the texture --------------------------------------------------------------------------------
def _pack(x, y, z, r):
    q = [np.clip(np.rint(F32(127.0) * v.astype(F32)), -128, 127).astype(np.int32) & 0xFF for v in (x, y, z, r)]
    return (q[0] | (q[1] << 8) | (q[2] << 16) | (q[3] << 24)).astype(np.uint32).view(np.int32)


def _np_texture(field, w, h, mode, rough, dst_diffuse, color):
    f = field.reshape(h, w)
    dx = np.roll(f, -1, axis=1) - np.roll(f, 1, axis=1)
    dy = np.roll(f, -1, axis=0) - np.roll(f, 1, axis=0)
    k = F32(1.0) / np.sqrt((dx * dx + dy * dy) + F32(1.0))
    nx, ny, nz = dx * k, dy * k, k
    packed = _pack(nx, ny, nz, np.full_like(nx, rough))
    ys, xs = np.mgrid[0:h, 0:w]
    c = color.reshape(h, w)
    if mode == A.WATER_COLOR_KEEP:
        diffuse = dst_diffuse.reshape(h, w)
    else:
        u = (xs + np.trunc(nx * F32(16.0)).astype(np.int64)) & (w - 1)
        v = (ys + np.trunc(ny * F32(16.0)).astype(np.int64)) & (h - 1) if mode == A.WATER_COLOR_DEFORMED else ys
        diffuse = c[v, u]
    return diffuse.reshape(-1), packed.reshape(-1)


def _np_mip(diffuse, packed, w, h):
    d = diffuse.reshape(h, w).astype(np.uint32)
    b = packed.reshape(h, w).view(np.uint32)
    q = [(b >> (8 * i) & 0xFF).astype(np.uint8).view(np.int8).astype(F32) * F32(1.0 / 127.0) for i in range(4)]
    sel = [(slice(0, None, 2), slice(0, None, 2)), (slice(1, None, 2), slice(0, None, 2)), (slice(0, None, 2), slice(1, None, 2)),
           (slice(1, None, 2), slice(1, None, 2))]
    s = [((q[c][sel[0]] + q[c][sel[1]]) + q[c][sel[2]]) + q[c][sel[3]] for c in range(4)]
    ln = np.maximum(np.sqrt((s[0] * s[0] + s[1] * s[1]) + s[2] * s[2]), F32(1e-6))
    dev = (F32(1.0) - ln * F32(0.25)) / F32(0.5)
    dr = (s[3].astype(np.float64) * 0.25 + dev.astype(np.float64)).astype(F32)  # one rounding, as fma
    dr = np.minimum(np.maximum(dr, F32(1.0) / F32(126.0)), F32(1.0))
    acc = [np.zeros(s[0].shape, np.uint32) for _ in range(4)]
    for sl in sel:
        c = d[sl]
        acc[0] += c & 0xFF
        acc[1] += c & 0xFF00
        acc[2] += c & 0xFF0000
        acc[3] += (c & 0xFF000000) >> 2
    avg = (acc[0] >> 2) | ((acc[1] >> 2) & 0xFF00) | ((acc[2] >> 2) & 0xFF0000) | (acc[3] & 0xFF000000)
    return avg.reshape(-1), _pack(s[0] / ln, s[1] / ln, s[2] / ln, dr).reshape(-1)


@pytest.fixture(scope="module")
def water_map():
    m = sw_host.Map(seed=3, grid=(10, 10, 4), texture_size=64, num_materials=3, num_decal_textures=1, num_model_textures=1)
    lv = m.data.levels
    assert lv[0]["width"] == 64 and lv[5]["width"] == 64 and [int(lv[i]["width"]) for i in range(5)] == [64, 32, 16, 8, 4]
    return m


def _params(mode, init=A.WATER_INIT_NONE, rough=0.4, base=0x80604020):
    return A.WaterTextureParams(0, 5, mode, init, rough, base)


@pytest.mark.parametrize("mode", [A.WATER_COLOR_KEEP, A.WATER_COLOR_DEFORMED, A.WATER_COLOR_DEFORMED_X])
def test_texture_matches_rolled_arrays(water_map, mode):
    m = water_map
    lv = m.data.levels
    rng = np.random.default_rng(mode)
    field = (rng.normal(size=64 * 64) * 0.7).astype(F32)
    out = oracle.water_make_texture(m.data, field, 6, 6, _params(mode))
    tex = m.data.texels
    d0 = tex[lv[0]["offset"]:lv[0]["offset"] + 4096]
    col = tex[lv[5]["offset"]:lv[5]["offset"] + 4096]
    diffuse, packed = _np_texture(field, 64, 64, mode, F32(0.4), d0["diffuse"], col["diffuse"])
    o0 = out[lv[0]["offset"]:lv[0]["offset"] + 4096]
    assert np.array_equal(o0["diffuse"], diffuse) and np.array_equal(o0["packed_normal_roughness"], packed)
    w = 64
    for mip in range(1, 5):
        diffuse, packed = _np_mip(diffuse, packed, w, w)
        w //= 2
        om = out[lv[mip]["offset"]:lv[mip]["offset"] + w * w]
        assert np.array_equal(om["diffuse"], diffuse), mip
        assert np.array_equal(om["packed_normal_roughness"], packed), mip
    # the rest of the texels are untouched
    mask = np.ones(len(tex), bool)
    for i in range(5):
        mask[lv[i]["offset"]:lv[i]["offset"] + int(lv[i]["width"]) * int(lv[i]["height"])] = False
    assert np.array_equal(out[mask], tex[mask])


def test_flat_field_and_init(water_map):
    """A flat field gives the +z normal everywhere: bytes (0, 0, 127, rint(127 r)); the first update's
    fills take the base colour or the colour level; roughness is clamped to [1/126, 1]."""
    m = water_map
    lv = m.data.levels
    flat = np.zeros(64 * 64, F32)
    out = oracle.water_make_texture(m.data, flat, 6, 6, _params(A.WATER_COLOR_KEEP, A.WATER_INIT_BASE_COLOR, rough=0.5))
    o0 = out[lv[0]["offset"]:lv[0]["offset"] + 4096]
    assert np.all(o0["diffuse"] == 0x80604020)
    assert np.all(o0["packed_normal_roughness"].view(np.uint32) == (127 << 16) | (64 << 24))  # rint(63.5) = 64
    o4 = out[lv[4]["offset"]:lv[4]["offset"] + 16]
    assert np.all(o4["diffuse"] == 0x80604020) and np.all(o4["packed_normal_roughness"].view(np.uint32) == (127 << 16) | (64 << 24))
    out = oracle.water_make_texture(m.data, flat, 6, 6, _params(A.WATER_COLOR_KEEP, A.WATER_INIT_SOURCE, rough=0.0))
    o0 = out[lv[0]["offset"]:lv[0]["offset"] + 4096]
    col = m.data.texels[lv[5]["offset"]:lv[5]["offset"] + 4096]
    assert np.array_equal(o0["diffuse"], col["diffuse"])
    assert np.all(o0["packed_normal_roughness"].view(np.uint32) >> 24 == 1)  # rint(127 / 126) = 1
