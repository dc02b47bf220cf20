"""Pins the CPU oracle (oracle/) against what the mathematics fixes, without a GPU.

The reference has no tests and no golden vectors (its CI only builds), so every check here is a
closed form, an exhaustive integer enumeration, an invariant or a brute-force model, chosen so that a
plausible mistake (wrong rounding, a dropped clamp, a swapped operand, a transposed lane) fails.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
import scenes
import sw_abi as A

lib = oracle.load()


# ---------------------------------------------------------------------------------------------
# Colour conversion (fast_math.rs:131-160): min(x, 255) -> cvtps2dq (half-even) -> low byte
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("x,expected", [
    (0.0, 0x00), (0.5, 0x00), (1.5, 0x02), (2.5, 0x02), (3.5, 0x04), (254.5, 0xFE), (254.51, 0xFF), (255.0, 0xFF),
    (300.0, 0xFF), (1e30, 0xFF), (float("inf"), 0xFF), (float("nan"), 0xFF),
    (-0.4, 0x00), (-0.5, 0x00), (-0.6, 0xFF), (-1.0, 0xFF), (-1.5, 0xFE), (-300.0, 0xD4), (-3e9, 0x00),
])
def test_color32_conversion_closed_form(x, expected):
    for lane in range(4):
        v = [0.0, 0.0, 0.0, 0.0]
        v[lane] = x
        got = lib.oracle_color32_from_f32x4((C.c_float * 4)(*v))
        assert (got >> (8 * lane)) & 0xFF == expected, (x, lane, hex(got))
        assert got & ~(0xFF << (8 * lane)) & 0xFFFFFFFF == 0


@pytest.mark.parametrize("x,expected", [(0.5, 0), (1.5, 2), (65534.5, 65534), (65535.0, 65535), (70000.0, 65535),
                                        (float("nan"), 65535), (-1.0, 0xFFFF), (-2.5, 0xFFFE)])
def test_color64_conversion_closed_form(x, expected):
    got = lib.oracle_color64_from_f32x4((C.c_float * 4)(0.0, x, 0.0, 0.0))
    assert (got >> 16) & 0xFFFF == expected


# ---------------------------------------------------------------------------------------------
# Blending (abstract_color.rs, fast_math.rs:24-62, rasterizer.rs:1511-1536), exhaustive per channel
# ---------------------------------------------------------------------------------------------
def _blend_table(mode):
    a = np.arange(256, dtype=np.uint32)
    dst, src = np.meshgrid(a, a, indexing="ij")
    dst, src = dst.ravel(), src.ravel()
    out = np.zeros_like(dst)
    for i in range(dst.size):
        out[i] = lib.oracle_blend32(mode, int(dst[i]) * 0x010101, int(src[i]) * 0x010101 | 0x80000000)
    return dst, src, out


def test_blend_average_and_additive_exhaustive():
    a = np.arange(256, dtype=np.uint32)
    dst, src = np.meshgrid(a, a, indexing="ij")
    dst, src = dst.ravel(), src.ravel()
    f = np.vectorize(lambda d, s, m: lib.oracle_blend32(m, int(d), int(s)), otypes=[np.uint32])
    for lane in range(4):
        sh = 8 * lane
        avg = f(dst << sh, src << sh, A.BLEND_AVERAGE) >> sh
        assert np.array_equal(avg, (dst + src + 1) >> 1)  # pavgb
        add = f(dst << sh, src << sh, A.BLEND_ADDITIVE) >> sh
        assert np.array_equal(add, np.minimum(dst + src, 255))  # paddusb


def test_blend_alpha_exhaustive():
    a = np.arange(256, dtype=np.uint32)
    dst, alpha = np.meshgrid(a, a, indexing="ij")
    dst, alpha = dst.ravel(), alpha.ravel()
    src_rgb = 200
    f = np.vectorize(lambda d, al, m: lib.oracle_blend32(m, int(d) * 0x010101, (int(al) << 24) | src_rgb * 0x010101),
                     otypes=[np.uint32])
    out = f(dst, alpha, A.BLEND_ALPHA_BLEND)
    expect = (dst * (255 - alpha) + src_rgb * alpha) >> 8
    assert np.array_equal(out & 0xFF, expect)
    assert np.array_equal((out >> 16) & 0xFF, expect)
    # alpha lane: (dst_a * (255 - a) + a * a) >> 8 with dst_a = 0
    assert np.array_equal(out >> 24, (alpha * alpha) >> 8)
    t = f(dst, alpha, A.BLEND_ALPHA_TEST)
    assert np.array_equal(t & 0xFF, np.where(alpha > 0x7F, src_rgb, dst))


def test_blend_64_random():
    rng = np.random.default_rng(0)
    for _ in range(3000):
        d = [int(x) for x in rng.integers(0, 65536, size=4)]
        s = [int(x) for x in rng.integers(0, 65536, size=4)]
        dv = sum(x << (16 * i) for i, x in enumerate(d))
        sv = sum(x << (16 * i) for i, x in enumerate(s))
        avg = lib.oracle_blend64(A.BLEND_AVERAGE, dv, sv)
        add = lib.oracle_blend64(A.BLEND_ADDITIVE, dv, sv)
        for i in range(4):
            assert (avg >> (16 * i)) & 0xFFFF == (d[i] + s[i] + 1) >> 1
            assert (add >> (16 * i)) & 0xFFFF == min(d[i] + s[i], 65535)


# ---------------------------------------------------------------------------------------------
# PackedNormalRoughness (textures.rs:31-62, fast_math.rs:64-96), exhaustive per i8 lane
# ---------------------------------------------------------------------------------------------
def test_pack_unpack_exhaustive():
    out = (C.c_float * 4)()
    for b in range(-128, 128):
        packed = (b & 0xFF) | ((b & 0xFF) << 8) | ((b & 0xFF) << 16) | ((b & 0xFF) << 24)
        lib.oracle_unpack_normal_roughness(C.c_int32(np.int32(np.uint32(packed)).item()), out)
        for i in range(4):
            assert out[i] == np.float32(b) * np.float32(1.0 / 127.0)
        # re-packing returns the byte: x * 127 / 127 rounds back, and packs_epi16 saturates to
        # [-128, 127] so -128 survives (the scalar fallback would clamp it to -127, fast_math.rs:424)
        v = (C.c_float * 3)(out[0], out[1], out[2])
        rp = lib.oracle_pack_normal_roughness(v, out[3]) & 0xFFFFFFFF
        assert rp == packed


@pytest.mark.parametrize("x", [0.5 / 127, 1.5 / 127, -0.5 / 127, 2.0, -2.0, 0.999, -1.0])
def test_pack_rounding_and_saturation(x):
    v = (C.c_float * 3)(x, 0.0, 0.0)
    p = lib.oracle_pack_normal_roughness(v, 1.0) & 0xFF
    q = int(np.clip(np.rint(np.float32(x) * np.float32(127.0)), -128, 127))
    assert p == (q & 0xFF)


# ---------------------------------------------------------------------------------------------
# Fixed16 (common/fixed_math.rs)
# ---------------------------------------------------------------------------------------------
def _i32(x):
    return ((x + 2**31) % 2**32) - 2**31


def _trunc_div(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def test_fixed16_ops_against_integer_arithmetic():
    rng = np.random.default_rng(1)
    vals = [0, 1, -1, 65536, -65536, 2**31 - 1, -2**31, 32768, -32768, 3, -3] + [int(v) for v in rng.integers(-2**31, 2**31, 400)]
    for a in vals:
        for b in vals[:40]:
            assert lib.oracle_fixed16_mul(a, b) == _i32((a * b) >> 16)
            if b != 0:
                assert lib.oracle_fixed16_div(a, b) == _i32(_trunc_div(a * 65536, b))
                assert lib.oracle_fixed16_mul_div(a, 12345, b) == _i32(_trunc_div(a * 12345, b))


# ---------------------------------------------------------------------------------------------
# Specular closed forms (surfaces.rs:1088-1154)
# ---------------------------------------------------------------------------------------------
def test_specular_closed_forms():
    rng = np.random.default_rng(2)
    for _ in range(2000):
        c = float(rng.uniform(-1, 1))
        ir = float(rng.uniform(0.75, 50))
        x = (c - 1.0) * ir
        expect = ir / (x * x * math.pi * math.sqrt(7) / 2 + 2 * math.pi / math.sqrt(7))
        got = lib.oracle_specular_intensity(oracle.MODE_EXACT, c, ir)
        assert abs(got - expect) <= 2e-6 * max(1.0, abs(expect))
        got_sse = lib.oracle_specular_intensity(oracle.MODE_SSE, c, ir)
        assert abs(got_sse - expect) <= 1.5 * 2**-12 * abs(expect) + 1e-6
        f = lib.oracle_fresnel_factor_base(c)
        assert abs(f - max(1 - c, 0) ** 5) <= 1e-6 * max(1, max(1 - c, 0) ** 5)
        r = float(rng.uniform(0, 1))
        kd = lib.oracle_specular_k_dielectric(f, r)
        kd_ref = (f * 0.96 + 0.04) * (1 - r) + 0.12 * r
        assert abs(kd - kd_ref) < 1e-6 * max(1.0, abs(kd_ref))
        km = lib.oracle_specular_k_metal(f, r)
        km_ref = f * (1 - r) + 0.5 * r
        assert abs(km - km_ref) < 1e-6 * max(1.0, abs(km_ref))


# ---------------------------------------------------------------------------------------------
# Mips (textures.rs:288-362)
# ---------------------------------------------------------------------------------------------
def _mips(src, w, h):
    out = np.zeros(w * h, dtype=A.TEXEL_DTYPE)
    n = lib.oracle_build_texture_mips(C.c_void_p(src.ctypes.data), w, h, C.c_void_p(out.ctypes.data))
    return out[:n]


def test_mips_constant_texture_stays_constant():
    src = np.zeros(64 * 64, dtype=A.TEXEL_DTYPE)
    src["diffuse"] = 0x80C04020
    src["packed_normal_roughness"] = scenes.pack_normal_roughness(np.tile([[0, 0, 1.0]], (64 * 64, 1)), np.full(64 * 64, 0.5))
    out = _mips(src, 64, 64)
    assert len(out) == 32 * 32 + 16 * 16 + 8 * 8 + 4 * 4
    assert np.all(out["diffuse"] == 0x80C04020)
    assert np.all(out["packed_normal_roughness"] == src["packed_normal_roughness"][0])


def test_mips_average4_by_hand():
    src = np.zeros(16 * 16, dtype=A.TEXEL_DTYPE)
    src["packed_normal_roughness"] = 0x7F7F0000  # normal (0, 0, 1), roughness 1
    src["diffuse"][[0, 1, 16, 17]] = [0xFF000000, 0x01020304, 0x03040506, 0x05060708]
    out = _mips(src, 16, 16)
    # per channel floor(sum / 4); alpha sums (a >> 2) then masks
    assert out["diffuse"][0] == ((0 + 4 + 6 + 8) >> 2) | (((0 + 3 + 5 + 7) >> 2) << 8) | (((0 + 2 + 4 + 6) >> 2) << 16) | \
        ((((0xFF000000 >> 2) + (0x01000000 >> 2) + (0x03000000 >> 2) + (0x05000000 >> 2)) & 0xFF000000))


def test_front_end_mips_match_oracle():
    """The front-end's build_texture_mips (product, load time) is bit-exact with the oracle's."""
    import sw_host
    hl = sw_host.load()
    rng = np.random.default_rng(3)
    src = np.zeros(64 * 32, dtype=A.TEXEL_DTYPE)
    src["diffuse"] = rng.integers(0, 2**32, size=src.size, dtype=np.uint64).astype(np.uint32)
    src["packed_normal_roughness"] = scenes.pack_normal_roughness(scenes._rand_normals(rng, src.size),
                                                                  rng.uniform(1 / 126, 1, size=src.size))
    ref = _mips(src, 64, 32)
    out = np.zeros_like(ref)
    n = hl.swh_build_texture_mips(C.c_void_p(src.ctypes.data), 64, 32, C.c_void_p(out.ctypes.data))
    assert n == len(ref)
    assert np.array_equal(out, ref)


# ---------------------------------------------------------------------------------------------
# Surface builder closed forms (surfaces.rs:433-971)
# ---------------------------------------------------------------------------------------------
def _single_surface_scene(texels, tw, th, lm_simple=None, lm_dir=None, surf_kw=None, lights=None, color_bits=32,
                          lite=None, tex_flags=0):
    levels = [(0, tw, th, tex_flags, 0)]
    if lite is not None:
        levels.append((0, lite.shape[1], lite.shape[0], A.TEX_LITE_LEVEL, 0))
    m = A.MapData(texels, lite.ravel() if lite is not None else np.zeros(1, np.uint32),
                  lm_simple if lm_simple is not None else np.zeros((1, 3), np.float32),
                  lm_dir if lm_dir is not None else np.zeros(1, A.DIRLM_DTYPE), np.array(levels, dtype=A.LEVEL_DTYPE))
    s = np.zeros(1, dtype=A.SURFACE_DTYPE)
    s["basis_u"] = (1, 0, 0)
    s["basis_v"] = (0, 1, 0)
    s["basis_normal"] = (0, 0, 1)
    s["size"] = (tw, th)
    s["flags"] = A.SURF_DIFFUSE_LAYER
    s["emissive_level"] = A.NO_TEXTURE
    for k, v in (surf_kw or {}).items():
        s[k] = v
    v = np.zeros(1, dtype=A.VIEW_DTYPE)
    v["width"], v["height"], v["num_surfaces"] = 4, 4, 1
    ls = lights if lights is not None else np.zeros(0, A.LIGHT_DTYPE)
    b = A.BatchData(v, s, ls, np.arange(len(ls), dtype=np.uint32), np.zeros(0, A.DRAW_DTYPE), np.zeros(0, np.int32),
                    color_bits)
    arena, _ = oracle.build_surfaces(m, b)
    return arena


def _texels(diffuse):
    t = np.zeros(diffuse.size, dtype=A.TEXEL_DTYPE)
    t["diffuse"] = diffuse.ravel()
    t["packed_normal_roughness"] = 0x7F7F0000
    return t


def test_surface_without_lightmap_is_black_with_zero_alpha():
    rng = np.random.default_rng(4)
    tex = _texels(rng.integers(0, 2**32, size=64, dtype=np.uint64).astype(np.uint32))
    arena = _single_surface_scene(tex, 8, 8)
    assert np.all(arena[:64] == 0)


@pytest.mark.parametrize("L", [0.0, 0.37, 1.0, 1.5, 3.0])
def test_constant_lightmap_scales_texel(L):
    rng = np.random.default_rng(5)
    diffuse = rng.integers(0, 2**32, size=16 * 16, dtype=np.uint64).astype(np.uint32)
    lm = np.full((1 + 3 * 3, 3), np.float32(L), dtype=np.float32)
    for mode in (oracle.MODE_EXACT, oracle.MODE_SSE):
        arena = _single_surface_scene(_texels(diffuse), 16, 16, lm_simple=lm,
                                      surf_kw={"lightmap_offset": 1, "lightmap_size": (3, 3), "lightmap_scale_log2": 4})
        for lane in range(3):
            ch = (diffuse >> (8 * lane)) & 0xFF
            expect = np.rint(np.minimum(ch.astype(np.float32) * np.float32(L), 255)).astype(np.uint32)
            assert np.array_equal((arena[:256] >> (8 * lane)) & 0xFF, expect)
        assert np.array_equal(arena[:256] >> 24, diffuse >> 24)  # alpha * 1.0


def test_lightmap_ramp_interpolation():
    """Lightmap linear in u: the light at texel u is linear with k = (base & mask + 0.5) / scale."""
    tex = _texels(np.full(64 * 4, 0x00FFFFFF, dtype=np.uint32))
    lm = np.zeros((1 + 5 * 3, 3), dtype=np.float32)
    for j in range(3):
        for i in range(5):
            lm[1 + j * 5 + i] = i * 0.25  # 0, 0.25, 0.5, 0.75, 1.0
    arena = _single_surface_scene(tex, 64, 4, lm_simple=lm,
                                  surf_kw={"lightmap_offset": 1, "lightmap_size": (5, 3), "lightmap_scale_log2": 4,
                                           "lightmap_tc_shift": (0, 0)})
    u = np.arange(64)
    light = ((u >> 4) + ((u & 15) + 0.5) / 16.0) * 0.25
    expect = np.rint(np.minimum(255.0 * light, 255.0))
    got = (arena[:64] & 0xFF).astype(np.float64)
    assert np.max(np.abs(got - expect)) <= 1
    assert np.all(np.diff(got) >= 0)


def test_single_dynamic_light_lambert():
    """No lightmap, one unshadowed light above a flat z-normal surface: Lambert * (1/d^2 - 1/r^2)+."""
    tex = _texels(np.full(16 * 16, 0x00C8C8C8, dtype=np.uint32))
    lights = np.zeros(1, A.LIGHT_DTYPE)
    lights["position"] = (3.0, 5.0, 10.0)
    lights["radius"] = 30.0
    lights["inv_square_radius"] = np.float32(1 / 900.0)
    lights["color"] = (500.0, 500.0, 500.0)
    arena = _single_surface_scene(tex, 16, 16, lights=lights, surf_kw={"num_lights": 1})
    x, y = np.meshgrid(np.arange(16) + 0.5, np.arange(16) + 0.5)
    d = np.stack([3.0 - x, 5.0 - y, np.full_like(x, 10.0)], -1)
    d2 = (d ** 2).sum(-1)
    light = 500.0 * (10.0 / np.sqrt(d2)) * np.maximum(1 / d2 - 1 / 900.0, 0)
    expect = np.rint(np.minimum(200.0 * light, 255)).ravel()
    got = (arena[:256] & 0xFF).astype(np.float64)
    assert np.max(np.abs(got - expect)) <= 1
    assert np.all(arena[:256] >> 24 == 0)  # alpha stays 0 without a lightmap


@pytest.mark.parametrize("blend", range(5))
def test_emissive_layer_integer_formulas(blend):
    rng = np.random.default_rng(6 + blend)
    lite = rng.integers(0, 2**32, size=(8, 8), dtype=np.uint64).astype(np.uint32)
    tex = _texels(np.full(64, 0x00404040, dtype=np.uint32))
    lm = np.full((1 + 9, 3), np.float32(1.0), dtype=np.float32)
    light = (1.25, 0.5, 2.0)
    arena = _single_surface_scene(tex, 8, 8, lm_simple=lm, lite=lite,
                                  surf_kw={"lightmap_offset": 1, "lightmap_size": (3, 3), "lightmap_scale_log2": 4,
                                           "flags": A.SURF_DIFFUSE_LAYER | A.SURF_EMISSIVE, "emissive_level": 1,
                                           "emissive_light": light, "emissive_blend": blend})
    base = np.uint32(0x00404040)
    lv = [int(np.rint(np.float32(l) * 256)) for l in light]  # [R, G, B]
    t = lite.ravel()
    for lane, li in ((0, lv[2]), (1, lv[1]), (2, lv[0])):
        tc = ((t >> (8 * lane)) & 0xFF).astype(np.int64)
        mod = (tc * li) >> 8
        dst = (int(base) >> (8 * lane)) & 0xFF
        if blend == A.BLEND_NONE:
            e = mod
        elif blend == A.BLEND_AVERAGE:
            e = (mod + dst) >> 1
        elif blend == A.BLEND_ADDITIVE:
            e = mod + dst
        elif blend == A.BLEND_ALPHA_TEST:
            e = np.where((t >> 24) > 0x7F, mod, dst)
        else:
            al = (t >> 24).astype(np.int64)
            e = (mod * al + dst * (255 - al)) >> 8
        e = np.where((e > 0) & (e < 32768), np.minimum(e, 255), 0)
        assert np.array_equal(((arena[:64] >> (8 * lane)) & 0xFF).astype(np.int64), e), lane


def _dir_lightmap(ambient, vec, color, deviation=0.25, n=1 + 9):
    lm = np.zeros(n, dtype=A.DIRLM_DTYPE)
    lm["ambient_light"] = ambient
    lm["light_direction_vector_scaled"] = vec
    lm["directional_light_deviation"] = deviation
    lm["directional_light_color"] = color
    return lm


def _wrap_u8(x):
    """cvtps2dq (half-even) of min(x, 255), low byte (fast_math.rs:131-144)."""
    return (np.rint(np.minimum(x, 255.0)).astype(np.int64)) & 0xFF


@pytest.mark.parametrize("dz", [0.8, 0.0, -0.35])
def test_directional_lightmap_without_normal_map(dz):
    """Constant directional element, no normal map, no specular (surfaces.rs:551-572): light =
    ambient + color * vector_scaled.z, with z NOT clamped, so a negative z darkens below ambient and
    a negative total wraps to the low byte. Distinct R, G, B values pin the lane order
    (fast_math.rs:167-175: [R, G, B] arrays land in lanes R, G, B of the B, G, R, A vector)."""
    rng = np.random.default_rng(11)
    diffuse = rng.integers(0, 2**32, size=16 * 16, dtype=np.uint64).astype(np.uint32)
    ambient = np.array([0.30, 0.55, 0.10], np.float32)
    color = np.array([1.40, 0.35, 0.90], np.float32)
    vec = np.array([0.3, -0.6, dz], np.float32)
    lm = _dir_lightmap(ambient, vec, color)
    arena = _single_surface_scene(_texels(diffuse), 16, 16, lm_dir=lm,
                                  surf_kw={"lightmap_offset": 1, "lightmap_size": (3, 3), "lightmap_scale_log2": 4,
                                           "flags": A.SURF_DIFFUSE_LAYER | A.SURF_DIRECTIONAL})
    light_rgb = ambient.astype(np.float64) + color.astype(np.float64) * float(dz)
    for lane, li in ((0, light_rgb[2]), (1, light_rgb[1]), (2, light_rgb[0])):
        ch = ((diffuse >> (8 * lane)) & 0xFF).astype(np.float64)
        expect = _wrap_u8(ch * li)
        got = ((arena[:256] >> (8 * lane)) & 0xFF).astype(np.int64)
        diff = np.minimum((got - expect) % 256, (expect - got) % 256)
        assert diff.max() <= 1, (lane, dz)
    assert np.array_equal(arena[:256] >> 24, diffuse >> 24)  # alpha * the lane-3 1.0


def test_directional_lightmap_with_normal_map():
    """With a normal map the directional term is color * max(dot(vector_scaled, n), 0), n = the
    texel's packed signed bytes / 127 (textures.rs:41-55), not renormalised."""
    rng = np.random.default_rng(12)
    n = 16 * 16
    diffuse = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    normals = scenes._rand_normals(rng, n, z_min=-0.5)
    tex = np.zeros(n, dtype=A.TEXEL_DTYPE)
    tex["diffuse"] = diffuse
    tex["packed_normal_roughness"] = scenes.pack_normal_roughness(normals, np.ones(n))
    ambient = np.array([0.05, 0.10, 0.20], np.float32)
    color = np.array([0.9, 0.6, 1.3], np.float32)
    vec = np.array([0.5, 0.2, 0.7], np.float32)
    lm = _dir_lightmap(ambient, vec, color)
    arena = _single_surface_scene(tex, 16, 16, lm_dir=lm, tex_flags=A.TEX_HAS_NORMAL_MAP,
                                  surf_kw={"lightmap_offset": 1, "lightmap_size": (3, 3), "lightmap_scale_log2": 4,
                                           "flags": A.SURF_DIFFUSE_LAYER | A.SURF_DIRECTIONAL})
    packed = tex["packed_normal_roughness"].view(np.uint32)
    nb = np.stack([((packed >> (8 * i)) & 0xFF).astype(np.int8).astype(np.float64) for i in range(3)], -1) / 127.0
    dot = np.maximum(nb @ vec.astype(np.float64), 0.0)
    assert np.any(dot == 0.0) and np.any(dot > 0.5)  # both sides of the clamp are exercised
    for lane, c in ((0, 2), (1, 1), (2, 0)):
        ch = ((diffuse >> (8 * lane)) & 0xFF).astype(np.float64)
        expect = _wrap_u8(ch * (ambient[c] + color[c] * dot))
        got = ((arena[:256] >> (8 * lane)) & 0xFF).astype(np.int64)
        assert np.max(np.abs(got - expect)) <= 1, lane


# ---------------------------------------------------------------------------------------------
# Rasteriser invariants (rasterizer.rs)
# ---------------------------------------------------------------------------------------------
def _raster(draw_specs, W, H, textures, color_bits=32, clip=None):
    """draw_specs: list of (vertices_fixed16 flat, de, te, tex_index, blend, mode)."""
    arena = np.concatenate([t.ravel() for t in textures]).astype(np.uint32 if color_bits == 32 else np.uint64)
    offs = np.cumsum([0] + [t.size for t in textures])
    draws = np.zeros(len(draw_specs), dtype=A.DRAW_DTYPE)
    vdata = []
    vi = 0
    surfs = np.zeros(len(textures), dtype=A.SURFACE_DTYPE)
    for k, t in enumerate(textures):
        surfs[k]["size"] = (t.shape[1], t.shape[0])
        surfs[k]["pixels_offset"] = offs[k]
        surfs[k]["emissive_level"] = A.NO_TEXTURE
    for i, (verts, de, te, ti, blend, mode) in enumerate(draw_specs):
        d = draws[i]
        d["kind"], d["blend"], d["mode"] = A.DRAW_POLYGON, blend, mode
        d["first_vertex_int"], d["num_vertices"] = vi, len(verts) // 2
        d["depth_equation"], d["tc_equation"] = de, te
        d["texture_offset"] = offs[ti]
        d["texture_size"] = (textures[ti].shape[1], textures[ti].shape[0])
        d["clip_rect"] = clip if clip is not None else (0, 0, W, H)
        vdata.append(np.asarray(verts, np.int32))
        vi += len(verts)
    v = np.zeros(1, dtype=A.VIEW_DTYPE)
    v["width"], v["height"], v["num_draws"], v["num_surfaces"] = W, H, len(draws), 0
    b = A.BatchData(v, np.zeros(0, A.SURFACE_DTYPE), np.zeros(0, A.LIGHT_DTYPE), np.zeros(0, np.uint32), draws,
                    np.concatenate(vdata), color_bits, arena_texels=arena.size)
    fb, _, st, oob = oracle.render(A.MapData(np.zeros(1, A.TEXEL_DTYPE), np.zeros(1, np.uint32), np.zeros((1, 3), np.float32),
                                              np.zeros(1, A.DIRLM_DTYPE), np.zeros(0, A.LEVEL_DTYPE)), b, arena=arena.copy())
    return fb.reshape(H, W), st, oob


AFFINE_FLAT = (np.array([0.0, 0.0, 1.0], np.float32), np.array([1.0, 0.0, 0.0, 1.0, 0.0, 0.0], np.float32))


def _fx(pts):
    return [int(np.int32(np.float32(c) * np.float32(65536.0))) for p in pts for c in p]


def test_axis_aligned_rectangle_exact_coverage():
    tex = np.full((64, 64), 7, dtype=np.uint32)
    de, te = AFFINE_FLAT
    for mode in (A.TC_AFFINE, A.TC_PERSPECTIVE, A.TC_LINE_Z):
        # vertices at k + 0.5 cover pixels [x0, x1) x [y0, y1) exactly (clockwise on screen)
        pts = [(3.5, 2.5), (20.5, 2.5), (20.5, 11.5), (3.5, 11.5)]
        te2 = te.copy()
        fb, st, oob = _raster([(_fx(pts), de, te2, 0, A.BLEND_NONE, mode)], 32, 16, [tex])
        expect = np.zeros((16, 32), np.uint32)
        expect[3:12, 4:21] = 7
        assert oob == 0
        assert np.array_equal(fb, expect), mode


def test_identity_texture_coordinates_fetch_pixel_index():
    """tc(x, y) = (x, y): every pixel fetches the texel with its own coordinates. The texture is
    wider than the span: the affine path clamps edge coordinates to size - 1/65536
    (rasterizer.rs:622-643), which would shorten the step of a texture exactly as wide."""
    W, H = 40, 24
    tex = np.arange((W + 8) * (H + 8), dtype=np.uint32).reshape(H + 8, W + 8)
    de, te = AFFINE_FLAT
    pts = [(0.0, 0.0), (float(W), 0.0), (float(W), float(H)), (0.0, float(H))]
    fb, st, oob = _raster([(_fx(pts), de, te, 0, A.BLEND_NONE, A.TC_AFFINE)], W, H, [tex])
    assert oob == 0
    assert np.array_equal(fb, tex[:H, :W])


def test_constant_depth_all_modes_agree():
    """With d_inv_z_dx = d_inv_z_dy = 0 the three interpolation modes agree."""
    rng = np.random.default_rng(7)
    tex = rng.integers(0, 2**32, size=(32, 32), dtype=np.uint64).astype(np.uint32)
    for _ in range(20):
        pts = sorted_poly(rng, 48, 40)
        de = np.array([0.0, 0.0, 1.0 / 16.0], np.float32)  # z = 16 everywhere (power of two scale)
        a, b = rng.uniform(0.1, 0.5, size=2)
        te = np.array([a / 16, 0, 0, b / 16, 1 / 16, 1 / 16], np.float32)
        outs = []
        for mode in (A.TC_AFFINE, A.TC_LINE_Z, A.TC_PERSPECTIVE):
            fb, _, oob = _raster([(_fx(pts), de, te, 0, A.BLEND_NONE, mode)], 48, 40, [tex])
            assert oob == 0
            outs.append(fb)
        cover = [o != 0 for o in outs]
        assert np.array_equal(cover[0], cover[1]) and np.array_equal(cover[0], cover[2])
        # texels may differ by the modes' fixed-point rounding, but only by one texel step
        assert np.mean(outs[1] == outs[2]) > 0.9


def sorted_poly(rng, W, H):
    cx, cy = rng.uniform(8, W - 8), rng.uniform(8, H - 8)
    n = int(rng.integers(3, 8))
    ang = np.sort(rng.uniform(-np.pi, np.pi, size=n))
    r = rng.uniform(3, min(W, H) / 2 - 2)
    return [(cx + r * np.cos(t), cy + r * np.sin(t) * 0.8) for t in ang]


def test_shared_edge_is_watertight():
    """Two convex polygons sharing an edge: every pixel of the union is written exactly once."""
    rng = np.random.default_rng(8)
    one = np.full((4, 4), 1, dtype=np.uint32)
    de, te = AFFINE_FLAT
    te = np.zeros(6, np.float32)
    for _ in range(300):
        W, H = 64, 48
        # the shared edge runs from above the screen to below it, so the union covers the screen
        a = (rng.uniform(0, W), -10.0)
        b = (rng.uniform(0, W), H + 10.0)
        left = [(-10.0, -10.0), a, b, (-10.0, H + 10.0)]  # clockwise on screen
        right = [a, (W + 10.0, -10.0), (W + 10.0, H + 10.0), b]
        specs = [(_fx(left), de, te, 0, A.BLEND_ADDITIVE, A.TC_AFFINE), (_fx(right), de, te, 0, A.BLEND_ADDITIVE, A.TC_AFFINE)]
        fb, st, oob = _raster(specs, W, H, [one])
        assert np.all(fb == 1), np.unique(fb, return_counts=True)


def test_writes_stay_inside_clip_rect():
    rng = np.random.default_rng(9)
    tex = np.full((8, 8), 5, dtype=np.uint32)
    de, te = AFFINE_FLAT
    te = np.zeros(6, np.float32)
    for _ in range(50):
        pts = sorted_poly(rng, 64, 48)
        clip = (int(rng.integers(0, 30)), int(rng.integers(0, 20)), int(rng.integers(31, 64)), int(rng.integers(21, 48)))
        fb, st, oob = _raster([(_fx(pts), de, te, 0, A.BLEND_NONE, A.TC_AFFINE)], 64, 48, [tex], clip=clip)
        mask = np.zeros_like(fb, dtype=bool)
        mask[clip[1]:clip[3], clip[0]:clip[2]] = True
        assert not np.any(fb[~mask])


def test_painter_order_last_writer_wins():
    rng = np.random.default_rng(10)
    W, H = 48, 40
    texs = [np.full((4, 4), i + 1, dtype=np.uint32) for i in range(6)]
    te = np.zeros(6, np.float32)
    de = AFFINE_FLAT[0]
    for _ in range(30):
        polys = [sorted_poly(rng, W, H) for _ in range(6)]
        specs = [(_fx(p), de, te, i, A.BLEND_NONE, A.TC_AFFINE) for i, p in enumerate(polys)]
        fb, _, _ = _raster(specs, W, H, texs)
        expect = np.zeros((H, W), np.uint32)
        for i, p in enumerate(polys):
            alone, _, _ = _raster([specs[i]], W, H, texs)
            expect = np.where(alone != 0, alone, expect)
        assert np.array_equal(fb, expect)


# ---------------------------------------------------------------------------------------------
# Triangles (rasterizer.rs:731-1292): textured Gouraud over lite textures
# ---------------------------------------------------------------------------------------------
def _raster_tris(tris, W, H, lite, blend=A.BLEND_NONE, color_bits=32, bg=0):
    """tris: list of 3-vertex lists of (x, y, tc_u, tc_v, light_r, light_g, light_b) in floats."""
    draws = np.zeros(len(tris), dtype=A.DRAW_DTYPE)
    vdata = []
    for i, tri in enumerate(tris):
        d = draws[i]
        d["kind"], d["blend"] = A.DRAW_TRIANGLE, blend
        d["first_vertex_int"], d["num_vertices"] = 21 * i, 3
        d["texture_offset"] = 0
        d["texture_size"] = (lite.shape[1], lite.shape[0])
        d["clip_rect"] = (0, 0, W, H)
        vdata.append(np.array(_fx(tri), np.int32))
    v = np.zeros(1, dtype=A.VIEW_DTYPE)
    v["width"], v["height"], v["num_draws"] = W, H, len(draws)
    v["clear_background"], v["background_color"] = 1, bg
    b = A.BatchData(v, np.zeros(0, A.SURFACE_DTYPE), np.zeros(0, A.LIGHT_DTYPE), np.zeros(0, np.uint32), draws,
                    np.concatenate(vdata), color_bits, arena_texels=1)
    m = A.MapData(np.zeros(1, A.TEXEL_DTYPE), lite.ravel().astype(np.uint32), np.zeros((1, 3), np.float32),
                  np.zeros(1, A.DIRLM_DTYPE), np.zeros(0, A.LEVEL_DTYPE))
    fb, _, st, oob = oracle.render(m, b)
    return fb.reshape(H, W), st, oob


@pytest.mark.parametrize("light", [(1.0, 1.0, 1.0), (1.5, 0.25, 0.75), (3.0, 2.0, 0.0)])
def test_triangle_constant_light_closed_form(light):
    """Constant texel T and constant light L: every covered pixel is min((T_c * L_c) >> 16, 255)
    per channel with alpha 0 (rasterizer.rs:1239-1240, ColorVecI saturating pack)."""
    T = 0x80C864F0
    lite = np.full((8, 8), T, dtype=np.uint32)
    lr, lg, lb = light
    quad = [(3.5, 2.5), (20.5, 2.5), (20.5, 11.5), (3.5, 11.5)]
    v = [(x, y, 4.0, 4.0, lr, lg, lb) for x, y in quad]
    fb, st, oob = _raster_tris([[v[0], v[1], v[2]], [v[0], v[2], v[3]]], 32, 16, lite, bg=0x11111111)
    assert oob == 0
    exp = 0
    for lane, lc in ((0, lb), (1, lg), (2, lr)):
        c = (T >> (8 * lane)) & 0xFF
        exp |= min((c * int(lc * 65536)) >> 16, 255) << (8 * lane)
    expect = np.full((16, 32), 0x11111111, np.uint32)
    expect[3:12, 4:21] = exp
    assert np.array_equal(fb, expect)


def test_triangle_pair_is_watertight():
    """Two triangles sharing a random edge, together covering the screen: with additive blending of
    a texel 1 at light 1.0 every pixel is written exactly once."""
    rng = np.random.default_rng(13)
    lite = np.full((4, 4), 1, dtype=np.uint32)
    W, H = 48, 40
    for _ in range(200):
        a = (rng.uniform(-5, W + 5), rng.uniform(-5, H + 5))
        big = [(-20.0, -20.0), (W + 20.0, -20.0), (W + 20.0, H + 20.0), (-20.0, H + 20.0)]
        # split the big quad into a fan around `a`: 4 triangles sharing edges
        tris = []
        for j in range(4):
            q0, q1 = big[j], big[(j + 1) % 4]
            tris.append([(q0[0], q0[1], 1.0, 1.0, 1.0, 1.0, 1.0), (q1[0], q1[1], 1.0, 1.0, 1.0, 1.0, 1.0),
                         (a[0], a[1], 1.0, 1.0, 1.0, 1.0, 1.0)])
        fb, st, oob = _raster_tris(tris, W, H, lite, blend=A.BLEND_ADDITIVE)
        assert oob == 0
        assert np.all(fb == 1), np.unique(fb, return_counts=True)


# ---------------------------------------------------------------------------------------------
# Tonemap (postprocessor.rs:1003-1024): Reinhard, exhaustive over 16-bit channel values
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("exposure", [0.125, 1.0, 7.5])
def test_tonemap_closed_form_exhaustive(exposure):
    vals = np.arange(65536, dtype=np.uint64)
    hdr = (vals | (vals << 16) | (vals << 32)).reshape(256, 256)
    out, sums = oracle.tonemap(hdr, exposure, (1.0, 1.0, 1.0))
    c = vals.astype(np.float64)
    mapped = c / (c / 255.0 + 1.0 / exposure)
    expect = np.rint(np.minimum(mapped, 255.0))
    got = (out.ravel() & 0xFF).astype(np.float64)
    assert np.max(np.abs(got - expect)) <= 1
    assert np.mean(got == expect) > 0.999
    assert np.all((out >> 24) == 255)  # alpha lane is c / 0 -> +inf or NaN -> 255


# ---------------------------------------------------------------------------------------------
# Numeric modes: the x86 (rsqrtss / rcpss) mode stays within +-1 LSB of the exact mode
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("color_bits", [32, 64])
def test_sse_mode_within_one_lsb_of_exact(color_bits):
    m, b = scenes.make_batch(21, color_bits=color_bits, view_sizes=((160, 120), (64, 64)))
    _, a1, _, _ = oracle.render(m, b, mode=oracle.MODE_EXACT)
    _, a2, _, _ = oracle.render(m, b, mode=oracle.MODE_SSE)
    ch = np.uint16 if color_bits == 64 else np.uint8
    d = np.abs(a1.view(ch).astype(np.int64) - a2.view(ch).astype(np.int64))
    assert d.max() <= 1
