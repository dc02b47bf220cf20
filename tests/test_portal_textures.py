"""Portal texture overlays (build_view_portals_textures' texture mix, partial_renderer.rs:2205-2276):
mix_surface_with_texture (surfaces.rs:903-971) of a lite texture level over each textured portal's
rect of the portal textures.

The oracle's mix is pinned against a per-channel integer model written from the definition: the
source texel of rect texel (x, y) is ((tc_min[0] + x) mod w, (tc_min[1] + y) mod h) of the level,
each colour channel is multiplied by rint(light * 256) and shifted right by 8 (alpha by 0), then
blended into the destination by the mode's integer rule and packed with unsigned saturation
(for Color32 through the two packus steps, so a 16-bit intermediate above 32767 packs to 0). The
front-end's overlays are checked for their level, blend and placement in the level layout.
"""
import numpy as np
import pytest

import oracle
import sw_abi as A
import sw_host


def _levels_map(rng):
    """A map with two lite levels (8x4 and 5x3, the second not a power of two) and random texels."""
    lite = rng.integers(0, 2 ** 32, size=8 * 4 + 5 * 3, dtype=np.uint64).astype(np.uint32)
    levels = np.zeros(3, A.LEVEL_DTYPE)
    levels[0] = (0, 4, 4, 0, 0)  # a Texture level (not lite)
    levels[1] = (0, 8, 4, A.TEX_LITE_LEVEL, 0)
    levels[2] = (32, 5, 3, A.TEX_LITE_LEVEL, 0)
    texels = np.zeros(16, A.TEXEL_DTYPE)
    return A.MapData(texels, lite, np.zeros((1, 3), np.float32), np.zeros(1, A.DIRLM_DTYPE), levels), lite, levels


def _model_mix(dst, lite, levels, mix, bits):
    """The overlay of one mix over a copy of dst, from the definition."""
    out = dst.copy()
    lv = levels[int(mix["level"])]
    w, h, off = int(lv["width"]), int(lv["height"]), int(lv["offset"])
    tex = lite[off:off + w * h].reshape(h, w)
    rw, rh = (int(v) for v in mix["size"])
    tu, tv = (int(v) for v in mix["tc_start"])
    ys, xs = np.meshgrid(np.arange(rh), np.arange(rw), indexing="ij")
    t = tex[(tv + ys) % h, (tu + xs) % w].astype(np.int64).ravel()
    li = [int(np.rint(np.float32(mix["light"][c]) * np.float32(256.0))) for c in range(3)]  # R, G, B
    lanes_light = (li[2], li[1], li[0], 0)  # B, G, R, A
    d0 = int(mix["dst_offset"])
    d = out[d0:d0 + rw * rh].astype(np.uint64)
    shift, mask = (8, 0xFF) if bits == 32 else (16, 0xFFFF)
    res = np.zeros(rw * rh, np.uint64)
    blend = int(mix["blend"])
    al = (t >> 24) & 0xFF
    for lane in range(4):
        mod = (((t >> (8 * lane)) & 0xFF) * lanes_light[lane]) >> 8
        dc = ((d >> np.uint64(shift * lane)) & np.uint64(mask)).astype(np.int64)
        if blend == A.BLEND_NONE:
            e = mod
        elif blend == A.BLEND_AVERAGE:
            e = (mod + dc) >> 1
        elif blend == A.BLEND_ADDITIVE:
            e = mod + dc
        elif blend == A.BLEND_ALPHA_TEST:
            e = np.where(al > 0x7F, mod, dc)
        else:
            e = (mod * al + dc * (255 - al)) >> 8
        if bits == 32:
            e = np.where((e > 0) & (e < 32768), np.minimum(e, 255), 0)
        else:
            e = np.clip(e, 0, 65535)
        res |= e.astype(np.uint64) << np.uint64(shift * lane)
    out[d0:d0 + rw * rh] = res.astype(out.dtype)
    return out


def _random_mixes(rng, n_texels, count):
    mixes = np.zeros(count, A.PORTAL_MIX_DTYPE)
    base = 0
    for i in range(count):
        rw, rh = int(rng.integers(1, 23)), int(rng.integers(1, 9))
        base += int(rng.integers(0, 5))
        mixes[i]["dst_offset"] = base
        mixes[i]["size"] = (rw, rh)
        mixes[i]["tc_start"] = (int(rng.integers(-40, 40)), int(rng.integers(-40, 40)))
        mixes[i]["level"] = 1 + i % 2
        mixes[i]["blend"] = i % 5
        mixes[i]["light"] = rng.uniform(0.0, 3.0, size=3).astype(np.float32)
        base += rw * rh
    assert base <= n_texels
    return mixes


@pytest.mark.parametrize("bits", [32, 64])
def test_mix_follows_the_integer_model(bits):
    rng = np.random.default_rng(11 + bits)
    md, lite, levels = _levels_map(rng)
    n = 2000
    dtype = np.uint32 if bits == 32 else np.uint64
    mixes = _random_mixes(rng, n, 10)
    # light values on the half-way points of the rounding, and one big enough to overflow the pack
    mixes[0]["light"] = (0.5 / 256, 1.5 / 256, 2.5 / 256)
    mixes[3]["light"] = (200.0, 0.0, 1.0)
    dst = rng.integers(0, 2 ** 63, size=n, dtype=np.uint64).astype(dtype)
    got = dst.copy()
    oracle.mix_portal_textures(md, bits, mixes, got)
    want = dst.copy()
    for mx in mixes:
        want = _model_mix(want, lite, levels, mx, bits)
    assert np.array_equal(got, want)


def test_mix_model_detects_a_transposed_wrap():
    """The model's addressing is not symmetric in u and v: swapping them changes the result."""
    rng = np.random.default_rng(3)
    md, lite, levels = _levels_map(rng)
    mix = np.zeros(1, A.PORTAL_MIX_DTYPE)
    mix[0] = (0, (7, 3), (-3, 2), 2, A.BLEND_NONE, (1.0, 1.0, 1.0), 0)
    got = np.zeros(21, np.uint32)
    oracle.mix_portal_textures(md, 32, mix, got)
    swapped = mix.copy()
    swapped[0]["tc_start"] = (2, -3)
    assert not np.array_equal(got, _model_mix(np.zeros(21, np.uint32), lite, levels, swapped[0], 32))


@pytest.fixture(scope="module")
def scene():
    m = sw_host.large_map(4, grid=(10, 10, 4), num_view_portals=12)
    fb = sw_host.FrameBuilder()
    f = fb.build(m, m.camera_path(8, 48)[13], width=480, height=300, color_bits=32, use_directional_lightmaps=1,
                 num_dynamic_lights=4, portals_depth=2)
    return m, f


def test_generator_gives_every_second_portal_a_texture(scene):
    m, _ = scene
    vp = m.view_portals()
    for k, p in enumerate(vp):
        if k % 2 == 0:
            assert p["texture_level"] == A.NO_TEXTURE
        else:
            assert p["texture_level"] != A.NO_TEXTURE
            assert p["texture_blending_mode"] == (A.BLEND_AVERAGE if k % 4 == 1 else A.BLEND_ALPHA_BLEND)
            assert p["texture_light_scale"] == 1.0 and np.all(p["texture_light_add"] == 0.0)


def test_overlays_sit_on_the_portal_rects(scene):
    m, f = scene
    vp = m.view_portals()
    levels = sw_host.portal_levels(f)
    mixes = sw_host.portal_level_mixes(f)
    assert len(mixes) == len(levels) + 1 and len(mixes[0]) == 0
    assert sum(len(x) for x in mixes) > 0
    # level 1's overlays: the main frame's textured, visible portals at their own rects
    main = [(i, p) for i, p in enumerate(f["portals"]) if p["visible"] and vp[i]["texture_level"] != A.NO_TEXTURE]
    assert len(mixes[1]) == len(main)
    for mx, (i, p) in zip(mixes[1], main):
        assert mx["dst_offset"] == p["texture_offset"]
        assert tuple(mx["size"]) == tuple(p["resolution"]) and tuple(mx["tc_start"]) == tuple(p["tc_min"])
        assert mx["level"] == vp[i]["texture_level"] + p["mip"]
        assert mx["blend"] == vp[i]["texture_blending_mode"]
        assert np.all(np.isfinite(mx["light"])) and np.all(mx["light"] >= 0.0)
    # every level's rects lie inside the textures that level samples and do not overlap
    for k in range(1, len(mixes)):
        total = sum(int(g["portal_texels"]) for g in levels[k - 1])
        spans = sorted((int(x["dst_offset"]), int(x["dst_offset"]) + int(x["size"][0]) * int(x["size"][1])) for x in mixes[k])
        for (a0, a1), (b0, _) in zip(spans, spans[1:]):
            assert a1 <= b0
        assert all(b <= total for _, b in spans)


def test_overlaid_levels_equal_mixing_each_level(scene):
    """render_portal_levels with overlays equals, level by level, the plain levels' textures with the
    overlay applied before the level above samples them; the deepest textures are the overlay
    over black."""
    m, f = scene
    batches = sw_host.portal_level_batches(f)
    mixes = sw_host.portal_level_mixes(f)
    outs, oob = oracle.render_portal_levels(m.data, batches, mixes=mixes)
    assert oob == 0
    below = np.zeros(max(batches[-1].portal_texels, 1), np.uint32)
    oracle.mix_portal_textures(m.data, 32, mixes[len(batches)], below)
    for k in reversed(range(len(batches))):
        fb, _, _, _ = oracle.render(m.data, batches[k], portal=below)
        assert np.array_equal(fb, outs[k])
        if k:
            below = fb.copy()
            oracle.mix_portal_textures(m.data, 32, mixes[k], below)
    plain, _ = oracle.render_portal_levels(m.data, batches)
    assert not np.array_equal(plain[0], outs[0])  # the overlays show in the main view
