"""The CPU front-end (libsw_host.so) that turns seeded maps and camera poses into command buffers:
structural checks of what it emits (no GPU). Its output is the input of both the oracle and the
HIP path, so GPU parity does not depend on it; these tests pin its own contracts."""
import numpy as np
import pytest

import sw_abi as A
import sw_host


@pytest.fixture(scope="module")
def scene():
    return sw_host.large_map_alpha(13, grid=(10, 10, 4))


def _frames(m, **kw):
    fb = sw_host.FrameBuilder()
    return [fb.build(m, p, **kw) for p in m.camera_path(6, 6)]


def test_frames_are_deterministic(scene):
    kw = dict(width=320, height=200, color_bits=64, use_directional_lightmaps=1, num_dynamic_lights=8, num_models=12,
              num_decals=40, seed=3, time=0.7)
    a = _frames(scene, **kw)
    b = _frames(scene, **kw)
    for fa, fb in zip(a, b):
        for k in ("surfaces", "draws", "vertex_data", "lights", "light_indices"):
            assert np.array_equal(fa[k], fb[k]), k


def test_triangle_draws_stay_on_screen_and_in_texture(scene):
    W, H = 320, 200
    frames = _frames(scene, width=W, height=H, color_bits=32, use_directional_lightmaps=1, num_dynamic_lights=8,
                     num_models=12, num_decals=60, seed=5, time=1.3)
    levels = scene.data.levels
    ntri = 0
    for f in frames:
        d = f["draws"]
        v = f["vertex_data"]
        for t in d[d["kind"] == A.DRAW_TRIANGLE]:
            ntri += 1
            o = int(t["first_vertex_int"])
            pts = v[o:o + 21].reshape(3, 7).astype(np.float64) / 65536.0
            # clipped to the leaf bounds, which lie inside the screen
            assert np.all(pts[:, 0] >= -1e-3) and np.all(pts[:, 0] <= W + 1e-3)
            assert np.all(pts[:, 1] >= -1e-3) and np.all(pts[:, 1] <= H + 1e-3)
            # the texture is a lite level of the map and the range lies inside the lite texels
            lv = levels[(levels["offset"] == t["texture_offset"]) & (levels["flags"] & A.TEX_LITE_LEVEL != 0)]
            assert len(lv) >= 1
            assert int(t["texture_offset"]) + int(t["texture_size"][0]) * int(t["texture_size"][1]) <= len(scene.data.lite)
            assert t["blend"] <= A.BLEND_ALPHA_BLEND
    assert ntri > 0


def test_decal_texture_coordinates_are_clamped(scene):
    """Decal vertices clamp their texture coordinates to [1/32, size - 1/32]
    (partial_renderer.rs:2987-2990), so the rasteriser never wraps a decal."""
    frames = _frames(scene, width=320, height=200, color_bits=32, use_directional_lightmaps=1, num_models=0,
                     num_decals=80, seed=9, time=0.0)
    n = 0
    for f in frames:
        d = f["draws"]
        v = f["vertex_data"]
        for t in d[d["kind"] == A.DRAW_TRIANGLE]:
            assert t["blend"] == A.BLEND_ALPHA_BLEND  # every decal in these scenes alpha-blends
            o = int(t["first_vertex_int"])
            pts = v[o:o + 21].reshape(3, 7)
            for i in range(2):
                lo = int(np.int32(np.float32(1.0 / 32.0) * np.float32(65536.0)))
                hi = int(np.int32(np.float32(float(t["texture_size"][i]) - 1.0 / 32.0) * np.float32(65536.0)))
                assert np.all(pts[:, 2 + i] >= lo) and np.all(pts[:, 2 + i] <= hi)
            n += 1
    assert n > 0


def test_models_are_animated_by_time(scene):
    kw = dict(width=320, height=200, color_bits=32, use_directional_lightmaps=1, num_models=12, num_decals=0, seed=3)
    fb = sw_host.FrameBuilder()
    pose = scene.camera_path(6, 6)[0]
    a = fb.build(scene, pose, time=0.0, **kw)
    b = fb.build(scene, pose, time=0.05, **kw)
    ta = a["draws"]["kind"] == A.DRAW_TRIANGLE
    tb = b["draws"]["kind"] == A.DRAW_TRIANGLE
    assert ta.sum() > 0 and tb.sum() > 0
    # the polygons of the map do not move, the models' vertices do
    assert np.array_equal(a["surfaces"], b["surfaces"])
    assert not np.array_equal(a["vertex_data"], b["vertex_data"])


def test_dynamic_lights_only_reach_nearby_surfaces(scene):
    f = _frames(scene, width=320, height=200, color_bits=32, use_directional_lightmaps=1, num_dynamic_lights=16,
                seed=2, time=0.0)[0]
    s = f["surfaces"]
    assert np.all(s["num_lights"] <= 6)
    li = f["light_indices"]
    for d in s:
        for k in range(int(d["num_lights"])):
            assert li[int(d["first_light"]) + k] < len(f["lights"])
