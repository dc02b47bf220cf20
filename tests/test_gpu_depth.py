"""GPU parity of sw_render_depth (DepthRasterizer, rasterizer.rs:1295-1474): depth buffers of many
targets (cube sides and projector maps of several sizes, spans longer than one wave) from random
convex polygons with rounding-heavy depth equations, overwriting and depth-tested, bit-exact to the
oracle; then the maps shadow a frame's lights."""
import numpy as np
import pytest

import oracle
import scenes
import sw_abi as A

pytestmark = pytest.mark.gpu


def _depth_scene(rng, sizes, draws_per_target):
    targets = np.zeros(len(sizes), dtype=A.DEPTH_TARGET_DTYPE)
    draws, vdata = [], []
    off = 0
    for t, (w, h) in enumerate(sizes):
        targets[t]["offset"], targets[t]["width"], targets[t]["height"] = off, w, h
        targets[t]["first_draw"], targets[t]["num_draws"] = len(draws), draws_per_target
        targets[t]["flags"] = A.DEPTH_CIRCLE if (w == h and t % 2 == 1) else 0  # projector maps' circle
        off += w * h
        for i in range(draws_per_target):
            fx, de, _ = scenes.make_polygon(rng, w, h, 16, 16, A.TC_PERSPECTIVE, thin=(i % 7 == 3))
            d = np.zeros(1, dtype=A.DEPTH_DRAW_DTYPE)[0]
            d["first_vertex_int"], d["num_vertices"] = sum(len(v) for v in vdata), len(fx) // 2
            d["depth_equation"] = de
            d["depth_test"] = 1 if i % 3 == 2 else 0
            draws.append(d)
            vdata.append(fx)
    return targets, np.array(draws, dtype=A.DEPTH_DRAW_DTYPE), np.concatenate(vdata).astype(np.int32), off


@pytest.mark.parametrize("seed", [1, 2])
def test_depth_buffers_bit_exact(gpu_ctx, seed):
    rng = np.random.default_rng(seed)
    sizes = [(64, 64)] * 6 + [(256, 256), (200, 130), (17, 300), (1, 1)]
    targets, draws, vdata, n = _depth_scene(rng, sizes, 24)
    init = rng.uniform(0.0, 0.01, size=n).astype(np.float32)
    print(f"depth seed {seed}: {len(draws)} draws, {n} texels", flush=True)
    out = gpu_ctx.render_depth(targets, draws, vdata, init)
    ref = oracle.render_depth(targets, draws, vdata, init)
    bad = np.flatnonzero(out.view(np.uint32) != ref.view(np.uint32))
    assert bad.size == 0, f"{bad.size} texels differ, first {bad[:5].tolist()}"
    assert np.count_nonzero(out != init) > n // 4  # the polygons cover much of the buffers


def test_depth_validation(gpu_ctx):
    rng = np.random.default_rng(3)
    targets, draws, vdata, n = _depth_scene(rng, [(32, 32)], 4)
    with pytest.raises(A.SwError, match="outside the depth texels"):
        gpu_ctx.render_depth(targets, draws, vdata, np.zeros(n - 1, np.float32))
    bad = draws.copy()
    bad[1]["num_vertices"] = 30
    with pytest.raises(A.SwError, match="vertices out of range"):
        gpu_ctx.render_depth(targets, bad, vdata, np.zeros(n, np.float32))


def test_depth_maps_shadow_a_frame(gpu_ctx):
    """Cube maps drawn by sw_render_depth into the shadow texels, sampled by the surface builder: the
    frame equals the oracle's with the oracle's own depth maps."""
    rng = np.random.default_rng(7)
    m, b = scenes.make_batch(31, color_bits=32, view_sizes=((160, 120),), num_lights=4, surfaces_per_view=40)
    size = 32
    targets, draws, vdata, n = _depth_scene(rng, [(size, size)] * (6 * len(b.lights)), 6)
    # the polygons' 1/z spans the scene's distances: scale the equations so the maps are decisive
    draws["depth_equation"] *= np.float32(4.0)
    init = np.zeros(n, np.float32)
    depth = gpu_ctx.render_depth(targets, draws, vdata, init)
    ref_depth = oracle.render_depth(targets, draws, vdata, init)
    assert np.array_equal(depth.view(np.uint32), ref_depth.view(np.uint32))
    sh = np.zeros(len(b.lights), dtype=A.LIGHT_SHADOW_DTYPE)
    sh["kind"], sh["size"] = A.SHADOW_CUBE, size
    sh["offset"] = np.arange(len(b.lights)) * 6 * size * size
    try:
        gpu_ctx.set_light_shadows(sh, depth)
        oracle.set_light_shadows(sh, ref_depth)
        mh = gpu_ctx.upload_map(m)
        try:
            fb = gpu_ctx.render_plain(mh, b)
        finally:
            gpu_ctx.release_map(mh)
        ref_fb, _, _, _ = oracle.render(m, b)
        assert np.array_equal(fb.reshape(-1), ref_fb)
    finally:
        gpu_ctx.set_light_shadows(None, None)
        oracle.set_light_shadows(None, None)


@pytest.mark.parametrize("color_bits", [32, 64])
def test_front_end_cube_shadows_frame(gpu_ctx, color_bits):
    """The whole shadow path of a frame: the front-end's cube draw lists for the frame's dynamic lights
    (swh_shadow_add_cube), drawn by sw_render_depth, sampled by the surface builder; depth maps and
    the frame bit-exact to the oracle's, and the shadows darken the frame."""
    import sw_host
    m = sw_host.large_map(5, grid=(12, 12, 4))
    pose = m.camera_path(9, 1)[0]
    f = sw_host.FrameBuilder().build(m, pose, width=240, height=160, color_bits=color_bits, num_dynamic_lights=8,
                                     use_directional_lightmaps=1, seed=4)
    b = sw_host.batch_from_frames([f])
    assert len(b.lights) > 0
    size = 64
    targets, draws, vdata, n, sh = sw_host.cube_shadows_for_lights(m, b.lights, size)
    print(f"front-end shadows {color_bits}: {len(b.lights)} lights, {len(draws)} draws", flush=True)
    init = np.zeros(n, np.float32)
    depth = gpu_ctx.render_depth(targets, draws, vdata, init)
    ref_depth = oracle.render_depth(targets, draws, vdata, init)
    assert np.array_equal(depth.view(np.uint32), ref_depth.view(np.uint32))
    mh = gpu_ctx.upload_map(m.data)
    try:
        unshadowed = gpu_ctx.render_plain(mh, b).reshape(-1)
        gpu_ctx.set_light_shadows(sh, depth)
        oracle.set_light_shadows(sh, ref_depth)
        fb = gpu_ctx.render_plain(mh, b).reshape(-1)
        ref_fb, _, _, _ = oracle.render(m.data, b)
        assert np.array_equal(fb, ref_fb)
        assert not np.array_equal(fb, unshadowed)
    finally:
        gpu_ctx.set_light_shadows(None, None)
        oracle.set_light_shadows(None, None)
        gpu_ctx.release_map(mh)


def test_front_end_projector_shadows_frame(gpu_ctx):
    """Projector maps (swh_shadow_add_projector: a rotated camera, the circle mask) for every dynamic
    light of a frame, drawn by sw_render_depth and sampled through the returned bases: depth maps and
    the shadowed frame bit-exact to the oracle's."""
    import sw_host
    m = sw_host.large_map(5, grid=(12, 12, 4))
    pose = m.camera_path(11, 1)[0]
    f = sw_host.FrameBuilder().build(m, pose, width=200, height=150, color_bits=32, num_dynamic_lights=8,
                                     use_directional_lightmaps=1, seed=6)
    b = sw_host.batch_from_frames([f])
    rng = np.random.default_rng(12)
    size = 64
    sb = sw_host.ShadowBuilder()
    sh = np.zeros(len(b.lights), dtype=A.LIGHT_SHADOW_DTYPE)
    for i, lt in enumerate(b.lights):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        basis = sb.add_projector(m, lt["position"], q, 1.3, size)
        sh[i]["kind"], sh[i]["size"], sh[i]["offset"] = A.SHADOW_PROJECTOR, size, i * size * size
        sh[i]["basis_x"], sh[i]["basis_y"], sh[i]["basis_z"] = basis[0], basis[1], basis[2]
    targets, draws, vdata, n = sb.arrays()
    sb.close()
    print(f"front-end projectors: {len(b.lights)} lights, {len(draws)} draws", flush=True)
    init = np.zeros(n, np.float32)
    depth = gpu_ctx.render_depth(targets, draws, vdata, init)
    ref_depth = oracle.render_depth(targets, draws, vdata, init)
    assert np.array_equal(depth.view(np.uint32), ref_depth.view(np.uint32))
    assert np.count_nonzero(depth == np.float32(1e24)) > 0
    mh = gpu_ctx.upload_map(m.data)
    try:
        gpu_ctx.set_light_shadows(sh, depth)
        oracle.set_light_shadows(sh, ref_depth)
        fb = gpu_ctx.render_plain(mh, b).reshape(-1)
        ref_fb, _, _, _ = oracle.render(m.data, b)
        assert np.array_equal(fb, ref_fb)
    finally:
        gpu_ctx.set_light_shadows(None, None)
        oracle.set_light_shadows(None, None)
        gpu_ctx.release_map(mh)
