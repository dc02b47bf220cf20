"""Parity of the HIP path with the oracle on frames produced by the renderer front-end from seeded
maps (the benchmark configurations at reduced sizes), through the plain drop-in path and the
batched path."""
import numpy as np
import pytest

import oracle
import sw_abi
import sw_host

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def map0():
    return sw_host.small_map(3)


@pytest.fixture(scope="module")
def map1():
    return sw_host.large_map(5, grid=(12, 12, 4))


def test_small_map_frames_bit_exact(gpu_ctx, map0):
    fb = sw_host.FrameBuilder()
    mh = gpu_ctx.upload_map(map0.data)
    try:
        for i, pose in enumerate(map0.camera_path(2, 6)):
            _say(f"case small map pose {i}")
            f = fb.build(map0, pose, width=320, height=200)
            b = sw_host.batch_from_frames([f])
            img = gpu_ctx.render_plain(mh, b).reshape(-1)
            ref, _, st, oob = oracle.render(map0.data, b)
            assert oob == 0
            assert np.array_equal(img, ref), f"pose {i}: {np.count_nonzero(img != ref)} pixels differ"
    finally:
        gpu_ctx.release_map(mh)


@pytest.mark.parametrize("color_bits", [32, 64])
def test_large_map_frames_bit_exact(gpu_ctx, map1, color_bits):
    fb = sw_host.FrameBuilder()
    mh = gpu_ctx.upload_map(map1.data)
    try:
        frames = []
        for i, pose in enumerate(map1.camera_path(3, 4)):
            _say(f"case large map pose {i} bits {color_bits}")
            f = fb.build(map1, pose, width=640, height=360, color_bits=color_bits, use_directional_lightmaps=1,
                         num_dynamic_lights=6, seed=i)
            frames.append(f)
            b = sw_host.batch_from_frames([f])
            img = gpu_ctx.render_plain(mh, b).reshape(-1)
            ref, _, st, oob = oracle.render(map1.data, b)
            assert oob == 0
            assert np.array_equal(img, ref), f"pose {i}: {np.count_nonzero(img != ref)} pixels differ"
        # the same views batched into one launch give the same pixels
        _say("case large map batch")
        b = sw_host.batch_from_frames(frames)
        out = gpu_ctx.render_batch(mh, b)
        ref, _, _, _ = oracle.render(map1.data, b)
        assert np.array_equal(out, ref)
    finally:
        gpu_ctx.release_map(mh)


@pytest.fixture(scope="module")
def map2():
    return sw_host.large_map_alpha(7, grid=(12, 12, 4))


@pytest.mark.parametrize("color_bits", [64, 32])
def test_models_decals_hdr_frames(gpu_ctx, map2, color_bits):
    """The HDR + models scene at reduced size: 16 dynamic lights, animated triangle models,
    decals, alpha-blended materials; 64-bit frames are tonemapped. Framebuffers are bit-exact
    against the oracle, the tonemapped image too."""
    fb = sw_host.FrameBuilder()
    mh = gpu_ctx.upload_map(map2.data)
    tri_pixels = 0
    try:
        for i, pose in enumerate(map2.camera_path(4, 6)):
            _say(f"case models pose {i} bits {color_bits}")
            f = fb.build(map2, pose, width=640, height=360, color_bits=color_bits, use_directional_lightmaps=1,
                         num_dynamic_lights=16, num_models=24, num_decals=96, seed=11, time=0.37 * i)
            b = sw_host.batch_from_frames([f])
            ref, _, st, oob = oracle.render(map2.data, b)
            assert oob == 0
            tri_pixels += st["triangle_pixels"]
            if color_bits == 64:
                img, out, sums = gpu_ctx.render_plain(mh, b, tonemap=(1.0, (1.0, 1.0, 1.0)))
                tref, tsums = oracle.tonemap(ref.reshape(img.shape), 1.0, (1.0, 1.0, 1.0))
                assert np.array_equal(out, tref)
            else:
                img = gpu_ctx.render_plain(mh, b)
            img = img.reshape(-1)
            assert np.array_equal(img, ref), f"pose {i}: {np.count_nonzero(img != ref)} pixels differ"
    finally:
        gpu_ctx.release_map(mh)
    assert tri_pixels > 0, "the path drew no model or decal pixels"


def test_frames_of_a_loaded_sqwm_map(gpu_ctx, map1, tmp_path):
    """A map saved as .sqwm and loaded back (lightmaps and light grid through the format's
    compression) renders bit-exact against the oracle on the loaded data."""
    path = str(tmp_path / "map1.sqwm")
    map1.save_sqwm(path)
    loaded = sw_host.Map.load_sqwm(path)
    fb = sw_host.FrameBuilder()
    mh = gpu_ctx.upload_map(loaded.data)
    try:
        for i, pose in enumerate(loaded.camera_path(4, 3)):
            _say(f"case loaded sqwm map pose {i}")
            f = fb.build(loaded, pose, width=480, height=270, use_directional_lightmaps=1, num_models=8, seed=i)
            b = sw_host.batch_from_frames([f])
            img = gpu_ctx.render_plain(mh, b).reshape(-1)
            ref, _, _, oob = oracle.render(loaded.data, b)
            assert oob == 0
            assert np.array_equal(img, ref), f"pose {i}: {np.count_nonzero(img != ref)} pixels differ"
    finally:
        gpu_ctx.release_map(mh)


@pytest.mark.parametrize("threads", [4, 16])
def test_virtual_thread_frames_bit_exact(gpu_ctx, map1, threads):
    """The reference's N-thread image (split_rect's N rects, each pass with its own clip rect and
    half-pixel-widened viewport): the front-end's N runs of draws render bit-exact against the oracle
    through the plain and the batched path."""
    fb = sw_host.FrameBuilder()
    mh = gpu_ctx.upload_map(map1.data)
    try:
        frames = []
        for i, pose in enumerate(map1.camera_path(9, 3)):
            _say(f"case virtual threads {threads} pose {i}")
            f = fb.build(map1, pose, width=640, height=360, use_directional_lightmaps=1, num_dynamic_lights=4,
                         num_models=6, num_decals=24, seed=i, virtual_threads=threads)
            frames.append(f)
            b = sw_host.batch_from_frames([f])
            img = gpu_ctx.render_plain(mh, b).reshape(-1)
            ref, _, _, oob = oracle.render(map1.data, b)
            assert oob == 0
            assert np.array_equal(img, ref), f"pose {i}: {np.count_nonzero(img != ref)} pixels differ"
        b = sw_host.batch_from_frames(frames)
        out = gpu_ctx.render_batch(mh, b)
        ref, _, _, _ = oracle.render(map1.data, b)
        assert np.array_equal(out, ref)
    finally:
        gpu_ctx.release_map(mh)


@pytest.mark.parametrize("color_bits", [32, 64])
def test_inline_models_and_view_model_frames(gpu_ctx, color_bits):
    """Frames with moving inline (brush) models in their leaves, triangle models sorted with them,
    decals, dynamic lights and a view model: bit-exact against the oracle, plain and batched."""
    m = sw_host.large_map(8, grid=(12, 12, 4), num_submodels=14)
    fb = sw_host.FrameBuilder()
    mh = gpu_ctx.upload_map(m.data)
    try:
        frames = []
        for i, pose in enumerate(m.camera_path(5, 4)):
            _say(f"case inline models pose {i} bits {color_bits}")
            f = fb.build(m, pose, width=640, height=360, color_bits=color_bits, use_directional_lightmaps=1,
                         num_dynamic_lights=6, num_models=16, num_decals=40, seed=i, time=0.6 * i, view_model=1)
            assert f["first_submodel_surface"] <= len(f["surfaces"])
            frames.append(f)
            b = sw_host.batch_from_frames([f])
            img = gpu_ctx.render_plain(mh, b).reshape(-1)
            ref, _, _, oob = oracle.render(m.data, b)
            assert oob == 0
            assert np.array_equal(img, ref), f"pose {i}: {np.count_nonzero(img != ref)} pixels differ"
        assert sum(len(f["surfaces"]) - f["first_submodel_surface"] for f in frames) > 0
        b = sw_host.batch_from_frames(frames)
        out = gpu_ctx.render_batch(mh, b)
        ref, _, _, _ = oracle.render(m.data, b)
        assert np.array_equal(out, ref)
    finally:
        gpu_ctx.release_map(mh)


@pytest.mark.parametrize("color_bits", [32, 64])
def test_portal_frames_bit_exact(gpu_ctx, color_bits):
    """Mirrors, camera and parallax portals two levels deep: every level's views on the GPU (each
    level one batch sampling the level below's framebuffer in HBM, with the portal textures'
    overlays mixed in by sw_mix_portal_textures) against the oracle's levels; the main view again
    through the plain path with the portal textures set by sw_set_portal_textures."""
    m = sw_host.large_map(4, grid=(10, 10, 4), num_view_portals=12)
    fb = sw_host.FrameBuilder()
    mh = gpu_ctx.upload_map(m.data)
    try:
        done = overlays = 0
        poses = m.camera_path(8, 48)
        for i in (13, 7, 17, 23):  # poses that see portal views two levels deep
            f = fb.build(m, poses[i], width=480, height=300, color_bits=color_bits, use_directional_lightmaps=1,
                         num_dynamic_lights=4, num_models=8, time=0.3 * i, portals_depth=2)
            levels = sw_host.portal_level_batches(f)
            assert len(levels) == 3
            mixes = sw_host.portal_level_mixes(f)
            overlays += sum(len(x) for x in mixes)
            _say(f"case portals pose {i}: {[len(b.views) for b in levels]} views per level, "
                 f"{[len(x) for x in mixes]} overlays per level, color {color_bits}")
            outs = gpu_ctx.render_portal_levels(mh, levels, mixes=mixes)
            refs, oob = oracle.render_portal_levels(m.data, levels, mixes=mixes)
            assert oob == 0
            for k, (o, r) in enumerate(zip(outs, refs)):
                assert np.array_equal(o, r), f"pose {i} level {k}: {np.count_nonzero(o != r)} pixels differ"
            # the main view through the plain path, sampling level 1 with its overlays
            tex = refs[1].copy()
            oracle.mix_portal_textures(m.data, color_bits, mixes[1], tex)
            plain = gpu_ctx.render_plain(mh, levels[0], portal=tex).reshape(-1)
            assert np.array_equal(plain, refs[0])
            done += 1
        assert done == 4 and overlays > 0
    finally:
        gpu_ctx.release_map(mh)


@pytest.mark.parametrize("color_bits", [32, 64])
def test_portal_texture_mix_bit_exact(gpu_ctx, color_bits):
    """sw_mix_portal_textures on random rects of a texture buffer (every blend mode, negative and
    wrapping texture coordinates, a non-power-of-two level, light values that overflow the pack)
    against the oracle; overlapping rects and a non-lite level are errors that leave the buffer."""
    m = sw_host.large_map(2, grid=(6, 6, 4))
    rng = np.random.default_rng(31 + color_bits)
    lite_levels = [i for i, lv in enumerate(m.data.levels) if lv["flags"] & sw_abi.TEX_LITE_LEVEL]
    assert lite_levels
    n = 1 << 16
    dtype = np.uint32 if color_bits == 32 else np.uint64
    mixes = np.zeros(24, sw_abi.PORTAL_MIX_DTYPE)
    base = 0
    for i in range(len(mixes)):
        w, h = int(rng.integers(1, 90)), int(rng.integers(1, 40))
        mixes[i] = (base, (w, h), (int(rng.integers(-300, 300)), int(rng.integers(-300, 300))),
                    lite_levels[int(rng.integers(0, len(lite_levels)))], i % 5, rng.uniform(0.0, 2.5, 3), 0)
        base += w * h + int(rng.integers(0, 9))
    mixes[5]["light"] = (150.0, 0.75, 0.0)
    assert base <= n
    _say(f"case portal mix: {len(mixes)} rects, {base} texels, color {color_bits}")
    texels = rng.integers(0, 2 ** 63, size=n, dtype=np.uint64).astype(dtype)
    mh = gpu_ctx.upload_map(m.data)
    buf = gpu_ctx.to_device(texels)
    try:
        gpu_ctx.mix_portal_textures(mh, color_bits, mixes, buf, n)
        got = buf.read(dtype, n)
        want = texels.copy()
        oracle.mix_portal_textures(m.data, color_bits, mixes, want)
        assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} texels differ"
        bad = mixes[:2].copy()
        bad[1]["dst_offset"] = bad[0]["dst_offset"] + 1
        with pytest.raises(RuntimeError):
            gpu_ctx.mix_portal_textures(mh, color_bits, bad, buf, n)
        bad = mixes[:1].copy()
        bad[0]["level"] = [i for i in range(len(m.data.levels)) if i not in lite_levels][0]
        with pytest.raises(RuntimeError):
            gpu_ctx.mix_portal_textures(mh, color_bits, bad, buf, n)
        assert np.array_equal(buf.read(dtype, n), want)
    finally:
        buf.free()
        gpu_ctx.release_map(mh)


@pytest.mark.parametrize("color_bits", [32, 64])
def test_device_model_frames_bit_exact(gpu_ctx, color_bits):
    """Frames whose models are prepared on the device: sw_prepare_models, sw_expand_mesh_runs (the
    meshes clipped by near, leaf and 2D box planes, fans in place) and one sw_render_batch, against
    the oracle's preparation, expansion and frame; the fan sizes equal the oracle's."""
    m = sw_host.large_map(0, grid=(12, 12, 4))
    poses = m.camera_path(3, 40)
    mesh = m.model_mesh()
    fbld = sw_host.FrameBuilder()
    mh = gpu_ctx.upload_map(m.data)
    dm = sw_host.DeviceModels(gpu_ctx, m)
    try:
        clipped = 0
        for i in (7, 21, 31, 0):
            _say(f"case device models pose {i} bits {color_bits}")
            f = fbld.build(m, poses[i], width=640, height=360, color_bits=color_bits, use_directional_lightmaps=1,
                           num_models=24, num_decals=10, seed=i, time=0.3 * i, view_model=1, device_models=1)
            md = f["models"]
            clipped += int(np.sum(md["runs"]["num_planes3d"] > 0))
            img, fans = dm.render(mh, f)
            verts, tris, counts, ok = oracle.prepare_models([mesh], md["instances"])
            assert ok
            ints, ofans = oracle.expand_mesh_runs(md["runs"], verts, tris, counts)
            assert np.array_equal(fans, ofans), f"pose {i}: fan sizes differ"
            draws = sw_host.assemble_model_draws(f, ofans, len(f["vertex_data"]))
            f2 = dict(f, draws=draws, vertex_data=np.concatenate([f["vertex_data"], ints]).astype(np.int32))
            ref, _, st, oob = oracle.render(m.data, sw_host.batch_from_frames([f2]))
            assert oob == 0 and st["triangle_pixels"] > 0
            assert np.array_equal(img, ref), f"pose {i}: {np.count_nonzero(img != ref)} pixels differ"
        assert clipped >= 3
        # a run whose count index is out of range is an error
        bad = f["models"]["runs"][:1].copy()
        bad["count_index"] = 1000
        vb, tb, cb = gpu_ctx.alloc(1 << 16), gpu_ctx.alloc(1 << 16), gpu_ctx.alloc(64)
        ib = gpu_ctx.alloc(1 << 20)
        try:
            with pytest.raises(RuntimeError):
                gpu_ctx.expand_mesh_runs(bad, vb, 1, tb, 1, cb, 1, ib.ptr.value, (1 << 20) // 4)
        finally:
            for b in (vb, tb, cb, ib):
                b.free()
    finally:
        dm.close()
        gpu_ctx.release_map(mh)
