"""Parity of the HIP path (libsw_gpu.so through the C ABI) with the CPU oracle.

Bar: bit-exact against the oracle's exact numeric mode (1/sqrt and 1/x correctly rounded, the
reference's own non-SSE definitions, fast_math.rs:376-384), for surfaces, framebuffers and work
counters; within +-1 LSB per channel against the oracle's SSE mode (rsqrtss / rcpss, the x86
reference build, fast_math.rs:13-22).
"""
import numpy as np
import pytest

import oracle
import scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _channel_maxdiff(a, b):
    """Largest per-channel difference (8-bit channels of Color32, 16-bit channels of Color64)."""
    ch = np.uint16 if a.dtype == np.uint64 else np.uint8
    return int(np.abs(a.view(ch).astype(np.int64) - b.view(ch).astype(np.int64)).max(initial=0))


@pytest.mark.parametrize("color_bits", [32, 64])
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_batch_bit_exact(gpu_ctx, color_bits, seed):
    _say(f"case batch seed={seed} bits={color_bits}")
    m, b = scenes.make_batch(seed, color_bits=color_bits, view_sizes=((160, 120), (97, 61), (64, 200)))
    mh = gpu_ctx.upload_map(m)
    try:
        fb = gpu_ctx.render_batch(mh, b)
        arena = gpu_ctx.read_surfaces(b.arena_texels, color_bits)
        st = gpu_ctx.stats()
    finally:
        gpu_ctx.release_map(mh)
    ofb, oarena, ost, oob = oracle.render(m, b, mode=oracle.MODE_EXACT)
    assert np.array_equal(arena, oarena[:b.arena_texels]), "surface arena differs from the oracle"
    assert np.array_equal(fb, ofb), f"framebuffer differs in {np.count_nonzero(fb != ofb)} pixels"
    for k in ("polygon_pixels", "triangle_pixels", "blend_pixels", "surface_texels", "emissive_texels"):
        assert st[k] == ost[k], (k, st[k], ost[k])
    # +-1 LSB against the x86 (rsqrtss / rcpss) numeric mode
    _, sarena, _, _ = oracle.render(m, b, mode=oracle.MODE_SSE)
    assert _channel_maxdiff(arena, sarena[:b.arena_texels]) <= 1


@pytest.mark.parametrize("color_bits", [32, 64])
def test_plain_path_matches_oracle(gpu_ctx, color_bits):
    _say(f"case plain bits={color_bits}")
    m, b = scenes.make_batch(11, color_bits=color_bits, view_sizes=((200, 150),), draws_per_view=60)
    mh = gpu_ctx.upload_map(m)
    try:
        fb = gpu_ctx.render_plain(mh, b)
    finally:
        gpu_ctx.release_map(mh)
    ofb, _, _, _ = oracle.render(m, b, mode=oracle.MODE_EXACT)
    assert np.array_equal(fb.reshape(-1), ofb)


def test_no_clear_draws_over_previous_frame(gpu_ctx):
    """Without clear_background the draws land on what the framebuffer holds (the reference draws
    into the caller's buffer and clears only on request, partial_renderer.rs:249-257)."""
    import sw_abi
    _say("case no-clear")
    m, b1 = scenes.make_batch(5, color_bits=32, view_sizes=((128, 96),))
    views2 = b1.views.copy()
    views2["clear_background"] = 0
    draws2 = b1.draws[::-1].copy()
    b2 = sw_abi.BatchData(views2, b1.surfaces, b1.lights, b1.light_indices, draws2, b1.vertex_data, 32)
    mh = gpu_ctx.upload_map(m)
    try:
        gpu_ctx.render_plain(mh, b1)
        fb2 = gpu_ctx.render_plain(mh, b2)
    finally:
        gpu_ctx.release_map(mh)
    ofb1, _, _, _ = oracle.render(m, b1)
    import ctypes as C
    lib = oracle.load()
    arena = np.zeros(max(b2.arena_texels, 1), dtype=np.uint32)
    ofb = ofb1.copy()
    md, cb = m.desc(), b2.ctypes_batch()
    assert lib.oracle_render_batch(oracle.MODE_EXACT, C.byref(md), C.byref(cb), C.c_void_p(arena.ctypes.data),
                                   C.c_void_p(ofb.ctypes.data), 1, None, None) == 0
    assert np.array_equal(fb2.reshape(-1), ofb)


def test_graph_replay_matches_eager(gpu_ctx):
    """Replays of the captured batch equal the eager run, with eager work in between."""
    import sw_abi
    _say("case graph replay")
    m, b = scenes.make_batch(7, color_bits=32, view_sizes=((160, 120), (320, 200)))
    mh = gpu_ctx.upload_map(m)
    dev = sw_abi.DeviceCommandBuffers(gpu_ctx, b)
    out = gpu_ctx.alloc(b.fb_texels * 4)
    try:
        eager = gpu_ctx.render_batch(mh, b, device_inputs=dev, out=out)
        ref, _, _, _ = oracle.render(m, b)
        assert np.array_equal(eager, ref)
        for i in range(4):
            g = gpu_ctx.render_batch(mh, b, graph=True, device_inputs=dev, out=out)
            assert np.array_equal(g, eager), f"replay {i} differs"
            e2 = gpu_ctx.render_batch(mh, b, device_inputs=dev, out=out)
            assert np.array_equal(e2, eager)
        for i in range(3):
            g = gpu_ctx.render_batch(mh, b, graph=True, device_inputs=dev, out=out)
            assert np.array_equal(g, eager), f"back-to-back replay {i} differs"
    finally:
        out.free()
        dev.free()
        gpu_ctx.release_map(mh)


@pytest.mark.parametrize("exposure", [0.25, 1.0, 3.0])
@pytest.mark.parametrize("size", [(256, 160), (255, 161), (2560, 1440)])
def test_tonemap_bit_exact(gpu_ctx, exposure, size):
    """Tonemapped frames of an even, an odd and the HDR workload's 2560x1440 size equal the
    oracle's bit for bit (alpha 255), colour sums within float tolerance."""
    _say(f"case tonemap exposure={exposure} size={size}")
    m, b = scenes.make_batch(9, color_bits=64, view_sizes=(size,))
    mh = gpu_ctx.upload_map(m)
    try:
        fb, out, sums = gpu_ctx.render_plain(mh, b, tonemap=(exposure, (1.0, 0.9, 1.1)))
    finally:
        gpu_ctx.release_map(mh)
    ref, rsums = oracle.tonemap(fb, exposure, (1.0, 0.9, 1.1))
    assert np.array_equal(out, ref)
    assert np.all((out >> 24) == 255)
    np.testing.assert_allclose(sums, rsums, rtol=1e-5)


@pytest.mark.parametrize("color_bits", [32, 64])
@pytest.mark.parametrize("rows", [4, 8, 16])
@pytest.mark.parametrize("scalar_spans", [False, True])
def test_every_tile_height_bit_exact(gpu_ctx, rows, color_bits, scalar_spans):
    """The raster tile height and the way span records are loaded change the launch only, never a
    pixel."""
    _say(f"case tile rows={rows} bits={color_bits} scalar_spans={scalar_spans}")
    m, b = scenes.make_batch(30 + rows, color_bits=color_bits, view_sizes=((160, 120), (97, 61), (64, 200)))
    prev = gpu_ctx.tile_rows(), gpu_ctx.scalar_spans()
    mh = gpu_ctx.upload_map(m)
    try:
        gpu_ctx.set_tile_rows(rows)
        gpu_ctx.set_scalar_spans(scalar_spans)
        assert gpu_ctx.tile_rows() == rows and gpu_ctx.scalar_spans() == scalar_spans
        fb = gpu_ctx.render_batch(mh, b)
        st = gpu_ctx.stats()
    finally:
        gpu_ctx.set_tile_rows(prev[0])
        gpu_ctx.set_scalar_spans(prev[1])
        gpu_ctx.release_map(mh)
    ofb, _, ost, _ = oracle.render(m, b, mode=oracle.MODE_EXACT)
    assert np.array_equal(fb, ofb)
    for k in ("polygon_pixels", "triangle_pixels", "blend_pixels"):
        assert st[k] == ost[k], (k, st[k], ost[k])


def test_autotune_keeps_output(gpu_ctx):
    """The autotuner times every tile height, keeps one of them, and the output stays exact."""
    import sw_abi
    _say("case autotune")
    m, b = scenes.make_batch(41, color_bits=32, view_sizes=((320, 200), (160, 120)))
    prev = gpu_ctx.tile_rows()
    mh = gpu_ctx.upload_map(m)
    dev = sw_abi.DeviceCommandBuffers(gpu_ctx, b)
    out = gpu_ctx.alloc(b.fb_texels * 4)
    try:
        rows, ms = gpu_ctx.autotune(mh, b, dev, out)
        assert rows in sw_abi.TILE_ROWS and gpu_ctx.tile_rows() == rows
        assert all(v > 0 for v in ms.values()), ms
        assert ms[rows] == min(ms.values())
        fb = gpu_ctx.render_batch(mh, b, device_inputs=dev, out=out)
        g = gpu_ctx.render_batch(mh, b, graph=True, device_inputs=dev, out=out)
    finally:
        gpu_ctx.set_tile_rows(prev)
        out.free()
        dev.free()
        gpu_ctx.release_map(mh)
    ofb, _, _, _ = oracle.render(m, b)
    assert np.array_equal(fb, ofb)
    assert np.array_equal(g, ofb)


def test_autotune_over_several_batches_keeps_output(gpu_ctx):
    """sw_autotune_batches times every tile height over a sample of batches (frames of one map at
    different sizes, into one framebuffer), keeps the fastest, and every batch stays exact."""
    import sw_abi
    _say("case autotune over batches")
    # the same seed gives the same map; the view sizes differ
    pairs = [scenes.make_batch(43, color_bits=32, view_sizes=(sz,)) for sz in ((320, 200), (160, 120), (256, 256))]
    m = pairs[0][0]
    bs = [b for _, b in pairs]
    prev, prev_sspan = gpu_ctx.tile_rows(), gpu_ctx.scalar_spans()
    mh = gpu_ctx.upload_map(m)
    devs = [sw_abi.DeviceCommandBuffers(gpu_ctx, b) for b in bs]
    out = gpu_ctx.alloc(max(b.fb_texels for b in bs) * 4)
    try:
        rows, ms = gpu_ctx.autotune_many(mh, bs, devs, out)
        assert rows in sw_abi.TILE_ROWS and gpu_ctx.tile_rows() == rows
        assert all(v > 0 for v in ms.values()) and ms[rows] == min(ms.values())
        fbs = [gpu_ctx.render_batch(mh, b, device_inputs=d, out=out) for b, d in zip(bs, devs)]
    finally:
        gpu_ctx.set_tile_rows(prev)
        gpu_ctx.set_scalar_spans(prev_sspan)
        out.free()
        for d in devs:
            d.free()
        gpu_ctx.release_map(mh)
    for b, fb in zip(bs, fbs):
        ofb, _, _, _ = oracle.render(m, b)
        assert np.array_equal(fb, ofb)


@pytest.mark.parametrize("color_bits", [32, 64])
def test_surfaces_read_before_raster(gpu_ctx, color_bits):
    """sw_read_surfaces between sw_build_surfaces and sw_raster_draws builds the surfaces on the
    spot (they are otherwise built by the raster call's launch sequence); both the arena and the
    frame equal the oracle's."""
    _say(f"case peek surfaces bits={color_bits}")
    m, b = scenes.make_batch(43, color_bits=color_bits, view_sizes=((200, 150),), draws_per_view=50)
    mh = gpu_ctx.upload_map(m)
    try:
        fb, arena = gpu_ctx.render_plain(mh, b, peek_surfaces=True)
    finally:
        gpu_ctx.release_map(mh)
    ofb, oarena, _, _ = oracle.render(m, b, mode=oracle.MODE_EXACT)
    assert np.array_equal(arena, oarena[:len(arena)])
    assert np.array_equal(fb.reshape(-1), ofb)


@pytest.mark.parametrize("color_bits", [32, 64])
def test_single_view_device_inputs(gpu_ctx, color_bits):
    """A one-view batch with device-resident command buffers: the surface blocks search their
    surfaces themselves and run in the setup launch; eager and graph replay equal the oracle."""
    import sw_abi
    _say(f"case single view device inputs bits={color_bits}")
    m, b = scenes.make_batch(47, color_bits=color_bits, view_sizes=((320, 240),), draws_per_view=90)
    mh = gpu_ctx.upload_map(m)
    dev = sw_abi.DeviceCommandBuffers(gpu_ctx, b)
    out = gpu_ctx.alloc(b.fb_texels * (8 if color_bits == 64 else 4))
    try:
        eager = gpu_ctx.render_batch(mh, b, device_inputs=dev, out=out)
        arena = gpu_ctx.read_surfaces(b.arena_texels, color_bits)
        graphs = [gpu_ctx.render_batch(mh, b, graph=True, device_inputs=dev, out=out) for _ in range(3)]
    finally:
        out.free()
        dev.free()
        gpu_ctx.release_map(mh)
    ofb, oarena, _, _ = oracle.render(m, b, mode=oracle.MODE_EXACT)
    assert np.array_equal(arena, oarena[:b.arena_texels])
    assert np.array_equal(eager, ofb)
    for g in graphs:
        assert np.array_equal(g, ofb)


@pytest.mark.parametrize("views", [1, 3])
@pytest.mark.parametrize("device_inputs", [False, True])
def test_invalid_surface_is_an_error_and_recovers(gpu_ctx, views, device_inputs):
    """A descriptor naming a texture level past the table is reported (status code and message),
    nothing reads outside the map; the next valid batch renders exactly."""
    import sw_abi
    _say(f"case invalid surface views={views} device={device_inputs}")
    sizes = ((160, 120), (97, 61), (64, 200))[:views]
    m, b = scenes.make_batch(49, color_bits=32, view_sizes=sizes)
    bad_surfaces = b.surfaces.copy()
    bad_surfaces[len(bad_surfaces) // 2]["texture_level"] = 10 ** 6
    bad_surfaces[len(bad_surfaces) // 2]["flags"] |= sw_abi.SURF_DIFFUSE_LAYER
    bad = sw_abi.BatchData(b.views, bad_surfaces, b.lights, b.light_indices, b.draws, b.vertex_data, 32,
                           arena_texels=b.arena_texels)
    mh = gpu_ctx.upload_map(m)
    try:
        if device_inputs:
            dev_bad = sw_abi.DeviceCommandBuffers(gpu_ctx, bad)
            dev_ok = sw_abi.DeviceCommandBuffers(gpu_ctx, b)
            out = gpu_ctx.alloc(b.fb_texels * 4)
            try:
                with pytest.raises(sw_abi.SwError, match="surface"):
                    gpu_ctx.render_batch(mh, bad, device_inputs=dev_bad, out=out)
                fb = gpu_ctx.render_batch(mh, b, device_inputs=dev_ok, out=out)
            finally:
                out.free()
                dev_bad.free()
                dev_ok.free()
        else:
            with pytest.raises(sw_abi.SwError, match="surface"):
                gpu_ctx.render_batch(mh, bad)
            fb = gpu_ctx.render_batch(mh, b)
    finally:
        gpu_ctx.release_map(mh)
    ofb, _, _, _ = oracle.render(m, b)
    assert np.array_equal(fb, ofb)


@pytest.mark.parametrize("case", ["tiny", "large", "wrap_overflow"])
def test_surface_block_layouts(gpu_ctx, case):
    """The surface kernel's block structure: blocks overlapping more surfaces than one pass of
    per-surface records holds (many 1..3-texel surfaces), arenas large enough for several texels
    per thread (> 5.2 Mtexels), and texture coordinates whose tc_start + v overflows i32 (the
    reference's wrapping expression is kept); arena and frame equal the oracle's."""
    _say(f"case surface blocks {case}")
    if case == "tiny":
        m, b = scenes.make_batch(61, view_sizes=((96, 64), (80, 60)), surfaces_per_view=3000, draws_per_view=20,
                                 max_surface_size=3)
    elif case == "large":
        m, b = scenes.make_batch(62, view_sizes=((128, 96), (96, 64), (64, 64)), surfaces_per_view=170, draws_per_view=20,
                                 max_surface_size=300, num_lm=8000)
        assert b.arena_texels > 5_300_000, b.arena_texels
    else:
        m, b = scenes.make_batch(63, view_sizes=((160, 120),), surfaces_per_view=40, draws_per_view=30)
        s = b.surfaces
        s["tc_start"][::3, 1] = 2**31 - 7
        s["tc_start"][1::3, 0] = -2**31 + 2
        s["tc_start"][2::3, 1] = -2**31 + 1
    mh = gpu_ctx.upload_map(m)
    try:
        fb = gpu_ctx.render_batch(mh, b)
        arena = gpu_ctx.read_surfaces(b.arena_texels, 32)
    finally:
        gpu_ctx.release_map(mh)
    ofb, oarena, _, _ = oracle.render(m, b, mode=oracle.MODE_EXACT)
    assert np.array_equal(arena, oarena[:b.arena_texels]), \
        f"surface arena differs in {np.count_nonzero(arena != oarena[:b.arena_texels])} texels"
    assert np.array_equal(fb, ofb)


@pytest.mark.parametrize("device_inputs", [False, True])
def test_many_views_large_arena(gpu_ctx, device_inputs):
    """Nine views over 4.6 Mtexels of surfaces (two texels per thread in the surface blocks, the
    block table from the host or from the setup launch); arena, frame and graph replays equal the
    oracle's."""
    import sw_abi
    _say(f"case many views large arena device={device_inputs}")
    sizes = ((96, 64), (80, 60), (128, 72), (64, 64), (100, 50), (72, 96), (90, 90), (120, 64), (60, 40))
    m, b = scenes.make_batch(71, view_sizes=sizes, surfaces_per_view=40, draws_per_view=30, max_surface_size=300,
                             num_lm=8000)
    assert b.arena_texels > (1 << 22), b.arena_texels
    mh = gpu_ctx.upload_map(m)
    try:
        if device_inputs:
            dev = sw_abi.DeviceCommandBuffers(gpu_ctx, b)
            out = gpu_ctx.alloc(b.fb_texels * 4)
            try:
                fb = gpu_ctx.render_batch(mh, b, device_inputs=dev, out=out)
                arena = gpu_ctx.read_surfaces(b.arena_texels, 32)
                graphs = [gpu_ctx.render_batch(mh, b, graph=True, device_inputs=dev, out=out) for _ in range(3)]
            finally:
                out.free()
                dev.free()
        else:
            fb = gpu_ctx.render_batch(mh, b)
            arena = gpu_ctx.read_surfaces(b.arena_texels, 32)
            graphs = []
    finally:
        gpu_ctx.release_map(mh)
    ofb, oarena, _, _ = oracle.render(m, b, mode=oracle.MODE_EXACT)
    assert np.array_equal(arena, oarena[:b.arena_texels])
    assert np.array_equal(fb, ofb)
    for g in graphs:
        assert np.array_equal(g, ofb)


@pytest.mark.parametrize("color_bits", [32, 64])
@pytest.mark.parametrize("seed", [81, 82])
def test_lightmap_lines_staged(gpu_ctx, color_bits, seed):
    """Surface blocks of four or more texels per thread (arenas above 786,432 texels) stage each surface's
    lightmap lines (the vertical mix of two lightmap rows per surface row and column) in LDS;
    surfaces whose lines do not fit the block's share fetch their elements directly. Lightmap
    scales 1..16, simple and directional, surfaces up to 150 texels wide (many blocks hold both
    kinds): arena and frame equal the oracle's, eager and graph-replayed."""
    import sw_abi
    _say(f"case lightmap lines bits={color_bits} seed={seed}")
    sizes = ((96, 64), (80, 60), (128, 72), (64, 64), (100, 50), (72, 96))
    m, b = scenes.make_batch(seed + color_bits, color_bits=color_bits, view_sizes=sizes, surfaces_per_view=60,
                             draws_per_view=30, max_surface_size=150, num_lm=8000)
    assert b.arena_texels > 786_432, b.arena_texels
    mh = gpu_ctx.upload_map(m)
    dev = sw_abi.DeviceCommandBuffers(gpu_ctx, b)
    out = gpu_ctx.alloc(b.fb_texels * color_bits // 8)
    try:
        fb = gpu_ctx.render_batch(mh, b, device_inputs=dev, out=out)
        arena = gpu_ctx.read_surfaces(b.arena_texels, color_bits)
        graphs = [gpu_ctx.render_batch(mh, b, graph=True, device_inputs=dev, out=out) for _ in range(2)]
    finally:
        out.free()
        dev.free()
        gpu_ctx.release_map(mh)
    ofb, oarena, _, _ = oracle.render(m, b, mode=oracle.MODE_EXACT)
    assert np.array_equal(arena, oarena[:b.arena_texels]), \
        f"surface arena differs in {np.count_nonzero(arena != oarena[:b.arena_texels])} texels"
    assert np.array_equal(fb, ofb)
    for g in graphs:
        assert np.array_equal(g, ofb)


def test_decreasing_arena_offsets_are_an_error(gpu_ctx):
    import sw_abi
    _say("case decreasing arena offsets")
    m, b = scenes.make_batch(73, view_sizes=((64, 48), (64, 48)), surfaces_per_view=4, draws_per_view=4)
    views = b.views.copy()
    views["surface_arena_offset"] = views["surface_arena_offset"][::-1]
    bad = sw_abi.BatchData(views, b.surfaces, b.lights, b.light_indices, b.draws, b.vertex_data, 32,
                           arena_texels=b.arena_texels)
    mh = gpu_ctx.upload_map(m)
    try:
        with pytest.raises(sw_abi.SwError, match="non-decreasing"):
            gpu_ctx.render_batch(mh, bad)
    finally:
        gpu_ctx.release_map(mh)
