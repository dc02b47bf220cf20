"""GPU parity of shadowed dynamic lights (sw_set_light_shadows; light.rs:185-294 in
surfaces.rs:750-760): surfaces and framebuffers bit-exact to the oracle with cube and projector maps
whose depths straddle the scene's, through the batched path and the plain path."""
import numpy as np
import pytest

import oracle
import scenes
import sw_abi as A

pytestmark = pytest.mark.gpu


def _random_shadows(rng, n, size=16):
    sh = np.zeros(n, dtype=A.LIGHT_SHADOW_DTYPE)
    off = 0
    for i in range(n):
        kind = [A.SHADOW_CUBE, A.SHADOW_PROJECTOR, A.SHADOW_NONE][i % 3]
        sh[i]["kind"], sh[i]["size"], sh[i]["offset"] = kind, size, off
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        sh[i]["basis_x"], sh[i]["basis_y"], sh[i]["basis_z"] = q[0], q[1], q[2]
        off += size * size * (6 if kind == A.SHADOW_CUBE else 1)
    # depths (1 / z) around the scene's distances (tens to hundreds of units): both tests' outcomes
    tex = rng.uniform(0.0, 0.03, size=max(off, 1)).astype(np.float32)
    return sh, tex


@pytest.fixture
def shadow_ctx(gpu_ctx):
    yield gpu_ctx
    gpu_ctx.set_light_shadows(None, None)
    oracle.set_light_shadows(None, None)


@pytest.mark.parametrize("color_bits", [32, 64])
def test_shadowed_batch_bit_exact(shadow_ctx, color_bits):
    gpu_ctx = shadow_ctx
    rng = np.random.default_rng(color_bits)
    m, b = scenes.make_batch(21, color_bits=color_bits, view_sizes=((160, 120), (97, 61)), num_lights=9, surfaces_per_view=40)
    sh, tex = _random_shadows(rng, len(b.lights))
    gpu_ctx.set_light_shadows(sh, tex)
    oracle.set_light_shadows(sh, tex)
    mh = gpu_ctx.upload_map(m)
    try:
        print(f"shadowed batch bits {color_bits}", flush=True)
        fb = gpu_ctx.render_batch(mh, b)
        arena = gpu_ctx.read_surfaces(b.arena_texels, color_bits)
        ref_fb, ref_arena, _, _ = oracle.render(m, b)
        assert np.array_equal(arena, ref_arena[:b.arena_texels]), np.count_nonzero(arena != ref_arena[:b.arena_texels])
        assert np.array_equal(fb, ref_fb)
        # the shadows matter here: without them the surfaces differ
        oracle.set_light_shadows(None, None)
        plain_arena, _ = oracle.build_surfaces(m, b)
        assert not np.array_equal(plain_arena[:b.arena_texels], ref_arena[:b.arena_texels])
    finally:
        gpu_ctx.release_map(mh)


def test_shadowed_plain_frame_bit_exact(shadow_ctx):
    gpu_ctx = shadow_ctx
    rng = np.random.default_rng(3)
    m, b = scenes.make_batch(22, color_bits=32, view_sizes=((200, 150),), num_lights=7, surfaces_per_view=50)
    sh, tex = _random_shadows(rng, len(b.lights), size=32)
    gpu_ctx.set_light_shadows(sh, tex)
    oracle.set_light_shadows(sh, tex)
    mh = gpu_ctx.upload_map(m)
    try:
        fb, arena = gpu_ctx.render_plain(mh, b, peek_surfaces=True)
        ref_fb, ref_arena, _, _ = oracle.render(m, b)
        assert np.array_equal(arena[:b.arena_texels], ref_arena[:b.arena_texels])
        assert np.array_equal(fb.reshape(-1), ref_fb)
        # switching the shadows off gives the unshadowed surfaces again
        gpu_ctx.set_light_shadows(None, None)
        oracle.set_light_shadows(None, None)
        fb2 = gpu_ctx.render_plain(mh, b)
        ref2, _, _, _ = oracle.render(m, b)
        assert np.array_equal(fb2.reshape(-1), ref2)
    finally:
        gpu_ctx.release_map(mh)


@pytest.mark.parametrize("color_bits", [32, 64])
def test_shadowed_large_arena_bit_exact(shadow_ctx, color_bits):
    """Shadowed surfaces over an arena of four or more texels per thread (blocks that stage their lightmap
    lines in LDS): arena and frame equal the oracle's."""
    gpu_ctx = shadow_ctx
    rng = np.random.default_rng(40 + color_bits)
    m, b = scenes.make_batch(23, color_bits=color_bits, view_sizes=((160, 120), (128, 96), (97, 61), (120, 80)),
                             num_lights=9, surfaces_per_view=60, max_surface_size=150, num_lm=8000)
    assert b.arena_texels > 786_432, b.arena_texels
    sh, tex = _random_shadows(rng, len(b.lights))
    gpu_ctx.set_light_shadows(sh, tex)
    oracle.set_light_shadows(sh, tex)
    mh = gpu_ctx.upload_map(m)
    try:
        print(f"shadowed large arena bits {color_bits}", flush=True)
        fb = gpu_ctx.render_batch(mh, b)
        arena = gpu_ctx.read_surfaces(b.arena_texels, color_bits)
        ref_fb, ref_arena, _, _ = oracle.render(m, b)
        assert np.array_equal(arena, ref_arena[:b.arena_texels]), np.count_nonzero(arena != ref_arena[:b.arena_texels])
        assert np.array_equal(fb, ref_fb)
    finally:
        gpu_ctx.release_map(mh)


def test_shadow_validation(shadow_ctx):
    sh = np.zeros(2, dtype=A.LIGHT_SHADOW_DTYPE)
    sh[1]["kind"], sh[1]["size"], sh[1]["offset"] = A.SHADOW_CUBE, 8, 10
    with pytest.raises(A.SwError, match="outside the shadow texels"):
        shadow_ctx.set_light_shadows(sh, np.zeros(6 * 64, np.float32))
    sh[1]["kind"] = 7
    with pytest.raises(A.SwError, match="unknown kind"):
        shadow_ctx.set_light_shadows(sh, np.zeros(6 * 64 + 10, np.float32))
