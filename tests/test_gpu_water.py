"""GPU parity of the water effect (sw_water_step, sw_water_make_texture;
generative_texture_effect_water.rs): fields of several sizes through steps with wave-source points
(repeated and wrapped coordinates), and the wavy texture with its mips in every colour mode, bit-exact
to the oracle."""
import numpy as np
import pytest

import oracle
import sw_abi as A
import sw_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("log2", [(4, 4), (6, 5), (9, 8), (10, 11)])
def test_water_steps_bit_exact(gpu_ctx, log2):
    lw, lh = log2
    n = 1 << (lw + lh)
    rng = np.random.default_rng(lw + 7 * lh)
    f = rng.normal(size=n).astype(np.float32)
    fo = rng.normal(size=n).astype(np.float32)
    att = np.float32(1.0 - 1.0 / 120.0)
    steps = []
    for s in range(10):
        k = int(rng.integers(0, 40))
        p = np.zeros(k, dtype=A.WAVE_POINT_DTYPE)
        p["x"] = rng.integers(0, 4 << lw, k)
        p["y"] = rng.integers(0, 4 << lh, k)
        p["value"] = rng.normal(size=k) * 3
        if k > 3:
            p[1]["x"], p[1]["y"] = p[0]["x"], p[0]["y"]  # a repeated address
        steps.append(p)
    print(f"water steps {lw}x{lh}", flush=True)
    a, b = gpu_ctx.water_steps(f, fo, lw, lh, att, steps)
    ra, rb = f, fo
    for p in steps:
        ra, rb = oracle.water_step(ra, rb, lw, lh, att, p)
    assert np.array_equal(a.view(np.uint32), ra.view(np.uint32))
    assert np.array_equal(b.view(np.uint32), rb.view(np.uint32))


@pytest.mark.parametrize("mode,init", [(A.WATER_COLOR_KEEP, A.WATER_INIT_NONE), (A.WATER_COLOR_KEEP, A.WATER_INIT_BASE_COLOR),
                                       (A.WATER_COLOR_KEEP, A.WATER_INIT_SOURCE), (A.WATER_COLOR_DEFORMED, A.WATER_INIT_NONE),
                                       (A.WATER_COLOR_DEFORMED_X, A.WATER_INIT_NONE)])
def test_water_texture_bit_exact(gpu_ctx, mode, init):
    m = sw_host.Map(seed=3, grid=(10, 10, 4), texture_size=64, num_materials=3, num_decal_textures=1, num_model_textures=1)
    lv = m.data.levels
    lv["flags"][0:5] |= 1  # the water texture has a normal map (SW_TEX_HAS_NORMAL_MAP), as the effect sets
    rng = np.random.default_rng(mode * 3 + init)
    field = (rng.normal(size=64 * 64) * 0.8).astype(np.float32)
    params = A.WaterTextureParams(0, 5, mode, init, 0.3, 0x80604020)
    mh = gpu_ctx.upload_map(m.data)
    try:
        gpu_ctx.water_make_texture(mh, field, 6, 6, params)
        got = gpu_ctx.map_read(mh, 0, 0, len(m.data.texels))
    finally:
        gpu_ctx.release_map(mh)
    ref = oracle.water_make_texture(m.data, field, 6, 6, params)
    bad = np.flatnonzero(got.view(np.uint64) != ref.view(np.uint64))
    assert bad.size == 0, f"{bad.size} texels differ, first {bad[:5].tolist()}"


def _water_levels(m, lw, lh, rng):
    """Levels 0-4 of the map rewritten as a 2^lw x 2^lh texture and its four mips (with a normal map)
    in texels appended to the map, level 5 as a colour texture of the same size."""
    d = m.data
    w, h = 1 << lw, 1 << lh
    sizes = [(max(w >> k, 1), max(h >> k, 1)) for k in range(5)]
    extra = sum(a * b for a, b in sizes) + w * h
    base = len(d.texels)
    add = np.zeros(extra, dtype=d.texels.dtype)
    add.view(np.uint32).reshape(-1, 2)[:, 0] = rng.integers(0, 2 ** 32, extra, dtype=np.uint64).astype(np.uint32)
    d.texels = np.concatenate([d.texels, add])
    off = base
    for k, (a, b) in enumerate(sizes):
        d.levels[k]["offset"], d.levels[k]["width"], d.levels[k]["height"] = off, a, b
        d.levels[k]["flags"] |= 1
        off += a * b
    d.levels[5]["offset"], d.levels[5]["width"], d.levels[5]["height"] = off, w, h
    d.levels[5]["flags"] &= ~np.uint32(A.TEX_LITE_LEVEL if hasattr(A, "TEX_LITE_LEVEL") else 0)


@pytest.mark.parametrize("log2", [(4, 4), (7, 5), (5, 8), (10, 10)])
@pytest.mark.parametrize("mode", [A.WATER_COLOR_KEEP, A.WATER_COLOR_DEFORMED, A.WATER_COLOR_DEFORMED_X])
def test_water_texture_sizes_bit_exact(gpu_ctx, log2, mode):
    """Non-square and single-tile sizes and the bench's 1024 x 1024: the tile split and every mip's
    write stride, all five levels against the oracle."""
    lw, lh = log2
    m = sw_host.Map(seed=3, grid=(10, 10, 4), texture_size=64, num_materials=3, num_decal_textures=1, num_model_textures=1)
    rng = np.random.default_rng(lw * 31 + lh * 7 + mode)
    _water_levels(m, lw, lh, rng)
    field = (rng.normal(size=1 << (lw + lh)) * 0.8).astype(np.float32)
    params = A.WaterTextureParams(0, 5, mode, A.WATER_INIT_NONE, 0.3, 0x80604020)
    print(f"water texture {lw}x{lh} mode {mode}", flush=True)
    mh = gpu_ctx.upload_map(m.data)
    try:
        gpu_ctx.water_make_texture(mh, field, lw, lh, params)
        got = gpu_ctx.map_read(mh, 0, 0, len(m.data.texels))
    finally:
        gpu_ctx.release_map(mh)
    ref = oracle.water_make_texture(m.data, field, lw, lh, params)
    lv = m.data.levels
    for k in range(5):
        o, n = int(lv[k]["offset"]), int(lv[k]["width"]) * int(lv[k]["height"])
        bad = np.flatnonzero(got[o:o + n].view(np.uint64) != ref[o:o + n].view(np.uint64))
        assert bad.size == 0, f"mip {k}: {bad.size} texels differ, first {bad[:5].tolist()}"
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_water_validation(gpu_ctx):
    m = sw_host.Map(seed=3, grid=(10, 10, 4), texture_size=64, num_materials=3, num_decal_textures=1, num_model_textures=1)
    m.data.levels["flags"][0:5] &= ~np.uint32(1)
    mh = gpu_ctx.upload_map(m.data)
    try:
        with pytest.raises(A.SwError, match="size out of range"):
            gpu_ctx.water_steps(np.zeros(64, np.float32), np.zeros(64, np.float32), 3, 3, 0.9, [np.zeros(0, A.WAVE_POINT_DTYPE)])
        with pytest.raises(A.SwError, match="lacks SW_TEX_HAS_NORMAL_MAP"):
            gpu_ctx.water_make_texture(mh, np.zeros(64 * 64, np.float32), 6, 6, A.WaterTextureParams(0, 5, 0, 0, 0.5, 0))
        m.data.levels["flags"][0:5] |= 1
        gpu_ctx.release_map(mh)
        mh = gpu_ctx.upload_map(m.data)
        with pytest.raises(A.SwError, match="must be the field's size"):
            gpu_ctx.water_make_texture(mh, np.zeros(128 * 64, np.float32), 7, 6, A.WaterTextureParams(0, 5, 0, 0, 0.5, 0))
        with pytest.raises(A.SwError, match="one of the destinations"):
            gpu_ctx.water_make_texture(mh, np.zeros(64 * 64, np.float32), 6, 6, A.WaterTextureParams(0, 0, 1, 0, 0.5, 0))
    finally:
        gpu_ctx.release_map(mh)
