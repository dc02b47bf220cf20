"""GPU parity of the fire effect (sw_fire_step, sw_fire_make_texture; generative_texture_effect_fire.rs):
heat maps of several sizes through steps of both update variants with set_heat points (repeated and
wrapped), and the palette texture with its lite mips, bit-exact to the oracle."""
import numpy as np
import pytest

import oracle
import sw_abi as A
import sw_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("slow", [False, True])
@pytest.mark.parametrize("log2", [(4, 4), (6, 7), (10, 9)])
def test_fire_steps_bit_exact(gpu_ctx, slow, log2):
    lw, lh = log2
    n = 1 << (lw + lh)
    rng = np.random.default_rng(lw + 5 * lh + int(slow))
    heat = rng.integers(0, 256, n).astype(np.uint8)
    steps = []
    for _ in range(8):
        k = int(rng.integers(0, 100))
        p = np.zeros(k, A.HEAT_POINT_DTYPE)
        p["x"], p["y"], p["heat"] = rng.integers(0, 3 << lw, k), rng.integers(0, 3 << lh, k), rng.integers(0, 256, k)
        if k > 3:
            p[2]["x"], p[2]["y"] = p[0]["x"], p[0]["y"]
        steps.append(p)
    print(f"fire steps {lw}x{lh} slow={slow}", flush=True)
    got = gpu_ctx.fire_steps(heat, lw, lh, 0.985, slow, steps)
    ref = heat
    for p in steps:
        ref = oracle.fire_step(ref, lw, lh, 0.985, slow, p)
    assert np.array_equal(got, ref), np.count_nonzero(got != ref)


def test_fire_texture_bit_exact(gpu_ctx):
    m = sw_host.Map(seed=3, grid=(10, 10, 4), texture_size=64, num_materials=3, num_decal_textures=1, num_model_textures=1)
    rng = np.random.default_rng(9)
    heat = rng.integers(0, 256, 64 * 64).astype(np.uint8)
    palette = rng.integers(0, 2 ** 32, 256, dtype=np.uint64).astype(np.uint32)
    mh = gpu_ctx.upload_map(m.data)
    try:
        gpu_ctx.fire_make_texture(mh, heat, 6, 6, 20, palette)
        got = gpu_ctx.map_read(mh, 1, 0, len(m.data.lite))
        with pytest.raises(A.SwError, match="lite levels"):
            gpu_ctx.fire_make_texture(mh, heat, 6, 6, 0, palette)
    finally:
        gpu_ctx.release_map(mh)
    ref = oracle.fire_make_texture(m.data, heat, 6, 6, 20, palette)
    assert np.array_equal(got, ref), np.count_nonzero(got != ref)


@pytest.mark.parametrize("log2", [(4, 4), (7, 5), (5, 8), (10, 10)])
def test_fire_texture_sizes_bit_exact(gpu_ctx, log2):
    """Non-square and single-tile sizes and 1024 x 1024: lite levels 20-24 rewritten as the heat map's
    size and its mips in lite texels appended to the map, all five levels against the oracle."""
    lw, lh = log2
    m = sw_host.Map(seed=3, grid=(10, 10, 4), texture_size=64, num_materials=3, num_decal_textures=1, num_model_textures=1)
    d = m.data
    w, h = 1 << lw, 1 << lh
    sizes = [(max(w >> k, 1), max(h >> k, 1)) for k in range(5)]
    base = len(d.lite)
    d.lite = np.concatenate([d.lite, np.zeros(sum(a * b for a, b in sizes), np.uint32)])
    off = base
    for k, (a, b) in enumerate(sizes):
        d.levels[20 + k]["offset"], d.levels[20 + k]["width"], d.levels[20 + k]["height"] = off, a, b
        off += a * b
    rng = np.random.default_rng(lw * 13 + lh)
    heat = rng.integers(0, 256, w * h).astype(np.uint8)
    palette = rng.integers(0, 2 ** 32, 256, dtype=np.uint64).astype(np.uint32)
    print(f"fire texture {lw}x{lh}", flush=True)
    mh = gpu_ctx.upload_map(d)
    try:
        gpu_ctx.fire_make_texture(mh, heat, lw, lh, 20, palette)
        got = gpu_ctx.map_read(mh, 1, 0, len(d.lite))
    finally:
        gpu_ctx.release_map(mh)
    ref = oracle.fire_make_texture(d, heat, lw, lh, 20, palette)
    for k in range(5):
        o, n = int(d.levels[20 + k]["offset"]), int(d.levels[20 + k]["width"]) * int(d.levels[20 + k]["height"])
        assert np.array_equal(got[o:o + n], ref[o:o + n]), f"mip {k}: {np.count_nonzero(got[o:o + n] != ref[o:o + n])} differ"
    assert np.array_equal(got, ref)
