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
