"""The fire effect's heat map and texture (generative_texture_effect_fire.rs) on the CPU: the oracle's
in-place update in the reference's row order (update_heat_map, :357-468) pinned by a rolled-array
formulation (fast variant), a row-chain formulation (slow variant), constant maps and the cleared last
row; set_heat's maximum with wrapped and repeated points; the palette texture and the lite mips
against per-channel floor averages. GPU parity is in test_gpu_fire.py."""
import numpy as np
import pytest

import oracle
import sw_abi as A
import sw_host


def _att(a):
    return int(np.float32(min(max(a, 0.0), 1.0)) * np.float32(1 << 18))


def _np_fast(hm, w, h, att):
    m = hm.reshape(h, w).astype(np.uint32)
    r1 = np.roll(m, -1, axis=0)
    r2 = np.roll(m, -2, axis=0)
    r2[h - 2] = m[h - 1]
    s = np.roll(r1, 1, axis=1) + r1 + np.roll(r1, -1, axis=1) + r2
    out = ((s * np.uint32(_att(att))) >> 20).astype(np.uint8)
    out[h - 1] = 0
    return out.reshape(-1)


def _py_slow(hm, w, h, att):
    a = _att(att)
    m = hm.reshape(h, w).astype(np.int64)
    out = np.zeros((h, w), np.uint8)
    for y in range(h - 1):
        row, below = m[y], m[y + 1]
        first = ((row[w - 1] + row[0] + row[1] + below[0]) * a) >> 20
        prev = first
        out[y, 0] = first
        for x in range(1, w - 1):
            prev = ((prev + row[x] + row[x + 1] + below[x]) * a) >> 20
            out[y, x] = prev
        out[y, w - 1] = ((prev + row[w - 1] + first + below[w - 1]) * a) >> 20
    return out.reshape(-1)


@pytest.mark.parametrize("log2", [(4, 4), (5, 6), (7, 4)])
@pytest.mark.parametrize("att", [0.97, 1.0])
def test_update_variants(log2, att):
    lw, lh = log2
    w, h = 1 << lw, 1 << lh
    rng = np.random.default_rng(lw * 7 + lh)
    hm = rng.integers(0, 256, w * h).astype(np.uint8)
    assert np.array_equal(oracle.fire_step(hm, lw, lh, att, False, None), _np_fast(hm, w, h, att))
    assert np.array_equal(oracle.fire_step(hm, lw, lh, att, True, None), _py_slow(hm, w, h, att))


def test_constant_map_and_points():
    w = h = 16
    hm = np.full(w * h, 100, np.uint8)
    a = _att(0.9)
    out = oracle.fire_step(hm, 4, 4, 0.9, False, None).reshape(h, w)
    assert np.all(out[:h - 1] == (400 * a) >> 20) and np.all(out[h - 1] == 0)
    pts = np.zeros(4, A.HEAT_POINT_DTYPE)
    pts[:] = [(3, 5, 200), (3 + 16, 5 + 32, 250), (3, 5, 10), (7, 15, 1)]  # wrapped repeats: max 250
    out = oracle.fire_step(hm, 4, 4, 0.9, False, pts).reshape(h, w)
    assert out[5, 3] == 250 and out[15, 7] == 1
    assert out[0, 0] == (400 * a) >> 20


def test_texture_and_lite_mips():
    m = sw_host.Map(seed=3, grid=(10, 10, 4), texture_size=64, num_materials=3, num_decal_textures=1, num_model_textures=1)
    lv = m.data.levels
    assert [int(lv[i]["width"]) for i in range(20, 25)] == [64, 32, 16, 8, 4] and all(lv[20:25]["flags"] & 256)
    rng = np.random.default_rng(5)
    heat = rng.integers(0, 256, 64 * 64).astype(np.uint8)
    palette = rng.integers(0, 2 ** 32, 256, dtype=np.uint64).astype(np.uint32)
    out = oracle.fire_make_texture(m.data, heat, 6, 6, 20, palette)
    lite = out[lv[20]["offset"]:lv[20]["offset"] + 4096]
    assert np.array_equal(lite, palette[heat])
    cur, w = lite.reshape(64, 64), 64
    for mip in range(1, 5):
        ch = [((cur >> (8 * c)) & 0xFF).astype(np.uint32) for c in range(4)]
        avg = [(q[0::2, 0::2] + q[1::2, 0::2] + q[0::2, 1::2] + q[1::2, 1::2]) >> 2 for q in ch]
        cur = avg[0] | (avg[1] << 8) | (avg[2] << 16) | (avg[3] << 24)
        w //= 2
        got = out[lv[20 + mip]["offset"]:lv[20 + mip]["offset"] + w * w].reshape(w, w)
        assert np.array_equal(got, cur), mip
    mask = np.ones(len(out), bool)
    for i in range(20, 25):
        mask[lv[i]["offset"]:lv[i]["offset"] + int(lv[i]["width"]) ** 2] = False
    assert np.array_equal(out[mask], m.data.lite[mask])
