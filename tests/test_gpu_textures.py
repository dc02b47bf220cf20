"""GPU parity of sw_apply_texture_ops (generative_texture_effect_turb.rs, generative_texture_effect_layered.rs,
surfaces.rs:903-971): the map's device textures after a frame's turbulence and layer operations equal
the oracle's on host copies, bit for bit, then the rendered frame does."""
import numpy as np
import pytest

import oracle
import sw_abi as A
import sw_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tmap():
    # 12 materials of 64x64 with normal maps (5 mips each), 4 decal (32x32) and 4 model (64x64) lite textures
    return sw_host.Map(seed=11, grid=(10, 10, 4), texture_size=64, num_materials=12, normal_maps=1, specular_fraction=0.5,
                       num_decal_textures=4, num_model_textures=4)


def _material_levels(m, k):
    """The five mip levels of material k (Texture levels in table order, lite levels excluded)."""
    lv = m.data.levels
    tex = [i for i in range(len(lv)) if not (lv[i]["flags"] & A.TEX_LITE_LEVEL)]
    return tex[k * 5:(k + 1) * 5]


def _lite_chains(m):
    lv = m.data.levels
    lite = [i for i in range(len(lv)) if lv[i]["flags"] & A.TEX_LITE_LEVEL]
    chains, cur = [], []
    for i in lite:
        if cur and lv[i]["width"] * 2 != lv[cur[-1]]["width"]:
            chains.append(cur)
            cur = []
        cur.append(i)
    chains.append(cur)
    return chains


def _frame_ops(m, time_s):
    """One frame of effects: turbulence of material 0 into material 1's levels, a three-layer stack
    into material 2 (none, additive modulated and shifted, alpha blend), turbulence of a model skin
    into another and an emissive layer stack over a model skin, every mip."""
    ops, shifts = [], []
    turb = sw_host.TurbShifts()
    src, dst = _material_levels(m, 0), _material_levels(m, 1)
    for mip, (s, d) in enumerate(zip(src, dst)):
        w, h = int(m.data.levels[d]["width"]), int(m.data.levels[d]["height"])
        table = turb.compute(4.0, 32.0, 0.7, time_s, w, h, mip)
        ops.append((A.TEXOP_TURB, d, s, 0, (0, 0), (1, 1, 1), len(shifts)))
        shifts.extend(table.tolist())
    l0, l1, l2, l3 = (_material_levels(m, k) for k in (3, 4, 5, 2))
    for mip in range(5):
        sx = sw_host.layer_shift(sw_host.evaluate_function(1, [13.0, 0.0], time_s), mip)
        sy = sw_host.layer_shift(sw_host.evaluate_function(2, [0.5, 0.3, 20.0, 0.0], time_s), mip)
        mod = sw_host.evaluate_function(3, [0.25, 0.0, 0.6, 0.8], time_s)
        ops.append((A.TEXOP_LAYER, l3[mip], l0[mip], A.BLEND_NONE, (0, 0), (1, 1, 1), 0))
        ops.append((A.TEXOP_LAYER, l3[mip], l1[mip], A.BLEND_ADDITIVE, (sx, sy), (mod, mod * 0.5, 0.25), 0))
        ops.append((A.TEXOP_LAYER, l3[mip], l2[mip], A.BLEND_ALPHA_BLEND, (-sy, 3), (1, 1, 1), 0))
    chains = _lite_chains(m)
    skins = [c for c in chains if m.data.levels[c[0]]["width"] == 64]
    decals = [c for c in chains if m.data.levels[c[0]]["width"] == 32]
    for mip, (s, d) in enumerate(zip(skins[0], skins[1])):
        w, h = int(m.data.levels[d]["width"]), int(m.data.levels[d]["height"])
        table = turb.compute(2.0, 16.0, 1.3, time_s, w, h, mip)
        ops.append((A.TEXOP_TURB, d, s, 0, (0, 0), (1, 1, 1), len(shifts)))
        shifts.extend(table.tolist())
    for mip in range(len(skins[2])):
        ops.append((A.TEXOP_LAYER_LITE, skins[2][mip], skins[3][mip], A.BLEND_NONE, (0, 0), (1, 1, 1), 0))
        if mip < len(decals[0]):
            ops.append((A.TEXOP_LAYER_LITE, skins[2][mip], decals[0][mip], A.BLEND_ALPHA_TEST, (5 >> mip, -9), (1.5, 1.0, 0.5), 0))
            ops.append((A.TEXOP_LAYER_LITE, skins[2][mip], decals[1][mip], A.BLEND_AVERAGE, (0, 7), (0.9, 0.9, 0.9), 0))
    arr = np.zeros(len(ops), dtype=A.TEXTURE_OP_DTYPE)
    for i, (k, d, s, b, sh, l, t) in enumerate(ops):
        arr[i]["kind"], arr[i]["dst_level"], arr[i]["src_level"], arr[i]["blend"] = k, d, s, b
        arr[i]["shift"], arr[i]["light"], arr[i]["turb_shifts"] = sh, l, t
    return arr, np.array(shifts, dtype=np.int32)


def test_texture_ops_bit_exact(gpu_ctx, tmap):
    mh = gpu_ctx.upload_map(tmap.data)
    try:
        data = tmap.data
        for frame, t in enumerate((0.0, 0.37, 1.9)):
            print(f"texture ops frame {frame} t {t}", flush=True)
            ops, shifts = _frame_ops(tmap, t)
            gpu_ctx.apply_texture_ops(mh, ops, shifts)
            tex, lite = oracle.apply_texture_ops(data, ops, shifts)
            gtex = gpu_ctx.map_read(mh, 0, 0, len(tex))
            glite = gpu_ctx.map_read(mh, 1, 0, len(lite))
            assert np.array_equal(gtex.view(np.uint64), tex.view(np.uint64)), np.count_nonzero(gtex.view(np.uint64) != tex.view(np.uint64))
            assert np.array_equal(glite, lite), np.count_nonzero(glite != lite)
            data = A.MapData(tex, lite, data.lm_simple, data.lm_dir, data.levels, data.sky32, data.sky64)
        # a frame drawn with the animated textures: the device's and the oracle's maps agree
        fb = sw_host.FrameBuilder()
        pose = tmap.camera_path(1, 1)[0]
        f = fb.build(tmap, pose, width=320, height=200)
        b = sw_host.batch_from_frames([f])
        img = gpu_ctx.render_plain(mh, b).reshape(-1)
        ref, _, _, _ = oracle.render(data, b)
        assert np.array_equal(img, ref)
    finally:
        gpu_ctx.release_map(mh)


@pytest.mark.parametrize("case", ["same_level", "src_is_dst", "not_consecutive", "turb_size", "kind"])
def test_texture_op_validation(gpu_ctx, tmap, case):
    mh = gpu_ctx.upload_map(tmap.data)
    try:
        a, b, c = _material_levels(tmap, 0)[0], _material_levels(tmap, 1)[0], _material_levels(tmap, 2)[0]
        ops = np.zeros(2, dtype=A.TEXTURE_OP_DTYPE)
        ops["kind"] = A.TEXOP_LAYER
        ops["light"] = 1.0
        if case == "same_level":
            ops = ops[:1]
            ops[0]["dst_level"] = ops[0]["src_level"] = a
        elif case == "src_is_dst":
            ops[0]["dst_level"], ops[0]["src_level"] = a, b
            ops[1]["dst_level"], ops[1]["src_level"] = c, a
        elif case == "not_consecutive":
            ops = np.concatenate([ops, ops[:1]])
            ops[0]["dst_level"], ops[0]["src_level"] = a, b
            ops[1]["dst_level"], ops[1]["src_level"] = c, b
            ops[2]["dst_level"], ops[2]["src_level"] = a, b
        elif case == "turb_size":
            ops = ops[:1]
            ops[0]["kind"], ops[0]["dst_level"], ops[0]["src_level"] = A.TEXOP_TURB, a, _material_levels(tmap, 1)[1]
        else:
            ops = ops[:1]
            ops[0]["kind"], ops[0]["dst_level"], ops[0]["src_level"] = A.TEXOP_LAYER_LITE, a, b
        with pytest.raises(A.SwError):
            gpu_ctx.apply_texture_ops(mh, ops, np.zeros(64, np.int32))
    finally:
        gpu_ctx.release_map(mh)


@pytest.mark.parametrize("bits", [32, 64])
def test_skybox_side_bit_exact(gpu_ctx, tmap, bits):
    """sw_make_skybox_side into the map's skybox buffer: the side and its mips equal the oracle's, and
    the texels around it are untouched."""
    rng = np.random.default_rng(40 + bits)
    dt, lane = (np.uint32, 8) if bits == 32 else (np.uint64, 16)
    img = rng.integers(0, 2 ** 63 if bits == 64 else 2 ** 32, size=(100, 70), dtype=np.uint64).astype(dt)
    mh = gpu_ctx.upload_map(tmap.data)
    try:
        which = 2 if bits == 32 else 3
        before = gpu_ctx.map_read(mh, which, 0, len(tmap.data.sky32 if bits == 32 else tmap.data.sky64))
        off = 1234
        size = gpu_ctx.make_skybox_side(mh, img, 2.3, off)
        assert size == 70
        ref = oracle.make_skybox_side(img, 2.3)
        after = gpu_ctx.map_read(mh, which, 0, len(before))
        assert np.array_equal(after[off:off + len(ref)], ref)
        assert np.array_equal(after[:off], before[:off]) and np.array_equal(after[off + len(ref):], before[off + len(ref):])
        with pytest.raises(A.SwError):
            gpu_ctx.make_skybox_side(mh, img, 1.0, len(before) - 10)
    finally:
        gpu_ctx.release_map(mh)
