"""Compiled-map files (common/bsp_map_save_load.rs:23-362, bsp_map_compact.rs:7-199,
lightmap_compression.rs) and the Quake MAP writer (map_file_q1.rs:24-115, map_polygonizer.rs:293-305),
on the CPU: the file is parsed here independently of the writer, with the element layouts of the
reference's #[repr(C)] structs written out below."""
import os

import numpy as np
import pytest

import sw_abi as A
import sw_host

# element layouts of bsp_map_compact.rs / lightmap_compression.rs (all #[repr(C)])
PLANE = [("vec", "<f4", 3), ("dist", "<f4")]
DT = {
    0: np.dtype([("children", "<u4", 2), ("plane", PLANE)]),  # BSPNode
    1: np.dtype([("first_polygon", "<u4"), ("num_polygons", "<u4"), ("first_leaf_portal", "<u4"),
                 ("num_leaf_portals", "<u4")]),  # BSPLeaf
    2: np.dtype([("first_vertex", "<u4"), ("num_vertices", "<u4"), ("plane", PLANE), ("tex_coord_equation", PLANE, 2),
                 ("tex_coord_min", "<i4", 2), ("tex_coord_max", "<i4", 2), ("lightmap_data_offset", "<u4"),
                 ("texture", "<u4")]),  # Polygon
    3: np.dtype([("leafs", "<u4", 2), ("plane", PLANE), ("first_vertex", "<u4"), ("num_vertices", "<u4")]),  # Portal
    4: np.dtype("<u4"),  # leafs_portals
    5: np.dtype(("<f4", 3)),  # vertices
    6: np.dtype("S64"),  # Texture = [u8; 64]
    7: np.dtype([("root_node", "<u4"), ("first_polygon", "<u4"), ("num_polygons", "<u4")]),  # Submodel
    8: np.dtype([("plane", PLANE), ("first_polygon", "<u4"), ("num_polygons", "<u4"), ("children", "<u4", 2)]),
    9: np.dtype([("first_key_value_pair", "<u4"), ("num_key_value_pairs", "<u4"), ("submodel_index", "<u4")]),
    10: np.dtype([("key", "<u2", 2), ("value", "<u2", 2)]),  # KeyValuePair of StringRef {offset, size}
    11: np.dtype("u1"),
    12: np.dtype([("components", "u1", 3), ("scale", "u1")]),  # CompressedColor
    13: np.dtype([("ambient", "u1", 4), ("dir", "i1", 3), ("dir_scale", "u1"), ("deviation", "u1"), ("color", "u1", 4)]),
    14: np.dtype([("grid_cell_size", "<f4", 3), ("grid_start", "<f4", 3), ("grid_size", "<u4", 3)]),
    15: np.dtype([("start_z", "<u4"), ("first_sample", "<u4"), ("num_samples", "<u4")]),
    16: np.dtype([("cube", "u1", (6, 4)), ("dir", "i1", 3), ("dir_scale", "u1"), ("color", "u1", 4)]),
}
SIZES = [24, 16, 80, 32, 4, 12, 64, 12, 32, 12, 8, 1, 4, 13, 36, 12, 32]


def parse(path):
    raw = open(path, "rb").read()
    assert raw[:4] == b"SqwM"
    version = int(np.frombuffer(raw, "<u4", 1, 4)[0])
    lumps = np.frombuffer(raw, "<u4", 96, 8).reshape(32, 3)
    out = {}
    for i in range(17):
        off, size, count = (int(x) for x in lumps[i])
        out[i] = np.frombuffer(raw, DT[i], count, off) if count else np.zeros(0, DT[i])
    return raw, version, lumps, out


def decompress_color(c):
    return c["components"].astype(np.float32) * (c["scale"].astype(np.float32) / np.float32(255))[..., None]


@pytest.fixture(scope="module")
def maps(tmp_path_factory):
    d = tmp_path_factory.mktemp("sqwm")
    m = sw_host.Map(seed=7, grid=(10, 9, 4), solid_probability=0.35, num_lights=6, sun=1, sky_fraction=0.3,
                    directional_lightmaps=1, texture_size=64, num_materials=6, num_decal_textures=2, num_model_textures=2)
    path = os.path.join(d, "grid.sqwm")
    m.save_sqwm(path)
    return m, path, d


def test_header_and_lump_layout(maps):
    """SqwM, version 11, 32 lumps; the used lumps carry the reference's element sizes and follow the
    header back to back in lump order (write_lump appends); unused lumps stay zero."""
    m, path, _ = maps
    raw, version, lumps, data = parse(path)
    assert version == 11
    assert [int(x) for x in lumps[:17, 1]] == SIZES
    off = 392
    for i in range(17):
        assert int(lumps[i, 0]) == off
        off += int(lumps[i, 1]) * int(lumps[i, 2])
    assert off == len(raw)
    assert not lumps[17:].any()
    info = m.info
    assert len(data[0]) == info["num_nodes"] and len(data[1]) == info["num_leafs"]
    assert len(data[2]) == info["num_polygons"] and len(data[3]) == info["num_portals"]
    assert int(lumps[14, 2]) == 1 and len(data[7]) == 0 and len(data[8]) == 0


def test_structure_is_consistent(maps):
    """Node children are nodes or FIRST_LEAF_INDEX + leaf; leaf ranges cover the polygons in order;
    polygon planes contain their vertices; textures name the materials; the light grid's columns
    stay inside the grid and the sample array."""
    _, path, _ = maps
    _, _, _, d = parse(path)
    nodes, leafs, polys, verts = d[0], d[1], d[2], d[5]
    FIRST_LEAF = 1 << 31
    for c in nodes["children"].ravel():
        assert (c - FIRST_LEAF < len(leafs)) if c >= FIRST_LEAF else (c < len(nodes))
    assert int(leafs["first_polygon"][0]) == 0
    assert np.all(leafs["first_polygon"][1:] == leafs["first_polygon"][:-1] + leafs["num_polygons"][:-1])
    for p in polys[:200]:
        v = verts[p["first_vertex"]:p["first_vertex"] + p["num_vertices"]]
        np.testing.assert_allclose(v @ p["plane"]["vec"], p["plane"]["dist"], atol=1e-3)
    names = [t.rstrip(b"\0").decode() for t in d[6]]
    assert names[-1] == "sw_gen/sky" and all(n.startswith("sw_gen/material_") for n in names[:-1])
    assert polys["texture"].max() < len(names)
    gh = d[14][0]
    gx, gy, gz = (int(x) for x in gh["grid_size"])
    cols = d[15]
    assert len(cols) == gx * gy
    assert np.all(cols["start_z"] + cols["num_samples"] <= gz)
    assert np.all(cols["first_sample"] + cols["num_samples"] <= len(d[16]))


def test_entities(maps):
    """worldspawn with the generator keys; one "light" entity (origin, light = intensity / 32) per light."""
    m, path, _ = maps
    _, _, _, d = parse(path)
    strings = d[11].tobytes()

    def s(ref):
        return strings[int(ref[0]):int(ref[0]) + int(ref[1])].decode()

    ents = []
    for e in d[9]:
        kv = d[10][e["first_key_value_pair"]:e["first_key_value_pair"] + e["num_key_value_pairs"]]
        ents.append({s(k["key"]): s(k["value"]) for k in kv})
    assert ents[0]["classname"] == "worldspawn" and ents[0]["sw_seed"] == "7" and ents[0]["sw_grid"] == "10 9 4"
    lights = [e for e in ents if e["classname"] == "light"]
    assert len(lights) == 6
    assert all(len(e["origin"].split()) == 3 and float(e["light"]) > 0 for e in lights)
    assert all(e["submodel_index"] == 0xFFFFFFFF for e in d[9])


@pytest.mark.parametrize("color,comps,scale", [((0.5, 0.25, 0.0), (127, 63, 0), 1), ((300.0, 10.0, -5.0), (255, 10, 0), 255),
                                                ((2.5, 1.0, 0.0), (212, 85, 0), 3), ((0.0, 0.0, 0.0), (0, 0, 0), 1),
                                                ((17.0, 17.0, 16.5), (255, 255, 247), 17)])
def test_color_compression_closed_form(color, comps, scale):
    """CompressedColor: scale = clamp(ceil(max), 1, 255); component = trunc(clamp(c * 255 / scale, 0, 255));
    decompress = component * scale / 255."""
    import ctypes as C
    lib = sw_host.load()
    out = (C.c_uint8 * 4)()
    lib.swh_compress_color((C.c_float * 3)(*color), out)
    assert tuple(out[:3]) == comps and out[3] == scale
    back = (C.c_float * 3)()
    lib.swh_decompress_color(out, back)
    np.testing.assert_array_equal(np.array(back[:], np.float32),
                                  np.array(comps, np.float32) * (np.float32(scale) / np.float32(255)))


def test_vector_compression_closed_form():
    import ctypes as C
    lib = sw_host.load()
    comps = (C.c_int8 * 3)()
    scale = C.c_uint8()
    lib.swh_compress_vector((C.c_float * 3)(-3.2, 1.0, 0.1), comps, C.byref(scale))
    assert tuple(comps[:]) == (-101, 31, 3) and scale.value == 4  # 127/4 = 31.75 per unit, truncated
    lib.swh_compress_vector((C.c_float * 3)(0.0, -0.5, 0.25), comps, C.byref(scale))
    assert tuple(comps[:]) == (0, -63, 31) and scale.value == 1


def test_round_trip(maps):
    """Loading gives the file's geometry and the decompressed lightmaps; saving the loaded map again
    reproduces every uncompressed lump byte for byte."""
    m, path, d = maps
    loaded = sw_host.Map.load_sqwm(path)
    _, _, lumps, data = parse(path)
    assert loaded.info == m.info
    # lightmaps: the loaded map's data are the file's, decompressed
    lm = decompress_color(data[12])
    np.testing.assert_array_equal(loaded.data.lm_simple, lm)
    dl = loaded.data.lm_dir
    np.testing.assert_array_equal(dl["ambient_light"], decompress_color(
        np.frombuffer(data[13]["ambient"].tobytes(), DT[12])))
    np.testing.assert_array_equal(dl["directional_light_deviation"], data[13]["deviation"].astype(np.float32) / np.float32(255))
    # the generated textures are the same
    np.testing.assert_array_equal(loaded.data.texels, m.data.texels)
    again = os.path.join(d, "again.sqwm")
    loaded.save_sqwm(again)
    raw2, _, lumps2, data2 = parse(again)
    for i in (0, 1, 2, 3, 4, 5, 6, 9, 10, 11, 14, 15):
        assert data2[i].tobytes() == data[i].tobytes(), i
    # compressing a decompressed value may pick another scale or component (the float round trip
    # can land just below an integer), but it represents the same light within one quantisation step
    a, b = decompress_color(data[12]), decompress_color(data2[12])
    step = np.maximum(data[12]["scale"], data2[12]["scale"]).astype(np.float32)[:, None] / 255
    assert np.all(np.abs(a - b) <= step * 1.0001)
    cube_a = np.frombuffer(data[16]["cube"].tobytes(), DT[12]).reshape(-1, 6)
    cube_b = np.frombuffer(data2[16]["cube"].tobytes(), DT[12]).reshape(-1, 6)
    step = np.maximum(cube_a["scale"], cube_b["scale"]).astype(np.float32)[..., None] / 255
    assert np.all(np.abs(decompress_color(cube_a) - decompress_color(cube_b)) <= step * 1.0001)
    va = data[16]["dir"] * (data[16]["dir_scale"].astype(np.float32) / 127)[:, None]
    vb = data2[16]["dir"] * (data2[16]["dir_scale"].astype(np.float32) / 127)[:, None]
    step = np.maximum(data[16]["dir_scale"], data2[16]["dir_scale"]).astype(np.float32)[:, None] / 127
    assert np.all(np.abs(va - vb) <= step * 1.0001)


def test_loaded_map_builds_the_same_frames(maps):
    """The frame builder on the loaded map: the same surfaces and draws (geometry from the file),
    lightmaps aside."""
    m, path, _ = maps
    loaded = sw_host.Map.load_sqwm(path)
    fb = sw_host.FrameBuilder()
    for pose in m.camera_path(3, 3):
        a = fb.build(m, pose, width=320, height=200)
        b = fb.build(loaded, pose, width=320, height=200)
        assert len(a["draws"]) == len(b["draws"]) and np.array_equal(a["draws"], b["draws"])
        assert np.array_equal(a["surfaces"], b["surfaces"])


def _patch(src, dst, offset, value, dtype="<u4"):
    raw = bytearray(open(src, "rb").read())
    raw[offset:offset + np.dtype(dtype).itemsize] = np.array([value], dtype).tobytes()
    open(dst, "wb").write(bytes(raw))


@pytest.mark.parametrize("what", ["truncated", "id", "version", "element_size", "polygon_vertex", "node_child", "texture_name",
                                  "lightmap_extent", "node_cycle", "node_shared"])
def test_bad_files_are_errors(maps, what):
    m, path, d = maps
    bad = os.path.join(d, f"bad_{what}.sqwm")
    raw = open(path, "rb").read()
    lumps = np.frombuffer(raw, "<u4", 96, 8).reshape(32, 3)
    if what == "truncated":
        open(bad, "wb").write(raw[:len(raw) // 2])
    elif what == "id":
        open(bad, "wb").write(b"QwsM" + raw[4:])
    elif what == "version":
        _patch(path, bad, 4, 10)
    elif what == "element_size":
        _patch(path, bad, 8 + 12 * 2 + 4, 76)
    elif what == "polygon_vertex":
        _patch(path, bad, int(lumps[2, 0]) + 80 * 5, 1 << 30)
    elif what == "node_child":
        _patch(path, bad, int(lumps[0, 0]), 0x7FFFFFFF)
    elif what == "lightmap_extent":
        # a lit polygon's lightmap starts at the last element: in range, but its lw * lh extent is not
        polys = np.frombuffer(raw, DT[2], int(lumps[2, 2]), int(lumps[2, 0]))  # lump: offset, element size, count
        k = int(np.nonzero(polys["lightmap_data_offset"])[0][0])
        count = max(int(lumps[12, 2]), int(lumps[13, 2]))
        _patch(path, bad, int(lumps[2, 0]) + 80 * k + DT[2].fields["lightmap_data_offset"][1], count - 1)
    elif what in ("node_cycle", "node_shared"):
        # the root (last node) points back at itself, or both its children name one subtree
        nn = int(lumps[0, 2])
        root = int(lumps[0, 0]) + 24 * (nn - 1)
        other = int(np.frombuffer(raw, "<u4", 1, root + 4)[0])
        _patch(path, bad, root, nn - 1 if what == "node_cycle" else other)
    else:
        raw2 = bytearray(raw)
        raw2[int(lumps[6, 0]):int(lumps[6, 0]) + 8] = b"unknown_"
        open(bad, "wb").write(bytes(raw2))
    with pytest.raises(RuntimeError, match="swh_map_load_sqwm"):
        sw_host.Map.load_sqwm(bad)


def _parse_quake_map(text):
    """A small parser of the format map_file_q1.rs reads: entities of "key" "value" pairs and brushes
    of ( x y z ) ( x y z ) ( x y z ) texture xoff yoff angle xscale yscale planes."""
    lines = [l.strip() for l in text.splitlines() if l.strip() and not l.strip().startswith("//")]
    ents, i = [], 0
    while i < len(lines):
        assert lines[i] == "{"
        i += 1
        ent = {"keys": {}, "brushes": []}
        while lines[i] != "}":
            if lines[i].startswith('"'):
                parts = lines[i].split('"')
                ent["keys"][parts[1]] = parts[3]
                i += 1
            else:
                assert lines[i] == "{"
                i += 1
                brush = []
                while lines[i] != "}":
                    t = lines[i].replace("(", " ").replace(")", " ").split()
                    pts = np.array([float(x) for x in t[:9]]).reshape(3, 3)
                    brush.append((pts, t[9], [float(x) for x in t[10:15]]))
                    i += 1
                ent["brushes"].append(brush)
                i += 1
        ents.append(ent)
        i += 1
    return ents


def test_quake_map(maps):
    """Every brush is an axis-aligned box: six planes whose (p0 - p1) x (p2 - p1) points out of the
    brush (map_polygonizer.rs:293-305 keeps points with dot(p, vec) <= dist); the boxes cover the
    solid and sky cells exactly once; lights and the generator keys are entities' keys."""
    m, _, d = maps
    qpath = os.path.join(d, "grid.map")
    m.save_quake_map(qpath)
    ents = _parse_quake_map(open(qpath).read())
    world = ents[0]
    assert world["keys"]["classname"] == "worldspawn" and world["keys"]["sw_seed"] == "7"
    volume = 0.0
    for brush in world["brushes"]:
        assert len(brush) == 6
        # the box from its planes: each plane is x_k = const for one axis k
        lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
        for pts, _, _ in brush:
            n = np.cross(pts[0] - pts[1], pts[2] - pts[1])
            k = int(np.argmax(np.abs(n)))
            assert np.count_nonzero(n) == 1
            lo[k], hi[k] = min(lo[k], pts[1][k]), max(hi[k], pts[1][k])
        assert np.all(hi > lo)
        center = (lo + hi) / 2
        for pts, tex, tc in brush:
            n = np.cross(pts[0] - pts[1], pts[2] - pts[1])
            dist = n @ pts[0]
            assert n @ center < dist  # the brush's centre is inside
            assert tex.startswith("sw_gen/") and tc == [0, 0, 0, 1, 1]
        volume += np.prod(hi - lo)
    gx, gy, gz = 10, 9, 4
    solid_cells = round(volume / 128.0 ** 3)
    assert solid_cells == gx * gy * gz - m.info["num_empty_cells"]
    assert sum(e["keys"].get("classname") == "light" for e in ents) == 6
    assert sum(e["keys"].get("classname") == "info_player_start" for e in ents) == 1
