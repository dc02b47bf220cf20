"""Inline (brush) models: the seeded doors and platforms of the map, placed every frame by the frame
time (InlineModelsIndex, inline_models_index.rs), drawn inside the leaves they touch after the
leaf's polygons in the reference's object order (partial_renderer.rs:593-664, 1362-1447,
3098-3288; draw_ordering.rs), and written to / read from .sqwm files (lumps 7 and 8).

The checks are geometric: the per-frame basis of a moved polygon is the model transform of its
static basis, and the pixels the models end up owning in the oracle's image are those whose camera
ray meets a model before any wall.
"""
import math

import numpy as np
import pytest

import oracle
import sw_abi as A
import sw_host
from test_sqwm import parse

W, H = 320, 200


@pytest.fixture(scope="module")
def scene(tmp_path_factory):
    m = sw_host.large_map(5, grid=(10, 10, 4), num_submodels=10)
    d = tmp_path_factory.mktemp("sub")
    path = str(d / "sub.sqwm")
    m.save_sqwm(path)
    return m, path


def _submodel_matrix(k, time, centre):
    """The front-end's entity rule: even models swing about the vertical axis through their centre,
    odd ones rise and sink; model matrix translation(position) * rotation * translation(-centre)."""
    R = np.eye(4)
    pos = np.array(centre, np.float64)
    if k % 2 == 0:
        a = 0.7 * math.sin(0.9 * time + 1.3 * k)
        c, s = math.cos(a), math.sin(a)
        R[:2, :2] = [[c, -s], [s, c]]
    else:
        pos = pos + [0.0, 0.0, 24.0 * math.sin(1.1 * time + 0.7 * k)]
    T1 = np.eye(4)
    T1[:3, 3] = pos
    T0 = np.eye(4)
    T0[:3, 3] = -np.asarray(centre, np.float64)
    return T1 @ R @ T0


def _boxes(data):
    """Per submodel: its polygons' bbox (model space) and polygon range."""
    out = []
    for sm in data[7]:
        polys = data[2][int(sm["first_polygon"]):int(sm["first_polygon"]) + int(sm["num_polygons"])]
        vs = np.concatenate([data[5][int(p["first_vertex"]):int(p["first_vertex"]) + int(p["num_vertices"])] for p in polys])
        out.append((vs.min(0).astype(np.float64), vs.max(0).astype(np.float64), polys))
    return out


def test_submodels_in_the_map_file(scene):
    m, path = scene
    _, _, lumps, data = parse(path)
    assert len(data[7]) == 10 and len(data[8]) == 60
    leaf_polys = sum(int(l["num_polygons"]) for l in data[1])
    for k, (mn, mx, polys) in enumerate(_boxes(data)):
        sm = data[7][k]
        assert int(sm["first_polygon"]) >= leaf_polys  # after every leaf's polygons
        centre = (mn + mx) / 2
        for p in polys:
            v = data[5][int(p["first_vertex"]):int(p["first_vertex"]) + 4].astype(np.float64)
            n = p["plane"]["vec"].astype(np.float64)
            assert np.allclose(v @ n, p["plane"]["dist"], atol=1e-3)  # vertices on the plane
            assert (v.mean(0) - centre) @ n > 0  # faces point out of the box
            assert p["lightmap_data_offset"] != 0  # lit at the initial position
        # the node chain from the root visits the six faces once
        seen, i = [], int(sm["root_node"])
        while i < len(data[8]):
            seen.append(int(data[8][i]["first_polygon"]))
            i = int(data[8][i]["children"][0])
        assert sorted(seen) == list(range(int(sm["first_polygon"]), int(sm["first_polygon"]) + 6))
    ents = data[9]
    assert sorted(int(e["submodel_index"]) for e in ents if e["submodel_index"] != 0xFFFFFFFF) == list(range(10))


def test_sqwm_round_trip_keeps_the_models(scene):
    m, path = scene
    loaded = sw_host.Map.load_sqwm(path)
    fb = sw_host.FrameBuilder()
    for i, pose in enumerate(m.camera_path(4, 6)):
        a = fb.build(m, pose, width=W, height=H, use_directional_lightmaps=1, time=0.7 * i)
        b = fb.build(loaded, pose, width=W, height=H, use_directional_lightmaps=1, time=0.7 * i)
        assert a["first_submodel_surface"] == b["first_submodel_surface"]
        assert np.array_equal(a["draws"], b["draws"]) and np.array_equal(a["vertex_data"], b["vertex_data"])
        for f in ("basis_u", "basis_v", "basis_start", "basis_normal", "size", "tc_start", "lightmap_offset"):
            assert np.array_equal(a["surfaces"][f], b["surfaces"][f])


def _static_basis(p):
    """u, v, start and normal of a polygon from its planes (surfaces.rs:18-47), in float64: the point
    with texture coordinates (t0, t1) on the plane is start + u t0 + v t1."""
    A3 = np.stack([p["tex_coord_equation"][0]["vec"], p["tex_coord_equation"][1]["vec"], p["plane"]["vec"]]).astype(np.float64)
    inv = np.linalg.inv(A3)
    d0, d1 = float(p["tex_coord_equation"][0]["dist"]), float(p["tex_coord_equation"][1]["dist"])
    n = p["plane"]["vec"].astype(np.float64)
    return inv[:, 0], inv[:, 1], inv @ np.array([-d0, -d1, float(p["plane"]["dist"])]), n / np.linalg.norm(n)


def test_moved_polygons_take_the_model_transform(scene):
    """A visible model polygon's surface basis is the static basis moved by the frame's model matrix:
    start -> M start, u, v and the normal rotated (the texture moves with the model)."""
    m, path = scene
    _, _, _, data = parse(path)
    boxes = _boxes(data)
    fb = sw_host.FrameBuilder()
    checked = 0
    for i, pose in enumerate(m.camera_path(6, 80)):
        t = 0.37 * i + 0.1
        f = fb.build(m, pose, width=W, height=H, use_directional_lightmaps=1, time=t)
        first = f["first_submodel_surface"]
        for s in f["surfaces"][first:]:
            # which polygon: the one whose moved plane carries the surface's start and normal (the
            # front-end's centre is that of the bbox extended by 1/4 on every side: the same centre)
            found = 0
            for k, (mn, mx, polys) in enumerate(boxes):
                M = _submodel_matrix(k, t, (mn + mx) / 2)
                for p in polys:
                    u, v, st, n = _static_basis(p)
                    if np.linalg.norm(M[:3, :3] @ n - s["basis_normal"]) > 1e-3:
                        continue
                    start = M[:3, :3] @ st + M[:3, 3]
                    if np.linalg.norm(start - s["basis_start"]) > 0.05:
                        continue
                    scale = np.linalg.norm(s["basis_u"]) / np.linalg.norm(u)
                    assert scale in (1, 2, 4, 8, 16) or abs(scale - round(scale)) < 1e-3
                    assert np.allclose(M[:3, :3] @ u * scale, s["basis_u"], atol=1e-3)
                    assert np.allclose(M[:3, :3] @ v * scale, s["basis_v"], atol=1e-3)
                    found += 1
            assert found == 1, (i, found)
            checked += 1
    assert checked >= 30, checked


# ---------------------------------------------------------------------------------------------
# Ray casts: the models' pixels in the image are those whose camera ray meets a model first
# ---------------------------------------------------------------------------------------------
def _camera(pose, fov=math.pi / 2):
    """build_view_matrix_with_full_rotation (matrix.rs:13-109) in float64; returns the 4x4 view
    matrix (world -> (x z, y z, z))."""
    x, y, z, yaw, pitch = (float(c) for c in pose[:5])

    def qmul(a, b):
        s1, v1 = a[0], np.array(a[1:])
        s2, v2 = b[0], np.array(b[1:])
        return np.concatenate([[s1 * s2 - v1 @ v2], s1 * v2 + s2 * v1 + np.cross(v1, v2)])
    az = yaw + math.pi / 2
    qz = np.array([math.cos(az / 2), 0, 0, math.sin(az / 2)])
    qy = np.array([math.cos(-pitch / 2), 0, math.sin(-pitch / 2), 0])
    q = qmul(qz, qy)
    w_, (a, b, c) = q[0], -q[1:]
    R = np.array([[1 - 2 * (b * b + c * c), 2 * (a * b - c * w_), 2 * (a * c + b * w_)],
                  [2 * (a * b + c * w_), 1 - 2 * (a * a + c * c), 2 * (b * c - a * w_)],
                  [2 * (a * c - b * w_), 2 * (b * c + a * w_), 1 - 2 * (a * a + b * b)]])
    B = np.array([[0, -1, 0], [0, 0, -1], [1, 0, 0]], np.float64)
    t = 1 / math.tan(fov / 2)
    P = np.diag([t / (W / H) * W / 2, t * H / 2, 1.0])
    P[0, 2], P[1, 2] = W / 2, H / 2
    V = np.eye(4)
    V[:3, :3] = P @ B @ R
    V[:3, 3] = -(P @ B @ R) @ np.array([x, y, z])
    return V, np.array([x, y, z])


def _rays(V):
    ys, xs = np.mgrid[0:H, 0:W]
    scr = np.stack([xs.ravel() + 0.5, ys.ravel() + 0.5, np.ones(W * H), np.ones(W * H)], 0)
    p = np.linalg.inv(V) @ scr
    return p[:3].T


def _hit_quads(o, d, quads):
    """Nearest hit distance (in units of d) of rays o + t d against convex quads, inf if none."""
    best = np.full(len(d), np.inf)
    for q in quads:
        n = np.cross(q[1] - q[0], q[2] - q[1])
        denom = d @ n
        with np.errstate(divide="ignore", invalid="ignore"):
            t = ((q[0] - o) @ n) / denom
        ok = np.isfinite(t) & (t > 1e-6)
        pts = o + d * t[:, None]
        for k in range(4):
            e = q[(k + 1) % 4] - q[k]
            ok &= (np.cross(e, pts - q[k]) @ n) * np.sign(n @ n) >= -1e-6 * np.linalg.norm(n)
        best = np.where(ok & (t < best), t, best)
    return best


def _hit_box(o, d, M, mn, mx):
    Mi = np.linalg.inv(M)
    o2 = (Mi[:3, :3] @ o) + Mi[:3, 3]
    d2 = d @ Mi[:3, :3].T
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = (mn - o2) / d2
        t2 = (mx - o2) / d2
    tmin = np.nanmax(np.minimum(t1, t2), axis=1)
    tmax = np.nanmin(np.maximum(t1, t2), axis=1)
    return np.where((tmax >= tmin) & (tmax > 0), np.maximum(tmin, 0), np.inf)


def test_model_pixels_match_ray_casts(scene):
    m, path = scene
    _, _, _, data = parse(path)
    boxes = _boxes(data)
    world = []
    for l in data[1]:
        for p in data[2][int(l["first_polygon"]):int(l["first_polygon"]) + int(l["num_polygons"])]:
            world.append(data[5][int(p["first_vertex"]):int(p["first_vertex"]) + int(p["num_vertices"])].astype(np.float64))
    fb = sw_host.FrameBuilder()
    total = bad = 0
    for i, pose in enumerate(m.camera_path(11, 40)):
        t = 0.53 * i
        f = fb.build(m, pose, width=W, height=H, use_directional_lightmaps=1, time=t)
        first = f["first_submodel_surface"]
        if first == len(f["surfaces"]):
            continue
        lo = int(f["surfaces"][first]["pixels_offset"]) if first < len(f["surfaces"]) else 1 << 62
        full, _, _, _ = oracle.render(m.data, sw_host.batch_from_frames([f]))
        g = dict(f)
        dr = f["draws"]
        model_draw = (dr["kind"] == A.DRAW_POLYGON) & (dr["texture_source"] == A.TEXSRC_SURFACE) & (dr["texture_offset"] >= lo)
        g["draws"] = dr[~model_draw]  # every draw of a model surface removed
        rest, _, _, _ = oracle.render(m.data, sw_host.batch_from_frames([g]))
        shown = (full != rest).reshape(H, W)
        V, o = _camera(pose)
        d = _rays(V) - o
        tw = _hit_quads(o, d, world)
        tm = np.full(W * H, np.inf)
        for k, (mn, mx, _) in enumerate(boxes):
            tm = np.minimum(tm, _hit_box(o, d, _submodel_matrix(k, t, (mn + mx) / 2), mn, mx))
        first_hit = (tm < tw).reshape(H, W)
        # compare away from silhouettes: pixels whose 3x3 neighbourhood agrees in the ray cast
        pad = np.pad(first_hit, 1, mode="edge")
        nb = [pad[1 + dy:1 + dy + H, 1 + dx:1 + dx + W] for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
        interior = np.all([x == first_hit for x in nb], axis=0)
        total += int(np.count_nonzero(first_hit & interior))
        bad += int(np.count_nonzero((shown != first_hit) & interior))
    assert total > 2000, total
    # a model texel equal to the wall texel behind it hides in `shown`; nothing else may differ
    assert bad <= 0.01 * total, (bad, total)


def test_view_model_is_drawn_last_over_the_viewport(scene):
    """draw_view_models (partial_renderer.rs:2284-2309): the view model's triangles follow the whole
    tree in the draw list, so every pixel they cover keeps their colour in the final image; with
    N virtual threads each rect ends with its own run of them, clipped to the rect."""
    m, _ = scene
    fb = sw_host.FrameBuilder()
    pose = m.camera_path(4, 3)[2]
    for n in (1, 4):
        f = fb.build(m, pose, width=W, height=H, use_directional_lightmaps=1, time=0.4, view_model=1, virtual_threads=n)
        g = fb.build(m, pose, width=W, height=H, use_directional_lightmaps=1, time=0.4, view_model=0, virtual_threads=n)
        tri = f["draws"]["kind"] == A.DRAW_TRIANGLE
        assert tri.sum() > 100 and np.count_nonzero(g["draws"]["kind"] == A.DRAW_TRIANGLE) == 0
        # the triangles come last in each rect's run of draws
        runs = {}
        for i, r in enumerate(f["draws"]["clip_rect"]):
            runs.setdefault(tuple(int(x) for x in r), []).append(i)
        for idx in runs.values():
            kinds = f["draws"]["kind"][idx]
            first_tri = int(np.argmax(kinds == A.DRAW_TRIANGLE)) if np.any(kinds == A.DRAW_TRIANGLE) else len(kinds)
            assert np.all(kinds[first_tri:] == A.DRAW_TRIANGLE)
        full, _, _, _ = oracle.render(m.data, sw_host.batch_from_frames([f]))
        h = dict(f)
        h["draws"] = f["draws"][tri]
        h["view"] = f["view"].copy()
        alone, _, _, _ = oracle.render(m.data, sw_host.batch_from_frames([h]))
        cover = alone != int(f["view"]["background_color"][0])
        assert cover.sum() > 500
        assert np.array_equal(full[cover], alone[cover])
        # held at the lower right: its pixels lie in the right half and below the top quarter
        ys, xs = np.nonzero(cover.reshape(H, W))
        assert xs.mean() > W * 0.55 and ys.mean() > H * 0.5
