"""View portals: mirrors, camera views and parallax portals (partial_renderer.rs:1834-2282, 3720-3776).

The front-end places each portal quad into the leaves it reaches, gives every visible portal a
texture rect at the mip its projection asks for, renders the view through it into that texture one
level down (portals_depth levels; below the last the textures are black) and draws the portal as a
leaf object sampling it. The checks are geometric, in float64, against what each view must show:
  - the texel a portal pixel of the main view samples is the portal's texture coordinate of the
    point where the pixel's camera ray meets the portal plane, over 2^mip, minus the rect's corner;
  - every pixel of a view seen through a portal shows the map polygon that the view's ray meets
    first: a mirror's rays leave the camera reflected in the mirror plane through the mirror point
    of the pixel, a camera portal's rays leave its camera, a parallax portal's rays leave the
    transformed camera through the transformed portal point; the polygon drawn last over a pixel
    lies in the plane of the polygon the ray meets first;
  - the texture rects of a frame lie end to end, each level's views are the visible portals of the
    level above in order, and below the last level the portal pixels are black.
"""
import math

import numpy as np
import pytest

import oracle
import sw_abi as A
import sw_host
from test_sqwm import parse

W, H = 640, 400


@pytest.fixture(scope="module")
def scene(tmp_path_factory):
    m = sw_host.Map(seed=3, grid=(8, 8, 4), solid_probability=0.3, num_lights=4, texture_size=64, num_materials=6,
                    normal_maps=0, specular_fraction=0.0, directional_lightmaps=0, sky_fraction=0.0, num_view_portals=12)
    path = str(tmp_path_factory.mktemp("portals") / "p.sqwm")
    m.save_sqwm(path)
    _, _, _, data = parse(path)
    polys, planes = [], []
    for p in data[2]:
        polys.append(data[5][int(p["first_vertex"]):int(p["first_vertex"]) + int(p["num_vertices"])].astype(np.float64))
        vec, dist = p["plane"]["vec"].astype(np.float64), float(p["plane"]["dist"])
        planes.append((vec / np.linalg.norm(vec), dist / np.linalg.norm(vec)))
    return m, (polys, planes), m.view_portals()


def _frames(m, depth=2, n=64, seed=1):
    fb = sw_host.FrameBuilder()
    out = []
    for pose in m.camera_path(seed, n):
        f = fb.build(m, pose, width=W, height=H, portals_depth=depth)
        if f["num_visible_portals"]:
            out.append((pose, f))
    return out


def _camera(pos, yaw, pitch, fov, w, h):
    """build_view_matrix_with_full_rotation (matrix.rs:13-109) in float64: world -> (x z, y z, z)."""
    def qmul(a, b):
        s1, v1 = a[0], np.array(a[1:])
        s2, v2 = b[0], np.array(b[1:])
        return np.concatenate([[s1 * s2 - v1 @ v2], s1 * v2 + s2 * v1 + np.cross(v1, v2)])
    az = yaw + math.pi / 2
    q = qmul(np.array([math.cos(az / 2), 0, 0, math.sin(az / 2)]), np.array([math.cos(-pitch / 2), 0, math.sin(-pitch / 2), 0]))
    w_, (a, b, c) = q[0], -q[1:]
    R = np.array([[1 - 2 * (b * b + c * c), 2 * (a * b - c * w_), 2 * (a * c + b * w_)],
                  [2 * (a * b + c * w_), 1 - 2 * (a * a + c * c), 2 * (b * c - a * w_)],
                  [2 * (a * c - b * w_), 2 * (b * c + a * w_), 1 - 2 * (a * a + b * b)]])
    B = np.array([[0, -1, 0], [0, 0, -1], [1, 0, 0]], np.float64)
    t = 1 / math.tan(fov / 2)
    P = np.diag([t / (w / h) * w / 2, t * h / 2, 1.0])
    P[0, 2], P[1, 2] = w / 2, h / 2
    V = np.eye(4)
    V[:3, :3] = P @ B @ R
    V[:3, 3] = -(P @ B @ R) @ np.asarray(pos, np.float64)
    return V


def _pixel_rays(V, w, h):
    """Points at depth 1 through the pixel centres of a w x h view (the ray directions from the eye)."""
    ys, xs = np.mgrid[0:h, 0:w]
    scr = np.stack([xs.ravel() + 0.5, ys.ravel() + 0.5, np.ones(w * h), np.ones(w * h)], 0)
    return (np.linalg.inv(V) @ scr)[:3].T


def _hit(o, d, poly, t_min):
    """Ray parameter of o + t d on a convex polygon (inf where the ray misses or t <= t_min)."""
    n = np.cross(poly[1] - poly[0], poly[2] - poly[1])
    nn = np.linalg.norm(n)
    if nn == 0:
        return np.full(len(d), np.inf)
    n = n / nn
    denom = d @ n
    with np.errstate(divide="ignore", invalid="ignore"):
        t = ((poly[0] - o) @ n) / denom
    ok = np.isfinite(t) & (t > t_min)
    pts = o + d * t[:, None]
    for k in range(len(poly)):
        e = poly[(k + 1) % len(poly)] - poly[k]
        ok &= np.cross(e, pts - poly[k]) @ n >= -1e-6 * np.linalg.norm(e)
    return np.where(ok, t, np.inf)


def _first_hits(o, d, polys, t_min):
    best = np.full(len(d), np.inf)
    idx = np.full(len(d), -1)
    for i, p in enumerate(polys):
        t = _hit(o, d, p, t_min)
        better = t < best
        best = np.where(better, t, best)
        idx = np.where(better, i, idx)
    return idx, best


def _last_draw(f, w, h):
    """Per pixel centre: the index of the last draw whose screen polygon covers it (-1: none)."""
    ys, xs = np.mgrid[0:h, 0:w]
    px, py = xs.ravel() + 0.5, ys.ravel() + 0.5
    out = np.full(w * h, -1)
    for di, d in enumerate(f["draws"]):
        v = f["vertex_data"][int(d["first_vertex_int"]):int(d["first_vertex_int"]) + 2 * int(d["num_vertices"])]
        v = v.reshape(-1, 2).astype(np.float64) / 65536.0
        side = np.ones(w * h, bool)
        for k in range(len(v)):
            a, b = v[k], v[(k + 1) % len(v)]
            side &= (b[0] - a[0]) * (py - a[1]) - (b[1] - a[1]) * (px - a[0]) >= 0
        out[side] = di
    return out


def _tc_point(portal, U, V):
    """The point of the portal plane with texture coordinates (U, V) (float64)."""
    e = portal["tex_coord_equation"].astype(np.float64)
    pl = portal["plane"].astype(np.float64)
    M = np.stack([e[0, :3], e[1, :3], pl[:3]])
    rhs = np.stack([U - e[0, 3], V - e[1, 3], np.full_like(U, pl[3])])
    return np.linalg.solve(M, rhs).T


def test_portal_rects_and_levels(scene):
    m, _, portals = scene
    frames = _frames(m)
    assert len(frames) >= 10
    deep = 0
    for _, f in frames:
        for level in sw_host.portal_levels(f):
            for g in level:
                off = 0
                for i, p in enumerate(g["portals"]):
                    if p["visible"]:
                        assert int(p["texture_offset"]) == off
                        off += int(p["resolution"][0]) * int(p["resolution"][1])
                    else:
                        assert p["resolution"][0] * p["resolution"][1] == 0
                assert off == g["portal_texels"]
                vis = [p for p in g["portals"] if p["visible"]]
                if g["portal_views"]:
                    assert len(g["portal_views"]) == len(vis)
                    for p, c in zip(vis, g["portal_views"]):
                        assert (int(c["view"]["width"][0]), int(c["view"]["height"][0])) == tuple(int(x) for x in p["resolution"])
                # every visible portal's rect covers the polygon's texture range at its mip
                for i, p in enumerate(g["portals"]):
                    if not p["visible"]:
                        continue
                    e = portals[i]["tex_coord_equation"].astype(np.float64)
                    vs = portals[i]["vertices"].astype(np.float64)
                    for a in range(2):
                        tc = vs @ e[a, :3] + e[a, 3]
                        s = 2.0 ** -int(p["mip"])
                        lo, hi = int(p["tc_min"][a]), int(p["tc_min"][a]) + int(p["resolution"][a])
                        assert lo <= math.floor(tc.min() * s) + 1 and hi >= math.ceil(tc.max() * s) - 1
        deep = max(deep, len(sw_host.portal_levels(f)))
    assert deep == 3
    # depth 1: one level of views, whose portal draws sample black textures
    _, f = frames[0]
    g = sw_host.FrameBuilder().build(m, frames[0][0], width=W, height=H, portals_depth=1)
    assert all(not c["portal_views"] for c in g["portal_views"])
    assert len(g["portal_views"]) == len(f["portal_views"])


def test_main_view_samples_the_portal_texture_of_the_ray_point(scene):
    m, _, portals = scene
    checked = 0
    for pose, f in _frames(m)[:12]:
        V = _camera(pose[:3], float(pose[3]), float(pose[4]), math.pi / 2, W, H)
        o = np.asarray(pose[:3], np.float64)
        d = _pixel_rays(V, W, H) - o
        last = _last_draw(f, W, H)
        for di, dr in enumerate(f["draws"]):
            if dr["texture_source"] != A.TEXSRC_PORTAL:
                continue
            pi = [i for i, p in enumerate(f["portals"]) if p["visible"] and int(p["texture_offset"]) == int(dr["texture_offset"])]
            assert len(pi) == 1
            p, portal = f["portals"][pi[0]], portals[pi[0]]
            sel = np.nonzero(last == di)[0]
            if len(sel) == 0:
                continue
            # the draw's equations at the pixel centres (their common z scale cancels)
            x = (sel % W) + 0.5
            y = (sel // W) + 0.5
            de = dr["depth_equation"].astype(np.float64)
            te = dr["tc_equation"].astype(np.float64)
            iz = de[0] * x + de[1] * y + de[2]
            u = (te[0] * x + te[2] * y + te[4]) / iz
            v = (te[1] * x + te[3] * y + te[5]) / iz
            # the ray's point on the portal plane and its texture coordinates
            n, dist = portal["plane"][:3].astype(np.float64), float(portal["plane"][3])
            t = (dist - o @ n) / (d[sel] @ n)
            X = o + d[sel] * t[:, None]
            e = portal["tex_coord_equation"].astype(np.float64)
            s = 2.0 ** -int(p["mip"])
            U = (X @ e[0, :3] + e[0, 3]) * s - int(p["tc_min"][0])
            Vv = (X @ e[1, :3] + e[1, 3]) * s - int(p["tc_min"][1])
            assert np.abs(u - U).max() < 0.02 and np.abs(v - Vv).max() < 0.02
            assert (U > -0.05).all() and (U < int(p["resolution"][0]) + 0.05).all()
            checked += len(sel)
    assert checked > 2000, checked


def _view_rays(kind, portal, info, main_pos):
    """Origin, per-pixel direction and the near limit of the rays of the view through a portal."""
    w, h = int(info["resolution"][0]), int(info["resolution"][1])
    if kind == sw_host.PORTAL_CAMERA:
        V = _camera(portal["camera_position"], float(portal["camera_yaw"]), float(portal["camera_pitch"]),
                    float(portal["camera_fov"]), w, h)
        o = portal["camera_position"].astype(np.float64)
        return o, _pixel_rays(V, w, h) - o, 1.0 - 1e-3
    ys, xs = np.mgrid[0:h, 0:w]
    s = 2.0 ** int(info["mip"])
    X = _tc_point(portal, (int(info["tc_min"][0]) + xs.ravel() + 0.5) * s, (int(info["tc_min"][1]) + ys.ravel() + 0.5) * s)
    c = np.asarray(main_pos, np.float64)
    pl = portal["plane"].astype(np.float64)
    if kind == sw_host.PORTAL_MIRROR:
        o = c - pl[:3] * (2.0 * (pl[:3] @ c - pl[3]) / (pl[:3] @ pl[:3]))
        return o, X - o, 1.0
    T = portal["transform"].astype(np.float64).reshape(4, 4).T
    o = T[:3, :3] @ c + T[:3, 3]
    Xt = X @ T[:3, :3].T + T[:3, 3]
    return o, Xt - o, 1.0


def test_portal_views_show_the_polygons_their_rays_meet(scene):
    m, (polys, planes), portals = scene
    quads = [portals[i]["vertices"].astype(np.float64) for i in range(len(portals))]
    counts = {sw_host.PORTAL_MIRROR: 0, sw_host.PORTAL_CAMERA: 0, sw_host.PORTAL_PARALLAX: 0}
    bads = dict(counts)
    for pose, f in _frames(m, depth=1, seed=6) + _frames(m, depth=1, seed=1, n=24):
        vis = [i for i, p in enumerate(f["portals"]) if p["visible"]]
        for i, g in zip(vis, f["portal_views"]):
            info, portal = f["portals"][i], portals[i]
            w, h = int(info["resolution"][0]), int(info["resolution"][1])
            if w * h < 200:
                continue
            kind = int(portal["kind"])
            o, d, t_min = _view_rays(kind, portal, info, pose[:3])
            idx, tw = _first_hits(o, d, polys, t_min)
            # portals seen inside the view cover the walls (a mirror's own quad lies on its near plane)
            others = [q for k, q in enumerate(quads) if k != i or kind != sw_host.PORTAL_MIRROR]
            _, tp = _first_hits(o, d, others, t_min)
            last = _last_draw(g, w, h)
            surf = {int(s["pixels_offset"]): s for s in g["surfaces"]}
            # compare away from edges: pixels whose 3x3 neighbourhood meets the same polygon
            img = idx.reshape(h, w)
            pad = np.pad(img, 1, mode="edge")
            interior = np.all([pad[1 + a:1 + a + h, 1 + b:1 + b + w] == img for a in (-1, 0, 1) for b in (-1, 0, 1)], axis=0)
            interior = interior.ravel() & (idx >= 0) & (tw < tp)
            bad = 0
            for k in np.nonzero(interior)[0]:
                di = last[k]
                if di < 0 or g["draws"][di]["texture_source"] != A.TEXSRC_SURFACE:
                    bad += 1
                    continue
                s = surf[int(g["draws"][di]["texture_offset"])]
                n1 = s["basis_normal"].astype(np.float64)
                d1 = n1 @ s["basis_start"].astype(np.float64)
                n2, d2 = planes[idx[k]]
                if np.abs(n1 - n2).max() > 1e-3 or abs(d1 - d2) > 0.05:
                    bad += 1
            counts[kind] += int(np.count_nonzero(interior))
            bads[kind] += bad
    assert min(counts.values()) > 1000, counts
    for kind in counts:
        assert bads[kind] <= 0.005 * counts[kind], (kind, bads, counts)


def test_depth_limit_leaves_portal_textures_black(scene):
    m, _, _ = scene
    pose, f = _frames(m, depth=0)[0]
    assert not f["portal_views"] and f["portal_texels"] > 0
    bs = sw_host.portal_level_batches(f)
    assert len(bs) == 1
    outs, oob = oracle.render_portal_levels(m.data, bs)
    assert oob == 0
    last = _last_draw(f, W, H)
    for di, dr in enumerate(f["draws"]):
        if dr["texture_source"] == A.TEXSRC_PORTAL and dr["blend"] == A.BLEND_NONE if hasattr(A, "BLEND_NONE") else 0:
            sel = last == di
            assert np.count_nonzero(sel) > 0
            assert (outs[0][sel] == 0).mean() > 0.98


def test_levels_feed_the_level_above(scene):
    """The main image's portal pixels change with the views below them, and only there."""
    m, _, _ = scene
    pose, f = [x for x in _frames(m) if len(sw_host.portal_levels(x[1])) == 3][0]
    bs = sw_host.portal_level_batches(f)
    outs, oob = oracle.render_portal_levels(m.data, bs)
    assert oob == 0
    alone, _, _, _ = oracle.render(m.data, bs[0], portal=np.zeros(max(bs[0].portal_texels, 1), bs[0].color_dtype))
    portal_px = np.zeros(W * H, bool)
    last = _last_draw(f, W, H)
    for di, dr in enumerate(f["draws"]):
        if dr["texture_source"] == A.TEXSRC_PORTAL:
            portal_px |= last == di
    diff = outs[0] != alone
    assert np.count_nonzero(diff) > 0.9 * np.count_nonzero(portal_px & (alone == 0))
    # differences stay on the portal draws (one pixel of rasterization edge allowed around them)
    pad = np.pad(portal_px.reshape(H, W), 1)
    near = np.any([pad[1 + a:1 + a + H, 1 + b:1 + b + W] for a in (-1, 0, 1) for b in (-1, 0, 1)], axis=0).ravel()
    assert not np.any(diff & ~near)
