"""Pins the oracle's hot-path arithmetic against float64 models of what each step computes.

These cover what exact-coverage and constant-depth tests leave open:
  - perspective spans with varying 1/z (rasterizer.rs:273-454): each texel is the floor of
    tc_eq / inv_z at the pixel centre, up to the fixed-point quantisation;
  - line-z spans (rasterizer.rs:456-581): exact texture coordinates at the span's first and last
    pixel, linear between them;
  - textured Gouraud triangles (rasterizer.rs:731-1292): texture coordinates and light are the
    barycentric interpolation at the pixel centre; vertices slightly outside the texture are pulled
    back by the clamp corrections and never fetch outside it;
  - the surface builder with a normal map, dielectric and metal specular and one to six dynamic
    lights (surfaces.rs:433-852, 1088-1154): every channel within half an LSB of a float64 shading
    model before rounding.
Each model is written from the geometry (pixel centres, barycentric weights, reflection vectors),
not from the oracle's integer steps, so a dropped term, a swapped operand or an off-by-half centre
fails it.
"""
import numpy as np
import pytest

import oracle
import scenes
import sw_abi as A
from test_oracle import _fx, _raster, _raster_tris, sorted_poly

oracle.load()


# ---------------------------------------------------------------------------------------------
# Polygon spans with varying depth
# ---------------------------------------------------------------------------------------------
def _index_texture(tw, th):
    """Texel value = 1 + index, so a framebuffer value names the texel it came from (0 = uncovered)."""
    return (np.arange(tw * th, dtype=np.uint32) + 1).reshape(th, tw)


def _random_perspective_plane(rng, W, H, tw, th, inv_z_range, margin=1.0, dx_rel=None):
    """Depth and texture equations (float32, the draw's layout) of a plane seen in perspective:
    1/z and tc/z linear in screen space. The values at three screen corners are drawn at random
    and the fourth corner is checked: a linear-fractional function takes its extremes over a
    convex region at the region's vertices, so tc stays inside [margin, size - margin] on the
    whole screen. With `dx_rel` the 1/z of the top-right corner differs from the top-left one's by
    at most that fraction (0: d_inv_z_dx = 0 exactly)."""
    corners = np.array([[0.0, 0.0], [W, 0.0], [0.0, H], [W, H]])
    lo, hi = inv_z_range
    while True:
        iz = rng.uniform(lo, hi, size=3)
        if dx_rel is not None:
            iz[1] = iz[0] * (1.0 + dx_rel * rng.uniform(-1.0, 1.0))
        tc = np.stack([rng.uniform(margin, tw - margin, size=3), rng.uniform(margin, th - margin, size=3)], 1)
        # solve a*x + b*y + c = value at corners 0, 1, 2
        M = np.array([[x, y, 1.0] for x, y in corners[:3]])
        de = np.linalg.solve(M, iz)
        tu = np.linalg.solve(M, tc[:, 0] * iz)
        tv = np.linalg.solve(M, tc[:, 1] * iz)
        de32 = de.astype(np.float32)
        te32 = np.array([tu[0], tv[0], tu[1], tv[1], tu[2], tv[2]], np.float32)
        iz4 = corners @ de32[:2].astype(np.float64) + np.float64(de32[2])
        if iz4.min() < lo * 0.5 or iz4.max() > hi * 1.5:
            continue
        u4 = (corners @ te32[[0, 2]].astype(np.float64) + np.float64(te32[4])) / iz4
        v4 = (corners @ te32[[1, 3]].astype(np.float64) + np.float64(te32[5])) / iz4
        if u4.min() < margin or u4.max() > tw - margin or v4.min() < margin or v4.max() > th - margin:
            continue
        return de32, te32


def _exact_tc(de, te, X, Y):
    de = de.astype(np.float64)
    te = te.astype(np.float64)
    iz = de[0] * X + de[1] * Y + de[2]
    return (te[0] * X + te[2] * Y + te[4]) / iz, (te[1] * X + te[3] * Y + te[5]) / iz


def _texel_coords(fb, tw):
    idx = fb.astype(np.int64) - 1
    return idx % tw, idx // tw


def _within_floor(got, t, delta):
    """got is floor(t') for some t' within delta of t."""
    return (got >= np.floor(t - delta)) & (got <= np.floor(t + delta))


@pytest.mark.parametrize("inv_z_range", [(2.0 ** -10, 2.0 ** -5), (0.01, 0.3)])
def test_perspective_spans_follow_projective_texture_coordinates(inv_z_range):
    """Random planes with nonzero d_inv_z_dx and d_inv_z_dy (the realistic z-scaled range, where
    1/z <= 2^-5, and a wide one inside the span clamps: 1/z below 1 = 2^29 / INV_Z_SCALE and
    tc * (1/z) below 2^5 = 2^29 / TC_SCALE, rasterizer.rs:344-366, 369-422): every covered pixel fetches floor(tc_eq / inv_z) at its centre
    (x + 1/2, y + 1/2), up to 0.03 texel of fixed-point error (z = 2^31 / (inv_z >> 8) keeps ~15
    significant bits; INV_Z_SCALE = 2^29, TC_SCALE = 2^24, rasterizer.rs:1573-1580)."""
    rng = np.random.default_rng(101)
    W, H = 96, 72
    tw, th = 64, 48
    tex = _index_texture(tw, th)
    worst = 0.0
    exact = total = 0
    for _ in range(40):
        de, te = _random_perspective_plane(rng, W, H, tw, th, inv_z_range)
        assert de[0] != 0 and de[1] != 0
        pts = sorted_poly(rng, W, H)
        fb, _, oob = _raster([(_fx(pts), de, te, 0, A.BLEND_NONE, A.TC_PERSPECTIVE)], W, H, [tex])
        assert oob == 0
        ys, xs = np.nonzero(fb)
        assert ys.size > 0
        gu, gv = _texel_coords(fb[ys, xs], tw)
        tu, tv = _exact_tc(de, te, xs + 0.5, ys + 0.5)
        assert np.all(_within_floor(gu, tu, 0.03)) and np.all(_within_floor(gv, tv, 0.03))
        worst = max(worst, np.max(np.abs(gu + 0.5 - tu)), np.max(np.abs(gv + 0.5 - tv)))
        exact += int(np.count_nonzero((gu == np.floor(tu)) & (gv == np.floor(tv))))
        total += ys.size
    assert worst < 0.53
    assert exact / total > 0.97


def test_perspective_spans_with_a_half_pixel_centre_error_would_fail():
    """The model above is sharp enough to see a half-pixel shift of the sampling point: evaluated
    at the pixel corner instead of its centre, a few percent of the texels disagree (the check
    above requires all of them to agree)."""
    rng = np.random.default_rng(102)
    W, H, tw, th = 96, 72, 64, 48
    tex = _index_texture(tw, th)
    de, te = _random_perspective_plane(rng, W, H, tw, th, (0.01, 0.3))
    fb, _, _ = _raster([(_fx([(2.0, 2.0), (90.0, 3.0), (85.0, 70.0), (4.0, 66.0)]), de, te, 0, A.BLEND_NONE,
                         A.TC_PERSPECTIVE)], W, H, [tex])
    ys, xs = np.nonzero(fb)
    gu, gv = _texel_coords(fb[ys, xs], tw)
    tu, tv = _exact_tc(de, te, xs + 0.0, ys + 0.0)
    ok = _within_floor(gu, tu, 0.03) & _within_floor(gv, tv, 0.03)
    assert np.mean(ok) < 0.98


def test_line_z_spans_interpolate_between_exact_span_ends():
    """Line-z mode (rasterizer.rs:456-581): the texture coordinates of the span's first and last
    pixel are exact (z from the clamped 1/z at those centres), the pixels between are the linear
    interpolation of the two, in 16.16 fixed point. Planes whose 1/z varies mostly down the screen
    (floors, ceilings) are what the mode is chosen for; with d_inv_z_dx = 0 it is exact."""
    rng = np.random.default_rng(103)
    W, H = 96, 72
    tw, th = 64, 48
    tex = _index_texture(tw, th)
    exact = total = 0
    for trial in range(40):
        de, te = _random_perspective_plane(rng, W, H, tw, th, (2.0 ** -10, 2.0 ** -5),
                                           dx_rel=0.0 if trial % 2 == 0 else 0.03)
        assert (de[0] == 0) == (trial % 2 == 0)
        pts = sorted_poly(rng, W, H)
        fb, _, oob = _raster([(_fx(pts), de, te, 0, A.BLEND_NONE, A.TC_LINE_Z)], W, H, [tex])
        assert oob == 0
        for y in np.unique(np.nonzero(fb)[0]):
            xs = np.nonzero(fb[y])[0]
            x0, x1 = xs[0], xs[-1]
            assert xs.size == x1 - x0 + 1  # a convex polygon covers one run per row
            gu, gv = _texel_coords(fb[y, xs], tw)
            u0, v0 = _exact_tc(de, te, np.array([x0 + 0.5, x1 + 0.5]), np.array([y + 0.5, y + 0.5]))
            f = (xs - x0) / max(x1 - x0, 1)
            lu = u0[0] + (u0[1] - u0[0]) * f
            lv = v0[0] + (v0[1] - v0[0]) * f
            assert np.all(_within_floor(gu, lu, 0.03)), (trial, y)
            assert np.all(_within_floor(gv, lv, 0.03)), (trial, y)
            eu, ev = _exact_tc(de, te, xs + 0.5, np.full(xs.size, y + 0.5))
            if de[0] == 0:
                assert np.all(_within_floor(gu, eu, 0.03)) and np.all(_within_floor(gv, ev, 0.03))
            exact += int(np.count_nonzero((gu == np.floor(lu)) & (gv == np.floor(lv))))
            total += xs.size
    assert total > 10000
    assert exact / total > 0.97


# ---------------------------------------------------------------------------------------------
# Textured Gouraud triangles
# ---------------------------------------------------------------------------------------------
def _fixed(v):
    return np.float64(int(np.int32(np.float32(v) * np.float32(65536.0)))) / 65536.0


def _random_triangle(rng, W, H, min_extent=3.0):
    while True:
        p = rng.uniform([-4.0, -4.0], [W + 4.0, H + 4.0], size=(3, 2))
        area = (p[1, 0] - p[0, 0]) * (p[2, 1] - p[0, 1]) - (p[2, 0] - p[0, 0]) * (p[1, 1] - p[0, 1])
        ys = np.sort(p[:, 1])
        # thin parts take the reference's minimum-dy hack; keep every part at least 3 rows tall
        if abs(area) > 40.0 and ys[1] - ys[0] >= min_extent and ys[2] - ys[1] >= min_extent:
            return p


def _barycentric(P, X, Y):
    (x0, y0), (x1, y1), (x2, y2) = P
    d = (y1 - y2) * (x0 - x2) + (x2 - x1) * (y0 - y2)
    w0 = ((y1 - y2) * (X - x2) + (x2 - x1) * (Y - y2)) / d
    w1 = ((y2 - y0) * (X - x2) + (x0 - x2) * (Y - y2)) / d
    return w0, w1, 1.0 - w0 - w1


def test_triangle_texture_coordinates_are_barycentric():
    """Light 1.0 at every vertex passes the texel through unchanged ((c * 2^16) >> 16), so the
    framebuffer names the fetched texel: it is floor(tc) of the barycentric interpolation of the
    vertices' texture coordinates at the pixel centre, up to the fixed16 slopes' rounding."""
    rng = np.random.default_rng(104)
    W, H, tw, th = 80, 64, 32, 24
    lite = _index_texture(tw, th)
    exact = total = 0
    for _ in range(60):
        P = _random_triangle(rng, W, H)
        tc = np.stack([rng.uniform(0.25, tw - 0.25, 3), rng.uniform(0.25, th - 0.25, 3)], 1)
        tri = [(P[k, 0], P[k, 1], tc[k, 0], tc[k, 1], 1.0, 1.0, 1.0) for k in range(3)]
        fb, _, oob = _raster_tris([tri], W, H, lite)
        assert oob == 0
        ys, xs = np.nonzero(fb)
        Pq = np.array([[_fixed(P[k, 0]), _fixed(P[k, 1])] for k in range(3)])
        tq = np.array([[_fixed(tc[k, 0]), _fixed(tc[k, 1])] for k in range(3)])
        w0, w1, w2 = _barycentric(Pq, xs + 0.5, ys + 0.5)
        assert np.all(np.minimum(np.minimum(w0, w1), w2) > -1e-3)  # only pixels whose centre is inside
        gy, gx = np.mgrid[0:H, 0:W]
        a0, a1, a2 = _barycentric(Pq, gx + 0.5, gy + 0.5)
        assert np.all(fb[np.minimum(np.minimum(a0, a1), a2) > 1e-3] != 0)  # and all of those
        tu = w0 * tq[0, 0] + w1 * tq[1, 0] + w2 * tq[2, 0]
        tv = w0 * tq[0, 1] + w1 * tq[1, 1] + w2 * tq[2, 1]
        gu, gv = _texel_coords(fb[ys, xs], tw)
        assert np.all(_within_floor(gu, tu, 0.01)) and np.all(_within_floor(gv, tv, 0.01))
        exact += int(np.count_nonzero((gu == np.floor(tu)) & (gv == np.floor(tv))))
        total += ys.size
    assert exact / total > 0.98


def test_triangle_gouraud_light_is_barycentric():
    """A white texel (255 per channel) under per-vertex lights below 1.0: each channel is
    floor(255 * light) of the barycentric light at the pixel centre, light in fixed16. The light
    steps by fixed16 slopes (fixed16_div truncates: below 2^-16 per step), across at most W pixels
    and down at most H rows, so the light is off by at most (W + H) * 2^-16 before the * 255."""
    rng = np.random.default_rng(105)
    W, H = 80, 64
    lite = np.full((4, 4), 0x00FFFFFF, np.uint32)
    delta = 255.0 * (W + H) / 65536.0
    total = exact = 0
    for _ in range(60):
        P = _random_triangle(rng, W, H)
        L = rng.uniform(0.0, 0.999, size=(3, 3))  # [vertex, (r, g, b)]
        tri = [(P[k, 0], P[k, 1], 1.5, 1.5, L[k, 0], L[k, 1], L[k, 2]) for k in range(3)]
        fb, _, oob = _raster_tris([tri], W, H, lite)
        assert oob == 0
        # a covered pixel may be black: coverage comes from the geometry (centres clearly inside)
        Pq = np.array([[_fixed(P[k, 0]), _fixed(P[k, 1])] for k in range(3)])
        gy, gx = np.mgrid[0:H, 0:W]
        w0, w1, w2 = _barycentric(Pq, gx + 0.5, gy + 0.5)
        inside = np.minimum(np.minimum(w0, w1), w2) > 1e-3
        for lane, ch in ((0, 2), (1, 1), (2, 0)):
            lq = np.array([_fixed(L[k, ch]) for k in range(3)])
            light = w0 * lq[0] + w1 * lq[1] + w2 * lq[2]
            got = ((fb >> (8 * lane)) & 0xFF).astype(np.float64)[inside]
            model = 255.0 * light[inside]
            assert np.all(got >= np.floor(model - delta)) and np.all(got <= np.floor(model + delta)), lane
            exact += int(np.count_nonzero(got == np.floor(model)))
            total += got.size
        assert np.all((fb[inside] >> 24) == 0)  # triangle output alpha is 0
    assert exact / total > 0.9


def test_triangle_clamp_corrections_keep_fetches_inside():
    """Vertices up to half a texel outside the texture (rasterizer.rs:1018-1192 pulls the left
    edge value or the x step back): no fetch leaves the texture, and each texel stays within one
    texel plus the overshoot of the clamped barycentric coordinate."""
    rng = np.random.default_rng(106)
    W, H, tw, th = 80, 64, 16, 12
    lite = _index_texture(tw, th)
    for _ in range(80):
        P = _random_triangle(rng, W, H)
        over = 0.5
        tc = np.stack([rng.uniform(-over, tw + over, 3), rng.uniform(-over, th + over, 3)], 1)
        k, axis = int(rng.integers(0, 3)), int(rng.integers(0, 2))
        tc[k, axis] = -over if rng.random() < 0.5 else (tw, th)[axis] + over  # one vertex surely outside
        tri = [(P[k, 0], P[k, 1], tc[k, 0], tc[k, 1], 1.0, 1.0, 1.0) for k in range(3)]
        fb, _, oob = _raster_tris([tri], W, H, lite)
        assert oob == 0
        ys, xs = np.nonzero(fb)
        gu, gv = _texel_coords(fb[ys, xs], tw)
        assert gu.min() >= 0 and gu.max() < tw and gv.min() >= 0 and gv.max() < th
        Pq = np.array([[_fixed(P[k, 0]), _fixed(P[k, 1])] for k in range(3)])
        w0, w1, w2 = _barycentric(Pq, xs + 0.5, ys + 0.5)
        tu = np.clip(w0 * tc[0, 0] + w1 * tc[1, 0] + w2 * tc[2, 0], 0, tw - 1e-9)
        tv = np.clip(w0 * tc[0, 1] + w1 * tc[1, 1] + w2 * tc[2, 1], 0, th - 1e-9)
        assert np.all(np.abs(gu + 0.5 - tu) <= 1.0 + 2 * over)
        assert np.all(np.abs(gv + 0.5 - tv) <= 1.0 + 2 * over)


# ---------------------------------------------------------------------------------------------
# Surface builder: normal map, specular, dynamic lights (surfaces.rs:433-852)
# ---------------------------------------------------------------------------------------------
SPEC_NONE, SPEC_DIELECTRIC, SPEC_METAL = 0, 1, 2


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _schlick5(cos):
    return np.maximum(1.0 - cos, 0.0) ** 5


def _specular_k(spec, cos, rough):
    f = _schlick5(cos)
    if spec == SPEC_DIELECTRIC:
        fresnel = 0.04 + 0.96 * f  # Schlick with zero-angle reflectivity 0.04
        return fresnel * (1.0 - rough) + 0.12 * rough  # rough limit: 3x the zero-angle reflectivity
    return f * (1.0 - rough) + 0.5 * rough  # metal: rough limit 0.5


def _lobe(cos_rl, inv_rough):
    """The specular lobe 1 / (x^2 pi sqrt7 / 2 + 2 pi / sqrt7) * inv_rough with x = (cos - 1) *
    inv_rough, normalised so that its integral over (-inf, 0] is 1/2 (surfaces.rs:1088-1114)."""
    x = (cos_rl - 1.0) * inv_rough
    s7 = np.sqrt(7.0)
    return inv_rough / (x * x * np.pi * s7 / 2.0 + 2.0 * np.pi / s7)


def _surface_model(s, tex, tw, th, tex_flags, lm_kind, lm, lights, cam_pos, swap_mix=False):
    """float64 shading of every texel of surface `s` (one value per lane B, G, R, A, unrounded)."""
    w, h = (int(x) for x in s["size"])
    U = s["basis_u"].astype(np.float64)
    V = s["basis_v"].astype(np.float64)
    N = s["basis_normal"].astype(np.float64)
    S = s["basis_start"].astype(np.float64)
    t0 = s["tc_start"].astype(np.int64)
    du, dv = np.meshgrid(np.arange(w), np.arange(h))
    # world position of the texel centre
    pos = S + U * (t0[0] + 0.5 + du)[..., None] + V * (t0[1] + 0.5 + dv)[..., None]
    tu = (t0[0] + du) % tw
    tv = (t0[1] + dv) % th
    tx = tex.reshape(th, tw)[tv, tu]
    diffuse = np.stack([(tx["diffuse"].astype(np.int64) >> (8 * k)) & 0xFF for k in range(4)], -1).astype(np.float64)
    packed = tx["packed_normal_roughness"].astype(np.int64) & 0xFFFFFFFF
    sb = [((packed >> (8 * k)) & 0xFF).astype(np.int64) for k in range(4)]
    sb = [np.where(b >= 128, b - 256, b).astype(np.float64) / 127.0 for b in sb]
    has_nm = bool(tex_flags & A.TEX_HAS_NORMAL_MAP)
    if tex_flags & A.TEX_HAS_NON_ONE_ROUGHNESS:
        spec = SPEC_METAL if tex_flags & A.TEX_IS_METAL else SPEC_DIELECTRIC
    else:
        spec = SPEC_NONE
    n_tex = np.stack(sb[:3], -1) if has_nm else np.broadcast_to(np.array([0.0, 0.0, 1.0]), (h, w, 3))
    rough = sb[3]
    Un, Vn = _unit(U), _unit(V)

    albedo = np.zeros((h, w, 4))
    direct = np.zeros((h, w, 4))
    if lm_kind != "none":
        L2 = int(s["lightmap_scale_log2"])
        lw, lh = (int(x) for x in s["lightmap_size"])
        sh = s["lightmap_tc_shift"].astype(np.int64)
        # the lightmap sample of element j lies at lightmap coordinate j; the texel centre at
        # (texel + shift + 1/2) / 2^scale_log2, interpolated bilinearly
        fu = (du + sh[0] + 0.5) / (1 << L2)
        fv = (dv + sh[1] + 0.5) / (1 << L2)
        iu, iv = np.floor(fu).astype(np.int64), np.floor(fv).astype(np.int64)
        au, av = fu - iu, fv - iv
        if swap_mix:  # the plausible mistake: each weight on the other element
            au, av = 1.0 - au, 1.0 - av
        base = int(s["lightmap_offset"])

        def bilinear(arr):
            arr = arr.reshape(-1, *arr.shape[1:]).astype(np.float64)

            def at(i, j):
                return arr[base + j * lw + i]
            wu, wv = au[..., None], av[..., None]
            if arr.ndim == 1:
                wu, wv = au, av
            return (at(iu, iv) * (1 - wu) * (1 - wv) + at(iu + 1, iv) * wu * (1 - wv) + at(iu, iv + 1) * (1 - wu) * wv
                    + at(iu + 1, iv + 1) * wu * wv)
        if lm_kind == "simple":
            amb = bilinear(lm)
        else:
            amb = bilinear(lm["ambient_light"])
            dvec = bilinear(lm["light_direction_vector_scaled"])
            dev = bilinear(lm["directional_light_deviation"])
            dcol = bilinear(lm["directional_light_color"])
        amb_bgr = amb[..., ::-1]
        if spec == SPEC_NONE:
            albedo[..., :3] = amb_bgr
            albedo[..., 3] = 1.0
            if lm_kind == "directional":
                dot = np.maximum((dvec * n_tex).sum(-1), 0.0) if has_nm else dvec[..., 2]
                albedo[..., :3] += dcol[..., ::-1] * dot[..., None]
        else:
            t = cam_pos - pos  # view vector in the texture's basis
            tts = np.stack([(t * Un).sum(-1), (t * Vn).sum(-1), (t * N).sum(-1)], -1)
            tn = (tts * n_tex).sum(-1)
            r = n_tex * (2.0 * tn)[..., None] - tts
            rlen = np.sqrt(np.maximum((r * r).sum(-1), 2.0 ** -30))
            cos = np.maximum(tn / rlen, 0.0)
            k = _specular_k(spec, cos, rough)
            albedo[..., :3] = amb_bgr * (1.0 - k)[..., None]
            albedo[..., 3] = 1.0
            direct[..., :3] = amb_bgr * k[..., None]
            if lm_kind == "directional":
                dlen = np.sqrt((dvec * dvec).sum(-1))
                cos_rl = (r * dvec).sum(-1) / (rlen * dlen)
                inv_r = np.maximum(1.0 / (rough + dev), 0.75)
                I = _lobe(cos_rl, inv_r)
                c = dcol[..., ::-1]
                if spec == SPEC_DIELECTRIC:
                    diff = np.maximum((dvec * n_tex).sum(-1), 0.0)
                    albedo[..., :3] += c * (diff * (1.0 - k))[..., None]
                    direct[..., :3] += c * (I * k * dlen)[..., None]
                else:
                    albedo[..., :3] += c * ((1.0 - k) * I * dlen)[..., None]
                    direct[..., :3] += c * (k * I * dlen)[..., None]

    if lights is not None and len(lights):
        # normal in world space: the texture basis (u, v normalised, polygon normal) applied to the
        # texel normal
        nw = (n_tex[..., 0:1] * Un + n_tex[..., 1:2] * Vn + n_tex[..., 2:3] * N) if has_nm else np.broadcast_to(N, (h, w, 3))
        if spec != SPEC_NONE:
            t = cam_pos - pos
            tn = (t * nw).sum(-1)
            r = nw * (2.0 * tn)[..., None] - t
            rlen2 = np.maximum((r * r).sum(-1), 2.0 ** -30)
            cos = np.maximum(tn / np.sqrt(rlen2), 0.0)
            k = _specular_k(spec, cos, rough)
            inv_r = 1.0 / rough
        for lt in lights:
            Lv = lt["position"].astype(np.float64) - pos
            d2 = np.maximum((Lv * Lv).sum(-1), 2.0 ** -30)
            att = np.maximum(1.0 / d2 - np.float64(lt["inv_square_radius"]), 0.0)
            col = lt["color"].astype(np.float64)[::-1]
            if spec == SPEC_METAL:
                diff = np.zeros_like(d2)
            else:
                diff = np.maximum((nw * Lv).sum(-1) / np.sqrt(d2), 0.0)
            if spec == SPEC_NONE:
                albedo[..., :3] += col * (diff * att)[..., None]
                continue
            I = _lobe((r * Lv).sum(-1) / np.sqrt(rlen2 * d2), inv_r)
            if spec == SPEC_DIELECTRIC:
                albedo[..., :3] += col * (diff * (1.0 - k) * att)[..., None]
                direct[..., :3] += col * (I * k * att)[..., None]
            else:
                albedo[..., :3] += col * ((1.0 - k) * I * att)[..., None]
                direct[..., :3] += col * (k * I * att)[..., None]
    out = diffuse * albedo
    if spec != SPEC_NONE:
        out = out + 255.0 * direct
    return out.reshape(-1, 4)


def _surface_case(rng, spec, lm_kind, num_lights, normal_map):
    tw, th = 32, 16
    w, h = 23, 19
    n = tw * th
    tex = np.zeros(n, dtype=A.TEXEL_DTYPE)
    tex["diffuse"] = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    normals = rng.normal(size=(n, 3))
    normals[:, 2] = np.abs(normals[:, 2]) + 0.6
    normals = _unit(normals)
    rough = rng.uniform(0.1, 0.6, size=n)
    tex["packed_normal_roughness"] = scenes.pack_normal_roughness(normals if normal_map else np.tile([0.0, 0.0, 1.0], (n, 1)),
                                                                   rough)
    flags = (A.TEX_HAS_NORMAL_MAP if normal_map else 0)
    if spec != SPEC_NONE:
        flags |= A.TEX_HAS_NON_ONE_ROUGHNESS | (A.TEX_IS_METAL if spec == SPEC_METAL else 0)
    # a plane through `start` with a skewed, non-unit texture basis
    Uu = _unit(rng.normal(size=3))
    Vv = _unit(np.cross(Uu, rng.normal(size=3)))
    nrm = _unit(np.cross(Uu, Vv))
    U = (Uu * rng.uniform(0.7, 2.0) + Vv * rng.uniform(-0.3, 0.3)).astype(np.float32)
    V = (Vv * rng.uniform(0.7, 2.0)).astype(np.float32)
    start = rng.uniform(-50, 50, size=3).astype(np.float32)
    tc0 = rng.integers(-40, 40, size=2)
    centre = start.astype(np.float64) + U * (tc0[0] + w / 2) + V * (tc0[1] + h / 2)
    cam = (centre + nrm * rng.uniform(15, 60) + rng.normal(size=3) * 10).astype(np.float32)
    surf_kw = {"basis_u": U, "basis_v": V, "basis_normal": nrm.astype(np.float32), "basis_start": start,
               "tc_start": tc0, "size": (w, h), "flags": A.SURF_DIFFUSE_LAYER}
    lm_simple = lm_dir = None
    if lm_kind != "none":
        L2 = 3
        shift = rng.integers(0, 8, size=2)
        lw, lh = ((w - 1 + shift[0]) >> L2) + 2, ((h - 1 + shift[1]) >> L2) + 2
        surf_kw.update({"lightmap_offset": 1, "lightmap_size": (lw, lh), "lightmap_scale_log2": L2,
                        "lightmap_tc_shift": shift})
        cnt = 1 + lw * lh
        if lm_kind == "simple":
            lm_simple = rng.uniform(0.1, 0.9, size=(cnt, 3)).astype(np.float32)
        else:
            surf_kw["flags"] |= A.SURF_DIRECTIONAL
            lm_dir = np.zeros(cnt, dtype=A.DIRLM_DTYPE)
            lm_dir["ambient_light"] = rng.uniform(0.05, 0.5, size=(cnt, 3))
            d = rng.normal(size=(cnt, 3))
            d[:, 2] = np.abs(d[:, 2]) + 0.3
            lm_dir["light_direction_vector_scaled"] = _unit(d) * rng.uniform(0.2, 0.8, size=(cnt, 1))
            lm_dir["directional_light_deviation"] = rng.uniform(0.0, 0.5, size=cnt)
            lm_dir["directional_light_color"] = rng.uniform(0.1, 0.6, size=(cnt, 3))
    lights = None
    if num_lights:
        lights = np.zeros(num_lights, A.LIGHT_DTYPE)
        for i in range(num_lights):
            p = centre + nrm * rng.uniform(8, 20) + (U * rng.uniform(-w, w) + V * rng.uniform(-h, h)) * 0.4
            r = rng.uniform(30, 80)  # most texels inside the radius, some outside for r = 30
            lights[i]["position"] = p
            lights[i]["radius"] = r
            lights[i]["inv_square_radius"] = np.float32(1.0 / (r * r))
            lights[i]["color"] = rng.uniform(60, 260, size=3) / np.sqrt(num_lights)
        surf_kw["num_lights"] = num_lights
    return tex, tw, th, flags, surf_kw, lm_simple, lm_dir, lights, cam


def _build_one(tex, tw, th, flags, surf_kw, lm_simple, lm_dir, lights, cam, mode):
    levels = [(0, tw, th, flags, 0)]
    m = A.MapData(tex, np.zeros(1, np.uint32), lm_simple if lm_simple is not None else np.zeros((1, 3), np.float32),
                  lm_dir if lm_dir is not None else np.zeros(1, A.DIRLM_DTYPE), np.array(levels, dtype=A.LEVEL_DTYPE))
    s = np.zeros(1, dtype=A.SURFACE_DTYPE)
    s["emissive_level"] = A.NO_TEXTURE
    for k, v in surf_kw.items():
        s[k] = v
    v = np.zeros(1, dtype=A.VIEW_DTYPE)
    v["width"], v["height"], v["num_surfaces"] = 4, 4, 1
    v["cam_pos"] = cam
    ls = lights if lights is not None else np.zeros(0, A.LIGHT_DTYPE)
    b = A.BatchData(v, s, ls, np.arange(len(ls), dtype=np.uint32), np.zeros(0, A.DRAW_DTYPE), np.zeros(0, np.int32), 32)
    arena, _ = oracle.build_surfaces(m, b, mode=mode)
    return arena, s[0]


CASES = [(spec, lm, nl, nm) for spec in (SPEC_NONE, SPEC_DIELECTRIC, SPEC_METAL) for lm in ("none", "simple", "directional")
         for nl, nm in ((0, 1), (1, 1), (3, 0), (6, 1)) if not (lm == "none" and nl == 0)]


@pytest.mark.parametrize("spec,lm_kind,num_lights,normal_map", CASES)
def test_surface_texels_match_float64_shading(spec, lm_kind, num_lights, normal_map):
    """Every channel of every texel within half an LSB (plus float32 noise) of the float64 model
    before rounding, i.e. the oracle rounds the model's value; the x86 mode (rsqrtss / rcpss,
    relative error up to 1.5 * 2^-12) within one LSB."""
    rng = np.random.default_rng(1000 + 97 * spec + 13 * num_lights + 7 * normal_map + len(lm_kind))
    case = _surface_case(rng, spec, lm_kind, num_lights, normal_map)
    tex, tw, th, flags, surf_kw, lm_simple, lm_dir, lights, cam = case
    arena, s = _build_one(*case, mode=oracle.MODE_EXACT)
    lm = lm_simple if lm_kind == "simple" else lm_dir
    model = _surface_model(s, tex, tw, th, flags, lm_kind, lm, lights, cam.astype(np.float64))
    n = int(s["size"][0]) * int(s["size"][1])
    got = np.stack([(arena[:n].astype(np.int64) >> (8 * k)) & 0xFF for k in range(4)], -1).astype(np.float64)
    clamped = np.minimum(model, 255.0)
    assert model.min() > -0.5  # positive light only: no negative wrap in this scene
    err = np.abs(got - clamped)
    assert err.max() <= 0.5 + 0.005, (err.max(), np.unravel_index(np.argmax(err), err.shape))
    # the scene exercises the range, not just black or saturated texels
    assert np.mean((clamped[..., :3] > 8) & (clamped[..., :3] < 247)) > 0.2
    sse, _ = _build_one(*case, mode=oracle.MODE_SSE)
    got2 = np.stack([(sse[:n].astype(np.int64) >> (8 * k)) & 0xFF for k in range(4)], -1).astype(np.float64)
    assert np.abs(got2 - clamped).max() <= 1.0


def test_surface_model_detects_a_swapped_lightmap_mix():
    """The model is sharp enough to see the lightmap mix's operands exchanged (the weight k on the
    first element instead of the second): the oracle then disagrees with it by many LSBs."""
    rng = np.random.default_rng(77)
    case = _surface_case(rng, SPEC_DIELECTRIC, "directional", 2, 1)
    tex, tw, th, flags, surf_kw, lm_simple, lm_dir, lights, cam = case
    arena, s = _build_one(*case, mode=oracle.MODE_EXACT)
    model = _surface_model(s, tex, tw, th, flags, "directional", lm_dir, lights, cam.astype(np.float64), swap_mix=True)
    n = int(s["size"][0]) * int(s["size"][1])
    got = np.stack([(arena[:n].astype(np.int64) >> (8 * k)) & 0xFF for k in range(4)], -1).astype(np.float64)
    assert np.abs(got - np.minimum(model, 255.0)).max() > 3
