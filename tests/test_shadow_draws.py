"""Shadow-map draw lists of the front-end (swh_shadow_add_cube: PartialRenderer::build_shadow_maps'
cube branch, partial_renderer.rs:541-566, with calculate_cube_shadow_map_side_matrices, light.rs:73-128,
and DepthRenderer::draw_map, depth_renderer.rs:47-287), drawn by the oracle's depth rasteriser and
checked against the map's geometry by ray casting: each side's centre texel holds the biased, scaled
1 / z of the wall its axis meets, a closed map leaves no texel uncovered, and the shadow lookup
(light.rs:185-233) sees points short of the first wall as lit and points beyond it as shadowed."""
import os

import numpy as np
import pytest

import oracle
import sw_abi as A
import sw_host
from test_sqwm import parse

AXES = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)], np.float64)


class Geometry:
    """The map's polygons from its .sqwm file, for ray casting."""

    def __init__(self, path):
        _, _, _, d = parse(path)
        polys, verts = d[2], d[5].astype(np.float64)
        self.normal = np.array([p["plane"]["vec"] for p in polys], np.float64)
        self.dist = np.array([p["plane"]["dist"] for p in polys], np.float64)
        nmax = int(polys["num_vertices"].max())
        pv = np.zeros((len(polys), nmax, 3))
        for i, p in enumerate(polys):
            v = verts[p["first_vertex"]:p["first_vertex"] + p["num_vertices"]]
            pv[i, :len(v)] = v
            pv[i, len(v):] = v[-1]  # degenerate closing edges
        self.v0 = pv
        self.v1 = np.roll(pv, -1, axis=1)
        for i, p in enumerate(polys):  # the last real edge closes back to vertex 0
            self.v1[i, p["num_vertices"] - 1] = pv[i, 0]

    def first_hit(self, o, d):
        """Nearest t > 0 where o + t d meets a polygon facing the ray's origin."""
        denom = self.normal @ d
        t = (self.dist - self.normal @ o) / np.where(denom == 0, np.nan, denom)
        ok = (t > 1e-6) & (denom < 0)
        p = o + t[:, None] * d
        e = self.v1 - self.v0
        c = np.cross(e, p[:, None, :] - self.v0)
        side = np.einsum("pkc,pc->pk", c, self.normal)
        scale = np.linalg.norm(self.normal, axis=1)[:, None] * np.maximum(np.linalg.norm(e, axis=2), 1e-9)
        r = side / scale  # either winding: all edges on one side
        inside = np.all(r >= -1e-4, axis=1) | np.all(r <= 1e-4, axis=1)
        cand = np.where(ok & inside, t, np.inf)
        return float(np.min(cand))


@pytest.fixture(scope="module")
def grid(tmp_path_factory):
    d = tmp_path_factory.mktemp("shadowmap")
    m = sw_host.Map(seed=11, grid=(10, 10, 4), solid_probability=0.3, num_lights=4, sun=0, sky_fraction=0.0,
                    texture_size=32, num_materials=4, num_decal_textures=1, num_model_textures=1)
    path = os.path.join(d, "grid.sqwm")
    m.save_sqwm(path)
    lights = [np.array(p[:3], np.float64) for p in m.camera_path(5, 4)]
    return m, Geometry(path), lights


def _maps(m, lights, size):
    sb = sw_host.ShadowBuilder()
    for lp in lights:
        sb.add_cube(m, lp, size)
    targets, draws, vdata, n = sb.arrays()
    sb.close()
    assert len(targets) == 6 * len(lights) and n == 6 * size * size * len(lights)
    assert np.array_equal(targets["offset"], np.arange(len(targets)) * size * size)
    tex = oracle.render_depth(targets, draws, vdata, np.zeros(n, np.float32))
    return targets, draws, tex


def test_side_centres_hold_the_axis_walls(grid):
    """The texel at the centre of side s (axis a_s) holds (1 - 1/1024) (1/D - 2^-20): D the distance along
    a_s to the first wall (the grid's walls are axis-aligned, so that wall faces the side and its depth
    equation has no slope terms)."""
    m, geo, lights = grid
    size = 32
    _, _, tex = _maps(m, lights, size)
    for li, lp in enumerate(lights):
        for s in range(6):
            dist = geo.first_hit(lp, AXES[s])
            assert np.isfinite(dist)
            want = (1.0 - 1.0 / 1024.0) * (1.0 / dist - 2.0 ** -20)
            side = tex[(6 * li + s) * size * size:(6 * li + s + 1) * size * size].reshape(size, size)
            got = side[size // 2, size // 2]
            assert abs(got - want) <= 2e-6 * want, (li, s, got, want)


def test_closed_map_covers_every_texel(grid):
    m, _, lights = grid
    _, draws, tex = _maps(m, lights, 16)
    assert np.all(tex > 0.0)
    assert np.all(draws["depth_test"] == 0)  # static geometry only: back to front, no depth test


def test_lookup_agrees_with_ray_casts(grid):
    """Random directions from each light: a point at 0.95 of the first hit is lit, one at 1.05 is in
    shadow (the map's side choice and texel addressing against the builder's side cameras)."""
    m, geo, lights = grid
    size = 128
    _, _, tex = _maps(m, lights, size)
    rng = np.random.default_rng(2)
    bad = flipped_bad = total = 0
    for li, lp in enumerate(lights):
        sh = np.zeros(1, dtype=A.LIGHT_SHADOW_DTYPE)
        sh[0]["kind"], sh[0]["size"] = A.SHADOW_CUBE, size
        t = tex[6 * li * size * size:6 * (li + 1) * size * size]
        for _ in range(60):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            hit = geo.first_hit(lp, d)
            if not np.isfinite(hit) or hit < 8.0:
                continue
            total += 1
            near = lp + 0.95 * hit * d
            far = lp + 1.05 * hit * d
            lit = oracle.shadow_factor(sh, t, tuple(lp - near))
            dark = oracle.shadow_factor(sh, t, tuple(lp - far))
            bad += int(lit != 1.0 or dark != 0.0)
            # the same points looked up through the opposite side: the test tells the two apart
            flipped_bad += int(oracle.shadow_factor(sh, t, tuple(near - lp)) != 1.0 or
                               oracle.shadow_factor(sh, t, tuple(far - lp)) != 0.0)
    assert total >= 150
    # rays grazing a wall: the slope bias (a texel's depth change) can exceed 5% of the depth; rays
    # passing a corner within a texel of the hit may see the neighbour's depth
    assert bad <= total // 20, (bad, total)
    assert flipped_bad >= total // 4, (flipped_bad, total)


def test_bad_arguments():
    sb = sw_host.ShadowBuilder()
    m = sw_host.Map(seed=1, grid=(10, 10, 4), texture_size=32, num_materials=2)
    with pytest.raises(ValueError):
        sb.add_cube(m, (0, 0, 0), 0)
    with pytest.raises(ValueError):
        sb.add_cube(m, (0, 0, 0), 70000)
    sb.close()


@pytest.mark.parametrize("size", [64, 33, 1])
def test_projector_circle(size):
    """make_shadow_map_circle (light.rs:130-154) on an undrawn target: 1e24 outside, the texels inside
    keep their values; the centre row is whole, the first row keeps at most its middle texel, each
    row's kept texels are one run centred on the square, and the kept area is the disc's."""
    tg = np.zeros(1, dtype=A.DEPTH_TARGET_DTYPE)
    tg[0]["width"], tg[0]["height"], tg[0]["flags"] = size, size, A.DEPTH_CIRCLE
    init = np.full(size * size, 0.5, np.float32)
    out = oracle.render_depth(tg, np.zeros(0, A.DEPTH_DRAW_DTYPE), np.zeros(2, np.int32), init).reshape(size, size)
    assert set(np.unique(out).tolist()) <= {0.5, np.float32(1e24)}
    kept = out == 0.5
    r = size // 2
    assert kept[r].all()
    assert kept[0].sum() == size % 2
    for y in range(size):
        xs = np.flatnonzero(kept[y])
        if len(xs):
            assert xs[-1] - xs[0] + 1 == len(xs) and xs[0] + xs[-1] == size - 1
    if size > 8:
        assert abs(kept.sum() - np.pi * r * r) < 4 * size


def test_projector_maps_agree_with_ray_casts(grid):
    """A projector looking along +x (identity rotation, 90 degrees): the centre texel holds the biased,
    scaled 1 / D of the wall ahead, and inside the circle the map equals the +x cube side's; through the
    returned basis, points short of the first wall along rays inside the beam are lit and points beyond
    it shadowed, while the mirrored basis fails far more often. The misses of the correct basis (about
    6%, the same rays as with the +x cube side) are points beyond walls met at grazing incidence, where
    the reference's slope bias (a whole texel's depth change) exceeds 5% of the depth."""
    m, geo, lights = grid
    size = 128
    bad = flipped_bad = total = 0
    for li, lp in enumerate(lights):
        sb = sw_host.ShadowBuilder()
        basis = sb.add_projector(m, lp, (1.0, 0.0, 0.0, 0.0), np.pi / 2, size)
        sb.add_cube(m, lp, size)
        targets, draws, vdata, n = sb.arrays()
        sb.close()
        assert len(targets) == 7 and targets[0]["flags"] == A.DEPTH_CIRCLE and n == 7 * size * size
        np.testing.assert_allclose(basis, [[0, 1, 0], [0, 0, 1], [-1, 0, 0]], atol=1e-6)
        tex = oracle.render_depth(targets, draws, vdata, np.zeros(n, np.float32))
        proj, side0 = tex[:size * size].reshape(size, size), tex[size * size:2 * size * size].reshape(size, size)
        inside = proj < 1e23
        assert inside.sum() > 0.75 * size * size and np.array_equal(proj[inside], side0[inside])
        dist = geo.first_hit(lp, np.array([1.0, 0.0, 0.0]))
        want = (1.0 - 1.0 / 1024.0) * (1.0 / dist - 2.0 ** -20)
        got = proj[size // 2, size // 2]
        assert abs(got - want) <= 2e-6 * want, (li, got, want)
        rng = np.random.default_rng(li)
        for flip in (False, True):
            sh = np.zeros(1, dtype=A.LIGHT_SHADOW_DTYPE)
            sh[0]["kind"], sh[0]["size"] = A.SHADOW_PROJECTOR, size
            sh[0]["basis_x"], sh[0]["basis_y"], sh[0]["basis_z"] = (-1 if flip else 1) * basis[0], (-1 if flip else 1) * basis[1], basis[2]
            for _ in range(100):
                d = np.array([1.0, *rng.uniform(-0.6, 0.6, 2)])  # inside the beam's circle (tan 45 deg = 1)
                d /= np.linalg.norm(d)
                hit = geo.first_hit(lp, d)
                if not np.isfinite(hit) or hit < 8.0:
                    continue
                miss = int(oracle.shadow_factor(sh, proj.reshape(-1), tuple(-0.95 * hit * d)) != 1.0 or
                           oracle.shadow_factor(sh, proj.reshape(-1), tuple(-1.05 * hit * d)) != 0.0)
                if flip:
                    flipped_bad += miss
                else:
                    bad += miss
                    total += 1
    assert total >= 200 and bad <= total // 8 and flipped_bad >= total // 5, (bad, flipped_bad, total)
