"""Shadow-map lookups of the surface builder (light.rs:185-294, surfaces.rs:750-760), on the CPU: the
oracle's get_light_shadow_factor pinned by closed forms (which cube side a direction reads, where in
the side, the depth test, the projector's frustum), and its effect on built surfaces (all lit = no
shadow maps; all shadowed = the light's colour at zero). GPU parity is in test_gpu_shadows.py."""
import numpy as np
import pytest

import oracle
import scenes
import sw_abi as A


def _cube(size, side_values):
    sh = np.zeros(1, dtype=A.LIGHT_SHADOW_DTYPE)
    sh[0]["kind"], sh[0]["size"], sh[0]["offset"] = A.SHADOW_CUBE, size, 0
    tex = np.concatenate([np.full(size * size, v, np.float32) for v in side_values])
    return sh, tex


# vec (towards the light) -> the side it reads: a point on the light's -x side has vec.x > 0 -> side 1
SIDE_OF = [((5.0, 1.0, -2.0), 1), ((-5.0, 1.0, 2.0), 0), ((1.0, 6.0, 2.0), 3), ((1.0, -6.0, 2.0), 2),
           ((1.0, 2.0, 7.0), 5), ((1.0, 2.0, -7.0), 4)]


@pytest.mark.parametrize("vec,side", SIDE_OF)
def test_cube_side_selection_and_depth_test(vec, side):
    """Only the selected side is in shadow (its depth above 1 / |major component|): the point is lit
    for every other side's values; the test is depth >= stored (equality lit)."""
    major = max(abs(c) for c in vec)
    for shadowed in range(6):
        vals = [0.0] * 6
        vals[shadowed] = 1.0 / major + 0.01
        sh, tex = _cube(4, vals)
        assert oracle.shadow_factor(sh, tex, vec) == (0.0 if shadowed == side else 1.0)
    vals = [0.0] * 6
    vals[side] = np.float32(1.0) / np.float32(major)  # equal: lit
    sh, tex = _cube(4, vals)
    assert oracle.shadow_factor(sh, tex, vec) == 1.0


def test_cube_texel_addressing():
    """On side 5 (+z major) the texel is (u, v) = trunc(clamp(c / (2 z) + 0.5, 0, 1 - 2^-16) * size)
    with (c_x, c_y) = (-vec.x, vec.y): mark one texel of the side and find exactly the directions that
    land on it."""
    size = 8
    for u in range(size):
        for v in range(size):
            vals = np.zeros(6 * size * size, np.float32)
            vals[5 * size * size + v * size + u] = 1e9
            sh = np.zeros(1, dtype=A.LIGHT_SHADOW_DTYPE)
            sh[0]["kind"], sh[0]["size"] = A.SHADOW_CUBE, size
            # the centre of texel (u, v): c = ((u + 0.5) / size - 0.5) * 2 z
            z = 10.0
            cx, cy = ((u + 0.5) / size - 0.5) * 2 * z, ((v + 0.5) / size - 0.5) * 2 * z
            assert oracle.shadow_factor(sh, vals, (-cx, cy, z)) == 0.0
            # a neighbour texel's centre is lit
            cx2 = ((((u + 1) % size) + 0.5) / size - 0.5) * 2 * z
            assert oracle.shadow_factor(sh, vals, (-cx2, cy, z)) == 1.0


def test_projector():
    """Behind the projector (z <= 0) and outside its square: shadow; inside: the depth test on the
    texel (u, v) = trunc((dot(basis, vec) / (2 z) + 0.5) * size)."""
    size = 16
    sh = np.zeros(1, dtype=A.LIGHT_SHADOW_DTYPE)
    sh[0]["kind"], sh[0]["size"] = A.SHADOW_PROJECTOR, size
    sh[0]["basis_x"], sh[0]["basis_y"], sh[0]["basis_z"] = (1, 0, 0), (0, 1, 0), (0, 0, 1)
    tex = np.zeros(size * size, np.float32)
    assert oracle.shadow_factor(sh, tex, (0.0, 0.0, -3.0)) == 0.0
    assert oracle.shadow_factor(sh, tex, (0.0, 0.0, 0.0)) == 0.0
    assert oracle.shadow_factor(sh, tex, (30.0, 0.0, 10.0)) == 0.0  # u beyond the square
    assert oracle.shadow_factor(sh, tex, (1.0, -2.0, 10.0)) == 1.0
    u, v = int((1.0 / 20 + 0.5) * size), int((-2.0 / 20 + 0.5) * size)
    tex[v * size + u] = 0.2  # 1 / 10 < 0.2: shadowed
    assert oracle.shadow_factor(sh, tex, (1.0, -2.0, 10.0)) == 0.0
    tex[v * size + u] = 0.1  # equal: lit
    assert oracle.shadow_factor(sh, tex, (1.0, -2.0, 10.0)) == 1.0


@pytest.fixture
def no_shadows():
    yield
    oracle.set_light_shadows(None, None)


def test_surfaces_all_lit_and_all_shadowed(no_shadows):
    """All maps at depth 0: the surfaces equal those without shadow maps; all at +inf: equal to the
    lights with zero colour (a 0 shadow factor zeroes the light's term exactly)."""
    m, b = scenes.make_batch(5, num_lights=6, surfaces_per_view=30)
    n = len(b.lights)
    size = 4
    shadows = np.zeros(n, dtype=A.LIGHT_SHADOW_DTYPE)
    shadows["kind"], shadows["size"] = A.SHADOW_CUBE, size
    shadows["offset"] = np.arange(n) * 6 * size * size
    ref, _ = oracle.build_surfaces(m, b)
    oracle.set_light_shadows(shadows, np.zeros(n * 6 * size * size, np.float32))
    lit, _ = oracle.build_surfaces(m, b)
    assert np.array_equal(lit, ref)
    oracle.set_light_shadows(shadows, np.full(n * 6 * size * size, np.inf, np.float32))
    dark, _ = oracle.build_surfaces(m, b)
    oracle.set_light_shadows(None, None)
    b.lights["color"] = 0.0
    zero, _ = oracle.build_surfaces(m, b)
    assert np.array_equal(dark, zero)
    assert not np.array_equal(dark, ref)
