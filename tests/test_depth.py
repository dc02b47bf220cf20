"""The depth rasteriser of shadow maps (DepthRasterizer, rasterizer.rs:1295-1474) on the CPU: the
oracle's restatement pinned by exact coverage of rectangles, closed-form depths where every float
step is exact, the sequential float chains where they are not, the depth test and watertight shared
edges. GPU parity of sw_render_depth is in test_gpu_depth.py."""
import numpy as np
import pytest

import oracle
import sw_abi as A

F = 65536


def _rect(x0, y0, x1, y1):
    """Clockwise on screen (y down): (x0, y0), (x1, y0), (x1, y1), (x0, y1), in Fixed16."""
    return (np.array([x0, y0, x1, y0, x1, y1, x0, y1]) * F).astype(np.int64).astype(np.int32)


def _run(polys, eqs, w, h, depth_test=0, init=None):
    vd = np.concatenate(polys)
    draws = np.zeros(len(polys), dtype=A.DEPTH_DRAW_DTYPE)
    off = 0
    for i, p in enumerate(polys):
        draws[i]["first_vertex_int"], draws[i]["num_vertices"] = off, len(p) // 2
        draws[i]["depth_equation"] = eqs[i]
        draws[i]["depth_test"] = depth_test[i] if isinstance(depth_test, (list, tuple)) else depth_test
        off += len(p)
    tg = np.zeros(1, dtype=A.DEPTH_TARGET_DTYPE)
    tg[0]["width"], tg[0]["height"], tg[0]["num_draws"] = w, h, len(polys)
    tex = np.zeros(w * h, np.float32) if init is None else init
    return oracle.render_depth(tg, draws, vd, tex).reshape(h, w)


@pytest.mark.parametrize("r", [(2.3, 1.7, 9.6, 6.2), (0.5, 0.5, 3.5, 2.5), (-3.0, -2.0, 5.4, 4.49), (1.0, 1.0, 1.4, 9.0)])
def test_rectangle_coverage(r):
    """Rows round(y0) .. round(y1), columns floor(x0 + 1/2) .. floor(x1 + 1/2), clipped to the buffer."""
    w, h = 12, 8
    out = _run([_rect(*r)], [(0.0, 0.0, 0.75)], w, h)
    x0, y0, x1, y1 = r
    rows = range(max(int(np.floor(y0 + 0.5)), 0), min(int(np.floor(y1 + 0.5)), h))
    cols = range(max(int(np.floor(x0 + 0.5)), 0), min(int(np.floor(x1 + 0.5)), w))
    ref = np.zeros((h, w), np.float32)
    for y in rows:
        for x in cols:
            ref[y, x] = 0.75
    assert np.array_equal(out, ref)


def test_exact_linear_depth():
    """With dyadic steps every addition is exact: 1/z = k + (y + 1/2) dy + (x + 1/2) dx at each pixel."""
    w, h = 40, 30
    dx, dy, k = 2.0 ** -10, 2.0 ** -8, 0.125
    out = _run([_rect(-1, -1, 41, 31)], [(dx, dy, k)], w, h)
    ys, xs = np.mgrid[0:h, 0:w]
    assert np.array_equal(out, (k + (ys + 0.5) * dy + (xs + 0.5) * dx).astype(np.float32))


def test_sequential_chains():
    """Rounding steps: the first pixel of a row is line_inv_z + x0 dx, then += dx per pixel; line_inv_z
    starts at y0 dy + (dx + dy) / 2 + k and steps by += dy per row (float32, in that order)."""
    w, h = 70, 9
    dx, dy, k = np.float32(0.1), np.float32(0.37), np.float32(1.3)
    out = _run([_rect(2.2, 1.1, 69.9, 8.6)], [(dx, dy, k)], w, h)
    y0, y1, x0, x1 = 1, 9, 2, 70
    line = np.float32(np.float32(np.float32(y0) * dy) + np.float32(np.float32(0.5) * np.float32(dx + dy))) + k
    for y in range(y0, y1):
        v = np.float32(line + np.float32(np.float32(x0) * dx))
        for x in range(x0, x1):
            assert out[y, x] == v, (x, y)
            v = np.float32(v + dx)
        line = np.float32(line + dy)
    assert not np.any(out[:y0]) and not np.any(out[:, :x0])


def test_depth_test_and_overwrite():
    """Overlapping constant-depth rectangles: with the test the larger 1/z stays, without the later one
    overwrites; uncovered pixels keep their values (no clear)."""
    w, h = 16, 12
    a, b = _rect(1, 1, 10, 8), _rect(5, 4, 15, 11)
    init = np.full(w * h, -1.0, np.float32)
    over = _run([a, b], [(0, 0, 0.5), (0, 0, 0.25)], w, h, depth_test=0, init=init.copy())
    test = _run([a, b], [(0, 0, 0.5), (0, 0, 0.25)], w, h, depth_test=[0, 1], init=init.copy())
    assert over[6, 7] == 0.25 and test[6, 7] == 0.5
    assert over[2, 2] == 0.5 and test[10, 12] == 0.25
    assert over[0, 0] == -1.0 and test[11, 0] == -1.0


def test_shared_edge_is_watertight():
    """A convex quad split along a diagonal: no pixel is covered twice and every row's covered pixels
    are one run (no gap along the shared edge); the quad drawn whole covers the same rows."""
    rng = np.random.default_rng(4)
    w, h = 64, 48
    for _ in range(20):
        c = rng.uniform([5, 5], [w - 5, h - 5])
        ang = np.sort(rng.uniform(-np.pi, np.pi, 4))
        rad = rng.uniform(4, 20)  # points on a circle: convex
        p = np.stack([c[0] + rad * np.cos(ang), c[1] + rad * np.sin(ang)], axis=1)
        fx = (p * F).astype(np.int64).astype(np.int32)
        quad = fx.reshape(-1)
        t1 = fx[[0, 1, 2]].reshape(-1)
        t2 = fx[[0, 2, 3]].reshape(-1)
        mq = _run([quad], [(0, 0, 1.0)], w, h) > 0
        m1 = _run([t1], [(0, 0, 1.0)], w, h) > 0
        m2 = _run([t2], [(0, 0, 1.0)], w, h) > 0
        assert not np.any(m1 & m2)
        u = m1 | m2
        for y in range(h):
            xs = np.flatnonzero(u[y])
            if len(xs):
                assert xs[-1] - xs[0] + 1 == len(xs), y
        assert np.array_equal(u.any(axis=1), mq.any(axis=1))
