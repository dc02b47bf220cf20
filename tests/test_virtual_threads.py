"""The reference's N-thread output (perform_rasterization, partial_renderer.rs:1192-1283).

With N rayon threads the reference splits the screen with split_rect (rect_splitting.rs:23-78) and
runs the whole draw traversal once per rect: each pass clips with the rect's integer clip rect and
rejects and clips leaves against the rect widened by half a pixel, so the image depends on N (the
spans cut at rect edges step differently). The front-end's `virtual_threads` emits exactly those N
passes as N runs of draws whose clip rects partition the screen; the kernels then reproduce the
N-thread image. These tests pin the partition and the runs' independence on the oracle; the GPU
parity of such frames is in tests/test_gpu_frames.py.
"""
import numpy as np
import pytest

import oracle
import sw_host


@pytest.fixture(scope="module")
def scene():
    m = sw_host.large_map(5, grid=(12, 12, 4))
    return m, m.camera_path(3, 4)[1]


def _frame(m, pose, n, W=640, H=360):
    fb = sw_host.FrameBuilder()
    return fb.build(m, pose, width=W, height=H, use_directional_lightmaps=1, virtual_threads=n, num_models=6,
                    num_decals=20, seed=3)


def _runs(draws):
    """The draw list's runs of equal clip rects, in order."""
    runs = []
    for i, r in enumerate(draws["clip_rect"]):
        r = tuple(int(x) for x in r)
        if not runs or runs[-1][0] != r:
            runs.append((r, [i]))
        else:
            runs[-1][1].append(i)
    return runs


@pytest.mark.parametrize("n", [2, 3, 4, 5, 7, 8, 16, 33, 64])
def test_rects_partition_the_screen(scene, n):
    m, pose = scene
    W, H = 640, 360
    f = _frame(m, pose, n, W, H)
    runs = _runs(f["draws"])
    rects = [r for r, _ in runs]
    assert len(rects) == n and len(set(rects)) == n  # one run per rect, every rect sees the room
    cover = np.zeros((H, W), np.int32)
    for x0, y0, x1, y1 in rects:
        assert 0 <= x0 < x1 <= W and 0 <= y0 < y1 <= H
        cover[y0:y1, x0:x1] += 1
    assert np.all(cover == 1)
    # the halving keeps the parts' areas in proportion to their counts: no rect is far from W*H/n
    areas = np.array([(x1 - x0) * (y1 - y0) for x0, y0, x1, y1 in rects], np.float64)
    assert areas.max() / areas.min() < 1.0 + 8.0 * n * (W + H) / (W * H)


def test_two_threads_split_at_the_middle(scene):
    """split_rect of a wide screen into 2: at max * 1/2 + min * 1/2 = W / 2, left part first."""
    m, pose = scene
    rects = [r for r, _ in _runs(_frame(m, pose, 2)["draws"])]
    assert rects == [(0, 0, 320, 360), (320, 0, 640, 360)]
    rects = [r for r, _ in _runs(_frame(m, pose, 4)["draws"])]
    assert rects == [(0, 0, 320, 180), (0, 180, 320, 360), (320, 0, 640, 180), (320, 180, 640, 360)]


@pytest.mark.parametrize("n", [3, 16])
def test_rect_runs_are_independent(scene, n):
    """Each run writes only inside its rect, and the full image is each run's image on its rect."""
    m, pose = scene
    W, H = 640, 360
    f = _frame(m, pose, n, W, H)
    b = sw_host.batch_from_frames([f])
    full, _, _, oob = oracle.render(m.data, b)
    assert oob == 0
    full = full.reshape(H, W)
    bg = int(f["view"]["background_color"][0])
    for (x0, y0, x1, y1), idx in _runs(f["draws"]):
        g = dict(f)
        g["draws"] = f["draws"][idx[0]:idx[-1] + 1]
        part, _, _, _ = oracle.render(m.data, sw_host.batch_from_frames([g]))
        part = part.reshape(H, W)
        assert np.array_equal(part[y0:y1, x0:x1], full[y0:y1, x0:x1])
        outside = np.ones((H, W), bool)
        outside[y0:y1, x0:x1] = False
        assert np.all(part[outside] == bg)


def test_thread_count_changes_the_image_only_slightly(scene):
    """The N-thread image differs from the one-thread image (the reference's output depends on N),
    in a small fraction of the pixels: spans cut at rect edges and leaf bounds clipped to the rects."""
    m, pose = scene
    W, H = 640, 360
    one, _, _, _ = oracle.render(m.data, sw_host.batch_from_frames([_frame(m, pose, 1, W, H)]))
    for n in (2, 4, 16):
        img, _, _, _ = oracle.render(m.data, sw_host.batch_from_frames([_frame(m, pose, n, W, H)]))
        frac = float(np.mean(img != one))
        assert 0.0 < frac < 0.05, (n, frac)
