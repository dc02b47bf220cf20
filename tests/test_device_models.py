"""Frames whose triangle models are prepared on the device (swh_frame_config.device_models): the
front-end lists the visible models as sw_prepare_models instances and each mesh drawn in a leaf
as an sw_mesh_run at its place in the draw list; sw_expand_mesh_runs then does draw_mesh's
per-triangle clipping and fans (partial_renderer.rs:3612-3716, clipping.rs:203-385).

Pins on the CPU:
  - the oracle's restatement of the expansion, fed the front-end's own prepared vertices and
    sorted triangles, reproduces the triangle draws the front-end's draw_mesh emitted for every run
    (vertex data bit for bit, fan sizes, order), over poses whose meshes are clipped by the near
    plane, the leaf planes and the 2D box planes;
  - a device_models frame has exactly the host frame's draws without the runs' triangles, the runs
    at the same places;
  - the oracle's device-model frame (its model preparation, its expansion, the triangles put in
    place) draws its models: a frame with models differs from the same frame without them exactly
    where the runs' triangles cover pixels;
  - the front-end's native draw-list assembly equals an independent numpy restatement, and rejects
    run positions out of order.
"""
import numpy as np
import pytest

import oracle
import sw_abi as A
import sw_host


@pytest.fixture(scope="module")
def scene():
    m = sw_host.large_map(0, grid=(12, 12, 4))
    return m, m.camera_path(3, 40)


def _build(m, pose, i, device):
    fb = sw_host.FrameBuilder()
    return fb.build(m, pose, width=640, height=360, color_bits=32, use_directional_lightmaps=1, num_models=24, num_decals=10,
                    seed=i, time=0.3 * i, view_model=1, device_models=device)


def _oracle_device_frame(m, f, mesh):
    md = f["models"]
    verts, tris, counts, ok = oracle.prepare_models([mesh], md["instances"])
    assert ok
    ints, fans = oracle.expand_mesh_runs(md["runs"], verts, tris, counts)
    base = len(f["vertex_data"])
    draws = sw_host.assemble_model_draws(f, fans, base)
    f2 = dict(f, draws=draws, vertex_data=np.concatenate([f["vertex_data"], ints]).astype(np.int32))
    return f2, fans


def test_expansion_reproduces_the_front_end_draw_mesh(scene):
    m, poses = scene
    planes3d = planes2d = checked = 0
    for i in (7, 11, 14, 21, 30, 31, 34, 0, 2, 3):
        f = _build(m, poses[i], i, 0)
        md = f["models"]
        runs = md["runs"]
        ints, fans = oracle.expand_mesh_runs(runs, md["prepared_vertices"], md["prepared_triangles"], md["prepared_counts"])
        layout, _, _ = A.mesh_run_layout(runs)
        for r, run in enumerate(runs):
            planes3d += int(run["num_planes3d"] > 0)
            planes2d += int(run["num_planes2d"] > 0)
            ib, fan_max, fb = layout[r]
            nt = int(run["max_triangles"])
            fan = fans[fb:fb + nt]
            p0, n = int(md["run_positions"][r]), int(md["run_num_draws"][r])
            assert int(fan.sum()) == n
            k = 0
            for t in range(nt):
                for j in range(int(fan[t])):
                    d = f["draws"][p0 + k]
                    assert d["kind"] == A.DRAW_TRIANGLE and d["blend"] == run["blend"]
                    assert d["texture_offset"] == run["texture_offset"]
                    host = f["vertex_data"][d["first_vertex_int"]:d["first_vertex_int"] + 21]
                    mine = ints[ib + (t * fan_max + j) * 21:ib + (t * fan_max + j) * 21 + 21]
                    assert np.array_equal(host, mine), (i, r, t, j)
                    k += 1
                    checked += 1
    assert planes3d >= 3 and planes2d >= 3 and checked > 5000


def test_device_frame_is_the_host_frame_without_model_triangles(scene):
    m, poses = scene
    for i in (7, 21, 31):
        h = _build(m, poses[i], i, 0)
        d = _build(m, poses[i], i, 1)
        hm, dm = h["models"], d["models"]
        assert len(hm["runs"]) == len(dm["runs"]) > 0
        keep = np.ones(len(h["draws"]), bool)
        for p, n in zip(hm["run_positions"], hm["run_num_draws"]):
            keep[int(p):int(p) + int(n)] = False
        hd = h["draws"][keep]
        assert len(d["draws"]) == len(hd)
        for name in A.DRAW_DTYPE.names:
            if name != "first_vertex_int":
                assert np.array_equal(d["draws"][name], hd[name]), name
        # the model triangles' vertex data is gone from the list; every other draw keeps its own
        for a, b in zip(d["draws"], hd):
            n = 7 * int(a["num_vertices"]) if a["kind"] == A.DRAW_TRIANGLE else 2 * int(a["num_vertices"])
            sa, sb = int(a["first_vertex_int"]), int(b["first_vertex_int"])
            assert np.array_equal(d["vertex_data"][sa:sa + n], h["vertex_data"][sb:sb + n])
        shifted = np.array([int(p) - int(np.sum(hm["run_num_draws"][:r])) for r, p in enumerate(hm["run_positions"])])
        assert np.array_equal(dm["run_positions"], shifted)
        # the same per-leaf decisions; the device runs draw every triangle slot of the mesh
        for a, b in zip(hm["runs"], dm["runs"]):
            for k in ("num_planes3d", "num_planes2d", "blend", "texture_offset", "planes3d", "planes2d", "clip_rect"):
                assert np.array_equal(a[k], b[k]), k
        assert np.all(dm["runs"]["max_triangles"] == len(m.model_mesh().triangles))
        assert len(dm["prepared_vertices"]) == 0


def test_oracle_device_frame_draws_its_models(scene):
    m, poses = scene
    mesh = m.model_mesh()
    for i in (7, 21):
        d = _build(m, poses[i], i, 1)
        f2, fans = _oracle_device_frame(m, d, mesh)
        assert int(fans.sum()) > 100
        with_models, _, st, oob = oracle.render(m.data, sw_host.batch_from_frames([f2]))
        assert oob == 0 and st["triangle_pixels"] > 0
        without, _, _, _ = oracle.render(m.data, sw_host.batch_from_frames([d]))
        assert np.count_nonzero(with_models != without) > 0


def _assemble_reference(frame, fans, int_base):
    """numpy restatement of the assembly: the draws split at the run positions, each run's fan
    triangles (slot t * fan_max + j for j below fan[t]) inserted there."""
    md = frame["models"]
    runs, pos = md["runs"], md["run_positions"]
    layout, _, _ = A.mesh_run_layout(runs)
    draws = frame["draws"]
    pieces, start = [], 0
    for r, run in enumerate(runs):
        p = int(pos[r])
        pieces.append(draws[start:p])
        start = p
        ib, fan_max, fb = layout[r]
        nt = int(run["max_triangles"])
        f = np.asarray(fans[fb:fb + nt], dtype=np.int64)
        t_idx = np.repeat(np.arange(nt), f)
        j_idx = np.arange(int(f.sum())) - np.repeat(np.cumsum(f) - f, f)
        d = np.zeros(len(t_idx), dtype=A.DRAW_DTYPE)
        d["kind"] = A.DRAW_TRIANGLE
        d["blend"] = run["blend"]
        d["first_vertex_int"] = int_base + ib + (t_idx * fan_max + j_idx) * 21
        d["num_vertices"] = 3
        d["texture_offset"] = run["texture_offset"]
        d["texture_size"] = run["texture_size"]
        d["clip_rect"] = run["clip_rect"]
        pieces.append(d)
    pieces.append(draws[start:])
    return np.concatenate(pieces)


def test_native_assembly_equals_a_numpy_restatement(scene):
    m, poses = scene
    mesh = m.model_mesh()
    rng = np.random.default_rng(2)
    for i in (7, 21, 31):
        d = _build(m, poses[i], i, 1)
        md = d["models"]
        _, _, slots = A.mesh_run_layout(md["runs"])
        fans = rng.integers(0, 3, size=slots).astype(np.uint8)
        got = sw_host.assemble_model_draws(d, fans, 1000)
        want = _assemble_reference(d, fans, 1000)
        assert np.array_equal(got, want)
    bad = dict(d, models=dict(d["models"], run_positions=d["models"]["run_positions"][::-1].copy()))
    if len(bad["models"]["run_positions"]) > 1 and bad["models"]["run_positions"][0] > bad["models"]["run_positions"][-1]:
        with pytest.raises(ValueError):
            sw_host.assemble_model_draws(bad, fans, 1000)
