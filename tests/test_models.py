"""Triangle-model preparation (triangle_models_rendering.rs:7-200, 324-401, 702-731; light.rs:296-340):
the oracle's restatement pinned on the CPU by closed forms, float64 geometry and invariants. GPU
parity of sw_prepare_models is in test_gpu_models.py."""
import math

import numpy as np
import pytest

import sw_model_scenes as S
import oracle
import sw_abi as A


def _one(inst_kwargs, mesh):
    inst = np.zeros(1, dtype=A.MODEL_INSTANCE_DTYPE)
    inst[0]["model_matrix"] = S.column_major(np.eye(4))
    inst[0]["model_view_matrix"] = S.column_major(np.eye(4))
    inst[0]["tc_scale"] = (1.0, 1.0)
    for k, v in inst_kwargs.items():
        inst[0][k] = v
    return oracle.prepare_models([mesh], inst)


def _cube_verts(normals):
    v = np.zeros((len(normals), 8), dtype=np.float32)
    v[:, :3] = np.arange(len(normals) * 3, dtype=np.float32).reshape(-1, 3) * 0.25 + 1.0
    v[:, 3:6] = normals
    v[:, 6:8] = np.arange(len(normals) * 2, dtype=np.float32).reshape(-1, 2) * 0.125
    return v


def test_identity_transform_and_axis_lights():
    """Identity matrices: positions pass through; tc = tc * scale + shift; an axis normal picks one
    light-cube side at weight 1 (negative axis: side 2k, positive: 2k + 1), plus each directional
    colour times max(n . vector, 0)."""
    normals = np.array([[1, 0, 0], [-1, 0, 0], [0, 3, 0], [0, -0.5, 0], [0, 0, 2], [0, 0, -1], [0, 0, 0]], dtype=np.float32)
    v = _cube_verts(normals)
    mesh = A.ModelMeshData(A.MODEL_NON_ANIMATED, v, np.zeros((0, 3)))
    cube = np.arange(18, dtype=np.float32).reshape(6, 3) * 10 + 1
    vec = np.array([[1, 0, 0], [0, 0, -2]], dtype=np.float32)
    col = np.array([[100, 200, 300], [7, 8, 9]], dtype=np.float32)
    verts, _, counts, ok = _one(dict(tc_scale=(2.0, 4.0), tc_shift=(0.5, -1.0), light_cube=cube, light_vector_scaled=vec,
                                     light_color=col), mesh)
    assert ok and counts[0] == 0
    assert np.array_equal(verts["pos"][:7], v[:, :3])
    assert np.array_equal(verts["tc"][:7], v[:, 6:8] * np.float32([2.0, 4.0]) + np.float32([0.5, -1.0]))
    sides = [1, 0, 3, 2, 5, 4]
    for i in range(6):
        n = normals[i] / np.linalg.norm(normals[i])
        expect = cube[sides[i]].astype(np.float64)
        for d in range(2):
            expect = expect + col[d] * max(float(n @ vec[d]), 0.0)
        np.testing.assert_allclose(verts["light"][i], expect, rtol=1e-6)
    # a zero normal: magnitude clamped to 1e-8, the normalised normal is 0: every cube side
    # contributes -0 * value (x <= 0 branch), no directional light
    assert np.array_equal(verts["light"][6], np.zeros(3, dtype=np.float32))


@pytest.mark.parametrize("lerp,expect", [(1.0, "f0"), (0.995, "f0"), (0.0, "f1"), (0.005, "f1"), (-3.0, "f1"),
                                         (7.0, "f0"), (float("nan"), "f1"), (0.5, "mix"), (0.25, "mix")])
def test_vertex_animation_lerp(lerp, expect):
    """lerp clamped to [0, 1] (NaN -> 0); within (0.01, 0.99) the frames are mixed, else the
    heavier frame is used as it is (:19-94)."""
    rng = np.random.default_rng(3)
    mesh = S.vertex_animated_mesh(rng, subdiv=1, frames=3)
    verts, _, _, _ = _one(dict(frames=(2, 0), lerp=lerp), mesh)
    f0, f1 = mesh.frame_vertices[2, :, :3], mesh.frame_vertices[0, :, :3]
    if expect == "f0":
        assert np.array_equal(verts["pos"], f0)
    elif expect == "f1":
        assert np.array_equal(verts["pos"], f1)
    else:
        l0 = np.float32(lerp)
        assert np.array_equal(verts["pos"], f0 * l0 + f1 * (np.float32(1) - l0))


def test_transform_against_float64():
    """Random rotation / scale / translation model and view matrices: positions match a float64
    transform, lights match a float64 evaluation with the inverse-transpose normal matrix."""
    rng = np.random.default_rng(11)
    mesh = S.non_animated_mesh(rng, subdiv=2)
    model = S.affine(S._rotation((1, 2, 3), 0.7), (10, -20, 300), 1.7)
    view = np.diag([500.0, 400.0, 1.0, 1.0]) @ S.affine(S._rotation((0, 1, 0), 0.3), (5, 0, 0))
    cube = rng.uniform(0, 100, size=(6, 3)).astype(np.float32)
    vec = rng.normal(size=(2, 3)).astype(np.float32)
    col = rng.uniform(0, 50, size=(2, 3)).astype(np.float32)
    verts, _, _, _ = _one(dict(model_matrix=S.column_major(model), model_view_matrix=S.column_major(view @ model),
                               light_cube=cube, light_vector_scaled=vec, light_color=col), mesh)
    mv = np.asarray(S.column_major(view @ model), dtype=np.float64).reshape(4, 4).T
    p = mesh.vertices[:, :3].astype(np.float64)
    ref = (mv[:3, :3] @ p.T).T + mv[:3, 3]
    np.testing.assert_allclose(verts["pos"], ref, rtol=2e-6, atol=1e-3)
    m3 = np.asarray(S.column_major(model), dtype=np.float64).reshape(4, 4).T[:3, :3]
    nmat = np.linalg.inv(m3).T
    for i in range(0, len(p), 7):
        n = nmat @ mesh.vertices[i, 3:6].astype(np.float64)
        n = n / np.linalg.norm(n)
        expect = np.zeros(3)
        for axis in range(3):
            expect += cube[axis * 2 + (0 if n[axis] <= 0 else 1)] * abs(n[axis])
        for d in range(2):
            expect += col[d] * max(float(n @ vec[d]), 0.0)
        np.testing.assert_allclose(verts["light"][i], expect, rtol=1e-4, atol=1e-3)


def test_singular_model_matrix_uses_identity_normals():
    """get_normals_matrix of a singular matrix: Matrix3::invert fails, the identity is used (:722-731)."""
    normals = np.array([[0, 1, 0], [0, 0, -1]], dtype=np.float32)
    mesh = A.ModelMeshData(A.MODEL_NON_ANIMATED, _cube_verts(normals), np.zeros((0, 3)))
    cube = np.arange(18, dtype=np.float32).reshape(6, 3)
    flat = S.affine(np.diag([1.0, 1.0, 0.0]))
    verts, _, _, _ = _one(dict(model_matrix=S.column_major(flat), light_cube=cube), mesh)
    assert np.array_equal(verts["light"][0], cube[3])
    assert np.array_equal(verts["light"][1], cube[4])


def _front_set_f64(mesh, verts, first):
    """Front faces decided in float64 from the same camera-space positions: the triangle's plane
    (through its vertices) lies on the positive side of the eye."""
    out = set()
    for t, (a, b, c) in enumerate(mesh.triangles):
        p = [verts["pos"][first + i].astype(np.float64) for i in (a, b, c)]
        n = np.cross(p[1] - p[0], p[2] - p[1])
        d = float(n @ p[0])
        if abs(d) > 1e-3 * np.linalg.norm(n) * np.linalg.norm(p[0]):
            if d > 0:
                out.add(t)
        else:
            out.add(("edge", t))
    return out


def test_back_faces_and_sort_invariants():
    """On closed spheres seen from outside: about half the triangles are front faces; the kept set
    matches the float64 decision away from edge-on triangles; the output is sorted by the largest
    vertex z, non-increasing, with equal depths in mesh order."""
    rng = np.random.default_rng(5)
    meshes = [S.vertex_animated_mesh(rng, subdiv=3), S.non_animated_mesh(rng, subdiv=2), S.skeleton_mesh(rng)]
    inst = S.instances(rng, meshes, 9)
    verts, tris, counts, ok = oracle.prepare_models(meshes, inst)
    assert ok
    for i, ins in enumerate(inst):
        m = meshes[int(ins["mesh"])]
        n = int(counts[i])
        assert 0.3 * len(m.triangles) < n < 0.7 * len(m.triangles)
        fv, ft = int(ins["first_vertex"]), int(ins["first_triangle"])
        out = tris[ft:ft + n]
        # map the output triangles back to mesh indices (triangles are distinct in a sphere)
        lookup = {tuple(t): k for k, t in enumerate(m.triangles.tolist())}
        idx = [lookup[tuple(t)] for t in out.tolist()]
        ref = _front_set_f64(m, verts, fv)
        sure = {t for t in ref if not isinstance(t, tuple)}
        edge = {t[1] for t in ref if isinstance(t, tuple)}
        assert sure <= set(idx) <= sure | edge
        z = verts["pos"][fv:fv + m.num_vertices, 2]
        key = [max(z[a], z[b], z[c]) for a, b, c in m.triangles[idx]]
        for j in range(1, n):
            assert key[j - 1] > key[j] or (key[j - 1] == key[j] and idx[j - 1] < idx[j])


def test_both_windings_split_exactly():
    """Each triangle together with its reversed copy: exactly one of the two is kept unless the
    plane passes through the eye (then neither)."""
    rng = np.random.default_rng(8)
    base = S.non_animated_mesh(rng, subdiv=2)
    both = np.concatenate([base.triangles, base.triangles[:, ::-1]])
    mesh = A.ModelMeshData(A.MODEL_NON_ANIMATED, base.vertices, both)
    inst = S.instances(rng, [mesh], 4)
    _, _, counts, _ = oracle.prepare_models([mesh], inst)
    assert all(int(c) == len(base.triangles) for c in counts)


def test_skeleton_single_bone_is_the_model_view():
    """One root bone, identity frames, weight 255: the vertex matrix is (model_view / 255 * I) * 255,
    the model-view matrix up to rounding (float64 check); with two bones the child's matrix is
    parent x local."""
    rng = np.random.default_rng(2)
    pos, tris = S.icosphere(1)
    v = np.zeros(len(pos), dtype=A.SKELETON_VERTEX_DTYPE)
    v["position"] = pos * 10
    v["normal"] = pos
    v["bones"][:, 0] = (0, 255)
    fb = np.tile(S.column_major(np.eye(4)), (2, 1, 1)).astype(np.float32)
    mesh = A.ModelMeshData(A.MODEL_SKELETON_ANIMATED, v, tris, bone_parents=[0xFFFFFFFF], frame_bones=fb[:, :1])
    inst = S.instances(rng, [mesh], 1, lerps=[1.0])
    verts, _, _, ok = oracle.prepare_models([mesh], inst)
    assert ok
    mv = np.asarray(inst[0]["model_view_matrix"], dtype=np.float64).reshape(4, 4).T
    ref = (mv[:3, :3] @ (pos * 10).astype(np.float64).T).T + mv[:3, 3]
    np.testing.assert_allclose(verts["pos"][:len(pos)], ref, rtol=1e-5, atol=1e-2)

    # parent 0 translates by (5, 0, 0), child 1 (parent 0) rotates: a vertex on bone 1 moves by both
    r = S.affine(S._rotation((0, 0, 1), 0.5))
    t = S.affine(trans=(5, 0, 0))
    fb2 = np.stack([S.column_major(t), S.column_major(r)])[None].repeat(2, axis=0)
    v2 = v.copy()
    v2["bones"][:, 0] = (1, 255)
    mesh2 = A.ModelMeshData(A.MODEL_SKELETON_ANIMATED, v2, tris, bone_parents=[0xFFFFFFFF, 0], frame_bones=fb2)
    verts2, _, _, ok = oracle.prepare_models([mesh2], inst)
    assert ok
    full = mv @ t @ r
    ref2 = (full[:3, :3] @ (pos * 10).astype(np.float64).T).T + full[:3, 3]
    np.testing.assert_allclose(verts2["pos"][:len(pos)], ref2, rtol=1e-5, atol=1e-2)


def test_bad_indices_are_reported():
    rng = np.random.default_rng(4)
    mesh = S.non_animated_mesh(rng, subdiv=1)
    bad = mesh.triangles.copy()
    bad[3, 1] = mesh.num_vertices + 7
    mesh_bad = A.ModelMeshData(A.MODEL_NON_ANIMATED, mesh.vertices, bad)
    inst = S.instances(rng, [mesh_bad], 1)
    _, _, _, ok = oracle.prepare_models([mesh_bad], inst)
    assert not ok


def test_sse_mode_differs_only_in_lights():
    """x86 rsqrtss (12-bit estimate) against 1/sqrt: positions and texture coordinates identical, light
    within the estimate's relative error."""
    rng = np.random.default_rng(6)
    meshes = [S.vertex_animated_mesh(rng, subdiv=2)]
    inst = S.instances(rng, meshes, 3)
    ve, te, ce, _ = oracle.prepare_models(meshes, inst, mode=oracle.MODE_EXACT)
    vs, ts, cs, _ = oracle.prepare_models(meshes, inst, mode=oracle.MODE_SSE)
    assert np.array_equal(ve["pos"], vs["pos"]) and np.array_equal(ve["tc"], vs["tc"])
    assert np.array_equal(ce, cs) and np.array_equal(te, ts)
    np.testing.assert_allclose(vs["light"], ve["light"], rtol=2 * 1.5 * 2 ** -12, atol=1e-3)
