"""Seeded triangle-model meshes and instances (synthetic workloads for the model-preparation tests
(geodesic spheres, so roughly half the triangles face any camera outside them) in the three vertex
layouts of triangle_model.rs:63-112, and instances with rotation / translation / scale matrices
placed in front of a camera looking down +z."""
import math

import numpy as np

import sw_abi as A


def icosphere(subdiv):
    """Unit geodesic sphere: positions [n, 3] and triangles [t, 3] wound so that (v1 - v0) x (v2 - v1)
    points inward (the reference's front faces have that plane at positive distance from the eye)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
             (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
             (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    pos = np.array(verts, dtype=np.float32)
    tris = np.array(faces, dtype=np.uint16)[:, ::-1].copy()  # inward-facing cross product
    return pos, tris


def _tex_coords(pos):
    u = np.arctan2(pos[:, 1], pos[:, 0]) / (2 * np.pi) + 0.5
    v = np.arccos(np.clip(pos[:, 2], -1, 1)) / np.pi
    return np.stack([u, v], axis=1).astype(np.float32)


def vertex_animated_mesh(rng, subdiv=3, frames=6, radius=24.0):
    pos, tris = icosphere(subdiv)
    fv = np.zeros((frames, len(pos), 6), dtype=np.float32)
    for f in range(frames):
        wobble = 1.0 + 0.15 * np.sin(3.0 * pos[:, 0] + f) * np.cos(2.0 * pos[:, 2] - 0.5 * f)
        p = pos * (radius * wobble)[:, None]
        n = pos + rng.normal(0.0, 0.05, size=pos.shape)
        fv[f, :, :3] = p
        fv[f, :, 3:] = n
    return A.ModelMeshData(A.MODEL_VERTEX_ANIMATED, _tex_coords(pos), tris, frame_vertices=fv)


def non_animated_mesh(rng, subdiv=2, radius=16.0):
    pos, tris = icosphere(subdiv)
    v = np.zeros((len(pos), 8), dtype=np.float32)
    v[:, :3] = pos * radius * rng.uniform(0.9, 1.1, size=(len(pos), 1))
    v[:, 3:6] = pos * rng.uniform(0.5, 2.0, size=(len(pos), 1))  # unnormalised normals
    v[:, 6:8] = _tex_coords(pos)
    return A.ModelMeshData(A.MODEL_NON_ANIMATED, v, tris)


def _rotation(axis, angle):
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    x, y, z = axis
    c, s = math.cos(angle), math.sin(angle)
    return np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                     [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                     [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])


def affine(rot=np.eye(3), trans=(0, 0, 0), scale=1.0):
    """Row-major 4x4 (numpy) of translate * rotate * scale."""
    m = np.eye(4)
    m[:3, :3] = np.asarray(rot) * scale
    m[:3, 3] = trans
    return m


def column_major(m):
    """Row-major numpy 4x4 -> the ABI's column-major 16 floats."""
    return np.asarray(m, dtype=np.float32).T.reshape(16)


def skeleton_mesh(rng, subdiv=2, frames=4, bones=12, radius=20.0):
    pos, tris = icosphere(subdiv)
    n = len(pos)
    v = np.zeros(n, dtype=A.SKELETON_VERTEX_DTYPE)
    v["tex_coord"] = _tex_coords(pos)
    v["position"] = pos * radius
    v["normal"] = pos
    # up to four bones per vertex, weights summing to 255 (the first may be 0)
    for i in range(n):
        k = int(rng.integers(1, 5))
        idx = rng.choice(bones, size=4, replace=False)
        w = np.zeros(4, dtype=np.int64)
        cuts = np.sort(rng.integers(0, 256, size=k - 1))
        parts = np.diff(np.concatenate([[0], cuts, [255]]))
        w[:k] = parts
        v["bones"][i, :, 0] = idx
        v["bones"][i, :, 1] = w
    parents = np.array([0xFFFFFFFF] + [int(rng.integers(0, b)) if rng.random() < 0.8 else 0xFFFFFFFF
                                       for b in range(1, bones)], dtype=np.uint32)
    fb = np.zeros((frames, bones, 16), dtype=np.float32)
    for f in range(frames):
        for b in range(bones):
            m = affine(_rotation(rng.normal(size=3), rng.uniform(-0.4, 0.4)), rng.normal(0, 2.0, size=3),
                       rng.uniform(0.9, 1.1))
            fb[f, b] = column_major(m)
    return A.ModelMeshData(A.MODEL_SKELETON_ANIMATED, v, tris, bone_parents=parents, frame_bones=fb)


def instances(rng, meshes, count, lerps=None):
    """`count` instances cycling over the meshes, each at its own place in front of the camera."""
    inst = np.zeros(count, dtype=A.MODEL_INSTANCE_DTYPE)
    fv = ft = 0
    for i in range(count):
        mi = i % len(meshes)
        m = meshes[mi]
        model = affine(_rotation(rng.normal(size=3), rng.uniform(0, 2 * np.pi)),
                       (rng.uniform(-200, 200), rng.uniform(-200, 200), rng.uniform(60, 400)), rng.uniform(0.5, 2.0))
        # camera: a mild rotation and a screen-like scale on x/y (the reference's projection-view
        # matrix maps camera space to x/z, y/z screen coordinates)
        view = affine(_rotation((0.1, 1.0, 0.2), rng.uniform(-0.2, 0.2)))
        view = np.diag([500.0, 500.0, 1.0, 1.0]) @ view
        inst[i]["mesh"] = mi
        nf = max(m.num_frames, 1)
        inst[i]["frames"] = (int(rng.integers(0, nf)), int(rng.integers(0, nf)))
        inst[i]["lerp"] = lerps[i % len(lerps)] if lerps is not None else rng.uniform(0, 1)
        inst[i]["model_matrix"] = column_major(model)
        inst[i]["model_view_matrix"] = column_major(view @ model)
        inst[i]["tc_scale"] = (64.0, 128.0)
        inst[i]["tc_shift"] = rng.uniform(-1, 1, size=2)
        inst[i]["light_cube"] = rng.uniform(0, 200, size=(6, 3))
        inst[i]["light_vector_scaled"] = rng.normal(0, 1, size=(2, 3))
        inst[i]["light_color"] = rng.uniform(0, 100, size=(2, 3))
        inst[i]["first_vertex"] = fv
        inst[i]["first_triangle"] = ft
        fv += m.num_vertices
        ft += len(m.triangles)
    return inst
