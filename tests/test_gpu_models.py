"""GPU parity of sw_prepare_models (triangle_models_rendering.rs:7-200, 324-401, 702-731): transformed
vertices, the kept and sorted triangles and their counts are bit-exact to the oracle's exact mode."""
import numpy as np
import pytest

import sw_model_scenes as S
import oracle
import sw_abi as A

pytestmark = pytest.mark.gpu


def _compare(gpu_ctx, meshes, inst):
    verts, tris, counts = gpu_ctx.prepare_models(meshes, inst)
    rv, rt, rc, ok = oracle.prepare_models(meshes, inst)
    assert ok
    assert np.array_equal(counts, rc), (counts, rc)
    for name in ("pos", "tc", "light"):
        a, b = verts[name].view(np.uint32), rv[name].view(np.uint32)
        bad = np.argwhere(a != b)
        assert bad.size == 0, f"{name}: {len(bad)} mismatches, first {bad[:3].tolist()}"
    for i, ins in enumerate(inst):
        ft, n = int(ins["first_triangle"]), int(rc[i])
        assert np.array_equal(tris[ft:ft + n], rt[ft:ft + n]), f"instance {i}"


@pytest.mark.parametrize("seed", [1, 2])
def test_mixed_kinds(gpu_ctx, seed):
    rng = np.random.default_rng(seed)
    meshes = [S.vertex_animated_mesh(rng, subdiv=3), S.non_animated_mesh(rng, subdiv=3), S.skeleton_mesh(rng, subdiv=3),
              S.vertex_animated_mesh(rng, subdiv=2, frames=2)]
    lerps = [0.5, 0.0, 1.0, 0.005, 0.995, 0.37, float("nan"), 2.0]
    inst = S.instances(rng, meshes, 24, lerps=lerps)
    print(f"models seed {seed}: {len(inst)} instances", flush=True)
    _compare(gpu_ctx, meshes, inst)


def test_many_instances_frame_workload(gpu_ctx):
    """64 instances of one 1280-triangle vertex-animated mesh (the HDR frame's model count)."""
    rng = np.random.default_rng(7)
    meshes = [S.vertex_animated_mesh(rng, subdiv=3, frames=8)]
    inst = S.instances(rng, meshes, 64)
    _compare(gpu_ctx, meshes, inst)


def test_large_mesh_sorts_in_global_scratch(gpu_ctx):
    """5120 and 20480 triangles: above the LDS sort's 4096 keys."""
    rng = np.random.default_rng(9)
    meshes = [S.vertex_animated_mesh(rng, subdiv=4, frames=2), S.non_animated_mesh(rng, subdiv=5)]
    inst = S.instances(rng, meshes, 4)
    _compare(gpu_ctx, meshes, inst)


def test_deep_bone_chain(gpu_ctx):
    """A chain of 255 bones (each the previous one's child): one bone per resolution round."""
    rng = np.random.default_rng(12)
    m = S.skeleton_mesh(rng, subdiv=2, frames=2, bones=255)
    m.bone_parents = np.array([0xFFFFFFFF] + list(range(254)), dtype=np.uint32)
    inst = S.instances(rng, [m], 2)
    _compare(gpu_ctx, [m], inst)


def test_empty_and_degenerate(gpu_ctx):
    """A mesh with no triangles, a singular model matrix, and skeleton vertices without bones."""
    rng = np.random.default_rng(13)
    base = S.non_animated_mesh(rng, subdiv=1)
    empty = A.ModelMeshData(A.MODEL_NON_ANIMATED, base.vertices, np.zeros((0, 3)))
    sk = S.skeleton_mesh(rng, subdiv=1)
    nobones = A.ModelMeshData(A.MODEL_SKELETON_ANIMATED, sk.vertices, sk.triangles)
    meshes = [empty, base, nobones]
    inst = S.instances(rng, meshes, 3)
    inst[1]["model_matrix"] = S.column_major(S.affine(np.diag([1.0, 0.0, 1.0])))
    _compare(gpu_ctx, meshes, inst)


def test_bad_vertex_index_reported_then_recovers(gpu_ctx):
    rng = np.random.default_rng(14)
    mesh = S.non_animated_mesh(rng, subdiv=1)
    bad = mesh.triangles.copy()
    bad[5, 2] = mesh.num_vertices + 100
    mesh_bad = A.ModelMeshData(A.MODEL_NON_ANIMATED, mesh.vertices, bad)
    inst = S.instances(rng, [mesh_bad], 2)
    with pytest.raises(A.SwError, match="out of range"):
        gpu_ctx.prepare_models([mesh_bad], inst)
    _compare(gpu_ctx, [mesh], S.instances(rng, [mesh], 2))


def test_host_validation(gpu_ctx):
    rng = np.random.default_rng(15)
    mesh = S.vertex_animated_mesh(rng, subdiv=1, frames=2)
    inst = S.instances(rng, [mesh], 1)
    inst[0]["frames"] = (0, 5)
    with pytest.raises(A.SwError, match="frame out of range"):
        gpu_ctx.prepare_models([mesh], inst)
    inst = S.instances(rng, [mesh], 1)
    inst[0]["mesh"] = 3
    with pytest.raises(A.SwError, match="mesh index"):
        gpu_ctx.prepare_models([mesh], inst)
