"""Host BVH builder and scene packing (CPU only).

The C++ builder (pulse_build_blas / pulse_build_tlas, restating blas.rs and
tlas.rs) is pinned against an independent pure-Python restatement
(oracle/bvh_ref.py) byte for byte on small inputs, and against structural
invariants on large ones."""
import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import bvh_ref  # noqa: E402

from pulse_amd import _native as N  # noqa: E402
from pulse_amd import procedural as P  # noqa: E402


def cpp_blas(prims):
    n = len(prims)
    nodes = np.zeros(max(2 * n - 1, 1), N.BLAS_NODE_DTYPE)
    idx = np.zeros(n, np.uint32)
    cnt = C.c_uint32()
    N.check(N.lib().pulse_build_blas(prims.ctypes.data, n, nodes.ctypes.data, len(nodes), C.byref(cnt), idx.ctypes.data))
    return nodes[: cnt.value], idx


def cpp_tlas(bmin, bmax, cen):
    n = len(cen)
    dt = np.dtype([("bounds_min", "<f4", 3), ("bounds_max", "<f4", 3), ("center", "<f4", 3)])
    arr = np.zeros(n, dt)
    arr["bounds_min"], arr["bounds_max"], arr["center"] = bmin, bmax, cen
    nodes = np.zeros(max(2 * n - 1, 1), N.TLAS_NODE_DTYPE)
    idx = np.zeros(n, np.uint32)
    cnt = C.c_uint32()
    N.check(N.lib().pulse_build_tlas(arr.ctypes.data if n else None, n, nodes.ctypes.data, len(nodes), C.byref(cnt),
                                     idx.ctypes.data if n else None))
    return nodes[: cnt.value], idx


def prims_from_mesh(mesh):
    f = mesh.indices.reshape(-1, 3)
    p = np.zeros(len(f), N.PRIMITIVE_DTYPE)
    p["p_first"], p["p_second"], p["p_third"] = mesh.positions[f[:, 0]], mesh.positions[f[:, 1]], mesh.positions[f[:, 2]]
    return p


def random_soup(n, rng, spread=10.0, size=1.0):
    c = rng.uniform(-spread, spread, (n, 1, 3))
    v = (c + rng.uniform(-size, size, (n, 3, 3))).astype(np.float32)
    p = np.zeros(n, N.PRIMITIVE_DTYPE)
    p["p_first"], p["p_second"], p["p_third"] = v[:, 0], v[:, 1], v[:, 2]
    return p


CASES = {
    "soup_37": lambda rng: random_soup(37, rng),
    "soup_300": lambda rng: random_soup(300, rng),
    "thin_slab": lambda rng: random_soup(200, rng, spread=5.0, size=0.01),
    "cornell_walls": lambda rng: prims_from_mesh(P.cornell_walls_mesh()),
    "box": lambda rng: prims_from_mesh(P.box_mesh()),
    "icosphere_2": lambda rng: prims_from_mesh(P.displaced_icosphere(2, 2.0, P.CounterRng(3), 0)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_blas_matches_python_restatement(case):
    rng = np.random.default_rng(11)
    prims = CASES[case](rng)
    if case == "thin_slab":
        prims["p_first"][:, 1] = prims["p_second"][:, 1] = prims["p_third"][:, 1] = np.float32(0.5)  # one flat axis
    nodes, idx = cpp_blas(prims)
    ref_nodes, ref_idx = bvh_ref.build_blas(prims, N.BLAS_NODE_DTYPE)
    assert nodes.tobytes() == ref_nodes.tobytes()
    np.testing.assert_array_equal(idx, ref_idx)


def test_blas_identical_centroids_stay_a_leaf():
    """Every axis flat: split (axis 0, 0.0); a partition that sends everything
    to one side leaves the node a leaf (blas.rs:124-128).  The reference's root
    would underflow its u32 index and panic when everything goes right; here it
    stays a leaf as every other node does."""
    p = np.zeros(20, N.PRIMITIVE_DTYPE)
    p["p_first"], p["p_second"], p["p_third"] = [1, 0, 0], [2, 0, 0], [1.5, 1, 0]
    nodes, idx = cpp_blas(p)
    assert len(nodes) == 1 and nodes[0]["tri_count"] == 20
    p["p_first"], p["p_second"], p["p_third"] = [-2, 0, 0], [-1, 0, 0], [-1.5, 1, 0]  # all go left
    nodes, _ = cpp_blas(p)
    assert len(nodes) == 1 and nodes[0]["tri_count"] == 20


def _check_tree(nodes, idx, points_of, leaf_max, count_name, first_name):
    n_items = len(idx)
    assert sorted(idx.tolist()) == list(range(n_items))
    seen = np.zeros(n_items, int)
    stack = [0]
    while stack:
        k = stack.pop()
        nd = nodes[k]
        if nd[count_name] > 0:
            items = idx[nd[first_name]: nd[first_name] + nd[count_name]]
            seen[nd[first_name]: nd[first_name] + nd[count_name]] += 1
            pts = np.concatenate([points_of(i) for i in items])
            assert (pts >= nd["aabb_min"]).all() and (pts <= nd["aabb_max"]).all()
        else:
            a = nd[first_name]
            for c in (a, a + 1):
                assert (nodes[c]["aabb_min"] >= nd["aabb_min"]).all() and (nodes[c]["aabb_max"] <= nd["aabb_max"]).all()
                stack.append(c)
    assert (seen == 1).all()


def test_blas_invariants_large_mesh():
    mesh = P.displaced_icosphere(5, 3.0, P.CounterRng(1), 7)
    prims = prims_from_mesh(mesh)
    nodes, idx = cpp_blas(prims)
    pts = np.stack([prims["p_first"], prims["p_second"], prims["p_third"]], 1)
    _check_tree(nodes, idx, lambda i: pts[i], 8, "tri_count", "a_or_first_tri")
    leaves = nodes[nodes["tri_count"] > 0]
    assert leaves["tri_count"].max() <= 8 and len(nodes) == 2 * len(leaves) - 1


@pytest.mark.parametrize("n", [0, 1, 2, 3, 17, 120])
def test_tlas_matches_python_restatement(n):
    rng = np.random.default_rng(n)
    c = rng.uniform(-50, 50, (n, 3)).astype(np.float32)
    h = rng.uniform(0.5, 4, (n, 3)).astype(np.float32)
    bmin, bmax = (c - h).astype(np.float32), (c + h).astype(np.float32)
    cen = (bmin + np.float32(0.5) * (bmax - bmin)).astype(np.float32)
    nodes, idx = cpp_tlas(bmin, bmax, cen)
    ref_nodes, ref_idx = bvh_ref.build_tlas(bmin, bmax, cen, N.TLAS_NODE_DTYPE)
    assert nodes.tobytes() == ref_nodes.tobytes()
    np.testing.assert_array_equal(idx, ref_idx)
    if n == 0:
        assert nodes[0]["a_or_first_instance"] == 0 and nodes[0]["instance_count"] == 0


def test_scene_packing_offsets_and_lights():
    """prepare_mesh_data / prepare_mesh_instances (scene/mod.rs:471-697):
    per-mesh offsets, instance bounds containing the transformed vertices,
    light CDFs ending at 1 with total area = sum of Heron areas."""
    scene, _ = P.icosphere_field(n_meshes=3, subdiv=2, n_instances=9, n_emissive=2, box=20.0)
    b = scene.buffers
    inst = b["instances"]
    assert scene.instance_count == 9 and scene.light_count == 2
    assert (inst["triangle_offset"] == inst["index_offset"]).all()
    mesh_tris = 320
    assert set(inst["triangle_offset"].tolist()) == {0, mesh_tris, 2 * mesh_tris}
    for k, li in enumerate(b["light_indices"]):
        mi = inst[li["mesh_instance_index"]]
        cdf = b["light_triangle_area_cdfs"][li["cdf_offset"]: li["cdf_offset"] + mi["triangle_count"]]
        assert abs(cdf[-1] - 1.0) < 1e-5 and (np.diff(cdf) >= 0).all()
        prims = b["primitives"][mi["triangle_offset"]: mi["triangle_offset"] + mi["triangle_count"]]
        e = lambda x, y: np.linalg.norm(x.astype(np.float64) - y, axis=1)
        a_, b_, c_ = e(prims["p_second"], prims["p_first"]), e(prims["p_third"], prims["p_first"]), e(prims["p_third"], prims["p_second"])
        s = (a_ + b_ + c_) / 2
        area = np.sqrt(s * (s - a_) * (s - b_) * (s - c_)).sum()
        assert abs(b["light_mesh_areas"][k] - area) / area < 1e-4
        assert b["materials"][mi["material_index"]]["emissive"][:3].max() > 0
    # world_object is the inverse of object_world
    for mi in inst:
        ow = mi["object_world"].reshape(4, 4).T
        wo = mi["world_object"].reshape(4, 4).T
        np.testing.assert_allclose(ow @ wo, np.eye(4), atol=2e-5)
    # TLAS leaves cover the transformed mesh vertices
    tl = b["tlas_nodes"]
    for node in tl[tl["instance_count"] > 0]:
        for j in range(node["instance_count"]):
            mi = inst[b["instance_indices"][node["a_or_first_instance"] + j]]
            pr = b["primitives"][mi["triangle_offset"]: mi["triangle_offset"] + mi["triangle_count"]]
            v = np.concatenate([pr["p_first"], pr["p_second"], pr["p_third"]])
            w = v @ mi["object_world"].reshape(4, 4).T[:3, :3].T + mi["object_world"].reshape(4, 4).T[:3, 3]
            assert (w >= node["aabb_min"] - 1e-4).all() and (w <= node["aabb_max"] + 1e-4).all()
