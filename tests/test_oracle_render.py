"""CPU oracle pins for traversal and the path loop.

* BVH closest hit == brute force over every instanced triangle (R3-R6);
* the kernels' culling + any-hit traversal gives the reference's results;
* f32 vs f64 agreement: hit t, and whole images under decision replay
  (relative RMSE <= 1e-4), with the decision-flip rate reported;
* NEE against a closed-form single-emitter case (including the reference's
  |n0 - nl|^2 divisor, Q3);
* invariants: black emitters give a black image, resuming accumulation equals
  one multi-sample call, the linear CDF scan equals the binary search.
"""
import numpy as np
import pytest

from oracle_lib import OracleScene, f3, lib, random_rays, same_bits
from pulse_amd import _native as N
from pulse_amd import procedural as P
from pulse_amd.path_tracer import make_params
from pulse_amd.scene import MaterialDesc, Mesh, SceneBuilder


@pytest.fixture(scope="module")
def cornell():
    return P.cornell()


@pytest.fixture(scope="module")
def field():
    return P.icosphere_field(n_meshes=6, subdiv=2, n_instances=40, n_emissive=5, box=24.0, radius=(1.0, 3.0))


@pytest.mark.parametrize("which", ["cornell", "field"])
def test_bvh_closest_hit_equals_bruteforce(cornell, field, which):
    scene, _ = cornell if which == "cornell" else field
    o = OracleScene(scene)
    rays = random_rays(6000, np.random.default_rng(1), (0, 0, 0), 0.95 if which == "cornell" else 16.0)
    h, c = o.trace_rays(rays)
    b, _ = o.trace_rays(rays, bruteforce=True)
    same_t = same_bits(h["t"], b["t"])
    assert same_t.all()
    ids_equal = (h["instance_index"] == b["instance_index"]) & (h["triangle_index"] == b["triangle_index"])
    # ids may differ only on exact t-ties (edge-shared hits: first found wins)
    assert ids_equal.mean() > 0.995
    assert (h["t"] < 1e30).mean() > 0.3
    assert c["max_stack"] < 32 and c["iteration_cap_hits"] == 0


def test_culling_traversal_is_exact(field):
    scene, camf = field
    o = OracleScene(scene)
    rays = random_rays(6000, np.random.default_rng(2), (0, 0, 0), 16.0)
    h0, c0 = o.trace_rays(rays)
    t = np.where(h0["t"] < 1e30, h0["t"], 5.0).astype(np.float32)
    sh = rays.copy()
    sh["t_max"] = t * np.random.default_rng(3).uniform(0.5, 1.5, len(t)).astype(np.float32)
    occ0 = o.shadow_rays(sh)
    o.set_cull_world(True)
    h1, c1 = o.trace_rays(rays)
    occ1 = o.shadow_rays(sh)
    for k in ("t", "u", "v"):
        assert same_bits(h0[k], h1[k]).all()
    assert (h0["instance_index"] == h1["instance_index"]).all() and (h0["triangle_index"] == h1["triangle_index"]).all()
    assert (occ0 == occ1).all()
    assert c1["blas_interior"] < c0["blas_interior"]
    cam = camf(64, 40)
    for mode in (N.MODE_REF_GGX, N.MODE_REF_NEE_LAMBERT, N.MODE_GGX_NEE):
        o.set_cull_world(False)
        a0, _, _ = o.render(cam, make_params(max_depth=4, mode=mode))
        o.set_cull_world(True)
        a1, _, _ = o.render(cam, make_params(max_depth=4, mode=mode))
        assert same_bits(a0, a1).all()


def test_f32_hits_agree_with_f64(field):
    scene, _ = field
    o = OracleScene(scene)
    rays = random_rays(4000, np.random.default_rng(4), (0, 0, 0), 16.0)
    h, _ = o.trace_rays(rays)
    tuv, ids = o.trace_rays_f64(rays)
    hit = (h["t"] < 1e30) & (tuv[:, 0] < 1e29)
    same_id = (ids[:, 0] == h["instance_index"]) & (ids[:, 1] == h["triangle_index"])
    assert (same_id | ~hit).mean() > 0.995
    m = hit & same_id
    np.testing.assert_allclose(h["t"][m], tuv[m, 0], rtol=2e-5, atol=1e-5)


@pytest.mark.parametrize("mode,nee,depth", [(N.MODE_REF_GGX, 0, 5), (N.MODE_REF_NEE_LAMBERT, 0, 5),
                                            (N.MODE_GGX_NEE, 0, 4), (N.MODE_GGX_NEE, 1, 4)])
def test_image_f32_vs_f64_decision_replay(cornell, mode, nee, depth):
    """rRMSE of the f32 image against the f64 image computed along the same
    decisions is <= 1e-4 (float64 everywhere except the random numbers)."""
    scene, camf = cornell
    o = OracleScene(scene)
    cam = camf(64, 64)
    a32, a64, flips = o.render_pair(cam, make_params(spp=2, max_depth=depth, mode=mode, nee_eval=nee), replay=True)
    ok = np.isfinite(a64[..., :3]) & np.isfinite(a32[..., :3])
    assert ok.mean() > 0.999
    err = np.sqrt(np.sum((a32[..., :3][ok] - a64[..., :3][ok]) ** 2) / np.sum(a64[..., :3][ok] ** 2))
    assert err <= 1e-4, err
    # decisions the f64 arithmetic would take differently: a small fraction of pixels
    assert flips[1] / (64 * 64) < 0.05


def test_black_emitters_give_black_image():
    scene, camf = P.icosphere_field(n_meshes=4, subdiv=2, n_instances=20, n_emissive=3, box=16.0, emission=0.0)
    o = OracleScene(scene)
    for mode in (N.MODE_REF_GGX, N.MODE_GGX_NEE, N.MODE_REF_NEE_LAMBERT):
        a, _, _ = o.render(camf(48, 32), make_params(max_depth=4, mode=mode))
        assert (a[..., :3] == 0).all()


def test_directly_seen_emitter_radiance():
    """REF_GGX with an emitter-only scene and max_depth 1: a pixel whose
    primary ray hits the emitter gets exactly its emissive value."""
    sb = SceneBuilder()
    m = sb.add_mesh(P.quad_mesh(4.0))
    mat = sb.add_material(MaterialDesc(base_color=(0.5, 0.5, 0.5, 1), emissive=(2.5, 1.5, 0.5, 1)))
    sb.add_instance(m, mat, P.trs((0, 0, 0)))
    scene = sb.build(1)
    cam = P.make_camera(16, 16, eye=(0, -3, 0.01), target=(0, 0, 0))
    a, _, _ = OracleScene(scene).render(cam, make_params(max_depth=1, mode=N.MODE_REF_GGX))
    lit = a[..., 0] > 0
    assert lit.mean() > 0.5
    np.testing.assert_array_equal(a[lit][:, :3], np.float32([2.5, 1.5, 0.5])[None].repeat(lit.sum(), 0))


def test_nee_closed_form_small_emitter():
    """Lambert NEE (utilities.wgsl:500-558) towards a small emitter triangle
    straight above the receiver: mean = base/pi * Le * cosR * cosE * area /
    |n0 - nl|^2 with |n0 - nl|^2 = 4 for facing normals (Q3), integrated
    numerically over the triangle."""
    h = 2.0
    tri = np.array([[0.0, h, 0.0], [0.05, h, 0.0], [0.0, h, 0.05]], np.float32)
    sb = SceneBuilder()
    mid = sb.add_mesh(Mesh(tri, np.tile(np.float32([0, -1, 0]), (3, 1)), np.array([0, 2, 1], np.uint32)))
    mat = sb.add_material(MaterialDesc(emissive=(10.0, 10.0, 10.0, 1)))
    sb.add_instance(mid, mat, np.eye(4))
    scene = sb.build(1)
    o = OracleScene(scene)
    p0, n0, base = f3([0.01, 0, 0.01]), f3([0, 1, 0]), f3([0.5, 0.5, 0.5])
    out = np.zeros(3, np.float32)
    vals = []
    import ctypes as C
    for k in range(20000):
        rng = C.c_uint32(k * 2654435761 & 0xFFFFFFFF)
        lib().oracle_direct_light_f32(o.h, p0.ctypes.data, n0.ctypes.data, base.ctypes.data, C.byref(rng), out.ctypes.data)
        vals.append(out[0])
    est = np.mean(vals)
    # numerical integral of the same integrand over the triangle
    area = 0.5 * 0.05 * 0.05
    uv = np.random.default_rng(0).random((400000, 2))
    uv = np.where(uv.sum(1, keepdims=True) < 1, uv, 1 - uv)
    pts = tri[0] + uv[:, :1] * (tri[1] - tri[0]) + uv[:, 1:] * (tri[2] - tri[0])
    d = pts - (p0 + 0.001 * n0)
    dn = d / np.linalg.norm(d, axis=1, keepdims=True)
    cos_r = dn @ n0
    cos_e = dn[:, 1]  # nl = (0, -1, 0) facing the receiver
    integrand = 0.5 / np.pi * 10.0 * cos_r * cos_e * area / 4.0
    exact = integrand.mean()
    assert abs(est - exact) / exact < 0.01


def test_resume_equals_multi_sample_call(cornell):
    scene, camf = cornell
    o = OracleScene(scene)
    cam = camf(32, 32)
    a = np.zeros((32, 32, 4), np.float32)
    for n in range(3):
        o.render(cam, make_params(accumulation_count=n, max_depth=4, mode=N.MODE_GGX_NEE), accum=a)
    b, _, _ = o.render(cam, make_params(accumulation_count=0, spp=3, max_depth=4, mode=N.MODE_GGX_NEE))
    assert same_bits(a, b).all()


def test_linear_cdf_scan_renders_identically(field):
    scene, camf = field
    o = OracleScene(scene)
    cam = camf(48, 32)
    p = make_params(max_depth=4, mode=N.MODE_GGX_NEE)
    a0, _, _ = o.render(cam, p)
    o.set_linear_scan(True)
    a1, _, _ = o.render(cam, p)
    assert same_bits(a0, a1).all()


def test_candidate_reintersection_reproduces_closest_hits(field):
    """The single-candidate test (the reference's triangle test of one given
    instance/triangle and its world-t conversion) reproduces every closest
    hit's t, u, v bit for bit: the tool the 4-wide traversal's tie checks
    use to re-intersect both sides of a difference."""
    scene, _ = field
    o = OracleScene(scene)
    rays = random_rays(4000, np.random.default_rng(21), (0, 0, 0), 16.0)
    h, _ = o.trace_rays(rays)
    hit = h["t"] < 1e30
    r = o.intersect_candidates(rays[hit], h["instance_index"][hit], h["triangle_index"][hit])
    for k in ("t", "u", "v"):
        assert same_bits(r[k], h[k][hit]).all()
    # a candidate the ray misses reports no hit
    other = (h["triangle_index"][hit] + 7) % 320
    m = o.intersect_candidates(rays[hit], h["instance_index"][hit], other)
    assert (m["t"] == np.float32(1e30)).mean() > 0.9
