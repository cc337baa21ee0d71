"""GPU adjoint (pulse_adjoint) against the oracle and finite differences.

* the replayed image equals the forward image bit for bit (same decisions);
* the gradient equals the oracle's: f64 central differences along the same
  decisions (all 7 parameters of every material);
* Euler's identity for emission (target 0): sum_j Le_j dL/dLe_j = 2 L;
* central finite differences of the f32 GPU renders (h = 1e-3, decisions
  free to change) on 16 random parameters: relative error <= 1e-2.

Two roughness derivatives: the default holds the sampled directions fixed
and differentiates the sampling weight together with the density it was
drawn from (detached / score function, unbiased for the derivative of the
expected image); PULSE_FLAG_ADJOINT_REPARAM carries tangents through the
moving directions instead, which misses the visibility changes they cause.
The oracle's f64 central differences along the f32 decisions pin each
(frozen half vectors with likelihood-ratio reweighting for the first).

The differentiable form renders with PULSE_FLAG_HORIZON_CLAMP: the
reference's GGX estimator keeps samples below the shading horizon (Q6) and
its pixel values are unbounded (1e16 at 8 spp on these scenes), which makes
an L2 loss meaningless.
"""
import numpy as np
import pytest

from oracle_lib import OracleScene, same_bits
from pulse_amd import _native as N
from pulse_amd import procedural as P
from pulse_amd.path_tracer import (Device, PulseAdjoint, apply_material_params, make_params,
                                   material_params)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = Device(0)
    d.set_option(N.OPT_TRAVERSAL_LOOP, 1)  # the reference's traversal order: bit-exact vs the oracle
    yield d
    d.close()


def _cornell_case(spp=2, depth=3, size=48, rr=None):
    scene, camf = P.cornell()
    cam = camf(size, size)
    prm = make_params(spp=spp, max_depth=depth, mode=N.MODE_GGX_NEE, rr_start_depth=depth if rr is None else rr,
                      flags=N.FLAG_HORIZON_CLAMP)
    return scene, cam, prm


MODES = [pytest.param(0, True, id="detached"), pytest.param(N.FLAG_ADJOINT_REPARAM, False, id="reparam")]


@pytest.mark.parametrize("loop", [1, 2])
@pytest.mark.parametrize("extra_flags,frozen", MODES)
def test_replay_matches_forward_and_oracle_gradient(dev, extra_flags, frozen, loop):
    """The replay walks the forward pass's BVHs (loop 2: the 4-wide ones), so
    it reproduces the forward image bit for bit either way."""
    scene, cam, prm = _cornell_case()
    prm.flags |= extra_flags
    dev.set_option(N.OPT_TRAVERSAL_LOOP, loop)
    try:
        _replay_case(dev, scene, cam, prm, frozen)
    finally:
        dev.set_option(N.OPT_TRAVERSAL_LOOP, 1)


def _replay_case(dev, scene, cam, prm, frozen):
    dev.upload_scene(scene)
    nm = len(scene.buffers["materials"])
    adj = PulseAdjoint(dev, cam.width, cam.height, nm)
    img = adj.forward(cam, prm)
    target = (img * np.float32(0.6) + np.float32(0.05)).astype(np.float32)
    target[..., 3] = 1.0
    adj.set_target(target)
    g, replay = adj.backward(cam, prm, want_replay=True)
    assert same_bits(replay, img).all()
    o = OracleScene(scene)
    mp = material_params(scene.buffers["materials"])
    ids = np.arange(mp.size, dtype=np.uint32)
    gref, loss_ref = o.grad_fd(cam, prm, target, mp.reshape(-1), ids, h=1e-6, frozen=frozen)
    gref = gref.reshape(nm, 7)
    assert abs(adj.loss() - loss_ref) <= 1e-4 * loss_ref
    scale = np.abs(gref).max()
    err = np.abs(g - gref)
    print("max err / scale", float((err / scale).max()), flush=True)
    assert (err <= 2e-3 * np.abs(gref) + 1e-4 * scale).all(), (g, gref)


@pytest.mark.parametrize("depth", [6, 8])
def test_deep_paths_replay_and_oracle_gradient(dev, depth):
    """More than 4 bounces (the deep kernels, fixed-direction roughness
    derivative): replay == forward and the gradient == the oracle's."""
    scene, cam, prm = _cornell_case(depth=depth)
    _replay_case(dev, scene, cam, prm, True)


@pytest.mark.parametrize("loop", [1, 2])
@pytest.mark.parametrize("extra_flags,frozen", MODES)
def test_roulette_replay_and_oracle_gradient(dev, extra_flags, frozen, loop):
    """Russian roulette from the first bounce on: the replay draws the same
    numbers and ends the same paths (replay == forward bit for bit), and the
    gradient equals the oracle's central differences along the same
    decisions, random survivals (p <= 1) reweighted by their likelihood
    ratio, certain ones (p > 1) divided by the perturbed p."""
    scene, cam, prm = _cornell_case(depth=4, rr=1)
    prm.flags |= extra_flags
    dev.set_option(N.OPT_TRAVERSAL_LOOP, loop)
    try:
        _replay_case(dev, scene, cam, prm, frozen)
    finally:
        dev.set_option(N.OPT_TRAVERSAL_LOOP, 1)


@pytest.mark.parametrize("loop", [1, 2])
@pytest.mark.parametrize("extra_flags,frozen", MODES)
def test_metallic_replay_and_oracle_gradient(dev, extra_flags, frozen, loop):
    """Metallic walls (0.6) and boxes (1.0), roulette from the first bounce:
    the albedo moves f0, so the GGX weights, the GGX light evaluator and
    the lobe probability F0max; the gradient equals the oracle's central
    differences along the same decisions, the lobe choice reweighted by its
    likelihood ratio like a random survival."""
    scene, cam, prm = _cornell_case(depth=4, rr=1)
    mats = scene.buffers["materials"]
    mats["metallic"][0] = 0.6
    mats["metallic"][1] = 1.0
    prm.flags |= extra_flags
    dev.set_option(N.OPT_TRAVERSAL_LOOP, loop)
    try:
        _replay_case(dev, scene, cam, prm, frozen)
    finally:
        dev.set_option(N.OPT_TRAVERSAL_LOOP, 1)


@pytest.mark.parametrize("extra_flags,frozen,depth", [
    pytest.param(0, True, 4, id="detached-4"), pytest.param(0, True, 5, id="detached-5"),
    pytest.param(0, True, 8, id="detached-8"), pytest.param(N.FLAG_ADJOINT_REPARAM, False, 4, id="reparam-4")])
def test_reference_path_trace_gradient(dev, extra_flags, frozen, depth):
    """The reference's own path_trace (PULSE_MODE_REF_GGX: emission at every
    hit, no light sampling; 5 bounces by default), roulette from the first
    bounce: replay == forward bit for bit, gradient == the oracle's central
    differences along the same decisions (reparameterised roughness up to 4
    bounces)."""
    scene, cam, prm = _cornell_case(depth=depth, rr=1)
    prm.mode = N.MODE_REF_GGX
    prm.flags |= extra_flags
    _replay_case(dev, scene, cam, prm, frozen)


@pytest.mark.parametrize("extra_flags,frozen,depth", [
    pytest.param(0, True, 3, id="detached-3"), pytest.param(0, True, 7, id="detached-7"),
    pytest.param(N.FLAG_ADJOINT_REPARAM, False, 3, id="reparam-3")])
def test_lambert_path_gradient(dev, extra_flags, frozen, depth):
    """The reference's path_traceBBB (PULSE_MODE_REF_NEE_LAMBERT: the first
    hit with its emission and Lambert light sample, then max_depth
    cosine-sampled bounces whose weight is the albedo), roulette from loop
    depth 1: replay == forward bit for bit, gradient == the oracle's
    central differences (no roughness dependence: that column is 0)."""
    scene, cam, prm = _cornell_case(depth=depth, rr=1)
    prm.mode = N.MODE_REF_NEE_LAMBERT
    prm.flags |= extra_flags
    _replay_case(dev, scene, cam, prm, frozen)


def _with_flags(prm, extra):
    out = make_params(spp=prm.spp, max_depth=prm.max_depth, mode=prm.mode, rr_start_depth=prm.rr_start_depth,
                      flags=prm.flags | extra)
    out.accumulation_count = prm.accumulation_count
    out.nee_eval = prm.nee_eval
    return out


@pytest.mark.parametrize("loop", [1, 2])
@pytest.mark.parametrize("extra_flags,frozen", MODES)
def test_recorded_paths_give_the_traced_gradient(dev, extra_flags, frozen, loop):
    """PULSE_FLAG_RECORD_PATHS: the forward keeps every path's closest hits
    and shadow visibilities and the adjoint reads them instead of tracing
    the paths again: the replay equals the forward bit for bit and the
    gradient equals the traced adjoint's (the same terms; only the order of
    the float atomics that sum them differs)."""
    scene, cam, prm = _cornell_case(depth=4, rr=1)
    prm.flags |= extra_flags
    dev.set_option(N.OPT_TRAVERSAL_LOOP, loop)
    try:
        dev.upload_scene(scene)
        nm = len(scene.buffers["materials"])
        adj = PulseAdjoint(dev, cam.width, cam.height, nm)
        img = adj.forward(cam, prm)
        target = (img * np.float32(0.6) + np.float32(0.05)).astype(np.float32)
        adj.set_target(target)
        g_traced = adj.backward(cam, prm)
        assert dev.get_option(N.OPT_ADJOINT_READ_RECORD) == 0
        img2 = adj.forward(cam, _with_flags(prm, N.FLAG_RECORD_PATHS))
        assert same_bits(img2, img).all()
        assert dev.get_option(N.OPT_RECORDED_SAMPLES) == prm.spp
        g_rec, replay = adj.backward(cam, prm, want_replay=True)
        assert dev.get_option(N.OPT_ADJOINT_READ_RECORD) == 1
        assert same_bits(replay, img).all()
        scale = np.abs(g_traced).max()
        np.testing.assert_allclose(g_rec, g_traced, rtol=1e-5, atol=1e-6 * scale)
        # forward + adjoint in one call records by itself
        loss, g_fa = adj.forward_backward(cam, prm)
        assert dev.get_option(N.OPT_ADJOINT_READ_RECORD) == 1
        np.testing.assert_allclose(g_fa, g_traced, rtol=1e-5, atol=1e-6 * scale)
    finally:
        dev.set_option(N.OPT_TRAVERSAL_LOOP, 1)


def test_path_record_lifetime(dev):
    """The record holds one image: accumulated calls with the same camera and
    parameters extend it (the adjoint of samples 0-5 reads it); a material
    update, another camera or a call that does not continue its samples
    leaves no valid record, and the adjoint then traces (same gradient)."""
    scene, cam, prm = _cornell_case(spp=2, depth=3, rr=1)
    dev.upload_scene(scene)
    nm = len(scene.buffers["materials"])
    adj = PulseAdjoint(dev, cam.width, cam.height, nm)
    rec = _with_flags(prm, N.FLAG_RECORD_PATHS)
    for acc in (0, 2, 4):
        rec.accumulation_count = acc
        img = adj.forward(cam, rec)
    assert dev.get_option(N.OPT_RECORDED_SAMPLES) == 6
    target = (img * np.float32(0.6) + np.float32(0.05)).astype(np.float32)
    adj.set_target(target)
    last = _with_flags(prm, 0)
    last.accumulation_count = 4
    g_rec, replay = adj.backward(cam, last, want_replay=True)
    assert dev.get_option(N.OPT_ADJOINT_READ_RECORD) == 1
    assert same_bits(replay, img).all()
    # the same adjoint traced: a material update (to the same values) drops the record
    adj.set_materials(scene.buffers["materials"])
    assert dev.get_option(N.OPT_RECORDED_SAMPLES) == 0
    g_traced = adj.backward(cam, last)
    assert dev.get_option(N.OPT_ADJOINT_READ_RECORD) == 0
    scale = np.abs(g_traced).max()
    np.testing.assert_allclose(g_rec, g_traced, rtol=1e-5, atol=1e-6 * scale)
    # a call that does not continue the record (accumulation_count 3 after 0-1)
    rec.accumulation_count = 0
    adj.forward(cam, rec)
    assert dev.get_option(N.OPT_RECORDED_SAMPLES) == 2
    rec.accumulation_count = 3
    adj.forward(cam, rec)
    assert dev.get_option(N.OPT_RECORDED_SAMPLES) == 0
    # another camera for the adjoint: the record is not read
    rec.accumulation_count = 0
    adj.forward(cam, rec)
    other = P.cornell()[1](48, 48)
    other.world_position[0] += 0.01
    adj.backward(other, prm)
    assert dev.get_option(N.OPT_ADJOINT_READ_RECORD) == 0


def test_roulette_deep_paths(dev):
    scene, cam, prm = _cornell_case(depth=8, rr=2)
    _replay_case(dev, scene, cam, prm, True)


def test_emission_euler_identity_gpu(dev):
    scene, cam, prm = _cornell_case(spp=2, depth=3, size=40)
    dev.upload_scene(scene)
    nm = len(scene.buffers["materials"])
    adj = PulseAdjoint(dev, cam.width, cam.height, nm)
    adj.forward(cam, prm)
    adj.set_target(np.zeros((cam.height, cam.width, 4), np.float32))
    g = adj.backward(cam, prm)
    loss = adj.loss()
    mp = material_params(scene.buffers["materials"])
    euler = float(np.sum(mp[:, 4:7] * g[:, 4:7]))
    assert abs(euler - 2 * loss) <= 2e-4 * loss


@pytest.mark.parametrize("light_power", [0, N.FLAG_LIGHT_POWER], ids=["uniform_lights", "light_power"])
@pytest.mark.parametrize("extra_flags,frozen", MODES)
def test_field_gradient_vs_oracle_and_rejections(dev, extra_flags, frozen, light_power):
    """The field's five emitters differ in area, so picking them by emission
    strength (PULSE_FLAG_LIGHT_POWER) gives each its own pick pdf; the
    replay holds it at its value, as does the oracle (its strength table is
    the scene's)."""
    scene, camf = P.icosphere_field(n_meshes=6, subdiv=3, n_instances=40, n_emissive=5, box=24.0, radius=(1.0, 3.0))
    cam = camf(64, 40)
    prm = make_params(spp=2, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=3,
                      flags=N.FLAG_HORIZON_CLAMP | extra_flags | light_power)
    dev.upload_scene(scene)
    nm = len(scene.buffers["materials"])
    adj = PulseAdjoint(dev, cam.width, cam.height, nm)
    img = adj.forward(cam, prm)
    target = (img * np.float32(1.3)).astype(np.float32)
    adj.set_target(target)
    g = adj.backward(cam, prm)
    o = OracleScene(scene)
    mp = material_params(scene.buffers["materials"])
    gref, _ = o.grad_fd(cam, prm, target, mp.reshape(-1), np.arange(mp.size, dtype=np.uint32), h=1e-6,
                        frozen=frozen)
    gref = gref.reshape(nm, 7)
    scale = np.abs(gref).max()
    print("field: max err / scale", float((np.abs(g - gref) / scale).max()), flush=True)
    assert (np.abs(g - gref) <= 2e-3 * np.abs(gref) + 1e-4 * scale).all()
    with pytest.raises(N.PulseError, match="UNSUPPORTED"):
        adj.backward(cam, make_params(spp=2, max_depth=8, mode=N.MODE_REF_NEE_LAMBERT, rr_start_depth=3))
    with pytest.raises(N.PulseError, match="UNSUPPORTED"):
        adj.backward(cam, make_params(spp=2, max_depth=9, mode=N.MODE_GGX_NEE, rr_start_depth=3,
                                      flags=N.FLAG_HORIZON_CLAMP))


@pytest.mark.parametrize("loop", [1, 2])
def test_accumulated_image_gradient(dev, loop):
    """Three pulse_trace calls accumulate samples 0-1, 2-3 and 4-5 into one
    image; the adjoint with the last call's parameters (accumulation_count
    4, spp 2) replays all six samples: its replay equals the accumulated
    image bit for bit and its gradient equals the oracle's central
    differences of the 6-sample image along the same decisions."""
    scene, cam, prm = _cornell_case(spp=2, depth=3, rr=1)
    dev.upload_scene(scene)
    dev.set_option(N.OPT_TRAVERSAL_LOOP, loop)
    try:
        nm = len(scene.buffers["materials"])
        adj = PulseAdjoint(dev, cam.width, cam.height, nm)
        for acc in (0, 2, 4):
            prm.accumulation_count = acc
            img = adj.forward(cam, prm)
        target = (img * np.float32(0.6) + np.float32(0.05)).astype(np.float32)
        adj.set_target(target)
        g, replay = adj.backward(cam, prm, want_replay=True)
        assert same_bits(replay, img).all()
        whole = make_params(spp=6, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=1, flags=N.FLAG_HORIZON_CLAMP)
        o = OracleScene(scene)
        oimg, _, _ = o.render(cam, whole)
        assert same_bits(oimg, img).all()
        mp = material_params(scene.buffers["materials"])
        gref, _ = o.grad_fd(cam, whole, target, mp.reshape(-1), np.arange(mp.size, dtype=np.uint32), h=1e-6,
                            frozen=True)
        gref = gref.reshape(nm, 7)
        scale = np.abs(gref).max()
        assert (np.abs(g - gref) <= 2e-3 * np.abs(gref) + 1e-4 * scale).all(), (g, gref)
    finally:
        dev.set_option(N.OPT_TRAVERSAL_LOOP, 1)


def _field_fd_case():
    scene, camf = P.icosphere_field(n_meshes=8, subdiv=4, n_instances=60, n_emissive=6, box=30.0, radius=(1.5, 3.5))
    return scene, camf


def test_gradient_vs_central_fd_of_gpu_renders(dev):
    """Central FD (h = 1e-3) of the f32 GPU loss, decisions free to change,
    on 16 random albedo and emission parameters (the sampled directions do
    not depend on them): relative error <= 1e-2."""
    scene, camf = _field_fd_case()
    cam = camf(128, 96)
    prm = make_params(spp=8, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=3, flags=N.FLAG_HORIZON_CLAMP)
    dev.upload_scene(scene)
    mats = scene.buffers["materials"]
    nm = len(mats)
    adj = PulseAdjoint(dev, cam.width, cam.height, nm)
    mp = material_params(mats)
    rng = np.random.default_rng(0)
    hidden = mp.copy()
    hidden[:, 0:3] = np.clip(hidden[:, 0:3] + rng.uniform(-0.3, 0.3, (nm, 3)), 0.05, 0.95)
    hidden[:, 3] = np.clip(hidden[:, 3] + rng.uniform(-0.3, 0.3, nm), 0.05, 1.0)
    adj.set_materials(apply_material_params(mats, hidden))
    target = adj.forward(cam, prm)
    adj.set_target(target)
    adj.set_materials(mats)
    adj.forward(cam, prm)
    g = adj.backward(cam, prm).reshape(-1)
    cand = np.array([7 * m + j for m in range(nm) for j in (0, 1, 2, 4, 5, 6)])
    ids = rng.choice(cand, 16, replace=False)
    h = 1e-3
    fd = []
    for pid in ids:
        vals = []
        for sgn in (1, -1):
            p = mp.copy().reshape(-1)
            p[pid] += sgn * h
            adj.set_materials(apply_material_params(mats, p.reshape(nm, 7)))
            adj.forward(cam, prm)
            vals.append(adj.loss())
        fd.append((vals[0] - vals[1]) / (2 * h))
    adj.set_materials(mats)
    fd = np.array(fd)
    scale = np.abs(g).max()
    rel = np.abs(g[ids] - fd) / np.maximum(np.abs(fd), 1e-2 * scale)
    print("ids", ids.tolist(), "rel", rel.tolist(), flush=True)
    assert (rel <= 1e-2).all()


def test_roughness_derivative_vs_free_fd(dev):
    """The roughness derivative against central FD of GPU renders whose
    sampled directions and visibility follow the perturbed roughness.  A
    statistical check: FD of a Monte Carlo render in roughness is noisy
    (paths jump across silhouettes), so it uses a linear functional of the
    image (fixed random pixel weights), 256x192 at 512 spp, h = 1e-2, the
    materials with roughness >= 0.3 (near-mirror lobes make the FD itself
    heavy-tailed), and bounds the summed error: sum |g - fd| <= 0.25 sum |fd|,
    with matching signs where |fd| >= 10 % of the largest."""
    scene, camf = _field_fd_case()
    cam = camf(256, 192)
    prm = make_params(spp=512, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=3, flags=N.FLAG_HORIZON_CLAMP)
    dev.upload_scene(scene)
    mats = scene.buffers["materials"]
    nm = len(mats)
    adj = PulseAdjoint(dev, cam.width, cam.height, nm)
    mp = material_params(mats)
    weights = np.random.default_rng(1).uniform(0.5, 1.5, (cam.height, cam.width, 4)).astype(np.float32)
    adj.set_target(np.zeros_like(weights))
    adj.forward(cam, prm)
    g = adj.backward(cam, prm, image=weights)
    ids = [7 * m + 3 for m in range(nm) if mp[m, 3] >= 0.3]
    assert len(ids) >= 5
    h = 1e-2

    def functional(p):
        adj.set_materials(apply_material_params(mats, p.reshape(nm, 7)))
        img = adj.forward(cam, prm)
        return 2.0 / (3 * cam.width * cam.height) * float(np.sum(weights[..., :3].astype(np.float64) * img[..., :3]))

    fd = []
    for pid in ids:
        pp, pm = mp.copy().reshape(-1), mp.copy().reshape(-1)
        pp[pid] += h
        pm[pid] -= h
        fd.append((functional(pp) - functional(pm)) / (2 * h))
    adj.set_materials(mats)
    fd = np.array(fd)
    gr = g.reshape(-1)[ids]
    print("adj", gr.tolist(), "fd", fd.tolist(), flush=True)
    assert np.abs(gr - fd).sum() <= 0.25 * np.abs(fd).sum()
    big = np.abs(fd) >= 0.1 * np.abs(fd).max()
    assert (np.sign(gr[big]) == np.sign(fd[big])).all()
