"""CPU pins of the oracle's gradient (f64 central differences along the f32
decisions), the reference the GPU adjoint is checked against."""
import numpy as np

from oracle_lib import OracleScene
from pulse_amd import _native as N
from pulse_amd import procedural as P
from pulse_amd.path_tracer import make_params, material_params


def _setup(spp=2, depth=3):
    scene, camf = P.cornell()
    cam = camf(24, 24)
    prm = make_params(spp=spp, max_depth=depth, mode=N.MODE_GGX_NEE, rr_start_depth=depth,
                      flags=N.FLAG_HORIZON_CLAMP)
    return scene, cam, prm


def test_fd_step_independence():
    """Along fixed decisions the loss is smooth: h = 1e-6 and 1e-7 agree.  (The
    reference's specular weight F G (o.m) / (F0max (n.o)(n.i)(n.m)) has no
    rejection of grazing n.i (Q6), so single paths can make the roughness
    derivative large and FD is linear only for small h.)"""
    scene, cam, prm = _setup()
    o = OracleScene(scene)
    img, _, _ = o.render(cam, prm)
    target = (img * 0.7).astype(np.float32)
    mp = material_params(scene.buffers["materials"])
    ids = np.arange(mp.size, dtype=np.uint32)
    g1, loss = o.grad_fd(cam, prm, target, mp.reshape(-1), ids, h=1e-6)
    g2, _ = o.grad_fd(cam, prm, target, mp.reshape(-1), ids, h=1e-7)
    scale = np.abs(g1).max()
    assert scale > 0 and loss > 0
    np.testing.assert_allclose(g1, g2, atol=1e-6 * scale, rtol=1e-4)


def test_emission_euler_identity():
    """With target 0, the image is homogeneous of degree 1 in the emissive
    parameters (emission is added at depth 0 and through NEE, each term
    linear in one emitter), so sum_j Le_j dL/dLe_j = 2 L."""
    scene, cam, prm = _setup()
    o = OracleScene(scene)
    target = np.zeros((cam.height, cam.width, 4), np.float32)
    mp = material_params(scene.buffers["materials"])
    ids = np.array([7 * m + 4 + c for m in range(len(mp)) for c in range(3)], np.uint32)
    g, loss = o.grad_fd(cam, prm, target, mp.reshape(-1), ids, h=1e-6)
    euler = float(np.sum(mp.reshape(-1)[ids] * g))
    assert abs(euler - 2 * loss) <= 1e-6 * loss


def test_albedo_gradient_sign_and_zero_for_unseen():
    """Raising every albedo brightens the image: with a target of 0 all albedo
    derivatives are >= 0; a material absent from the image has a zero row."""
    scene, cam, prm = _setup()
    o = OracleScene(scene)
    target = np.zeros((cam.height, cam.width, 4), np.float32)
    mp = material_params(scene.buffers["materials"])
    ids = np.array([7 * m + c for m in range(len(mp)) for c in range(3)], np.uint32)
    g, _ = o.grad_fd(cam, prm, target, mp.reshape(-1), ids, h=1e-6)
    assert (g >= -1e-12).all() and g.max() > 0


def test_horizon_clamp_bounds_the_estimator():
    """Without the clamp the reference's GGX estimator keeps below-horizon
    samples (Q6) and produces unbounded (and negative) pixel values; with it
    the image is non-negative and its mean is stable across sample counts."""
    scene, camf = P.icosphere_field(n_meshes=6, subdiv=3, n_instances=40, n_emissive=5, box=24.0, radius=(1.0, 3.0))
    o = OracleScene(scene)
    cam = camf(64, 48)
    means = []
    for spp in (2, 16):
        img, _, _ = o.render(cam, make_params(spp=spp, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=3,
                                              flags=N.FLAG_HORIZON_CLAMP))
        assert (img[..., :3] >= 0).all() and np.isfinite(img).all()
        means.append(img[..., :3].mean())
    assert abs(means[0] - means[1]) < 0.1 * means[1]
