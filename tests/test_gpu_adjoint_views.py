"""GPU: the batched adjoint (pulse_adjoint_views) over several views of one
scene, the loss being the sum of the views' losses.

* Against the oracle: three Cornell-box views from different cameras at
  ragged resolutions (not multiples of a block's 128 pixels), one call: every
  view's replay equals its forward image bit for bit, and the gradient equals
  the sum over the views of the oracle's f64 central differences along the
  same decisions (all 7 parameters of every material), in both roughness
  modes.
* Against pulse_adjoint: 20 views (two launches of 16 and 4) of an icosphere
  field with one material per instance (the gradient in device memory) and
  of the Cornell box (the gradient in LDS) equal the sum of the per-view
  single-view gradients up to the order of the f32 additions, for per-lane
  atomics and for 1, 2 and 16 wave-summed materials per add.
* Euler's identity for the emission summed over the views; the accumulate
  flag; argument errors.
"""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import OracleScene, same_bits
from pulse_amd import _native as N
from pulse_amd import procedural as P
from pulse_amd.path_tracer import Device, MultiViewAdjoint, PulseAdjoint, make_params, material_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = Device(0)
    d.set_option(N.OPT_TRAVERSAL_LOOP, 1)  # the reference's traversal order: bit-exact vs the oracle
    yield d
    d.set_option(N.OPT_ADJOINT_REDUCE_ROUNDS, -1)
    d.close()


def _params(depth=3, spp=2, flags=0):
    return make_params(spp=spp, max_depth=depth, mode=N.MODE_GGX_NEE, rr_start_depth=depth,
                       flags=N.FLAG_HORIZON_CLAMP | flags)


def _cornell_views():
    eyes = [(0.0, 0.0, 3.0), (0.6, 0.3, 2.6), (-0.7, -0.2, 2.8)]
    sizes = [(37, 23), (48, 40), (21, 30)]
    return [P.make_camera(w, h, eye=e, target=(0.0, 0.0, 0.0)) for e, (w, h) in zip(eyes, sizes)]


def _targets(imgs, a=0.6, b=0.05):
    out = []
    for img in imgs:
        t = (img * np.float32(a) + np.float32(b)).astype(np.float32)
        t[..., 3] = 1.0
        out.append(t)
    return out


@pytest.mark.parametrize("record", [False, True], ids=["traced", "recorded"])
@pytest.mark.parametrize("extra_flags,frozen", [pytest.param(0, True, id="detached"),
                                                pytest.param(N.FLAG_ADJOINT_REPARAM, False, id="reparam")])
def test_views_vs_oracle(dev, extra_flags, frozen, record):
    """`recorded`: the views' forward passes keep their paths
    (PULSE_FLAG_RECORD_PATHS, one record per camera) and the batched adjoint
    reads them for every view instead of tracing."""
    scene, _ = P.cornell()
    dev.upload_scene(scene)
    nm = len(scene.buffers["materials"])
    cams = _cornell_views()
    prm = _params(flags=extra_flags)
    mv = MultiViewAdjoint(dev, cams, nm)
    imgs = mv.forward(_params(flags=extra_flags | (N.FLAG_RECORD_PATHS if record else 0)))
    targets = _targets(imgs)
    mv.set_targets(targets)
    g, replays = mv.backward(prm, want_replay=True)
    assert dev.get_option(N.OPT_ADJOINT_READ_RECORD) == (1 if record else 0)
    for v, (img, rep) in enumerate(zip(imgs, replays)):
        assert same_bits(rep, img).all(), f"view {v}: replay differs from the forward image"
    o = OracleScene(scene)
    mp = material_params(scene.buffers["materials"])
    ids = np.arange(mp.size, dtype=np.uint32)
    gref = np.zeros(mp.size)
    for cam, tgt, loss in zip(cams, targets, mv.losses()):
        gv, loss_ref = o.grad_fd(cam, prm, tgt, mp.reshape(-1), ids, h=1e-6, frozen=frozen)
        assert abs(loss - loss_ref) <= 1e-4 * loss_ref
        gref += gv
    gref = gref.reshape(nm, 7)
    scale = np.abs(gref).max()
    print("views: max err / scale", float((np.abs(g - gref) / scale).max()), flush=True)
    assert (np.abs(g - gref) <= 2e-3 * np.abs(gref) + 1e-4 * scale).all(), (g, gref)


def _sum_of_single_views(dev, cams, imgs, targets, prm, nm):
    total = np.zeros((nm, 7))
    for cam, img, tgt in zip(cams, imgs, targets):
        adj = PulseAdjoint(dev, cam.width, cam.height, nm)
        adj.set_target(tgt)
        total += adj.backward(cam, prm, image=img).astype(np.float64)
    return total


def _field_views(camf_scale=70.0, n=20):
    cams = []
    for k in range(n):
        a = 2.0 * np.pi * k / n
        eye = (camf_scale * np.cos(a), 10.0 * np.sin(3 * a), camf_scale * np.sin(a))
        w, h = [(64, 48), (40, 40), (33, 17), (96, 64)][k % 4]
        cams.append(P.make_camera(w, h, eye=eye, target=(0.0, 0.0, 0.0)))
    return cams


@pytest.fixture(scope="module")
def per_instance_field():
    return P.icosphere_field(n_meshes=8, subdiv=3, n_instances=600, n_emissive=40, box=50.0, seed=5,
                             radius=(1.0, 3.0), roughness=(0.3, 1.0), per_instance_materials=True)


@pytest.mark.parametrize("case", ["per_instance_global", "cornell_lds"])
def test_views_equal_sum_of_single_views(dev, per_instance_field, case):
    if case == "per_instance_global":
        scene, _ = per_instance_field
        cams = _field_views()
        prm = _params(depth=4, spp=2)
    else:
        scene, _ = P.cornell()
        cams = [P.make_camera(w, h, eye=(0.3 * np.sin(k), 0.2 * np.cos(k), 2.8), target=(0.0, 0.0, 0.0))
                for k, (w, h) in enumerate([(32, 32), (19, 27), (48, 16), (8, 8), (1, 1)] * 4)]
        prm = _params(depth=3, spp=2)
    assert len(cams) == 20  # two launches: 16 views and 4
    dev.upload_scene(scene)
    nm = len(scene.buffers["materials"])
    assert (nm * 7 * 4 > 16384) == (case == "per_instance_global")
    mv = MultiViewAdjoint(dev, cams, nm)
    imgs = mv.forward(prm)
    targets = _targets(imgs, 0.7, 0.02)
    mv.set_targets(targets)
    dev.set_option(N.OPT_ADJOINT_REDUCE_ROUNDS, 0)
    total = _sum_of_single_views(dev, cams, imgs, targets, prm, nm)
    scale = np.abs(total).max()
    assert scale > 0
    for rounds in (0, 1, 2, 16):
        dev.set_option(N.OPT_ADJOINT_REDUCE_ROUNDS, rounds)
        g = mv.backward(prm).astype(np.float64)
        err = np.abs(g - total)
        print(f"{case} rounds {rounds}: max err / scale {float((err / scale).max()):.3g}", flush=True)
        assert (err <= 1e-5 * scale + 1e-4 * np.abs(total)).all()
    dev.set_option(N.OPT_ADJOINT_REDUCE_ROUNDS, -1)
    if case == "per_instance_global":
        assert np.count_nonzero(total[:600, 0]) > 100  # many per-instance parameters receive gradient


def test_views_euler_and_accumulate(dev, per_instance_field):
    scene, _ = per_instance_field
    cams = _field_views(n=8)
    prm = _params(depth=4, spp=2)
    dev.upload_scene(scene)
    nm = len(scene.buffers["materials"])
    mv = MultiViewAdjoint(dev, cams, nm)
    imgs = mv.forward(prm)
    mv.set_targets([np.zeros_like(i) for i in imgs])
    g = mv.backward(prm).astype(np.float64)
    losses = mv.losses()
    mp = material_params(scene.buffers["materials"])
    euler = float(np.sum(mp[:, 4:7] * g[:, 4:7]))
    assert sum(losses) > 0
    assert abs(euler - 2.0 * sum(losses)) <= 1e-4 * 2.0 * sum(losses), (euler, 2.0 * sum(losses))
    # accumulate: a second call adds the same gradient again
    p = N.Params.from_buffer_copy(prm)
    p.flags |= N.FLAG_ADJOINT_ACCUMULATE
    g2 = mv.backward(p).astype(np.float64)
    scale = np.abs(g).max()
    assert (np.abs(g2 - 2.0 * g) <= 1e-5 * scale + 1e-5 * np.abs(g)).all()


def test_views_argument_errors(dev):
    scene, _ = P.cornell()
    dev.upload_scene(scene)
    lib = N.lib()
    cams = _cornell_views()[:1]
    mv = MultiViewAdjoint(dev, cams, len(scene.buffers["materials"]))
    prm = _params()
    assert lib.pulse_adjoint_views(dev.handle, mv.views, 0, C.byref(prm), mv.grad.ptr) == -1  # PULSE_E_INVALID_ARG
    assert lib.pulse_adjoint_views(dev.handle, None, 1, C.byref(prm), mv.grad.ptr) == -1  # PULSE_E_INVALID_ARG
    bad = (N.AdjointView * 1)()
    bad[0].camera = cams[0]
    bad[0].image = mv.images[0].ptr  # no target
    assert lib.pulse_adjoint_views(dev.handle, bad, 1, C.byref(prm), mv.grad.ptr) == -1  # PULSE_E_INVALID_ARG
    bad[0].target = mv.targets[0].ptr
    bad[0].camera.width = 0
    assert lib.pulse_adjoint_views(dev.handle, bad, 1, C.byref(prm), mv.grad.ptr) == -1  # PULSE_E_INVALID_ARG
    ref = make_params(spp=2, max_depth=9, mode=N.MODE_GGX_NEE, rr_start_depth=3)
    assert lib.pulse_adjoint_views(dev.handle, mv.views, 1, C.byref(ref), mv.grad.ptr) == -7  # PULSE_E_UNSUPPORTED
    for bad_rounds in (17, -2):
        with pytest.raises(N.PulseError, match="INVALID_ARG"):
            dev.set_option(N.OPT_ADJOINT_REDUCE_ROUNDS, bad_rounds)
    assert dev.get_option(N.OPT_ADJOINT_REDUCE_ROUNDS) == -1
