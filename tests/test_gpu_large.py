"""GPU: the large workload's shape.

Forward: 256 displaced subdivision-6 icosphere BLAS (21M unique
triangles) under 1024 instances with 256 emissive ones, 3840x2160, 6
bounces with NEE, progressive accumulation to 64 samples per pixel: 64
frames with the reference's traversal (loop 1), one 64-sample call with the
4-wide BVHs (loop 2; equal to 64 frames by the running mean's definition).
Both equal the oracle's 64-frame progressive render of sampled rows bit for
bit.

Per-instance parameters on the same geometry (one material per instance):
the albedo gradient against central differences (h = 1e-3) of the GPU loss
on 16 parameters, relative error <= 1e-2.

Adjoint: one material per instance (1024 instances, 1025 materials x 7
parameters: the gradient buffer exceeds the LDS and goes to device
atomics, summed over each wave per material first), 6-bounce paths, 8
views in one batched call (pulse_adjoint_views).  Checked by identities
that need no reference render: with target 0, Euler's identity for the
emission (the image is linear in all emissions together, so
sum_j Le_j dL/dLe_j = 2 L) summed over the views; and the batched gradient
for nonzero targets equals the sum of the per-view gradients (pulse_adjoint
per view), as does the per-view calls' accumulation with
PULSE_FLAG_ADJOINT_ACCUMULATE.  The deep-path adjoint
against the oracle's central differences is in test_gpu_adjoint.py (the
Cornell box at 6 and 8 bounces, where the f32 and f64 renders agree along
the same decisions; in these dense fields some pixels do not, so f64
differences are no reference there).
"""
import numpy as np
import pytest

from oracle_lib import OracleScene, same_bits
from pulse_amd import _native as N
from pulse_amd import procedural as P
from pulse_amd.path_tracer import (Device, MultiViewAdjoint, PulseAdjoint, PulsePathTracerNode, apply_material_params,
                                   make_params, material_params)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def large():
    return P.icosphere_field(n_meshes=256, subdiv=6, n_instances=1024, n_emissive=256, box=110.0, seed=5,
                             threads=16)


def test_large_progressive_rows(dev, large):
    scene, camf = large
    assert scene.info["unique_triangles"] >= 16_000_000
    cam = camf(3840, 2160)
    dev.upload_scene(scene)
    rows = [0, 517, 1080, 1081, 1702, 2159]
    frames = 64
    images = {}
    for loop in (1, 2):
        dev.set_option(N.OPT_TRAVERSAL_LOOP, loop)
        node = PulsePathTracerNode(dev, cam.width, cam.height)
        if loop == 1:
            for f in range(frames):
                node.run(cam, make_params(accumulation_count=f, max_depth=6, mode=N.MODE_GGX_NEE))
        else:
            node.run(cam, make_params(accumulation_count=0, spp=frames, max_depth=6, mode=N.MODE_GGX_NEE))
        assert dev.stats().error_flags == 0
        images[loop] = node.readback()
        print(f"loop {loop}: {frames} samples rendered", flush=True)
    dev.set_option(N.OPT_TRAVERSAL_LOOP, 2)
    orc = OracleScene(scene)
    acc = np.zeros((cam.height, cam.width, 4), np.float32)
    for r in rows:
        for f in range(frames):
            orc.render(cam, make_params(accumulation_count=f, max_depth=6, mode=N.MODE_GGX_NEE), accum=acc,
                       threads=16, rows=(r, r + 1))
    o = acc[rows]
    # the reference's estimator is unbounded (GGX weights, Q6) and rand_f can
    # reach 1.0000005 (Q1): a few of the 64-sample means are not finite
    fin = np.isfinite(o[..., :3]).all(axis=-1)
    assert fin.mean() >= 0.999 and o[fin][:, :3].mean() > 0.0
    assert same_bits(images[1][rows], o).all()
    px = ~same_bits(images[2][rows], o).all(axis=-1)
    assert not px.any(), f"{px.sum()} of {px.size} pixels differ"


@pytest.fixture(scope="module")
def large_per_instance():
    return P.icosphere_field(n_meshes=256, subdiv=6, n_instances=1024, n_emissive=256, box=110.0, seed=5,
                             threads=16, roughness=(0.3, 1.0), per_instance_materials=True)


def test_large_per_instance_albedo_vs_fd(dev, large_per_instance):
    scene, camf = large_per_instance
    mats = scene.buffers["materials"]
    nm = len(mats)
    assert nm == 1025
    cam = camf(192, 128)
    prm = make_params(spp=4, max_depth=6, mode=N.MODE_GGX_NEE, rr_start_depth=6, flags=N.FLAG_HORIZON_CLAMP)
    dev.upload_scene(scene)
    adj = PulseAdjoint(dev, cam.width, cam.height, nm)
    start = material_params(mats)
    rng = np.random.default_rng(7)
    hidden = start.copy()
    hidden[:, 0:3] = np.clip(start[:, 0:3] + rng.uniform(-0.3, 0.3, (nm, 3)), 0.02, 0.98)
    adj.set_materials(apply_material_params(mats, hidden))
    adj.set_target(adj.forward(cam, make_params(spp=16, max_depth=6, mode=N.MODE_GGX_NEE, rr_start_depth=6,
                                                flags=N.FLAG_HORIZON_CLAMP)))
    adj.set_materials(mats)
    adj.forward(cam, prm)
    g = adj.backward(cam, prm)
    scale = np.abs(g[:1024, 0:3]).max()
    cand = [7 * m + j for m in range(1024) for j in (0, 1, 2) if abs(g[m, j]) > 0.02 * scale]
    assert len(cand) >= 16
    ids = rng.choice(cand, 16, replace=False)
    h = 1e-3
    fd = []
    for pid in ids:
        vals = []
        for sgn in (1, -1):
            p = start.copy().reshape(-1)
            p[pid] += sgn * h
            adj.set_materials(apply_material_params(mats, p.reshape(nm, 7)))
            adj.forward(cam, prm)
            vals.append(adj.loss())
        fd.append((vals[0] - vals[1]) / (2 * h))
    adj.set_materials(mats)
    fd = np.array(fd)
    rel = np.abs(g.reshape(-1)[ids] - fd) / np.maximum(np.abs(fd), 1e-2 * scale)
    print("per-instance ids", ids.tolist(), "rel", np.round(rel, 6).tolist(), flush=True)
    assert (rel <= 1e-2).all()


@pytest.fixture(scope="module")
def per_instance():
    return P.icosphere_field(n_meshes=16, subdiv=3, n_instances=1024, n_emissive=64, box=60.0, seed=9,
                             radius=(1.0, 3.0), roughness=(0.3, 1.0), per_instance_materials=True)


def _views(camf):
    """8 cameras on a circle around the field, looking at its centre."""
    cams = []
    for k in range(8):
        a = 2.0 * np.pi * k / 8
        eye = (70.0 * np.cos(a), 12.0 * np.sin(3 * a), 70.0 * np.sin(a))
        cams.append(P.make_camera(80, 64, eye=eye, target=(0.0, 0.0, 0.0)))
    return cams


def test_multiview_adjoint_per_instance_parameters(dev, per_instance):
    scene, camf = per_instance
    nm = len(scene.buffers["materials"])
    assert nm == 1024 + 1  # one per instance and the shared emitter's
    dev.upload_scene(scene)
    prm = make_params(spp=2, max_depth=6, mode=N.MODE_GGX_NEE, rr_start_depth=6, flags=N.FLAG_HORIZON_CLAMP)
    cams = _views(camf)
    adj = PulseAdjoint(dev, cams[0].width, cams[0].height, nm)
    mv = MultiViewAdjoint(dev, cams, nm)
    mp = material_params(scene.buffers["materials"])
    # the default: the device-memory gradient sums each wave's lanes per material first
    assert dev.get_option(N.OPT_ADJOINT_REDUCE_ROUNDS) == -1

    def pass_over_views(targets):
        """Gradient of the 8 views from one batched call, the same from
        per-view calls accumulated with PULSE_FLAG_ADJOINT_ACCUMULATE, the
        per-view gradients and the per-view losses."""
        imgs = mv.forward(prm)
        tgts = [targets(img) for img in imgs]
        mv.set_targets(tgts)
        batched = mv.backward(prm).astype(np.float64)
        per_view, losses, acc = [], [], None
        for v, cam in enumerate(cams):
            adj.set_target(tgts[v])
            per_view.append(adj.backward(cam, prm, image=imgs[v]).astype(np.float64))
            if v > 0:
                adj.grad.copy_from(acc.astype(np.float32).reshape(-1))
                p = N.Params.from_buffer_copy(prm)
                p.flags |= N.FLAG_ADJOINT_ACCUMULATE  # add this view's gradient to the others'
                acc = adj.backward(cam, p, image=imgs[v])
            else:
                acc = per_view[0].astype(np.float32)
        return batched, acc.astype(np.float64), per_view, mv.losses()

    # Euler's identity for the emission, summed over the views
    batched, acc, per_view, losses = pass_over_views(lambda img: np.zeros_like(img))
    euler = float(np.sum(mp[:, 4:7] * batched[:, 4:7]))
    assert sum(losses) > 0.0
    assert abs(euler - 2.0 * sum(losses)) <= 1e-4 * 2.0 * sum(losses), (euler, 2.0 * sum(losses))
    # the batched gradient and the accumulated one: the sum of the per-view gradients
    batched, acc, per_view, _ = pass_over_views(
        lambda img: (img * np.float32(0.7) + np.float32(0.02)).astype(np.float32))
    total = np.sum(per_view, axis=0)
    scale = np.abs(total).max()
    assert np.all(np.abs(batched - total) <= 1e-5 * scale + 1e-4 * np.abs(total))
    assert np.all(np.abs(acc - total) <= 1e-5 * scale + 1e-4 * np.abs(total))
    # every instance is a parameter: many of them receive gradient
    assert np.count_nonzero(total[:1024, 0]) > 200


def test_adjoint_depth_limits(dev, per_instance):
    scene, camf = per_instance
    nm = len(scene.buffers["materials"])
    dev.upload_scene(scene)
    cam = _views(camf)[0]
    adj = PulseAdjoint(dev, cam.width, cam.height, nm)
    ok = make_params(spp=1, max_depth=8, mode=N.MODE_GGX_NEE, rr_start_depth=8, flags=N.FLAG_HORIZON_CLAMP)
    adj.forward(cam, ok)
    adj.set_target(np.zeros((cam.height, cam.width, 4), np.float32))
    adj.backward(cam, ok)
    for bad in (make_params(spp=1, max_depth=9, mode=N.MODE_GGX_NEE, rr_start_depth=9),
                make_params(spp=1, max_depth=5, mode=N.MODE_GGX_NEE, rr_start_depth=5,
                            flags=N.FLAG_ADJOINT_REPARAM)):
        with pytest.raises(N.PulseError, match="UNSUPPORTED"):
            adj.backward(cam, bad)
