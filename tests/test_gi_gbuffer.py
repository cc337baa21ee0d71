"""The realtime GI pass over a G-buffer (src/pulse/gi.wgsl:56-126):
CPU pins of the oracle and GPU parity (pulse_trace_gbuffer)."""
import numpy as np
import pytest

from oracle_lib import OracleScene, same_bits
from pulse_amd import _native as N
from pulse_amd import procedural as P
from pulse_amd.path_tracer import make_params
from pulse_amd.scene import MaterialDesc, Mesh, SceneBuilder


def gbuffer(w, h, seed, lo=-0.8, hi=0.8):
    rng = np.random.default_rng(seed)
    pos = np.zeros((h, w, 4), np.float32)
    pos[..., :3] = rng.uniform(lo, hi, (h, w, 3))
    n = rng.normal(size=(h, w, 3))
    n /= np.linalg.norm(n, axis=2, keepdims=True)
    nrm = np.zeros((h, w, 4), np.float32)
    nrm[..., :3] = n
    base = np.zeros((h, w, 4), np.float32)
    base[..., :3] = rng.uniform(0.2, 0.8, (h, w, 3))
    emis = np.zeros((h, w, 4), np.float32)
    emis[::7, ::5, :3] = 2.0
    return {"position": pos, "normal": nrm, "base_color": base, "emissive": emis}


def _dark_scene():
    sb = SceneBuilder()
    walls = sb.add_mesh(P.cornell_walls_mesh())
    m = sb.add_material(MaterialDesc(base_color=(0.7, 0.6, 0.5, 1)))
    sb.add_instance(walls, m, np.eye(4))
    return sb.build(1)


@pytest.mark.parametrize("spp,depth", [(1, 0), (2, 5), (4, 3)])
def test_gi_without_lights_is_the_gbuffer_emission(spp, depth):
    """No emitters: no NEE term, and the GI pass adds no emission of hit
    surfaces (gi.wgsl:103-107), so every pixel is exactly its G-buffer
    emissive value (spp copies of E / spp sum to E for spp = 1, 2, 4)."""
    o = OracleScene(_dark_scene())
    g = gbuffer(24, 16, 1)
    out = o.render_gi(g, make_params(spp=spp, max_depth=depth))
    np.testing.assert_array_equal(out[..., :3], g["emissive"][..., :3])
    assert (out[..., 3] == 1).all()


def test_gi_direct_light_closed_form_and_rng_continuation():
    """max_depth 0: the pixel value is emission + Lambert NEE at the G-buffer
    point, averaged over the samples; over 10k pixels sharing one receiver
    it matches the integral of the same integrand over a small emitter
    triangle straight above (Q3 divisor 4).  The samples of a pixel draw
    from one RNG stream: with 2 samples the values differ from 1 sample."""
    h = 2.0
    tri = np.array([[0.0, h, 0.0], [0.05, h, 0.0], [0.0, h, 0.05]], np.float32)
    sb = SceneBuilder()
    mid = sb.add_mesh(Mesh(tri, np.tile(np.float32([0, -1, 0]), (3, 1)), np.array([0, 2, 1], np.uint32)))
    mat = sb.add_material(MaterialDesc(emissive=(10.0, 10.0, 10.0, 1)))
    sb.add_instance(mid, mat, np.eye(4))
    o = OracleScene(sb.build(1))
    w = hh = 100
    g = {k: np.zeros((hh, w, 4), np.float32) for k in ("position", "normal", "base_color", "emissive")}
    g["position"][..., :3] = [0.01, 0.0, 0.01]
    g["normal"][..., 1] = 1.0
    g["base_color"][..., :3] = 0.5
    one = o.render_gi(g, make_params(spp=1, max_depth=0))
    two = o.render_gi(g, make_params(spp=2, max_depth=0))
    p0, n0 = np.array([0.01, 0, 0.01]), np.array([0, 1.0, 0])
    area = 0.5 * 0.05 * 0.05
    uv = np.random.default_rng(0).random((400000, 2))
    uv = np.where(uv.sum(1, keepdims=True) < 1, uv, 1 - uv)
    pts = tri[0] + uv[:, :1] * (tri[1] - tri[0]) + uv[:, 1:] * (tri[2] - tri[0])
    d = pts - (p0 + 0.001 * n0)
    dn = d / np.linalg.norm(d, axis=1, keepdims=True)
    exact = (0.5 / np.pi * 10.0 * (dn @ n0) * dn[:, 1] * area / 4.0).mean()
    for img in (one, two):
        assert abs(img[..., 0].mean() - exact) / exact < 0.01
    assert (one[..., 0] != two[..., 0]).mean() > 0.9


@pytest.mark.gpu
@pytest.mark.parametrize("which,spp,depth,flags", [("cornell_lights", 2, 5, 0), ("cornell_lights", 1, 3, 16),
                                                   ("field", 2, 5, 0), ("field", 3, 2, 16)])
def test_gi_gpu_bit_exact(which, spp, depth, flags):
    from pulse_amd.path_tracer import Device

    if which == "field":
        scene, _ = P.icosphere_field(n_meshes=6, subdiv=3, n_instances=40, n_emissive=5, box=24.0,
                                     radius=(1.0, 3.0))
        g = gbuffer(72, 40, 3, -12.0, 12.0)
    else:
        scene, _ = P.cornell_lights()
        g = gbuffer(72, 40, 2)
    prm = make_params(spp=spp, max_depth=depth, flags=flags)
    want = OracleScene(scene).render_gi(g, prm)
    dev = Device(0)
    try:
        dev.upload_scene(scene)
        d = {k: dev.to_device(v) for k, v in g.items()}
        gb = N.GBuffer(d["position"].ptr.value, d["normal"].ptr.value, d["base_color"].ptr.value,
                       d["emissive"].ptr.value)
        out = dev.alloc(72 * 40 * 16)
        # traversal loop 1 (the reference's order) bit for bit, with the
        # shadow launches beside the extension launches (the default) or in
        # sequence; loop 2 (4-wide BVHs) the same up to pixels whose path met
        # a t-tie
        for loop, persistent, conc in ((1, 24, 1), (1, 24, 0), (1, 0, 1), (2, 24, 1)):
            dev.set_option(N.OPT_TRAVERSAL_LOOP, loop)
            dev.upload_scene(scene)
            dev.set_option(N.OPT_PERSISTENT_WAVES, persistent)
            dev.set_option(N.OPT_CONCURRENT_SHADOW, conc)
            N.check(N.lib().pulse_trace_gbuffer(dev.handle, gb, 72, 40, prm, out.ptr))
            got = out.copy_to(np.zeros((40, 72, 4), np.float32)) if dev.sync() is None else None
            if loop == 1:
                assert same_bits(got, want).all(), persistent
            else:
                assert (~same_bits(got, want).all(axis=-1)).mean() <= 1e-3
        with pytest.raises(N.PulseError, match="UNSUPPORTED"):
            N.check(N.lib().pulse_trace_gbuffer(dev.handle, gb, 72, 40, make_params(nee_eval=N.NEE_GGX), out.ptr))
    finally:
        dev.close()
