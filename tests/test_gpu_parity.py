"""GPU parity: the HIP kernels (through the C ABI) against the f32 CPU oracle.

The kernels and the oracle evaluate the reference's f32 arithmetic in the
same operation order with IEEE division/sqrt and the same transcendentals, so
closest-hit ids, t, u, v, occlusion and whole accumulated images are compared
bit for bit (NaN == NaN).
"""
import numpy as np
import pytest

from oracle_lib import OracleScene, random_rays, same_bits
from pulse_amd import _native as N
from pulse_amd import procedural as P
from pulse_amd.path_tracer import Device, PulsePathTracerNode, make_params

pytestmark = pytest.mark.gpu

DEFAULT_PERSISTENT, DEFAULT_LDS = 24, 24
AUTO = 0  # refill, gather and advance period: the traversal loop's tuned values

MODES = [(N.MODE_REF_GGX, N.NEE_LAMBERT, 5), (N.MODE_REF_NEE_LAMBERT, N.NEE_LAMBERT, 5),
         (N.MODE_GGX_NEE, N.NEE_LAMBERT, 4), (N.MODE_GGX_NEE, N.NEE_GGX, 4)]


@pytest.fixture(scope="module")
def dev():
    d = Device(0)
    d.set_option(N.OPT_TRAVERSAL_LOOP, 1)  # the reference's traversal order: bit-exact vs the oracle
    yield d
    d.close()


@pytest.fixture(scope="module")
def cornell():
    return P.cornell()


@pytest.fixture(scope="module")
def field():
    return P.icosphere_field(n_meshes=8, subdiv=3, n_instances=60, n_emissive=6, box=30.0, radius=(1.0, 3.0))


def _hits_equal(a, b):
    ok = same_bits(a["t"], b["t"]) & same_bits(a["u"], b["u"]) & same_bits(a["v"], b["v"])
    ok &= (a["instance_index"] == b["instance_index"]) & (a["triangle_index"] == b["triangle_index"])
    return ok


@pytest.mark.parametrize("which", ["cornell", "field"])
def test_trace_rays_bit_exact(dev, cornell, field, which):
    scene, _ = cornell if which == "cornell" else field
    rng = np.random.default_rng(1)
    rays = random_rays(20000, rng, (0, 0, 0), 0.95 if which == "cornell" else 20.0)
    dev.upload_scene(scene)
    g = dev.trace_rays(rays)
    o, _ = OracleScene(scene).trace_rays(rays)
    ok = _hits_equal(g, o)
    assert ok.all(), f"{(~ok).sum()} of {len(ok)} hits differ; first {np.flatnonzero(~ok)[:5]}"
    assert (g["t"] < 1e30).mean() > 0.3


def test_trace_rays_axis_parallel_and_tmax(dev, cornell):
    """Axis-parallel directions (slab 0*inf = NaN cases) and finite t_max."""
    scene, _ = cornell
    rng = np.random.default_rng(2)
    n = 4000
    rays = random_rays(n, rng, (0, 0, 0), 0.9)
    axes = np.eye(3, dtype=np.float32)
    rays["dir"][: n // 2] = axes[rng.integers(0, 3, n // 2)] * rng.choice([-1.0, 1.0], (n // 2, 1)).astype(np.float32)
    rays["t_max"][n // 2:] = rng.uniform(0.1, 2.0, n - n // 2).astype(np.float32)
    dev.upload_scene(scene)
    g = dev.trace_rays(rays)
    o, _ = OracleScene(scene).trace_rays(rays)
    assert _hits_equal(g, o).all()


@pytest.mark.parametrize("which", ["cornell", "field"])
def test_shadow_rays_match_closest_hit_threshold(dev, cornell, field, which):
    scene, _ = cornell if which == "cornell" else field
    rng = np.random.default_rng(3)
    rays = random_rays(20000, rng, (0, 0, 0), 0.95 if which == "cornell" else 20.0)
    orc = OracleScene(scene)
    hits, _ = orc.trace_rays(rays)
    t = np.where(hits["t"] < 1e30, hits["t"], 10.0).astype(np.float32)
    thr = t * rng.uniform(0.5, 1.5, len(t)).astype(np.float32)
    thr[::7] = t[::7]                      # exact ties: t < t is false
    thr[1::7] = np.nextafter(t[1::7], np.float32(np.inf))
    thr[2::13] = -1.0
    rays["t_max"] = thr
    dev.upload_scene(scene)
    g = dev.trace_shadow_rays(rays)
    o = orc.shadow_rays(rays)
    assert (g == o).all(), f"{(g != o).sum()} occlusion results differ"
    assert 0.05 < g.mean() < 0.95


def _render_gpu(dev, scene, camera, params):
    dev.upload_scene(scene)
    node = PulsePathTracerNode(dev, camera.width, camera.height)
    node.run(camera, params)
    return node.readback()


@pytest.mark.parametrize("mode,nee,depth", MODES)
def test_render_cornell_bit_exact(dev, cornell, mode, nee, depth):
    scene, camf = cornell
    cam = camf(96, 80)
    prm = make_params(spp=2, max_depth=depth, mode=mode, nee_eval=nee)
    g = _render_gpu(dev, scene, cam, prm)
    o, _, _ = OracleScene(scene).render(cam, prm)
    ok = same_bits(g, o)
    assert ok.all(), f"{(~ok).sum()} of {ok.size} channels differ, max |d| {np.nanmax(np.abs(g - o))}"
    assert np.nanmean(g[..., :3]) > 0.01


@pytest.mark.parametrize("mode,nee,depth", MODES)
def test_render_field_bit_exact(dev, field, mode, nee, depth):
    scene, camf = field
    cam = camf(120, 72)
    prm = make_params(spp=1, max_depth=depth, mode=mode, nee_eval=nee, accumulation_count=3)
    g = _render_gpu(dev, scene, cam, prm)
    o, _, _ = OracleScene(scene).render(cam, prm)
    ok = same_bits(g, o)
    assert ok.all(), f"{(~ok).sum()} of {ok.size} channels differ"


def test_accumulation_resume_equals_one_call(dev, cornell):
    """Three calls with accumulation counts 0, 1, 2 == one call with spp=3."""
    scene, camf = cornell
    cam = camf(64, 64)
    dev.upload_scene(scene)
    a = PulsePathTracerNode(dev, 64, 64)
    for n in range(3):
        a.run(cam, make_params(accumulation_count=n, spp=1, max_depth=4, mode=N.MODE_GGX_NEE))
    b = PulsePathTracerNode(dev, 64, 64)
    b.run(cam, make_params(accumulation_count=0, spp=3, max_depth=4, mode=N.MODE_GGX_NEE))
    assert same_bits(a.readback(), b.readback()).all()


def test_rr_start_depth_and_no_lights(dev):
    scene, camf = P.icosphere_field(n_meshes=4, subdiv=2, n_instances=30, n_emissive=0, box=20.0, emission=0.0)
    assert scene.light_count == 0
    cam = camf(48, 40)
    for prm in (make_params(max_depth=4, mode=N.MODE_GGX_NEE, rr_start_depth=2),
                make_params(max_depth=3, mode=N.MODE_REF_NEE_LAMBERT, rr_start_depth=1)):
        g = _render_gpu(dev, scene, cam, prm)
        o, _, _ = OracleScene(scene).render(cam, prm)
        assert same_bits(g, o).all()


def test_empty_scene_renders_black(dev):
    from pulse_amd.scene import SceneBuffers
    bufs = {name: np.zeros(0, dt) for name, dt in N.BUFFER_DTYPES.items()}
    bufs["tlas_nodes"] = np.zeros(1, N.TLAS_NODE_DTYPE)  # invalid root {0, 0} (tlas.rs:13-20)
    scene = SceneBuffers(0, 0, bufs)
    _, camf = P.cornell()
    cam = camf(32, 32)
    g = _render_gpu(dev, scene, cam, make_params(max_depth=5))
    assert (g[..., :3] == 0).all() and (g[..., 3] == 1).all()


def test_full_size_scene_frame(dev):
    """64 displaced subdivision-6 icospheres, 800 instances, 32 emitters at
    1920x1080 (the benchmark frame): the whole GPU frame equals the oracle's
    bit for bit."""
    scene, camf = P.icosphere_field()
    cam = camf(1920, 1080)
    prm = make_params(max_depth=4, mode=N.MODE_GGX_NEE, flags=N.FLAG_COLLECT_STATS)
    dev.upload_scene(scene)
    node = PulsePathTracerNode(dev, 1920, 1080)
    node.run(cam, prm)
    g = node.readback()
    st = dev.stats()
    assert st.error_flags == 0
    assert st.primary_rays == 1920 * 1080
    orc = OracleScene(scene)
    o, _, _ = orc.render(cam, prm, threads=16)
    ok = same_bits(g, o)
    assert ok.all(), f"{(~ok).sum()} channels differ, rows {np.unique(np.nonzero(~ok)[0])[:10]}"


@pytest.mark.parametrize("loop", [0, 1])
def test_traversal_variants_bit_exact(dev, field, loop):
    """Both traversal loops (nested as in the reference, single loop) give
    the oracle's hits, occlusion and image."""
    scene, camf = field
    dev.upload_scene(scene)
    dev.set_option(N.OPT_TRAVERSAL_LOOP, loop)
    try:
        rays = random_rays(8000, np.random.default_rng(9), (0, 0, 0), 20.0)
        g = dev.trace_rays(rays)
        orc = OracleScene(scene)
        o, _ = orc.trace_rays(rays)
        assert _hits_equal(g, o).all()
        cam = camf(80, 48)
        prm = make_params(spp=1, max_depth=4, mode=N.MODE_GGX_NEE)
        node = PulsePathTracerNode(dev, 80, 48)
        node.run(cam, prm)
        oi, _, _ = orc.render(cam, prm)
        assert same_bits(node.readback(), oi).all()
    finally:
        dev.set_option(N.OPT_TRAVERSAL_LOOP, 1)


@pytest.mark.parametrize("waves,refill,gather,lds,tps,period", [
    (0, 16, 1, 0, 1, 1), (4, 1, 1, 0, 1, 1), (16, 16, 1, 0, 2, 1), (16, 64, 8, 24, 1, 1), (64, 32, 64, 16, 2, 1),
    (16, 8, 16, 8, 1, 3), (8, 16, 32, 32, 1, 1), (24, 16, 1, 24, 2, 1), (24, 32, 64, 24, 2, 2)])
def test_persistent_traversal_bit_exact(dev, field, waves, refill, gather, lds, tps, period):
    """Persistent extend/shadow launches (lanes take new rays as theirs
    finish; instance changes gathered across lanes; the deep part of the
    stack in device memory) give the oracle's image for every grid size,
    refill and gather threshold and LDS stack size, in all three modes."""
    scene, camf = field
    dev.upload_scene(scene)
    dev.set_option(N.OPT_PERSISTENT_WAVES, waves)
    dev.set_option(N.OPT_REFILL_LANES, refill)
    dev.set_option(N.OPT_GATHER_LANES, gather)
    dev.set_option(N.OPT_LDS_STACK_ENTRIES, lds)
    dev.set_option(N.OPT_TRIS_PER_STEP, tps)
    dev.set_option(N.OPT_ADVANCE_PERIOD, period)
    try:
        assert dev.get_option(N.OPT_PERSISTENT_WAVES) == waves
        cam = camf(96, 56)
        orc = OracleScene(scene)
        for mode, rr in ((N.MODE_GGX_NEE, 1), (N.MODE_REF_GGX, 0), (N.MODE_REF_NEE_LAMBERT, 2)):
            prm = make_params(spp=2, max_depth=4, mode=mode, rr_start_depth=rr)
            node = PulsePathTracerNode(dev, cam.width, cam.height)
            node.run(cam, prm)
            oi, _, _ = orc.render(cam, prm)
            assert same_bits(node.readback(), oi).all(), (mode, waves, refill, gather)
    finally:
        dev.set_option(N.OPT_PERSISTENT_WAVES, DEFAULT_PERSISTENT)
        dev.set_option(N.OPT_REFILL_LANES, AUTO)
        dev.set_option(N.OPT_GATHER_LANES, AUTO)
        dev.set_option(N.OPT_ADVANCE_PERIOD, AUTO)
        dev.set_option(N.OPT_LDS_STACK_ENTRIES, DEFAULT_LDS)
        dev.set_option(N.OPT_TRIS_PER_STEP, 2)


@pytest.mark.parametrize("mode", [N.MODE_REF_GGX, N.MODE_GGX_NEE])
def test_horizon_clamp_bit_exact(dev, field, mode):
    scene, camf = field
    cam = camf(96, 64)
    prm = make_params(spp=2, max_depth=4, mode=mode, rr_start_depth=2, flags=N.FLAG_HORIZON_CLAMP)
    g = _render_gpu(dev, scene, cam, prm)
    o, _, _ = OracleScene(scene).render(cam, prm)
    assert same_bits(g, o).all()


@pytest.mark.parametrize("mode,nee,depth", [(N.MODE_GGX_NEE, N.NEE_LAMBERT, 4), (N.MODE_GGX_NEE, N.NEE_GGX, 4),
                                            (N.MODE_REF_NEE_LAMBERT, N.NEE_LAMBERT, 4)])
def test_light_power_bit_exact(dev, mode, nee, depth):
    """NEE with the light picked by emission strength: the oracle's image on
    a box with three emitters of different radiance."""
    scene, camf = P.cornell_lights()
    cam = camf(72, 64)
    prm = make_params(spp=2, max_depth=depth, mode=mode, nee_eval=nee, rr_start_depth=1, flags=N.FLAG_LIGHT_POWER)
    g = _render_gpu(dev, scene, cam, prm)
    o, _, _ = OracleScene(scene).render(cam, prm)
    assert same_bits(g, o).all()
    u = _render_gpu(dev, scene, cam, make_params(spp=2, max_depth=depth, mode=mode, nee_eval=nee, rr_start_depth=1))
    assert not same_bits(g, u).all()  # the flag changes the estimator
