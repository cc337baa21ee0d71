"""Batched form (pulse_batch_*): every view's image is bit-identical to
pulse_trace of that view alone (same scene, camera, params), for every
launch shape (persistent or not, chunk size, regroup frequency, traversal
loop, stack size, tile shape, graph replay or direct launches); graph
replays stay exact across frames with eager work in between; the autotuner
returns a valid shape that keeps the images."""
import numpy as np
import pytest

from oracle_lib import same_bits
from pulse_amd import _native as N
from pulse_amd import procedural as P
from pulse_amd.path_tracer import Batch, Device, make_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def views(dev):
    """Three scenes of different size, four views of uneven resolution."""
    cornell, cam_c = P.cornell()
    f1, cam_f1 = P.icosphere_field(n_meshes=4, subdiv=3, n_instances=30, n_emissive=4, box=24.0,
                                   radius=(1.0, 3.0), seed=5)
    f2, cam_f2 = P.icosphere_field(n_meshes=6, subdiv=4, n_instances=50, n_emissive=6, box=30.0,
                                   radius=(1.5, 3.5), seed=7)
    s_c, s_1, s_2 = dev.create_scene(cornell), dev.create_scene(f1), dev.create_scene(f2)
    return [(s_c, cam_c(64, 48)), (s_1, cam_f1(37, 29)), (s_2, cam_f2(96, 64)), (s_c, cam_c(17, 80))]


def _sequential(dev, views, params, acc_list, loop=2):
    """pulse_trace of each view alone with the batch's traversal loop (0 and
    1 give the same results)."""
    dev.set_option(N.OPT_TRAVERSAL_LOOP, loop)
    out = []
    for sc, cam in views:
        dev.bind_scene(sc)
        buf = dev.alloc(cam.width * cam.height * 16)
        for acc in acc_list:
            p = N.Params.from_buffer_copy(params)
            p.accumulation_count = acc
            N.check(N.lib().pulse_trace(dev.handle, cam, p, buf.ptr))
        dev.sync()
        out.append(buf.copy_to(np.zeros((cam.height, cam.width, 4), np.float32)))
    return out


PARAMS = [
    pytest.param(dict(spp=2, max_depth=4, mode=N.MODE_GGX_NEE, rr_start_depth=1), id="ggx_nee"),
    pytest.param(dict(spp=1, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=1, flags=N.FLAG_LIGHT_POWER),
                 id="ggx_nee_light_power"),
    pytest.param(dict(spp=1, max_depth=5, mode=N.MODE_REF_GGX), id="ref_ggx"),
    pytest.param(dict(spp=1, max_depth=4, mode=N.MODE_REF_NEE_LAMBERT, rr_start_depth=2), id="ref_nee_lambert"),
]


@pytest.mark.parametrize("pkw", PARAMS)
def test_batch_equals_sequential(dev, views, pkw):
    prm = make_params(**pkw)
    ref = _sequential(dev, views, prm, [0])
    b = Batch(dev, views, prm)
    assert b.config().traversal_loop == 2
    b.render(0)
    for i, (img, r) in enumerate(zip(b.images(), ref)):
        assert same_bits(img, r).all(), f"view {i}"
    b.close()


CONFIGS = [
    dict(persistent_waves_per_cu=0, chunk_rays=64),
    dict(persistent_waves_per_cu=24, traversal_loop=1),
    dict(persistent_waves_per_cu=0, chunk_rays=128, traversal_loop=1),
    dict(persistent_waves_per_cu=8, chunk_rays=128, traversal_loop=0),
    dict(persistent_waves_per_cu=16, chunk_rays=64, regroup_every=2),
    dict(persistent_waves_per_cu=0, chunk_rays=256, regroup_every=3, tile_width=16),
    dict(traversal_loop=0, tile_width=64),
    dict(stack_entries=64, tile_width=4, use_graph=0),
    dict(persistent_waves_per_cu=16, chunk_rays=1024, lds_stack_entries=8, refill_lanes=1),
    dict(persistent_waves_per_cu=32, chunk_rays=64, lds_stack_entries=0, refill_lanes=64, regroup_every=2),
    dict(persistent_waves_per_cu=4, chunk_rays=4096, lds_stack_entries=16, refill_lanes=32, use_graph=0),
]


@pytest.mark.parametrize("cfg", CONFIGS, ids=[str(i) for i in range(len(CONFIGS))])
def test_launch_shapes_keep_images(dev, views, cfg):
    prm = make_params(spp=2, max_depth=4, mode=N.MODE_GGX_NEE, rr_start_depth=1)
    b = Batch(dev, views, prm)
    b.set_config(**cfg)
    got = b.config().as_dict()
    for k, v in cfg.items():
        assert got[k] == v
    ref = _sequential(dev, views, prm, [0], loop=got["traversal_loop"])
    b.render(0)
    for i, (img, r) in enumerate(zip(b.images(), ref)):
        assert same_bits(img, r).all(), f"view {i} with {cfg}"
    b.close()


def test_graph_replays_accumulate_exactly(dev, views):
    """Three graph replays (frames 0, 2, 4 of a 2-spp progressive render)
    with eager launches and host reads between them equal six sequential
    samples; each intermediate frame is checked too."""
    prm = make_params(spp=2, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=1)
    b = Batch(dev, views, prm)
    assert b.config().use_graph == 0  # the default: direct launches
    b.set_config(use_graph=1)
    for frame in range(3):
        b.render(2 * frame)
        imgs = b.images()
        # eager work on the same stream between replays
        dev.trace_rays(np.zeros(4, N.RAY_DTYPE))
        ref = _sequential(dev, views, make_params(spp=1, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=1),
                          list(range(2 * frame + 2)))
        for i, (img, r) in enumerate(zip(imgs, ref)):
            assert same_bits(img, r).all(), f"frame {frame} view {i}"
    b.close()


def test_set_camera_and_stats(dev, views):
    prm = make_params(spp=1, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=3)
    b = Batch(dev, views, prm)
    b.render(0)
    st = b.stats()
    npix = sum(c.width * c.height for _, c in views)
    assert st.extension_rays >= npix and st.shadow_rays > 0
    # move view 1's camera: the batch renders the new view (graph re-used)
    sc, cam = views[1]
    cam2 = N.Camera.from_buffer_copy(cam)
    cam2.world_position[0] += 1.0
    cam2.inverse_view_proj[12] += 0.5
    b.set_camera(1, cam2)
    b.render(0)
    img = b.images()[1]
    ref = _sequential(dev, [(sc, cam2)], prm, [0])[0]
    assert same_bits(img, ref).all()
    bad = N.Camera.from_buffer_copy(cam2)
    bad.width += 1
    with pytest.raises(N.PulseError, match="resolution"):
        b.set_camera(1, bad)
    b.close()


def test_autotune_keeps_images(dev, views):
    prm = make_params(spp=1, max_depth=4, mode=N.MODE_GGX_NEE, rr_start_depth=1)
    b = Batch(dev, views, prm)
    cfg, ms = b.autotune(reps=1)
    assert ms > 0 and cfg.chunk_rays in (64, 128, 256, 512, 1024, 2048, 4096) and cfg.stack_entries >= 24
    ref = _sequential(dev, views, prm, [0], loop=cfg.traversal_loop)
    b.render(0)
    for i, (img, r) in enumerate(zip(b.images(), ref)):
        assert same_bits(img, r).all(), f"view {i} after autotune {cfg.as_dict()}"
    b.close()


def test_batch_rejects_bad_views(dev, views):
    prm = make_params(spp=1, max_depth=3, mode=N.MODE_GGX_NEE)
    with pytest.raises(N.PulseError):
        Batch(dev, [], prm)
    b = Batch(dev, views[:1], prm)
    with pytest.raises(N.PulseError, match="chunk_rays"):
        b.set_config(chunk_rays=100)
    with pytest.raises(N.PulseError, match="refill_lanes"):
        b.set_config(refill_lanes=0)
    with pytest.raises(N.PulseError, match="stack_entries"):
        b.set_config(stack_entries=8)
    b.close()


@pytest.mark.parametrize("loop", [1, 2])
@pytest.mark.parametrize("use_graph", [0, 1])
def test_concurrent_shadow_launches_keep_images(dev, views, loop, use_graph):
    """PULSE_OPT_CONCURRENT_SHADOW (default 1): the shadow launches of bounce
    k on the context's second stream beside bounce k+1's extension launch.
    pulse_trace and the batch (direct launches, with the pairs; a graph
    replay runs its launches in sequence) give the same images as with every
    launch in sequence, over two progressive frames."""
    prm = make_params(spp=2, max_depth=4, mode=N.MODE_GGX_NEE, rr_start_depth=1)
    assert dev.get_option(N.OPT_CONCURRENT_SHADOW) == 1
    with pytest.raises(N.PulseError):
        dev.set_option(N.OPT_CONCURRENT_SHADOW, 2)
    dev.set_option(N.OPT_CONCURRENT_SHADOW, 0)
    try:
        ref = _sequential(dev, views, prm, [0, 2], loop=loop)
    finally:
        dev.set_option(N.OPT_CONCURRENT_SHADOW, 1)
    conc = _sequential(dev, views, prm, [0, 2], loop=loop)
    # the waves each of a concurrent pair takes (PULSE_OPT_PAIR_WAVES) do not
    # change the images either
    assert dev.get_option(N.OPT_PAIR_WAVES) == dev.get_option(N.OPT_PERSISTENT_WAVES) // 2
    with pytest.raises(N.PulseError):
        dev.set_option(N.OPT_PAIR_WAVES, 65)
    for w in (1, 24):
        dev.set_option(N.OPT_PAIR_WAVES, w)
        try:
            other = _sequential(dev, views, prm, [0, 2], loop=loop)
        finally:
            dev.set_option(N.OPT_PAIR_WAVES, 0)
        for i, (o, r) in enumerate(zip(other, ref)):
            assert same_bits(o, r).all(), f"pair waves {w}, view {i}"
    b = Batch(dev, views, prm)
    b.set_config(use_graph=use_graph, traversal_loop=loop)
    b.render(0)
    b.render(2)
    for i, (img, c, r) in enumerate(zip(b.images(), conc, ref)):
        assert same_bits(c, r).all(), f"pulse_trace view {i}"
        assert same_bits(img, r).all(), f"batch view {i}"
    b.close()
