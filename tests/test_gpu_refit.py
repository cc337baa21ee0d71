"""Instance motion on the device (pulse_scene_refit_instances): new
transforms, their inverses and the 4-wide TLAS refit on the GPU.  A render
after the refit equals the oracle's render of the same scene rebuilt on the
host (pulse_sb_update_instances) up to t-ties -- the refit tree differs from
a rebuilt one, the closest hits do not -- and the instance records are the
host builder's (a different inverse would move every object-space ray and
show in every pixel).  The reference's binary TLAS is stale afterwards, so
traversal loops 0/1 are refused until a host update."""
import numpy as np
import pytest

from oracle_lib import OracleScene, same_bits
from pulse_amd import _native as N
from pulse_amd import procedural as P
from pulse_amd.path_tracer import Batch, Device, PulsePathTracerNode, make_params

from test_instance_update import _make, _transforms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = Device(0)
    d.set_option(N.OPT_TRAVERSAL_LOOP, 2)
    yield d
    d.close()


def _moved(idx=(0, 3, 5, 17, 23)):
    a = _make(_transforms(24, 1), keep=True)
    moved = _transforms(24, 2)
    for i in idx:
        a.builder.set_instance_transform(i, moved[i])
    return a, a.builder.update_instances(a)


def _cam():
    return P.make_camera(64, 48, eye=(0.0, 0.0, 40.0), target=(0.0, 0.0, 0.0))


def _render(dev, cam, prm):
    node = PulsePathTracerNode(dev, cam.width, cam.height)
    node.run(cam, prm)
    return node.readback()


def _few_differ(g, o):
    px = ~same_bits(g, o).all(axis=-1)
    assert px.mean() <= 1e-3, f"{px.sum()} of {px.size} pixels differ"


@pytest.mark.parametrize("idx", [(0, 3, 5, 17, 23), tuple(range(24))], ids=["five", "all"])
def test_refit_equals_host_rebuild_and_oracle(dev, idx):
    a, b = _moved(idx)
    cam = _cam()
    prm = make_params(spp=2, max_depth=4, mode=N.MODE_GGX_NEE, rr_start_depth=1)
    dev.upload_scene(a)
    before = _render(dev, cam, prm)
    dev.refit_instances(b.buffers["instances"]["object_world"])
    after = _render(dev, cam, prm)
    assert not same_bits(before, after).all()
    o, _, _ = OracleScene(b).render(cam, prm)
    _few_differ(after, o)
    dev.update_instances(b)  # the host's rebuilt TLAS: another tree, the same hits
    _few_differ(_render(dev, cam, prm), after)


def test_refit_stales_the_binary_tlas(dev):
    a, b = _moved()
    cam = _cam()
    prm = make_params(spp=1, max_depth=3, mode=N.MODE_GGX_NEE)
    dev.upload_scene(a)
    dev.refit_instances(b.buffers["instances"]["object_world"])
    dev.set_option(N.OPT_TRAVERSAL_LOOP, 1)
    try:
        with pytest.raises(N.PulseError, match="UNSUPPORTED"):
            _render(dev, cam, prm)
        dev.update_instances(b)  # the reference's rebuild: loop 1 runs again, bit-exact
        o, _, _ = OracleScene(b).render(cam, prm)
        assert same_bits(_render(dev, cam, prm), o).all()
    finally:
        dev.set_option(N.OPT_TRAVERSAL_LOOP, 2)


def test_refit_rejects_bad_input(dev):
    a, b = _moved()
    dev.upload_scene(a)
    m = b.buffers["instances"]["object_world"]
    with pytest.raises(N.PulseError, match="matrices"):
        dev.refit_instances(m[:-1])
    bad = m.copy()
    bad[4] = 0.0  # singular
    with pytest.raises(N.PulseError, match="singular"):
        dev.refit_instances(bad)
    with pytest.raises(N.PulseError):
        _render(dev, _cam(), make_params(max_depth=2))  # partly rewritten: upload again
    dev.upload_scene(a)
    _render(dev, _cam(), make_params(max_depth=2))


def test_batch_sees_the_refit(dev):
    a, b = _moved()
    cam = _cam()
    prm = make_params(spp=1, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=1)
    sa = dev.create_scene(a)
    batch = Batch(dev, [(sa, cam)], prm)
    batch.render(0)
    first = batch.images()[0]
    dev.refit_instances(b.buffers["instances"]["object_world"], target=sa)
    batch.render(0)
    got = batch.images()[0]
    assert not same_bits(first, got).all()
    dev.bind_scene(sa)
    assert same_bits(got, _render(dev, cam, prm)).all()
    batch.close()


def _projective(n=24, seed=1, which=0):
    t = _transforms(n, seed)
    t[which] = t[which].copy()
    t[which][3, 3] = 2.0  # w = 2: transform_position halves the point
    return t


def test_projective_instance_falls_back_to_the_reference_order():
    """The 4-wide traversal's instance records hold matrix rows 0-2 only:
    a scene with a projective instance matrix gets no 4-wide BVHs, the
    default traversal is then the reference's order (loop 1), bit-exact,
    and an explicit loop 2 is refused."""
    s = _make(_projective())
    cam = _cam()
    prm = make_params(spp=1, max_depth=3, mode=N.MODE_GGX_NEE)
    d = Device(0)
    try:
        d.upload_scene(s)
        o, _, _ = OracleScene(s).render(cam, prm)
        assert same_bits(_render(d, cam, prm), o).all()
        with pytest.raises(N.PulseError, match="UNSUPPORTED"):
            d.set_option(N.OPT_TRAVERSAL_LOOP, 2)
    finally:
        d.close()


def test_update_and_refit_refuse_a_projective_matrix(dev):
    a, b = _moved()
    cam = _cam()
    prm = make_params(spp=1, max_depth=3, mode=N.MODE_GGX_NEE)
    dev.upload_scene(a)
    before = _render(dev, cam, prm)
    p = _make(_projective())
    with pytest.raises(N.PulseError, match="bottom row"):
        dev.update_instances(p)
    # refused before anything was replaced: the scene renders as before
    assert same_bits(_render(dev, cam, prm), before).all()
    m = p.buffers["instances"]["object_world"]
    with pytest.raises(N.PulseError, match="bottom row"):
        dev.refit_instances(m)
    dev.upload_scene(a)
    assert same_bits(_render(dev, cam, prm), before).all()


# --- the 4-wide TLAS rebuilt on the device (pulse_scene_rebuild_instances) ---

@pytest.mark.parametrize("idx", [(0, 3, 5, 17, 23), tuple(range(24))], ids=["five", "all"])
def test_rebuild_equals_host_rebuild_and_oracle(dev, idx):
    """A new tree over the moved instances, built on the GPU: the render
    equals the oracle's render of the host-rebuilt scene up to t-ties, and
    the render after the host update likewise; two rebuilds from the same
    matrices give the same image bit for bit."""
    a, b = _moved(idx)
    cam = _cam()
    prm = make_params(spp=2, max_depth=4, mode=N.MODE_GGX_NEE, rr_start_depth=1)
    dev.upload_scene(a)
    m = b.buffers["instances"]["object_world"]
    dev.rebuild_instances(m)
    after = _render(dev, cam, prm)
    o, _, _ = OracleScene(b).render(cam, prm)
    _few_differ(after, o)
    dev.rebuild_instances(m)
    assert same_bits(_render(dev, cam, prm), after).all()
    dev.update_instances(b)
    _few_differ(_render(dev, cam, prm), after)


def _field_moved(n_instances=300, seed=9):
    scene, camf = P.icosphere_field(n_meshes=6, subdiv=3, n_instances=n_instances, n_emissive=8, box=40.0,
                                    radius=(1.0, 2.5), keep_builder=True)
    rng = np.random.default_rng(seed)
    inst = scene.buffers["instances"]
    for i in range(len(inst)):
        mt = np.asarray(inst[i]["object_world"], np.float64).reshape(4, 4).T.copy()
        mt[:3, 3] = rng.uniform(-20.0, 20.0, 3)
        scene.builder.set_instance_transform(i, mt)
    return scene, scene.builder.update_instances(scene), camf


def test_rebuild_restores_traversal_work(dev):
    """Every instance of a 300-instance field moved to a new place: the
    refit keeps the old tree (its boxes now overlap), the device rebuild
    makes a new one; node visits per ray (the kernels' census) after the
    rebuild are within 15 % of those of the host's tree and far below the
    refit's, and all three renders agree up to t-ties."""
    a, b, camf = _field_moved()
    cam = camf(96, 64)
    prm = make_params(spp=1, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=1)
    census = make_params(spp=1, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=1, flags=N.FLAG_COUNT_WORK)
    m = b.buffers["instances"]["object_world"]

    def visits():
        node = PulsePathTracerNode(dev, cam.width, cam.height)
        node.run(cam, census)
        w = dev.stats().work
        return sum(int(w[i][1]) for i in range(3)) / max(1, sum(int(w[i][0]) for i in range(3)))

    dev.upload_scene(a)
    dev.refit_instances(m)
    refit_img, refit_v = _render(dev, cam, prm), visits()
    dev.upload_scene(a)
    dev.rebuild_instances(m)
    rebuilt_img, rebuilt_v = _render(dev, cam, prm), visits()
    dev.upload_scene(b)
    host_img, host_v = _render(dev, cam, prm), visits()
    print("node visits per ray: refit", refit_v, "rebuilt", rebuilt_v, "host", host_v, flush=True)
    assert rebuilt_v <= 1.15 * host_v and rebuilt_v < 0.8 * refit_v
    _few_differ(rebuilt_img, host_img)
    _few_differ(refit_img, host_img)


def test_rebuild_single_instance_and_batch(dev):
    """One instance (a tree of one leaf) and a batch over a rebuilt scene."""
    t = _transforms(1, 3)
    s = _make(t, emissive=(), keep=True)
    moved = _transforms(1, 4)
    s.builder.set_instance_transform(0, moved[0])
    s2 = s.builder.update_instances(s)
    cam = P.make_camera(48, 40, eye=(0.0, 0.0, 30.0), target=(0.0, 0.0, 0.0))
    prm = make_params(spp=1, max_depth=2, mode=N.MODE_REF_GGX)
    dev.upload_scene(s)
    dev.rebuild_instances(s2.buffers["instances"]["object_world"])
    o, _, _ = OracleScene(s2).render(cam, prm)
    _few_differ(_render(dev, cam, prm), o)
    a, b = _moved()
    cam = _cam()
    prm = make_params(spp=1, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=1)
    sa = dev.create_scene(a)
    batch = Batch(dev, [(sa, cam)], prm)
    batch.render(0)
    dev.rebuild_instances(b.buffers["instances"]["object_world"], target=sa)
    batch.render(0)
    got = batch.images()[0]
    dev.bind_scene(sa)
    assert same_bits(got, _render(dev, cam, prm)).all()
    batch.close()


def test_rebuild_rejects_bad_input(dev):
    a, b = _moved()
    dev.upload_scene(a)
    m = b.buffers["instances"]["object_world"]
    with pytest.raises(N.PulseError, match="matrices"):
        dev.rebuild_instances(m[:-1])
    bad = m.copy()
    bad[4] = 0.0  # singular
    with pytest.raises(N.PulseError, match="singular"):
        dev.rebuild_instances(bad)
    with pytest.raises(N.PulseError):
        _render(dev, _cam(), make_params(max_depth=2))  # partly rewritten: upload again
    dev.upload_scene(a)
    _render(dev, _cam(), make_params(max_depth=2))
