"""Instance updates on the device (pulse_scene_update_instances): after
moving instances, a render equals the oracle's render of the rebuilt scene
and a render after a full upload; changes that need a full upload are
refused; batches pick the update up."""
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
    d.set_option(N.OPT_TRAVERSAL_LOOP, 1)  # the reference's traversal order: bit-exact vs the oracle
    yield d
    d.close()


def _moved():
    t0 = _transforms(24, 1)
    a = _make(t0, keep=True)
    moved = _transforms(24, 2)
    for i in (0, 3, 5, 17, 23):
        a.builder.set_instance_transform(i, moved[i])
    return a, a.builder.update_instances(a)


def _cam():
    return P.make_camera(64, 48, eye=(0.0, 0.0, 40.0), target=(0.0, 0.0, 0.0))


def _render(dev, cam, prm):
    node = PulsePathTracerNode(dev, cam.width, cam.height)
    node.run(cam, prm)
    return node.readback()


def test_update_equals_full_upload_and_oracle(dev):
    a, b = _moved()
    cam = _cam()
    prm = make_params(spp=2, max_depth=4, mode=N.MODE_GGX_NEE, rr_start_depth=1)
    dev.upload_scene(a)
    before = _render(dev, cam, prm)
    dev.update_instances(b)
    after = _render(dev, cam, prm)
    assert not same_bits(before, after).all()
    o, _, _ = OracleScene(b).render(cam, prm)
    assert same_bits(after, o).all()
    dev.upload_scene(b)
    assert same_bits(_render(dev, cam, prm), after).all()


def test_update_refuses_changed_lights(dev):
    a = _make(_transforms(24, 1))
    dev.upload_scene(a)
    c = _make(_transforms(24, 1), emissive=(3, 4))
    with pytest.raises(N.PulseError, match="LAYOUT"):
        dev.update_instances(c)
    # the resident scene is unchanged and still renders
    cam = _cam()
    prm = make_params(spp=1, max_depth=3, mode=N.MODE_GGX_NEE)
    o, _, _ = OracleScene(a).render(cam, prm)
    assert same_bits(_render(dev, cam, prm), o).all()


def test_batch_sees_scene_update(dev):
    a, b = _moved()
    ds = dev.create_scene(a)
    cam = _cam()
    prm = make_params(spp=1, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=1)
    batch = Batch(dev, [(ds, cam)], prm)
    batch.render(0)
    dev.update_instances(b, target=ds)
    batch.render(0)
    img = batch.images()[0]
    o, _, _ = OracleScene(b).render(cam, prm)
    assert same_bits(img, o).all()
    batch.close()
    ds.close()
