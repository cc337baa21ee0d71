"""GPU: pulse_trace_noalloc (no device allocation) equals pulse_trace.

Two calls in a row, the second on a smaller image, each compared bit for bit
with pulse_trace of the same camera and params through the C ABI; the ray
counts agree; sizes above the scratch and the flags that need buffers of
their own are refused; and the calls leave the device's free memory
unchanged while pulse_trace on a fresh context allocates.
"""
import ctypes as C

import numpy as np
import pytest

from pulse_amd import _native as N
from pulse_amd import procedural as P
from pulse_amd.path_tracer import Device, make_params

pytestmark = pytest.mark.gpu


def _hip_free_bytes() -> int:
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _render(dev, fn, cam, prm):
    buf = dev.alloc(cam.width * cam.height * 16)
    N.check(N.lib().pulse_memset_d(dev.handle, buf.ptr, 0, cam.width * cam.height * 16))
    N.check(fn(dev.handle, C.byref(cam), C.byref(prm), buf.ptr))
    dev.sync()
    st = dev.stats()
    img = buf.copy_to(np.zeros((cam.height, cam.width, 4), np.float32))
    return img, (st.primary_rays, st.extension_rays, st.shadow_rays)


def test_noalloc_equals_trace_twice_then_smaller():
    mx = C.c_uint64()
    N.check(N.lib().pulse_trace_noalloc_max_pixels(C.byref(mx)))
    assert mx.value == 65536
    scene, camf = P.cornell()
    field, fcamf = P.icosphere_field(n_meshes=6, subdiv=3, n_instances=40, n_emissive=5, box=24.0, radius=(1.0, 3.0))
    lib = N.lib()
    dev = Device(0)
    try:
        # two calls in a row on one scene, the second smaller (the scratch
        # keeps the first call's contents), then a third on another scene
        groups = [(scene, [(camf(256, 256), make_params(spp=2, max_depth=4, mode=N.MODE_GGX_NEE,
                                                         flags=N.FLAG_COLLECT_STATS)),
                           (camf(96, 64), make_params(accumulation_count=3, spp=1, max_depth=5, mode=N.MODE_REF_GGX,
                                                      flags=N.FLAG_COLLECT_STATS))]),
                  (field, [(fcamf(160, 100), make_params(spp=1, max_depth=4, mode=N.MODE_GGX_NEE, nee_eval=N.NEE_GGX,
                                                         rr_start_depth=1, flags=N.FLAG_COLLECT_STATS))])]
        for sc, cases in groups:
            dev.upload_scene(sc)
            got = []
            for cam, prm in cases:
                print(f"noalloc {cam.width}x{cam.height} mode {prm.mode}", flush=True)
                got.append(_render(dev, lib.pulse_trace_noalloc, cam, prm))
            for (cam, prm), (img, rays) in zip(cases, got):
                ref, ref_rays = _render(dev, lib.pulse_trace, cam, prm)
                assert np.array_equal(ref.view(np.uint32), img.view(np.uint32)), f"{cam.width}x{cam.height}"
                assert rays == ref_rays and rays[1] > 0
        # refused: above the scratch, and flags that need buffers of their own
        buf = dev.alloc(257 * 256 * 16)
        big = camf(257, 256)
        with pytest.raises(N.PulseError, match="UNSUPPORTED"):
            N.check(lib.pulse_trace_noalloc(dev.handle, C.byref(big), C.byref(make_params(max_depth=4)), buf.ptr))
        small = camf(32, 32)
        with pytest.raises(N.PulseError, match="UNSUPPORTED"):
            N.check(lib.pulse_trace_noalloc(dev.handle, C.byref(small),
                                            C.byref(make_params(max_depth=4, flags=N.FLAG_STAGE_TIMING)), buf.ptr))
    finally:
        dev.close()


def test_noalloc_allocates_nothing():
    scene, camf = P.cornell()
    cam = camf(256, 256)
    prm = make_params(spp=1, max_depth=4, mode=N.MODE_GGX_NEE)
    lib = N.lib()
    dev = Device(0)
    try:
        dev.upload_scene(scene)
        buf = dev.alloc(cam.width * cam.height * 16)
        N.check(lib.pulse_trace_noalloc(dev.handle, C.byref(cam), C.byref(prm), buf.ptr))  # loads the kernels
        dev.sync()
        before = _hip_free_bytes()
        for acc in range(3):
            prm.accumulation_count = acc
            N.check(lib.pulse_trace_noalloc(dev.handle, C.byref(cam), C.byref(prm), buf.ptr))
        dev.sync()
        assert _hip_free_bytes() == before
        # control: the allocating entry point sizes its buffers on first use
        N.check(lib.pulse_trace(dev.handle, C.byref(cam), C.byref(prm), buf.ptr))
        dev.sync()
        assert _hip_free_bytes() < before
    finally:
        dev.close()
