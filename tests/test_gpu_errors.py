"""GPU: device-side error conditions are reported once, then cleared.

The kernels OR condition bits (stack overflow, iteration cap) into a device
word of the context.  pulse_sync, pulse_get_stats and pulse_batch_get_stats
report the bits raised since the word was last read and clear it, so a
later call (a new frame, a new scene) starts clean.  The bits are set from
the host here (PULSE_OPT_INJECT_ERROR), as a kernel that met the condition
would set them; no kernel is made to fail.
"""
import numpy as np
import pytest

from pulse_amd import _native as N
from pulse_amd import procedural as P
from pulse_amd.path_tracer import Batch, Device, PulsePathTracerNode, make_params

pytestmark = pytest.mark.gpu


def test_error_word_is_reported_once_then_cleared():
    dev = Device(0)
    try:
        scene, camf = P.cornell()
        dev.upload_scene(scene)
        cam = camf(32, 32)
        node = PulsePathTracerNode(dev, 32, 32)
        node.run(cam, make_params(max_depth=4, mode=N.MODE_GGX_NEE))
        dev.sync()
        assert dev.stats().error_flags == 0

        dev.set_option(N.OPT_INJECT_ERROR, N.ERR_STACK_OVERFLOW)
        with pytest.raises(N.PulseError, match="STACK_OVERFLOW"):
            dev.sync()
        dev.sync()  # reported once
        assert dev.stats().error_flags == 0

        dev.set_option(N.OPT_INJECT_ERROR, N.ERR_ITERATION_CAP)
        assert dev.stats().error_flags == N.ERR_ITERATION_CAP
        assert dev.stats().error_flags == 0
        dev.sync()

        dev.set_option(N.OPT_INJECT_ERROR, N.ERR_ITERATION_CAP)
        with pytest.raises(N.PulseError, match="ITERATION_CAP"):
            dev.sync()
        # a later frame on the same context reports nothing
        node.run(cam, make_params(accumulation_count=1, max_depth=4, mode=N.MODE_GGX_NEE))
        dev.sync()
        assert dev.stats().error_flags == 0
        assert np.isfinite(node.readback()).all()

        # the batched form reads the same word
        sc = dev.create_scene(scene)
        batch = Batch(dev, [(sc, cam)], make_params(spp=1, max_depth=3, mode=N.MODE_GGX_NEE))
        batch.render(0)
        assert batch.stats().error_flags == 0
        dev.set_option(N.OPT_INJECT_ERROR, N.ERR_STACK_OVERFLOW)
        assert batch.stats().error_flags == N.ERR_STACK_OVERFLOW
        assert batch.stats().error_flags == 0
        batch.close()
        sc.close()
        with pytest.raises(N.PulseError, match="INVALID_ARG"):
            dev.set_option(N.OPT_INJECT_ERROR, -1)
    finally:
        dev.close()
