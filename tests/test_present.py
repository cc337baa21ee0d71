"""Presentation pass (pulse_present): CPU pins of the oracle and the GPU
kernel against it."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import present_ref as R  # noqa: E402

from pulse_amd import _native as N  # noqa: E402


def test_nearest_texel_known_answers():
    # same size: identity; 2x up: each texel twice; 2x down: the odd texels
    assert list(R.nearest_texel(np.arange(4), 4, 4)) == [0, 1, 2, 3]
    assert list(R.nearest_texel(np.arange(8), 4, 8)) == [0, 0, 1, 1, 2, 2, 3, 3]
    assert list(R.nearest_texel(np.arange(4), 8, 4)) == [1, 3, 5, 7]
    # 3 -> 2: centres at 0.75 and 2.25 source texels
    assert list(R.nearest_texel(np.arange(2), 3, 2)) == [0, 2]
    # never past the edge
    assert R.nearest_texel(np.arange(1000), 7, 1000).max() == 6


def test_srgb_known_answers():
    """Published anchors of the sRGB transfer function: 0 -> 0, 1 -> 255,
    18 % grey -> 118, 0.5 -> 188, the linear toe 12.92 c."""
    lin = np.array([0.0, 1.0, 0.18, 0.5, 0.001, -1.0, 7.0, np.nan])
    got = np.rint(R.srgb_encode(lin) * 255.0).astype(int).tolist()
    assert got == [0, 255, 118, 188, round(12.92 * 0.001 * 255), 0, 255, 0]
    # continuity at the junction of the two pieces
    a = 0.0031308
    assert abs(12.92 * a - (1.055 * a ** (1 / 2.4) - 0.055)) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", [N.PRESENT_RGBA32F, N.PRESENT_RGBA16F, N.PRESENT_RGBA8_SRGB, N.PRESENT_RGBA8_UNORM])
@pytest.mark.parametrize("dst", [(64, 48), (128, 96), (37, 23), (200, 51)])
def test_present_matches_oracle(fmt, dst):
    from pulse_amd.path_tracer import Device

    dev = Device(0)
    try:
        rng = np.random.default_rng(fmt * 100 + dst[0])
        src = (rng.random((48, 64, 4)) * 1.4 - 0.1).astype(np.float32)
        src[0, 0, 0] = np.nan
        src[1, 1, 1] = 70000.0  # above the half range: inf
        d_src = dev.to_device(src)
        item = {N.PRESENT_RGBA32F: 16, N.PRESENT_RGBA16F: 8}.get(fmt, 4)
        d_dst = dev.alloc(dst[0] * dst[1] * item)
        N.check(N.lib().pulse_present(dev.handle, d_src.ptr, 64, 48, d_dst.ptr, dst[0], dst[1], fmt))
        dev.sync()
        dt = {N.PRESENT_RGBA32F: np.float32, N.PRESENT_RGBA16F: np.float16}.get(fmt, np.uint8)
        got = d_dst.copy_to(np.zeros((dst[1], dst[0], 4), dt))
        want = R.present(src, dst[0], dst[1], fmt)
        if fmt == N.PRESENT_RGBA8_SRGB:
            # powf on the GPU vs f64 here: at most one code value apart
            assert np.abs(got.astype(int) - want.astype(int)).max() <= 1
            assert (got == want).mean() > 0.99
        elif fmt == N.PRESENT_RGBA8_UNORM:
            assert np.array_equal(got, want)
        else:  # floats: equal values, NaN where the oracle has NaN
            assert np.array_equal(got, want, equal_nan=True)
    finally:
        dev.close()
