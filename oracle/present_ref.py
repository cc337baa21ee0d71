"""Oracle (test infrastructure only) of the presentation pass: the
reference's upscaling node draws the rgba32f accumulation texture to the
view target with a fullscreen triangle sampling it through wgpu's default
sampler -- nearest filtering, clamp to edge -- at each target pixel centre
(src/path_tracer/pt_upscaling/pt_upscaling.wgsl:39-50, node.rs:29-97); the
target's format converts on store (IEEE half for the HDR Rgba16Float target,
the sRGB transfer function and 8-bit rounding for Rgba8UnormSrgb)."""
import numpy as np

RGBA32F, RGBA16F, RGBA8_SRGB, RGBA8_UNORM = 0, 1, 2, 3


def nearest_texel(x, src: int, dst: int):
    """floor((x + 1/2) * src / dst), exact, clamped to the edge."""
    x = np.asarray(x, np.int64)
    return np.minimum(((2 * x + 1) * src) // (2 * dst), src - 1)


def srgb_encode(c):
    """IEC 61966-2-1 linear -> sRGB on [0, 1] (NaN and negatives -> 0), f64."""
    c = np.nan_to_num(np.asarray(c, np.float64), nan=0.0)
    c = np.clip(c, 0.0, 1.0)
    return np.where(c <= 0.0031308, 12.92 * c, 1.055 * np.power(c, 1.0 / 2.4) - 0.055)


def present(accum: np.ndarray, dst_w: int, dst_h: int, fmt: int) -> np.ndarray:
    h, w = accum.shape[:2]
    img = accum[nearest_texel(np.arange(dst_h), h, dst_h)][:, nearest_texel(np.arange(dst_w), w, dst_w)]
    if fmt == RGBA32F:
        return img.astype(np.float32)
    if fmt == RGBA16F:
        return img.astype(np.float16)
    # unorm channels: the clamped f32 value times 255 in f32, rounded to
    # nearest even (what the target's store does with an f32 input)
    lin = np.clip(np.nan_to_num(img.astype(np.float32), nan=0.0), np.float32(0), np.float32(1))
    q = np.rint(lin * np.float32(255.0)).astype(np.uint8)
    if fmt == RGBA8_UNORM:
        return q
    out = q.copy()
    out[..., :3] = np.rint(srgb_encode(img[..., :3]) * 255.0).astype(np.uint8)
    return out
