// present.h -- the accumulation buffer drawn to a view target: the
// reference's upscaling pass (pt_upscaling.wgsl:39-50, node.rs:29-97) draws
// a fullscreen triangle that samples the rgba32f accumulation texture with
// wgpu's default sampler (nearest filtering, clamp to edge) at each target
// pixel centre, and the target's format does the conversion on store.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "pulse_hip.h"

namespace pt {

// Nearest source texel of target pixel x: floor((x + 1/2) * src / dst),
// evaluated exactly in integers, clamped to the edge.
__device__ __forceinline__ uint32_t nearest_texel(uint32_t x, uint32_t src, uint32_t dst) {
    uint64_t s = ((2ull * x + 1ull) * src) / (2ull * dst);
    return static_cast<uint32_t>(s < src ? s : src - 1u);
}

// Linear -> sRGB transfer function (IEC 61966-2-1), applied by an
// ...UnormSrgb target on store, then 8-bit quantisation (round to nearest).
__device__ __forceinline__ uint8_t srgb8(float c) {
    if (!(c > 0.0f)) c = 0.0f;  // NaN and negatives -> 0
    if (c > 1.0f) c = 1.0f;
    float e = c <= 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
    return static_cast<uint8_t>(__float2int_rn(e * 255.0f));
}
__device__ __forceinline__ uint8_t unorm8(float c) {
    if (!(c > 0.0f)) c = 0.0f;
    if (c > 1.0f) c = 1.0f;
    return static_cast<uint8_t>(__float2int_rn(c * 255.0f));
}

__global__ void __launch_bounds__(256) k_present(const float4 *src, uint32_t sw, uint32_t sh, void *dst, uint32_t dw,
                                                 uint32_t dh, int format) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= dw * dh) return;
    const uint32_t x = i % dw, y = i / dw;
    const float4 c = src[nearest_texel(y, sh, dh) * sw + nearest_texel(x, sw, dw)];
    switch (format) {
        case PULSE_PRESENT_RGBA32F:
            static_cast<float4 *>(dst)[i] = c;
            break;
        case PULSE_PRESENT_RGBA16F: {
            __half2 *o = static_cast<__half2 *>(dst) + 2 * i;
            o[0] = __halves2half2(__float2half_rn(c.x), __float2half_rn(c.y));
            o[1] = __halves2half2(__float2half_rn(c.z), __float2half_rn(c.w));
            break;
        }
        case PULSE_PRESENT_RGBA8_UNORM:
            static_cast<uchar4 *>(dst)[i] = make_uchar4(unorm8(c.x), unorm8(c.y), unorm8(c.z), unorm8(c.w));
            break;
        default:  // PULSE_PRESENT_RGBA8_SRGB: alpha is stored linearly
            static_cast<uchar4 *>(dst)[i] = make_uchar4(srgb8(c.x), srgb8(c.y), srgb8(c.z), unorm8(c.w));
            break;
    }
}

}  // namespace pt
