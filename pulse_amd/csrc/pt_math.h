// pt_math.h -- portable float32 transcendentals of the shading code.
//
// The reference calls WGSL acos/sin/cos/pow (src/utilities.wgsl:471-474,
// 733-737, 788, 836), whose results are left to the GPU driver.  The HIP
// kernels and the float32 oracle both call these functions instead, so the two
// agree bit for bit: they use only IEEE-754 single-precision + - * / sqrt and
// floor, no fused multiply-add (build with -ffp-contract=off) and no library
// calls.
//
// Method: Cody & Waite argument reduction by multiples of pi/2 with a
// three-part constant; polynomial cores fitted by weighted least squares on
// Chebyshev nodes (pulse_amd/csrc/gen_pt_math.py prints the coefficients and
// the measured error); acos via the half-angle identity
// acos(x) = 2 asin(sqrt((1-x)/2)) for |x| > 1/2.
// Measured max error against float64 on |x| <= 40: sin/cos 1.45 ulp where
// |result| >= 2^-6 and 8.5e-8 absolute everywhere (near the zeros the
// reduction error dominates); acos 1.10 ulp on [-1, 1].  sin/cos are meant for |x| <= 4096 (the kernels
// pass angles in [0, 2*pi*1.0000005]); outside that range and for inf/NaN they
// return NaN.
//
// This header is plain C++; HIP translation units include it inside
// `#pragma clang force_cuda_host_device begin/end`.
#ifndef PULSE_PT_MATH_H
#define PULSE_PT_MATH_H

#define PT_S0 -0x1.555556p-3f
#define PT_S1 0x1.11110ep-7f
#define PT_S2 -0x1.a01392p-13f
#define PT_S3 0x1.6db9d0p-19f
#define PT_C0 -0x1.000000p-1f
#define PT_C1 0x1.555556p-5f
#define PT_C2 -0x1.6c16b2p-10f
#define PT_C3 0x1.a00dc4p-16f
#define PT_C4 -0x1.239332p-22f
#define PT_A0 0x1.555556p-3f
#define PT_A1 0x1.3333f4p-4f
#define PT_A2 0x1.6d686cp-5f
#define PT_A3 0x1.fc8ccap-6f
#define PT_A4 0x1.1d4438p-6f
#define PT_A5 0x1.10b7a6p-5f
#define PT_PIO2_1 0x1.92p+0f
#define PT_PIO2_2 0x1.fb4p-12f
#define PT_PIO2_3 0x1.4442d2p-24f
#define PT_PIO2_HI 0x1.921fb6p+0f
#define PT_PIO2_LO -0x1.777a5cp-25f
#define PT_TWO_OVER_PI 0x1.45f306p-1f

inline float pt_sin_core(float r) {
    float z = r * r;
    float p = PT_S2 + z * PT_S3;
    p = PT_S1 + z * p;
    p = PT_S0 + z * p;
    return r + (r * z) * p;
}

inline float pt_cos_core(float r) {
    float z = r * r;
    float p = PT_C3 + z * PT_C4;
    p = PT_C2 + z * p;
    p = PT_C1 + z * p;
    p = PT_C0 + z * p;
    return 1.0f + z * p;
}

// Quadrant q (0..3) and remainder r in [-pi/4, pi/4] with x = q*pi/2 + r.
// Returns false outside the supported range (or for inf/NaN).
inline bool pt_reduce(float x, int &q, float &r) {
    float ax = x < 0.0f ? -x : x;
    if (!(ax <= 4096.0f)) return false;
    float kf = __builtin_floorf(x * PT_TWO_OVER_PI + 0.5f);
    r = ((x - kf * PT_PIO2_1) - kf * PT_PIO2_2) - kf * PT_PIO2_3;
    q = static_cast<int>(kf) & 3;
    return true;
}

inline float pt_sinf(float x) {
    int q;
    float r;
    if (!pt_reduce(x, q, r)) return (x - x) / (x - x);
    float s = pt_sin_core(r), c = pt_cos_core(r);
    return q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
}

inline float pt_cosf(float x) {
    int q;
    float r;
    if (!pt_reduce(x, q, r)) return (x - x) / (x - x);
    float s = pt_sin_core(r), c = pt_cos_core(r);
    return q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
}

// asin(x) for x in [0, 1/2] given z = x*x (or the exact (1-|y|)/2 of acos).
inline float pt_asin_core(float x, float z) {
    float p = PT_A4 + z * PT_A5;
    p = PT_A3 + z * p;
    p = PT_A2 + z * p;
    p = PT_A1 + z * p;
    p = PT_A0 + z * p;
    return x + (x * z) * p;
}

// acos on [-1, 1]; NaN outside (sqrt of a negative) and for NaN input.
inline float pt_acosf(float x) {
    float ax = x < 0.0f ? -x : x;
    if (ax <= 0.5f) {
        return PT_PIO2_HI - (pt_asin_core(x, x * x) - PT_PIO2_LO);
    }
    float z = (1.0f - ax) * 0.5f;
    float s = __builtin_sqrtf(z);
    float a = pt_asin_core(s, z);
    if (x > 0.0f) return 2.0f * a;
    return 2.0f * (PT_PIO2_HI - (a - PT_PIO2_LO));
}

// pow(x, 5) of fresnel_schlick (utilities.wgsl:787-789).  The base is clamped
// at 0: WGSL pow of a negative base is undefined, and 1 - cos can round below
// zero when cos rounds above 1.
inline float pt_pow5f(float x) {
    float b = x > 0.0f ? x : 0.0f;
    float b2 = b * b;
    return (b2 * b2) * b;
}

#endif  // PULSE_PT_MATH_H
