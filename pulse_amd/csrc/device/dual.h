// dual.h -- forward-mode dual numbers for the adjoint replay.
//
// Dual<N> carries a value and N tangents.  The value of every operation is
// the same single f32 operation the plain kernels execute (same operands,
// same order, same pt_math functions), so a path replayed in Dual<N>
// arithmetic takes exactly the decisions of the forward pass; the tangents
// follow the usual derivative rules.
#pragma once

#include <hip/hip_runtime.h>

#pragma clang force_cuda_host_device begin
#include "../pt_math.h"
#pragma clang force_cuda_host_device end

namespace pt {

template <int N>
struct Dual {
    float v;
    float d[N];
};

template <int N>
__device__ __forceinline__ Dual<N> dconst(float v) {
    Dual<N> r;
    r.v = v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = 0.0f;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> dvar(float v, int k) {
    Dual<N> r = dconst<N>(v);
    r.d[k] = 1.0f;
    return r;
}

#define PT_DUAL_BINOP(OP, VAL, TAN)                                                                  \
    template <int N>                                                                                 \
    __device__ __forceinline__ Dual<N> operator OP(const Dual<N> &a, const Dual<N> &b) {            \
        Dual<N> r;                                                                                   \
        r.v = VAL;                                                                                   \
        _Pragma("unroll") for (int i = 0; i < N; ++i) r.d[i] = TAN;                                  \
        return r;                                                                                    \
    }

PT_DUAL_BINOP(+, a.v + b.v, a.d[i] + b.d[i])
PT_DUAL_BINOP(-, a.v - b.v, a.d[i] - b.d[i])
PT_DUAL_BINOP(*, a.v * b.v, a.d[i] * b.v + a.v * b.d[i])
PT_DUAL_BINOP(/, a.v / b.v, (a.d[i] - r.v * b.d[i]) / b.v)
#undef PT_DUAL_BINOP

template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, float b) {
    Dual<N> r = a;
    r.v = a.v + b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(float a, const Dual<N> &b) {
    Dual<N> r = b;
    r.v = a + b.v;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, float b) {
    Dual<N> r = a;
    r.v = a.v - b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(float a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a - b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -b.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a) {
    Dual<N> r;
    r.v = -a.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, float b) {
    Dual<N> r;
    r.v = a.v * b;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(float a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a * b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a * b.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, float b) {
    Dual<N> r;
    r.v = a.v / b;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(float a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a / b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -r.v * b.d[i] / b.v;
    return r;
}

// Constants of scalar type S (zero tangents).
template <class S>
struct Lift;
template <>
struct Lift<float> {
    __device__ __forceinline__ static float c(float v) { return v; }
};
template <int N>
struct Lift<Dual<N>> {
    __device__ __forceinline__ static Dual<N> c(float v) { return dconst<N>(v); }
};
template <class S>
__device__ __forceinline__ S sconst(float v) { return Lift<S>::c(v); }

// Comparisons use the value.
template <int N>
__device__ __forceinline__ bool operator<(const Dual<N> &a, float b) { return a.v < b; }
template <int N>
__device__ __forceinline__ bool operator>(const Dual<N> &a, float b) { return a.v > b; }

// ---- scalar functions (overloaded for float and Dual<N>) ----
__device__ __forceinline__ float val(float x) { return x; }
template <int N>
__device__ __forceinline__ float val(const Dual<N> &x) { return x.v; }

__device__ __forceinline__ float sqrt_(float x) { return __builtin_sqrtf(x); }
template <int N>
__device__ __forceinline__ Dual<N> sqrt_(const Dual<N> &x) {
    Dual<N> r;
    r.v = __builtin_sqrtf(x.v);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = x.d[i] / (2.0f * r.v);
    return r;
}

__device__ __forceinline__ float sin_(float x) { return pt_sinf(x); }
__device__ __forceinline__ float cos_(float x) { return pt_cosf(x); }
__device__ __forceinline__ float acos_(float x) { return pt_acosf(x); }
__device__ __forceinline__ float pow5_(float x) { return pt_pow5f(x); }
template <int N>
__device__ __forceinline__ Dual<N> sin_(const Dual<N> &x) {
    Dual<N> r;
    r.v = pt_sinf(x.v);
    float c = pt_cosf(x.v);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = x.d[i] * c;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> cos_(const Dual<N> &x) {
    Dual<N> r;
    r.v = pt_cosf(x.v);
    float s = pt_sinf(x.v);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -x.d[i] * s;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> acos_(const Dual<N> &x) {
    Dual<N> r;
    r.v = pt_acosf(x.v);
    float k = -1.0f / __builtin_sqrtf(fmaxf(1.0f - x.v * x.v, 1e-30f));
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = x.d[i] * k;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> pow5_(const Dual<N> &x) {
    Dual<N> r;
    r.v = pt_pow5f(x.v);
    float b = x.v > 0.0f ? x.v : 0.0f;
    float k = 5.0f * (b * b) * (b * b);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = x.d[i] * k;
    return r;
}

__device__ __forceinline__ float fmax_(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ float fmin_(float a, float b) { return fminf(a, b); }
// fmaxf/fminf on the values (NaN operands ignored as fmaxf does); the result
// carries the tangent of the operand it equals.
template <int N>
__device__ __forceinline__ Dual<N> fmax_(const Dual<N> &a, const Dual<N> &b) {
    float m = fmaxf(a.v, b.v);
    Dual<N> r = (m == a.v) ? a : b;
    r.v = m;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> fmax_(const Dual<N> &a, float b) {
    float m = fmaxf(a.v, b);
    if (m == a.v) return a;
    return dconst<N>(m);
}

}  // namespace pt
