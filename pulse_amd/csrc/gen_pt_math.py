"""Fits the polynomial coefficients of pt_math.h and checks their error.

Method: Cody & Waite range reduction (sin/cos by multiples of pi/2; acos by
the half-angle identity acos(x) = 2 asin(sqrt((1-x)/2))) and polynomial cores
fitted by weighted least squares on Chebyshev nodes (relative error weight),
then rounded to float32.  The float32 evaluation is emulated op by op with
numpy float32 (IEEE single, no fused multiply-add), the same sequence the
header executes, and compared with float64 numpy to report the max ulp error.

Usage: python pulse_amd/csrc/gen_pt_math.py
"""
import numpy as np

f32 = np.float32


def cheb_nodes(a, b, n):
    k = np.arange(n)
    return 0.5 * (a + b) + 0.5 * (b - a) * np.cos((2 * k + 1) * np.pi / (2 * n))


def fit(fun, basis, a, b, n=4000):
    x = cheb_nodes(a, b, n)
    y = fun(x)
    A = np.stack([bf(x) for bf in basis], axis=1)
    w = 1.0 / np.maximum(np.abs(y), 1e-300)
    c, *_ = np.linalg.lstsq(A * w[:, None], y * w, rcond=None)
    return c


def ulp_err(got32, ref64):
    got = got32.astype(np.float64)
    ulp = np.spacing(np.abs(ref64).astype(np.float32)).astype(np.float64)
    return np.max(np.abs(got - ref64) / ulp)


# --- sin core: sin(r) = r + r^3 * (s0 + z*(s1 + z*(s2 + z*s3))), z = r^2 on |r|<=pi/4
R = np.pi / 4
sin_c = fit(lambda r: (np.sin(r) - r) / r**3, [lambda r, k=k: r ** (2 * k) for k in range(4)], 1e-4, R)
# --- cos core: cos(r) = 1 + z*(c0 + z*(c1 + z*(c2 + z*(c3 + z*c4))))
cos_c = fit(lambda r: (np.cos(r) - 1) / r**2, [lambda r, k=k: r ** (2 * k) for k in range(5)], 1e-4, R)
# --- asin core on [0, 0.5]: asin(x) = x + x*z*P(z), z = x^2, P degree 5
asin_c = fit(lambda x: (np.arcsin(x) - x) / x**3, [lambda x, k=k: x ** (2 * k) for k in range(6)], 1e-4, 0.5)

sin_c32 = sin_c.astype(f32)
cos_c32 = cos_c.astype(f32)
asin_c32 = asin_c.astype(f32)

PIO2_1 = f32(1.5703125)  # 8 significant bits: k*PIO2_1 exact for |k| < 2^16
rest = np.pi / 2 - float(PIO2_1)
PIO2_2 = f32(np.float32(rest))
# keep 12 significant bits in PIO2_2 so k*PIO2_2 is exact for |k| < 2^12
m, e = np.frexp(float(PIO2_2))
PIO2_2 = f32(np.ldexp(np.floor(m * 4096) / 4096, e))
PIO2_3 = f32(np.pi / 2 - float(PIO2_1) - float(PIO2_2))
PIO2_HI = f32(np.pi / 2)
PIO2_LO = f32(np.pi / 2 - float(PIO2_HI))
TWO_OVER_PI = f32(2 / np.pi)


def horner(z, cs):
    acc = cs[-1]
    for c in cs[-2::-1]:
        acc = f32(c + f32(z * acc))
    return acc


def sin_core(r):
    z = f32(r * r)
    return f32(r + f32(f32(r * z) * horner(z, sin_c32)))


def cos_core(r):
    z = f32(r * r)
    return f32(f32(1.0) + f32(z * horner(z, cos_c32)))


def reduce(x):
    k = np.floor(f32(f32(x * TWO_OVER_PI) + f32(0.5))).astype(f32)
    r = f32(f32(f32(x - f32(k * PIO2_1)) - f32(k * PIO2_2)) - f32(k * PIO2_3))
    return k.astype(np.int64) & 3, r


def pt_sin(x):
    q, r = reduce(x)
    s, c = sin_core(r), cos_core(r)
    return np.where(q == 0, s, np.where(q == 1, c, np.where(q == 2, -s, -c))).astype(f32)


def pt_cos(x):
    q, r = reduce(x)
    s, c = sin_core(r), cos_core(r)
    return np.where(q == 0, c, np.where(q == 1, -s, np.where(q == 2, -c, s))).astype(f32)


def asin_core(x, z):
    return f32(x + f32(f32(x * z) * horner(z, asin_c32)))


def pt_acos(x):
    ax = np.abs(x)
    # |x| <= 0.5
    z0 = f32(x * x)
    small = f32(PIO2_HI - f32(asin_core(x, z0) - PIO2_LO))
    # |x| > 0.5
    z1 = f32(f32(f32(1.0) - ax) * f32(0.5))
    s1 = np.sqrt(z1).astype(f32)
    a1 = asin_core(s1, z1)
    pos = f32(f32(2.0) * a1)
    neg = f32(f32(2.0) * f32(PIO2_HI - f32(a1 - PIO2_LO)))
    return np.where(ax <= 0.5, small, np.where(x > 0, pos, neg)).astype(f32)


def hexf(v):
    return float(v).hex() + "f"


if __name__ == "__main__":
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.linspace(0, 2 * np.pi * 1.0000005, 2_000_001), rng.uniform(-50, 50, 1_000_000)]).astype(f32)
    xs64 = xs.astype(np.float64)
    print("sin max ulp", ulp_err(pt_sin(xs), np.sin(xs64)))
    print("cos max ulp", ulp_err(pt_cos(xs), np.cos(xs64)))
    # relative-to-1 absolute error (sin/cos near zeros have huge ulp ratios)
    print("sin max abs", np.max(np.abs(pt_sin(xs).astype(np.float64) - np.sin(xs64))))
    print("cos max abs", np.max(np.abs(pt_cos(xs).astype(np.float64) - np.cos(xs64))))
    a = np.concatenate([np.linspace(-1, 1, 2_000_001), rng.uniform(-1, 1, 1_000_000)]).astype(f32)
    print("acos max ulp", ulp_err(pt_acos(a), np.arccos(a.astype(np.float64))))
    print("// sin core coefficients")
    for i, c in enumerate(sin_c32):
        print(f"#define PT_S{i} {hexf(c)} /* {float(c):.9g} */")
    print("// cos core coefficients")
    for i, c in enumerate(cos_c32):
        print(f"#define PT_C{i} {hexf(c)} /* {float(c):.9g} */")
    print("// asin core coefficients")
    for i, c in enumerate(asin_c32):
        print(f"#define PT_A{i} {hexf(c)} /* {float(c):.9g} */")
    for name, v in [("PT_PIO2_1", PIO2_1), ("PT_PIO2_2", PIO2_2), ("PT_PIO2_3", PIO2_3), ("PT_PIO2_HI", PIO2_HI),
                    ("PT_PIO2_LO", PIO2_LO), ("PT_TWO_OVER_PI", TWO_OVER_PI)]:
        print(f"#define {name} {hexf(v)} /* {float(v):.9g} */")
