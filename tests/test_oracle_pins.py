"""Mathematical pins of the oracle's estimators (CPU; test infrastructure).

Each test takes the Monte Carlo mean of an oracle estimator over its random
numbers and compares it with the integral the estimator must be unbiased
for, computed here by deterministic quadrature in float64 from the
reference's formulas and textbook microfacet terms.  A dropped term, a wrong
pdf, a normalised direction where the reference uses an unnormalised one, or
a missing 1/p in Russian roulette moves the mean by far more than the
tolerances.

* sample_direct_light_ggx (utilities.wgsl:562-655): at unit distance against
  textbook Cook-Torrance NEE (normalised directions) with the reference's
  |n0 - nl|^2 divisor (Q3); at distance 2 against the reference's integrand
  with the unnormalised to-light vector in the half vector and n.wi (Q7),
  which the normalised one misses by a factor of about two.
* importance_sample_ggx_d's weight (utilities.wgsl:816-872): the expected
  reflectance restricted to samples above the horizon equals the albedo
  (diffuse lobe: base / (1 - F0max) drawn with probability 1 - F0max) plus
  the integral of the specular BRDF over the upper hemisphere of incident
  directions (the reference's weight carries no cosine, Q6).  The quadrature
  runs over incident directions, so neither the sampling map nor its pdf
  enters the expected value.
* Russian roulette (path_tracer.wgsl:104-108): E[T' 1{alive}] = T for
  max(T) <= 1 and T / max(T) (survival certain) above 1 (Q17).
* trace_path: image means with and without Russian roulette agree within 3
  standard errors where the throughput cannot exceed 1 (path_traceBBB:
  throughput *= base colour); a one-bounce GGX_NEE pixel equals the hit's
  emission plus the NEE integral at the hit (Lambert and GGX evaluators).
"""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import OracleScene, f3, lib
from pulse_amd import _native as N
from pulse_amd.path_tracer import make_params
from pulse_amd.scene import MaterialDesc, Mesh, SceneBuilder, make_camera


# ---------------------------------------------------------------------------
# textbook terms (float64)
# ---------------------------------------------------------------------------
def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def d_ggx(ndm, r):
    """Trowbridge-Reitz / GGX normal distribution, alpha = roughness^2."""
    a2 = (r * r) ** 2
    return a2 / (np.pi * (ndm * ndm * (a2 - 1.0) + 1.0) ** 2)


def g1_schlick(ndv, r):
    """Schlick-GGX masking with k = alpha / 2."""
    k = r * r / 2.0
    return ndv / (ndv * (1.0 - k) + k)


def fresnel(cos, f0):
    return f0 + (1.0 - f0) * np.clip(1.0 - cos, 0.0, None)[..., None] ** 5


def f0_of(reflectance, metallic, base):
    return 0.16 * reflectance * reflectance * (1.0 - metallic) + np.asarray(base, np.float64) * metallic


def cook_torrance(n, o, i, r, f0):
    """F D G / (4 n.o n.i) for unit o, n and incident vector i (used as
    given: callers pass a unit vector or the reference's unnormalised one)."""
    m = _unit(0.5 * i + 0.5 * o)
    ndm = m @ n
    ndi = i @ n
    ndo = float(o @ n)
    return (fresnel(o @ m if m.ndim == 1 else m @ o, f0) * (d_ggx(ndm, r) * g1_schlick(ndi, r) * g1_schlick(ndo, r)
                                                            / (4.0 * ndo * ndi))[..., None])


def tri_quadrature(p0, p1, p2, n=300):
    """Centroids of the n^2 sub-triangles of a uniform subdivision and the
    area each stands for."""
    pts = []
    for i in range(n):
        j = np.arange(n - i)
        pts.append(np.stack([(i + 1 / 3) * np.ones_like(j) / n, (j + 1 / 3) / n], 1))
        j = np.arange(n - i - 1)
        pts.append(np.stack([(i + 2 / 3) * np.ones_like(j) / n, (j + 2 / 3) / n], 1))
    b = np.concatenate(pts)
    p0, p1, p2 = (np.asarray(p, np.float64) for p in (p0, p1, p2))
    x = p0 + b[:, :1] * (p1 - p0) + b[:, 1:] * (p2 - p0)
    area = 0.5 * np.linalg.norm(np.cross(p1 - p0, p2 - p0))
    return x, area / (n * n)


def _seeds(count, salt=0):
    i = np.arange(count, dtype=np.uint64) + np.uint64(salt) * np.uint64(0x9E3779B9)
    return ((i * np.uint64(2654435761) + np.uint64(12345)) & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def _material(base=(0.5, 0.5, 0.5), roughness=0.5, reflectance=0.5, metallic=0.0, emissive=(0, 0, 0)):
    return MaterialDesc(base_color=(*base, 1.0), emissive=(*emissive, 1.0), perceptual_roughness=roughness,
                        reflectance=reflectance, metallic=metallic)


# ---------------------------------------------------------------------------
# GGX NEE
# ---------------------------------------------------------------------------
def _emitter_scene(tri, le):
    sb = SceneBuilder()
    mid = sb.add_mesh(Mesh(np.asarray(tri, np.float32), np.tile(np.float32([0, -1, 0]), (3, 1)),
                           np.array([0, 1, 2], np.uint32)))
    mat = sb.add_material(_material(emissive=(le, le, le)))
    sb.add_instance(mid, mat, np.eye(4))
    return sb.build(1)


def _nee_mc(o, p0, n0, mat, wo, ggx, count=400_000):
    out = np.zeros((count, 3), np.float32)
    seeds = _seeds(count)
    m = mat.to_c()
    p0, n0, wo = f3(p0), f3(n0), f3(wo)  # held: the C call reads their memory
    lib().oracle_direct_light_batch_f32(o.h, p0.ctypes.data, n0.ctypes.data, C.byref(m), wo.ctypes.data,
                                        int(ggx), seeds.ctypes.data, count, out.ctypes.data, 8)
    v = out.astype(np.float64)
    return v.mean(0), v.std(0) / np.sqrt(count)


def _horizontal_emitter(center, side):
    c = np.asarray(center, np.float64)
    return np.array([c + (-side / 2, 0, -side / 3), c + (side / 2, 0, -side / 3), c + (0, 0, 2 * side / 3)])


def _nee_ggx_integral(tri, p0, n0, wo, r, f0, le, normalised_wi):
    """Integral over the emitter of the reference's NEE integrand: BRDF x Le
    x cos_r x cos_e / |n0 - nl|^2 (the estimator divides by the pdf
    1 / area of the one light), clamped at 0 per channel."""
    x, da = tri_quadrature(*tri)
    origin = p0 + 0.001 * n0
    to_light = x - p0
    d = _unit(x - origin)
    nl = _unit(np.cross(tri[1] - tri[0], tri[2] - tri[0]))
    nl = np.where((d @ -nl)[:, None] < 0, -nl, nl)
    cos_r = d @ n0
    cos_e = np.einsum("ij,ij->i", -d, nl)
    wi = _unit(to_light) if normalised_wi else to_light
    brdf = cook_torrance(n0, wo, wi, r, f0)
    dsq = np.sum((n0 - nl) ** 2, axis=1)
    val = brdf * le * (cos_r * cos_e / dsq)[:, None]
    return np.clip(val, 0.0, None).sum(0) * da


@pytest.mark.parametrize("roughness,alpha_deg", [(0.5, 30.0), (0.8, 55.0), (0.35, 10.0)])
def test_ggx_nee_unit_distance_is_cook_torrance(roughness, alpha_deg):
    """Emitter at unit distance in the mirror direction of wo: the
    unnormalised to-light vector is a unit vector to within 2 %, and the
    estimator's mean equals textbook Cook-Torrance NEE (normalised incident
    direction, divided by |n0 - nl|^2 = 4 instead of the squared distance)."""
    a = np.radians(alpha_deg)
    n0 = np.array([0.0, 1.0, 0.0])
    wo = np.array([-np.sin(a), np.cos(a), 0.0])
    tri = _horizontal_emitter((np.sin(a), np.cos(a), 0.0), 0.04)
    scene = _emitter_scene(tri, 10.0)
    mat = _material(base=(0.6, 0.4, 0.2), roughness=roughness)
    mean, se = _nee_mc(OracleScene(scene), np.zeros(3), n0, mat, wo, ggx=True)
    f0 = f0_of(0.5, 0.0, (0.6, 0.4, 0.2))
    x, da = tri_quadrature(*tri)
    d = _unit(x)
    brdf = cook_torrance(n0, wo, d, roughness, f0)  # textbook: unit incident direction
    cos_r = d @ n0
    cos_e = d[:, 1]  # nl = (0, -1, 0) faces the receiver
    exact = (brdf * 10.0 * (cos_r * cos_e / 4.0)[:, None]).sum(0) * da
    assert (se <= 2e-3 * exact).all(), (se, exact)
    np.testing.assert_allclose(mean, exact, rtol=1e-2)


@pytest.mark.parametrize("roughness,alpha_deg,metallic", [(0.5, 30.0, 0.0), (0.7, 45.0, 0.6)])
def test_ggx_nee_unnormalised_incident_vector(roughness, alpha_deg, metallic):
    """Emitter at distance 2: the reference evaluates the half vector and
    n.wi with the to-light vector of length ~2 (Q7).  The mean equals the
    integral of that integrand and differs from the normalised one by far
    more than the tolerance."""
    a = np.radians(alpha_deg)
    n0 = np.array([0.0, 1.0, 0.0])
    wo = np.array([-np.sin(a), np.cos(a), 0.0])
    tri = _horizontal_emitter((2 * np.sin(a), 2 * np.cos(a), 0.0), 0.08)
    base = (0.3, 0.5, 0.7)
    scene = _emitter_scene(tri, 10.0)
    mat = _material(base=base, roughness=roughness, metallic=metallic)
    mean, se = _nee_mc(OracleScene(scene), np.zeros(3), n0, mat, wo, ggx=True)
    f0 = f0_of(0.5, metallic, base)
    ref = _nee_ggx_integral(tri, np.zeros(3), n0, wo, roughness, f0, 10.0, normalised_wi=False)
    alt = _nee_ggx_integral(tri, np.zeros(3), n0, wo, roughness, f0, 10.0, normalised_wi=True)
    assert (se <= 2e-3 * ref).all()
    np.testing.assert_allclose(mean, ref, rtol=1e-2)
    assert (np.abs(alt - ref) > 0.2 * ref).all(), (alt, ref)


# ---------------------------------------------------------------------------
# GGX sampling weight
# ---------------------------------------------------------------------------
def _frame(n):
    t = _unit(np.cross(n, [0.0, 0.0, 1.0] if abs(n[2]) < 0.9 else [1.0, 0.0, 0.0]))
    return t, np.cross(n, t)


def _specular_albedo_no_cos(r, cos_o, f0, nt=1600, nphi=1600):
    """Integral over the upper hemisphere of incident directions of the
    Cook-Torrance BRDF, without the cosine (midpoint rule in (theta, phi))."""
    n = np.array([0.0, 0.0, 1.0])
    o = np.array([np.sqrt(1 - cos_o * cos_o), 0.0, cos_o])
    th = (np.arange(nt) + 0.5) * (np.pi / 2) / nt
    ph = (np.arange(nphi) + 0.5) * (2 * np.pi) / nphi
    T, P = np.meshgrid(th, ph, indexing="ij")
    i = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    f = cook_torrance(n, o, i, r, f0)
    w = (np.sin(T) * (np.pi / 2 / nt) * (2 * np.pi / nphi)).reshape(-1, 1)
    return (f * w).sum(0)


@pytest.mark.parametrize("r,cos_o,metallic,base", [
    (0.3, 0.9, 0.0, (0.6, 0.4, 0.2)),
    (0.6, 0.5, 0.0, (0.3, 0.5, 0.7)),
    (1.0, 0.2, 0.0, (0.5, 0.5, 0.5)),
    (0.45, 0.7, 0.5, (0.9, 0.6, 0.3)),
])
def test_ggx_sampling_weight_expectation(r, cos_o, metallic, base):
    nw = _unit(np.array([0.2, 0.9, 0.3]))  # a general normal: any orthonormal basis
    t, b = _frame(nw)
    wo = np.sqrt(1 - cos_o * cos_o) * t + cos_o * nw
    mat = _material(base=base, roughness=r, metallic=metallic).to_c()
    count = 2_000_000
    out = np.zeros((count, 6), np.float32)
    nw32, wo32, seeds = f3(nw), f3(wo), _seeds(count, 7)  # held: the C call reads their memory
    lib().oracle_sample_ggx_batch_f32(nw32.ctypes.data, wo32.ctypes.data, C.byref(mat), seeds.ctypes.data, count,
                                      out.ctypes.data)
    wi = out[:, :3].astype(np.float64)
    refl = out[:, 3:].astype(np.float64)
    finite = np.isfinite(refl).all(1)
    assert (~finite).mean() < 1e-5  # rand_f may reach 1.0000005 (Q1): NaN theta, ~3e-8 per draw
    above = finite & (wi @ nw > 0)
    contrib = np.where(above[:, None], refl, 0.0)
    contrib[~finite] = 0.0
    mean = contrib.mean(0)
    se = contrib.std(0) / np.sqrt(count)
    f0 = f0_of(0.5, metallic, base)
    exact = np.asarray(base) + _specular_albedo_no_cos(r, cos_o, f0)
    assert (se <= 3e-3 * exact).all(), (se, exact)
    np.testing.assert_allclose(mean, exact, rtol=1e-2)


# ---------------------------------------------------------------------------
# Russian roulette and the path loop
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("t", [(0.3, 0.6, 0.2), (0.9, 0.1, 0.5), (0.05, 0.02, 0.04), (1.7, 0.4, 0.2)])
def test_russian_roulette_expectation(t):
    t = np.asarray(t, np.float64)
    count = 200_000
    acc = np.zeros(3)
    alive_n = 0
    for s in _seeds(count, 3):
        v = f3(t)
        st = C.c_uint32(int(s))
        if lib().oracle_russian_roulette_f32(v.ctypes.data, C.byref(st)):
            acc += v
            alive_n += 1
    p = t.max()
    mean = acc / count
    if p <= 1.0:
        assert abs(alive_n / count - p) <= 4 * np.sqrt(p * (1 - p) / count) + 1e-6
        np.testing.assert_allclose(mean, t, rtol=4 * np.sqrt((1 - p) / (p * count)) + 1e-6)
    else:
        # rand_f <= 1.0000005 < p: the path always survives and is divided by p
        assert alive_n == count
        np.testing.assert_allclose(mean, t / p, rtol=1e-6)


def _cornell_diffuse():
    """Cornell box whose throughput stays <= 1 under path_traceBBB (every
    bounce multiplies by a base colour <= 1)."""
    from pulse_amd import procedural as P
    return P.cornell()


def test_russian_roulette_is_unbiased_in_the_path_loop():
    """REF_NEE_LAMBERT (path_traceBBB): throughput *= base colour <= 1, so p
    <= 1 at every roulette and the image mean with roulette from depth 0
    equals the mean without roulette.  Standard errors from 32 batches of
    32 samples per pixel (1024 spp over 24 x 24 pixels)."""
    scene, camf = _cornell_diffuse()
    o = OracleScene(scene)
    cam = camf(24, 24)
    means = {}
    for rr in (0, 5):
        batch = []
        for k in range(32):
            img, _, _ = o.render(cam, make_params(accumulation_count=32 * k, spp=32, max_depth=5,
                                                  mode=N.MODE_REF_NEE_LAMBERT, rr_start_depth=rr), threads=8)
            # the running mean over a zeroed buffer from count n0 holds
            # (sum of the 32 samples) / (n0 + 32)
            batch.append(img[..., :3].astype(np.float64).mean() * (32 * k + 32) / 32)
        means[rr] = np.array(batch)
    diff = means[0].mean() - means[5].mean()
    se = np.sqrt(means[0].var(ddof=1) / 32 + means[5].var(ddof=1) / 32)
    print("rr mean", means[0].mean(), "no-rr mean", means[5].mean(), "se", se)
    assert abs(diff) <= 3 * se, (diff, se)
    # the roulette is active: deeper bounces are dropped in a sizeable share
    # of paths, so a missing 1/p would lower the mean well beyond 3 se
    assert se < 0.01 * means[5].mean()


def _plane_and_emitter(le_recv, base, roughness):
    sb = SceneBuilder()
    h = 3.0
    pos = np.array([(-h, 0, -h), (h, 0, -h), (h, 0, h), (-h, 0, h)], np.float32)
    recv = sb.add_mesh(Mesh(pos, np.tile(np.float32([0, 1, 0]), (4, 1)), np.array([0, 2, 1, 0, 3, 2], np.uint32)))
    tri = _horizontal_emitter((0.4, 1.0, 0.1), 0.1).astype(np.float32)
    emit = sb.add_mesh(Mesh(tri, np.tile(np.float32([0, -1, 0]), (3, 1)), np.array([0, 1, 2], np.uint32)))
    m_recv = sb.add_material(_material(base=base, roughness=roughness, emissive=le_recv))
    m_emit = sb.add_material(_material(emissive=(10.0, 10.0, 10.0)))
    sb.add_instance(recv, m_recv, np.eye(4))
    sb.add_instance(emit, m_emit, np.eye(4))
    return sb.build(1), tri.astype(np.float64)


@pytest.mark.parametrize("nee", [N.NEE_LAMBERT, N.NEE_GGX])
def test_one_bounce_pixel_is_emission_plus_nee(nee):
    """GGX_NEE at max_depth 1: the pixel is the hit's emission (added at
    depth 0 only) plus NEE at the hit, nothing else.  Its mean over 8192
    frames equals emission + the NEE integral over the small emitter (the
    receiver is a light too, but NEE from a point of its plane to the plane
    gives cos_r < 0 and is clamped to 0)."""
    le_recv = (0.3, 0.2, 0.1)
    base = (0.6, 0.5, 0.4)
    rough = 0.6
    scene, tri = _plane_and_emitter(le_recv, base, rough)
    eye = np.array([-1.0, 2.5, 1.5])
    cam = make_camera(8, 8, eye=eye, target=(0.3, 0.0, 0.0))
    o = OracleScene(scene)
    frames = 8192
    img, _, _ = o.render(cam, make_params(spp=frames, max_depth=1, mode=N.MODE_GGX_NEE, nee_eval=nee), threads=8)
    ivp = np.asarray(cam.inverse_view_proj, np.float64).reshape(4, 4).T
    n0 = np.array([0.0, 1.0, 0.0])
    f0 = f0_of(0.5, 0.0, base)
    checked = 0
    for (x, y) in [(2, 3), (5, 5), (3, 6), (6, 2)]:
        c = ivp @ np.array([2 * x / 8 - 1, -(2 * y / 8 - 1), 1.0, 1.0])
        d = _unit(c[:3] / c[3] - eye)
        p0 = eye + (-eye[1] / d[1]) * d  # hit on the plane y = 0
        if np.abs(p0[[0, 2]]).max() > 2.9:
            continue
        wo = -d
        if nee == N.NEE_LAMBERT:
            xq, da = tri_quadrature(*tri)
            dd = _unit(xq - (p0 + 0.001 * n0))
            cos_r = dd @ n0
            cos_e = dd[:, 1]
            nee_int = np.asarray(base) / np.pi * 10.0 * np.sum(np.clip(cos_r * cos_e, 0, None) / 4.0) * da
        else:
            nee_int = _nee_ggx_integral(tri, p0, n0, wo, rough, f0, 10.0, normalised_wi=False)
        exact = np.asarray(le_recv) + nee_int
        got = img[y, x, :3].astype(np.float64)
        np.testing.assert_allclose(got, exact, rtol=1.5e-2, err_msg=f"pixel {x},{y}")
        checked += 1
    assert checked >= 3


def test_emission_only_at_depth_zero():
    """GGX_NEE adds emission at the first hit only.  Floor (emissive, grey)
    under an emissive ceiling with a black base colour, Lambert NEE: at depth
    1 the paths that reach the ceiling add NEE x base = 0 and no emission, so
    max_depth 2 renders the max_depth 1 image bit for bit, while adding the
    ceiling's emission at depth 1 would brighten it."""
    sb = SceneBuilder()
    h = 3.0
    quad = np.array([(-h, 0, -h), (h, 0, -h), (h, 0, h), (-h, 0, h)], np.float32)
    up = sb.add_mesh(Mesh(quad, np.tile(np.float32([0, 1, 0]), (4, 1)), np.array([0, 2, 1, 0, 3, 2], np.uint32)))
    down = sb.add_mesh(Mesh(quad, np.tile(np.float32([0, -1, 0]), (4, 1)), np.array([0, 1, 2, 0, 2, 3], np.uint32)))
    m_floor = sb.add_material(_material(base=(0.6, 0.5, 0.4), roughness=0.7, emissive=(0.2, 0.2, 0.2)))
    m_ceil = sb.add_material(_material(base=(0.0, 0.0, 0.0), roughness=0.7, emissive=(5.0, 5.0, 5.0)))
    sb.add_instance(up, m_floor, np.eye(4))
    sb.add_instance(down, m_ceil, P_translate((0.0, 1.0, 0.0)))
    scene = sb.build(1)
    # steeply down: every camera ray's first hit is on the floor
    cam = make_camera(16, 16, eye=(0.0, 0.9, 0.3), target=(0.0, 0.0, 0.0))
    o = OracleScene(scene)
    # the horizon clamp ends the paths whose sampled direction points below
    # the floor (the reference keeps them, Q6, and they would re-hit the floor
    # at depth 1 and add its NEE); it does not touch emission
    imgs = [o.render(cam, make_params(spp=64, max_depth=d, mode=N.MODE_GGX_NEE, nee_eval=N.NEE_LAMBERT,
                                      flags=N.FLAG_HORIZON_CLAMP), threads=8)[0] for d in (1, 2)]
    assert (imgs[0][..., :3] > 0.2).mean() > 0.5
    np.testing.assert_array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))


def P_translate(t):
    m = np.eye(4)
    m[:3, 3] = t
    return m
