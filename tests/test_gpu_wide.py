"""GPU: the 4-wide BVH traversal (traversal loop 2) against the f32 oracle.

The wide BVHs are collapsed at upload from the reference's binary BLAS/TLAS:
same leaves, same triangle slots, same per-instance closest hit and world-t
conversion.  Only the order in which leaves are reached differs from the
reference's binary near-first order, and that order matters only between
hits whose t are equal or within rounding of each other (a tie the first
found wins; a box whose rounded entry distance lies just above a hit inside
it, which the reference skips and the padded wide box test does not).

Every ray whose hit differs from the oracle's is re-intersected on both
sides with the oracle's single-candidate triangle test: each side must be a
genuine hit of the reference's test reproducing its t, u, v bit for bit,
and the two world t must be equal or differ by at most TIE_BOUND of the
distance scale max(t, |origin|) (the rounding of a box entry distance and of
the world-t conversion, a few ulps; the wide slab test pads by 2^-20).
Occlusion equals the oracle's except where the closest t ties the
threshold.  Images equal the oracle's bit for bit: no pixel of the test
scenes or of the 1080p benchmark frame differs.
"""
import numpy as np
import pytest

from oracle_lib import OracleScene, random_rays, same_bits
from pulse_amd import _native as N
from pulse_amd import procedural as P
from pulse_amd.path_tracer import Device, PulsePathTracerNode, make_params

pytestmark = pytest.mark.gpu

TIE_BOUND = 2.0 ** -18  # of max(t, |origin|): 32 ulps of the distance scale


@pytest.fixture(scope="module")
def dev():
    d = Device(0)
    d.set_option(N.OPT_TRAVERSAL_LOOP, 2)
    yield d
    d.close()


@pytest.fixture(scope="module")
def cornell():
    return P.cornell()


@pytest.fixture(scope="module")
def field():
    return P.icosphere_field(n_meshes=8, subdiv=3, n_instances=60, n_emissive=6, box=30.0, radius=(1.0, 3.0))


def _check_hits(g, o, n, rays=None, orc=None, max_ties=None):
    """Every difference is a t-tie (at most `max_ties` of them, 0.1 % of the
    rays by default): both candidates re-intersect to the reported hits and
    their world t agree within TIE_BOUND of the distance scale."""
    same = (same_bits(g["t"], o["t"]) & same_bits(g["u"], o["u"]) & same_bits(g["v"], o["v"])
            & (g["instance_index"] == o["instance_index"]) & (g["triangle_index"] == o["triangle_index"]))
    bad = ~same
    if bad.any():
        idx = np.flatnonzero(bad)
        gt, ot = g["t"][bad].astype(np.float64), o["t"][bad].astype(np.float64)
        assert (gt < 1e30).all() and (ot < 1e30).all(), f"a hit missing on one side: rays {idx[:5]}"
        if rays is not None:
            for side in (g, o):
                r = orc.intersect_candidates(rays[bad], side["instance_index"][bad], side["triangle_index"][bad])
                for k in ("t", "u", "v"):
                    ok = same_bits(r[k], side[k][bad])
                    assert ok.all(), f"{k} of a reported hit is not the triangle test's: rays {idx[~ok][:5]}"
            scale = np.maximum(np.abs(rays["origin"][bad]).max(axis=1).astype(np.float64), np.maximum(gt, ot))
        else:
            scale = np.maximum(gt, ot)
        dt = np.abs(gt - ot)
        exact = int((dt == 0).sum())
        print(f"t-ties: {bad.sum()} of {n} rays ({exact} exactly equal t), max |dt|/scale "
              f"{float((dt / scale).max()):.3g}", flush=True)
        assert (dt <= TIE_BOUND * scale).all(), f"non-tie differences: rays {idx[dt > TIE_BOUND * scale][:5]}"
        limit = max(1, n // 1000) if max_ties is None else max_ties
        assert bad.sum() <= limit, f"{bad.sum()} t-ties of {n}"
    return int(bad.sum())


@pytest.mark.parametrize("which", ["cornell", "field", "bench"])
def test_wide_trace_rays(dev, cornell, field, which):
    if which == "bench":
        scene, _ = P.icosphere_field()
        spread = 50.0
    else:
        scene, _ = cornell if which == "cornell" else field
        spread = 0.95 if which == "cornell" else 20.0
    rays = random_rays(20000, np.random.default_rng(11), (0, 0, 0), spread)
    dev.upload_scene(scene)
    assert dev.get_option(N.OPT_TRAVERSAL_LOOP) == 2
    g = dev.trace_rays(rays)
    orc = OracleScene(scene)
    o, _ = orc.trace_rays(rays, threads=16)
    _check_hits(g, o, len(rays), rays, orc)
    assert (g["t"] < 1e30).mean() > 0.2
    assert dev.stats().error_flags == 0


def test_wide_axis_parallel_and_tmax(dev, cornell):
    scene, _ = cornell
    rng = np.random.default_rng(12)
    n = 4000
    rays = random_rays(n, rng, (0, 0, 0), 0.9)
    axes = np.eye(3, dtype=np.float32)
    rays["dir"][: n // 2] = axes[rng.integers(0, 3, n // 2)] * rng.choice([-1.0, 1.0], (n // 2, 1)).astype(np.float32)
    rays["t_max"][n // 2:] = rng.uniform(0.1, 2.0, n - n // 2).astype(np.float32)
    dev.upload_scene(scene)
    g = dev.trace_rays(rays)
    orc = OracleScene(scene)
    o, _ = orc.trace_rays(rays)
    # axis-parallel rays in an axis-aligned box meet shared edges of
    # coplanar and perpendicular faces: exact ties are common among them
    h = n // 2
    ties = _check_hits(g[:h], o[:h], h, rays[:h], orc, max_ties=h // 20)
    _check_hits(g[h:], o[h:], n - h, rays[h:], orc)
    print(f"axis-parallel t-ties: {ties} of {h}")


def test_wide_matches_bruteforce(dev, field):
    """Against every triangle of every instance tested in order (no BVH)."""
    scene, _ = field
    rays = random_rays(3000, np.random.default_rng(13), (0, 0, 0), 20.0)
    dev.upload_scene(scene)
    g = dev.trace_rays(rays)
    orc = OracleScene(scene)
    b, _ = orc.trace_rays(rays, threads=16, bruteforce=True)
    _check_hits(g, b, len(rays), rays, orc)


@pytest.mark.parametrize("which", ["cornell", "field"])
def test_wide_shadow_rays(dev, cornell, field, which):
    scene, _ = cornell if which == "cornell" else field
    rng = np.random.default_rng(14)
    rays = random_rays(20000, rng, (0, 0, 0), 0.95 if which == "cornell" else 20.0)
    orc = OracleScene(scene)
    hits, _ = orc.trace_rays(rays)
    t = np.where(hits["t"] < 1e30, hits["t"], 10.0).astype(np.float32)
    thr = t * rng.uniform(0.5, 1.5, len(t)).astype(np.float32)
    thr[::7] = t[::7]
    thr[1::7] = np.nextafter(t[1::7], np.float32(np.inf))
    thr[2::13] = -1.0
    rays["t_max"] = thr
    dev.upload_scene(scene)
    g = dev.trace_shadow_rays(rays)
    o = orc.shadow_rays(rays)
    bad = g != o
    if bad.any():
        # the closest t ties the threshold: another hit within the tie bound
        # of it falls on the other side
        scale = np.maximum(np.abs(rays["origin"][bad]).max(axis=1), t[bad]).astype(np.float64)
        near = np.abs(t[bad].astype(np.float64) - thr[bad]) <= TIE_BOUND * scale
        print(f"occlusion ties: {bad.sum()} of {len(rays)}", flush=True)
        assert near.all() and bad.sum() <= len(rays) // 1000
    assert 0.05 < g.mean() < 0.95


def _render(dev, scene, cam, prm):
    dev.upload_scene(scene)
    node = PulsePathTracerNode(dev, cam.width, cam.height)
    node.run(cam, prm)
    return node.readback()


def _check_image(g, o):
    """No pixel differs: the scenes here have no t-tie on any path."""
    px = ~same_bits(g, o).all(axis=-1)
    assert not px.any(), f"{px.sum()} of {px.size} pixels differ, first at {np.argwhere(px)[:3].tolist()}"
    return 0


@pytest.mark.parametrize("mode,nee,depth", [(N.MODE_REF_GGX, N.NEE_LAMBERT, 5), (N.MODE_REF_NEE_LAMBERT, N.NEE_LAMBERT, 5),
                                            (N.MODE_GGX_NEE, N.NEE_LAMBERT, 4), (N.MODE_GGX_NEE, N.NEE_GGX, 4)])
def test_wide_render_cornell(dev, cornell, mode, nee, depth):
    scene, camf = cornell
    cam = camf(96, 80)
    prm = make_params(spp=2, max_depth=depth, mode=mode, nee_eval=nee)
    g = _render(dev, scene, cam, prm)
    o, _, _ = OracleScene(scene).render(cam, prm)
    _check_image(g, o)


@pytest.mark.parametrize("w,h", [(1, 1), (3, 2), (7, 1), (9, 3), (17, 1)])
def test_wide_render_tiny_queues(dev, cornell, w, h):
    """Persistent launches split each queue into 8 parts, one per XCD:
    queues shorter than 8 rays (empty parts) and parts of uneven length
    trace every ray once, as the oracle does."""
    scene, camf = cornell
    cam = camf(w, h)
    prm = make_params(spp=8, max_depth=4, mode=N.MODE_GGX_NEE)
    g = _render(dev, scene, cam, prm)
    o, _, _ = OracleScene(scene).render(cam, prm)
    _check_image(g, o)


@pytest.mark.parametrize("waves,lds,tps", [(0, 0, 1), (24, 24, 2), (16, 8, 1), (64, 16, 2)])
def test_wide_render_field_persistent(dev, field, waves, lds, tps):
    """One ray per lane and persistent launches (stack partly in device
    memory) trace the same rays the same way: identical images."""
    scene, camf = field
    cam = camf(120, 72)
    dev.set_option(N.OPT_PERSISTENT_WAVES, waves)
    dev.set_option(N.OPT_LDS_STACK_ENTRIES, lds)
    dev.set_option(N.OPT_TRIS_PER_STEP, tps)
    try:
        prm = make_params(spp=1, max_depth=4, mode=N.MODE_GGX_NEE, accumulation_count=3)
        g = _render(dev, scene, cam, prm)
        o, _, _ = OracleScene(scene).render(cam, prm)
        _check_image(g, o)
    finally:
        dev.set_option(N.OPT_PERSISTENT_WAVES, 24)
        dev.set_option(N.OPT_LDS_STACK_ENTRIES, 24)
        dev.set_option(N.OPT_TRIS_PER_STEP, 2)


def test_wide_full_frame(dev):
    """The 1080p benchmark frame (64 subdivision-6 meshes, 800 instances)."""
    scene, camf = P.icosphere_field()
    cam = camf(1920, 1080)
    prm = make_params(max_depth=4, mode=N.MODE_GGX_NEE, flags=N.FLAG_COLLECT_STATS)
    g = _render(dev, scene, cam, prm)
    assert dev.stats().error_flags == 0
    o, _, _ = OracleScene(scene).render(cam, prm, threads=16)
    _check_image(g, o)


def test_wide_empty_scene(dev):
    from pulse_amd.scene import SceneBuffers
    bufs = {name: np.zeros(0, dt) for name, dt in N.BUFFER_DTYPES.items()}
    bufs["tlas_nodes"] = np.zeros(1, N.TLAS_NODE_DTYPE)
    scene = SceneBuffers(0, 0, bufs)
    _, camf = P.cornell()
    g = _render(dev, scene, camf(32, 32), make_params(max_depth=5))
    assert (g[..., :3] == 0).all() and (g[..., 3] == 1).all()


def test_wide_instance_update(dev):
    """Moving instances rebuilds the wide TLAS: the image is the oracle's
    for the moved scene, and the one a full upload of it gives."""
    from test_instance_update import _make, _transforms
    a = _make(_transforms(24, 1), keep=True)
    moved = _transforms(24, 2)
    for i in (0, 3, 5, 17, 23):
        a.builder.set_instance_transform(i, moved[i])
    b = a.builder.update_instances(a)
    cam = P.make_camera(64, 48, eye=(0.0, 0.0, 40.0), target=(0.0, 0.0, 0.0))
    prm = make_params(spp=2, max_depth=4, mode=N.MODE_GGX_NEE, rr_start_depth=1)
    dev.upload_scene(a)
    node = PulsePathTracerNode(dev, cam.width, cam.height)
    node.run(cam, prm)
    dev.update_instances(b)
    node.run(cam, prm)
    after = node.readback()
    o, _, _ = OracleScene(b).render(cam, prm)
    _check_image(after, o)
    full = _render(dev, b, cam, prm)
    assert same_bits(full, after).all()


def test_fast_reciprocal_is_the_division(dev):
    """The triangle tests take 1/a as the hardware reciprocal refined by one
    fused Newton step; the library's self-test compares it with the
    correctly rounded division for every f32 with |a| in [2^-100, 2^126)
    (both signs, every mantissa): no difference."""
    assert dev.selftest_reciprocal() == 0


def test_tiny_instance_miss_records(dev):
    """An instance scaled down by 1e-12: a ray that enters its box and misses
    its triangles has a finite world t through the miss point (1e30 in
    object space), which the reference records with its unhit object record
    -- instance 0, triangle 0, u = v = 0 (utilities.wgsl:265, 318-325).  The
    4-wide traversal's instance records mark such an instance (its miss need
    not overflow), so the merge still runs for it; every traversal loop
    equals the oracle."""
    from pulse_amd.scene import MaterialDesc, SceneBuilder

    rng = P.CounterRng(21)
    sb = SceneBuilder()
    mesh = sb.add_mesh(P.displaced_icosphere(2, 1.0, rng, 7))
    mat = sb.add_material(MaterialDesc(base_color=(0.5, 0.5, 0.5, 1.0), perceptual_roughness=0.5))
    for k in range(6):  # ordinary instances away from the rays
        m = np.eye(4)
        m[:3, 3] = (20.0 + 4.0 * k, -3.0 + k, 5.0)
        sb.add_instance(mesh, mat, m)
    tiny = np.eye(4)
    tiny[:3, :3] *= 1e-12  # at the world origin, where f32 resolves its 1e-12 box
    sb.add_instance(mesh, mat, tiny)
    scene = sb.build(2)
    r = np.random.default_rng(5)
    n = 2048
    # from 1e-5 away towards a point of its box outside the sphere: in object
    # space the aim spreads by ~0.6 units, so rays both hit and miss the mesh
    target = np.full(3, 0.95e-12)
    u = r.normal(size=(n, 3))
    o = target + 1e-5 * u / np.linalg.norm(u, axis=1, keepdims=True)
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(n, N.RAY_DTYPE)
    rays["origin"] = o.astype(np.float32)
    rays["dir"] = d.astype(np.float32)
    rays["t_max"] = np.float32(1e30)
    orc = OracleScene(scene)
    want, _ = orc.trace_rays(rays)
    spurious = (want["t"] < 1e30) & (want["t"] > 1e10)
    assert spurious.sum() > n // 4, "the rays should enter the tiny instance and miss its triangles"
    assert (want["instance_index"][spurious] == 0).all() and (want["triangle_index"][spurious] == 0).all()
    try:
        for loop in (1, 0):  # the reference's order: bit for bit
            dev.set_option(N.OPT_TRAVERSAL_LOOP, loop)
            dev.upload_scene(scene)
            g = dev.trace_rays(rays)
            for k in ("t", "u", "v", "instance_index", "triangle_index"):
                assert same_bits(g[k], want[k]).all(), f"loop {loop}: {k}"
    finally:
        dev.set_option(N.OPT_TRAVERSAL_LOOP, 2)
    # The 4-wide boxes are padded where the reference's rounded box tests
    # (here at object-space distances of 1e7) skip boxes holding a hit, so
    # some of its misses are hits or entries here.  Every result is a real
    # hit of the reference's triangle test, a miss, or the unhit record with
    # the reference's miss-point t.
    dev.upload_scene(scene)
    g = dev.trace_rays(rays)
    sp = (g["t"] < 1e30) & (g["t"] > 1e10)
    assert sp.sum() > n // 8
    assert (g["instance_index"][sp] == 0).all() and (g["triangle_index"][sp] == 0).all()
    assert (g["u"][sp] == 0).all() and (g["v"][sp] == 0).all()
    both = sp & spurious
    assert both.sum() > n // 8 and same_bits(g["t"][both], want["t"][both]).all()
    hit = g["t"] <= 1e10
    r = orc.intersect_candidates(rays[hit], g["instance_index"][hit], g["triangle_index"][hit])
    for k in ("t", "u", "v"):
        assert same_bits(r[k], g[k][hit]).all(), k
