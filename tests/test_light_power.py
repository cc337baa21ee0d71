"""Light selection by emission strength (PULSE_FLAG_LIGHT_POWER): the
reference builds light_emission_strength_cdf (src/scene/mod.rs:726-742) and
has the pick and its pdf written but commented out (utilities.wgsl:501,
541-546, 659-673).  CPU pins of the oracle's restatement."""
import ctypes as C

import numpy as np

from oracle_lib import OracleScene, lib
from pulse_amd import _native as N
from pulse_amd import procedural as P
from pulse_amd.path_tracer import make_params


def test_strength_pick_is_clamped_upper_bound():
    """The reference's scan returns min(#{cdf <= e}, count - 1): the upper
    bound, clamped, also over flat steps (zero-strength lights) and for
    e >= 1 (rand_f reaches 1.0000005)."""
    rng = np.random.default_rng(3)
    for trial in range(300):
        n = int(rng.integers(1, 40))
        w = rng.random(n) * (rng.random(n) > 0.3)
        if w.sum() == 0:
            w[0] = 1.0
        cdf = np.cumsum((w / w.sum()).astype(np.float32), dtype=np.float32)
        for e in list(rng.random(8).astype(np.float32)) + [np.float32(0), np.float32(1.0000005), cdf[n // 2]]:
            got = lib().oracle_strength_pick_array(cdf.ctypes.data, n, C.c_float(float(e)))
            want = min(int(np.searchsorted(cdf, e, side="right")), n - 1)
            assert got == want, (trial, n, float(e), got, want)


def test_zero_strength_light_never_picked_below_one():
    cdf = np.array([0.25, 0.25, 0.75, 1.0], np.float32)
    for e in np.linspace(0, 0.999, 997, dtype=np.float32):
        assert lib().oracle_strength_pick_array(cdf.ctypes.data, 4, C.c_float(float(e))) != 1


def test_light_power_keeps_expectation_and_lowers_variance():
    """Lambert NEE (path_traceBBB mode) on a box lit by three emitters of
    radiance 2, 40 and 8: picking by strength leaves the 1024-spp image
    mean unchanged (within 0.5 %) and lowers the 16-spp error against it."""
    scene, camf = P.cornell_lights()
    o = OracleScene(scene)
    cam = camf(24, 24)
    res = {}
    for name, fl in (("uniform", 0), ("power", N.FLAG_LIGHT_POWER)):
        ref, _, _ = o.render(cam, make_params(spp=1024, max_depth=2, mode=N.MODE_REF_NEE_LAMBERT, rr_start_depth=3,
                                              flags=fl))
        lo, _, _ = o.render(cam, make_params(spp=16, max_depth=2, mode=N.MODE_REF_NEE_LAMBERT, rr_start_depth=3,
                                             flags=fl))
        res[name] = (ref[..., :3].astype(np.float64), lo[..., :3].astype(np.float64))
    mu, mp = res["uniform"][0].mean(), res["power"][0].mean()
    assert abs(mu - mp) <= 5e-3 * mu
    best = 0.5 * (res["uniform"][0] + res["power"][0])
    err = {k: float(((v[1] - best) ** 2).mean()) for k, v in res.items()}
    assert err["power"] < 0.9 * err["uniform"], err


def test_light_power_same_rng_consumption():
    """One random number picks the light either way, so with a single
    light (pdf 1 under both rules) the images are identical."""
    scene, camf = P.cornell()
    o = OracleScene(scene)
    cam = camf(16, 16)
    a, _, _ = o.render(cam, make_params(spp=2, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=1))
    b, _, _ = o.render(cam, make_params(spp=2, max_depth=3, mode=N.MODE_GGX_NEE, rr_start_depth=1,
                                        flags=N.FLAG_LIGHT_POWER))
    assert np.array_equal(a, b)
