"""CPU checks of the 4-wide BVHs the library derives at upload: its host code
(pulse_amd/csrc/host/wide_bvh.cpp) compiled into a test helper
(tests/wide_mirror.cpp) that walks the quantised layout as the kernels do,
with the oracle's triangle test.  Layout: every binary leaf's primitives
appear once in wide leaf entries of at most 32, every quantised box contains
its exact box.  Walk: the closest hits equal the oracle's (the reference's
binary walk) except on t-ties, and the stack never exceeds the layout's
bound."""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle_lib import OracleScene, random_rays, same_bits
from pulse_amd import _native as N
from pulse_amd import procedural as P

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def mirror(tmp_path_factory):
    out = tmp_path_factory.mktemp("wide") / "libwide_mirror.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                    "-I", str(ROOT / "include"), "-o", str(out), str(ROOT / "tests" / "wide_mirror.cpp"),
                    str(ROOT / "pulse_amd" / "csrc" / "host" / "wide_bvh.cpp")], check=True)
    L = C.CDLL(str(out))
    L.wide_mirror_trace.argtypes = [C.POINTER(N.SceneDesc), C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
    L.wide_mirror_check.argtypes = [C.POINTER(N.SceneDesc), C.c_void_p]
    return L


SCENES = {
    "cornell": (lambda: P.cornell(), 0.95),
    "field": (lambda: P.icosphere_field(n_meshes=8, subdiv=3, n_instances=60, n_emissive=6, box=30.0,
                                        radius=(1.0, 3.0)), 20.0),
    "dense": (lambda: P.icosphere_field(n_meshes=6, subdiv=4, n_instances=120, n_emissive=8, box=24.0), 16.0),
}


@pytest.mark.parametrize("which", list(SCENES))
def test_layout_invariants(mirror, which):
    scene, _ = SCENES[which][0]()
    d = scene.desc()
    out = np.zeros(3, np.uint64)
    assert mirror.wide_mirror_check(C.byref(d), out.ctypes.data) == 0
    assert out[0] > 0 and out[1] == 0 and out[2] == 0, out


@pytest.mark.parametrize("which", list(SCENES))
def test_walk_matches_oracle(mirror, which):
    make, spread = SCENES[which]
    scene, _ = make()
    rays = random_rays(6000, np.random.default_rng(21), (0, 0, 0), spread)
    d = scene.desc()
    hits = np.zeros(len(rays), N.HIT_DTYPE)
    stats = np.zeros(3, np.uint64)
    assert mirror.wide_mirror_trace(C.byref(d), rays.ctypes.data, hits.ctypes.data, len(rays),
                                    stats.ctypes.data) == 0
    o, _ = OracleScene(scene).trace_rays(rays)
    same = (same_bits(hits["t"], o["t"]) & same_bits(hits["u"], o["u"]) & same_bits(hits["v"], o["v"])
            & (hits["instance_index"] == o["instance_index"]) & (hits["triangle_index"] == o["triangle_index"]))
    bad = ~same
    dt = np.abs(hits["t"][bad].astype(np.float64) - o["t"][bad])
    assert (dt <= 1e-5 * np.abs(o["t"][bad])).all() and bad.sum() <= len(rays) // 1000
    assert (hits["t"] < 1e30).mean() > 0.2
    assert stats[1] <= stats[2], stats  # deepest stack within the layout's bound


def test_axis_parallel_rays(mirror):
    scene, _ = P.cornell()
    rng = np.random.default_rng(22)
    n = 2000
    rays = random_rays(n, rng, (0, 0, 0), 0.9)
    axes = np.eye(3, dtype=np.float32)
    rays["dir"] = axes[rng.integers(0, 3, n)] * rng.choice([-1.0, 1.0], (n, 1)).astype(np.float32)
    d = scene.desc()
    hits = np.zeros(n, N.HIT_DTYPE)
    stats = np.zeros(3, np.uint64)
    assert mirror.wide_mirror_trace(C.byref(d), rays.ctypes.data, hits.ctypes.data, n, stats.ctypes.data) == 0
    o, _ = OracleScene(scene).trace_rays(rays)
    bad = ~(same_bits(hits["t"], o["t"]) & (hits["instance_index"] == o["instance_index"])
            & (hits["triangle_index"] == o["triangle_index"]))
    dt = np.abs(hits["t"][bad].astype(np.float64) - o["t"][bad])
    assert (dt <= 1e-5 * np.abs(o["t"][bad])).all()  # only t-ties (common for axis-parallel rays here)
    assert same_bits(hits["t"], o["t"]).mean() > 0.95
