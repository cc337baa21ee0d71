"""ctypes binding of oracle/liboracle.so -- the CPU oracle (test infrastructure).

Only the tests, the smoke check and bench.py's CPU-baseline leg load it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from pulse_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
_P = C.c_void_p
_U32 = C.c_uint32
_lib = None

COUNTER_NAMES = ["tlas_interior", "tlas_leaf", "blas_interior", "blas_leaf", "instances_entered", "tri_tests",
                 "rays", "max_stack", "iteration_cap_hits", "shadow_rays", "primary_rays"]
CATEGORY_NAMES = ["primary", "bounce", "shadow"]
WORK_NAMES = ["tlas_interior", "tlas_leaf", "blas_interior", "blas_leaf", "instances_entered", "tri_tests", "rays"]
N_COUNTERS = 11 + 3 * 7


def parse_counters(cnt: np.ndarray) -> dict:
    out = dict(zip(COUNTER_NAMES, cnt[:11].tolist()))
    cat = cnt[11:].reshape(3, 7)
    out["by_category"] = {c: dict(zip(WORK_NAMES, cat[i].tolist())) for i, c in enumerate(CATEGORY_NAMES)}
    return out


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
        L = C.CDLL(str(ORACLE_SO), mode=os.RTLD_LOCAL)
        sig = [
            ("oracle_scene_create", _P, [C.POINTER(N.SceneDesc)]),
            ("oracle_scene_destroy", None, [_P]),
            ("oracle_scene_set_cdf_linear_scan", None, [_P, C.c_int]),
            ("oracle_scene_set_cull_world", None, [_P, C.c_int]),
            ("oracle_trace_rays_f32", C.c_int, [_P, _P, _P, _U32, C.c_int, C.c_int, _P]),
            ("oracle_trace_rays_f64", C.c_int, [_P, _P, _P, _P, _U32, C.c_int]),
            ("oracle_shadow_rays_f32", C.c_int, [_P, _P, _P, _U32, C.c_int]),
            ("oracle_intersect_candidates_f32", C.c_int, [_P, _P, _P, _P, _U32]),
            ("oracle_render_f32", C.c_double, [_P, C.POINTER(N.Camera), C.POINTER(N.Params), _P, C.c_int, _P, _U32, _U32]),
            ("oracle_render_pair", C.c_int, [_P, C.POINTER(N.Camera), C.POINTER(N.Params), _P, _P, C.c_int, C.c_int, _P]),
            ("oracle_grad_fd", C.c_int, [_P, C.POINTER(N.Camera), C.POINTER(N.Params), _P, _P, _U32, _P, _U32,
                                         C.c_double, _P, _P, C.c_int]),
            ("oracle_grad_fd_ex", C.c_int, [_P, C.POINTER(N.Camera), C.POINTER(N.Params), _P, _P, _U32, _P, _U32,
                                            C.c_double, _P, _P, C.c_int, C.c_int]),
            ("oracle_grad_fd_rows", C.c_int, [_P, C.POINTER(N.Camera), C.POINTER(N.Params), _P, _P, _U32, _P, _U32,
                                              C.c_double, _P, _P, C.c_int, C.c_int, _U32, _U32, C.c_int, _P]),
            ("oracle_strength_pick_array", _U32, [_P, _U32, C.c_float]),
            ("oracle_render_gi_f32", C.c_double, [_P, C.POINTER(N.Params), _P, _P, _P, _P, _U32, _U32, _P, C.c_int,
                                                  _P]),
            ("oracle_rand_u", _U32, [C.POINTER(_U32)]),
            ("oracle_rand_f", C.c_float, [C.POINTER(_U32)]),
            ("oracle_pixel_seed", _U32, [_U32, _U32]),
            ("oracle_sample_ggx_f32", None, [_P, _P, C.POINTER(N.Material), C.POINTER(_U32), _P]),
            ("oracle_sample_ggx_f64", None, [_P, _P, C.POINTER(N.Material), C.POINTER(_U32), _P]),
            ("oracle_sample_cosine_f32", None, [_P, C.c_float, C.c_float, _P]),
            ("oracle_ggx_half_vector_f32", None, [_P, C.c_float, C.c_float, C.c_float, _P]),
            ("oracle_d_ggx_f32", C.c_float, [C.c_float, C.c_float]),
            ("oracle_g_smith_f32", C.c_float, [C.c_float, C.c_float, C.c_float]),
            ("oracle_fresnel_f32", None, [C.c_float, _P, _P]),
            ("oracle_sample_triangle_f32", None, [C.c_float, C.c_float, _P, _P, _P, _P]),
            ("oracle_cdf_pick", _U32, [_P, C.c_float, _U32, _U32, C.c_int]),
            ("oracle_cdf_scan_array", _U32, [_P, _U32, C.c_float, C.c_int]),
            ("oracle_direct_light_f32", None, [_P, _P, _P, _P, C.POINTER(_U32), _P]),
            ("oracle_pt_math", None, [_P, _U32, _P, _P, _P]),
            ("oracle_direct_light_batch_f32", None, [_P, _P, _P, C.POINTER(N.Material), _P, C.c_int, _P, _U32, _P,
                                                     C.c_int]),
            ("oracle_sample_ggx_batch_f32", None, [_P, _P, C.POINTER(N.Material), _P, _U32, _P]),
            ("oracle_russian_roulette_f32", C.c_int, [_P, C.POINTER(_U32)]),
        ]
        for name, res, args in sig:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def f3(v) -> np.ndarray:
    return np.ascontiguousarray(v, dtype=np.float32)


class OracleScene:
    def __init__(self, scene):
        d = scene.desc()
        self.h = lib().oracle_scene_create(C.byref(d))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_scene_destroy(self.h)
            self.h = None

    def set_linear_scan(self, on: bool):
        lib().oracle_scene_set_cdf_linear_scan(self.h, int(on))

    def set_cull_world(self, on: bool):
        """Run the kernels' traversal (culling bounds, any-hit shadow rays)."""
        lib().oracle_scene_set_cull_world(self.h, int(on))

    def trace_rays(self, rays: np.ndarray, threads=8, bruteforce=False):
        rays = np.ascontiguousarray(rays, dtype=N.RAY_DTYPE)
        hits = np.zeros(len(rays), N.HIT_DTYPE)
        cnt = np.zeros(N_COUNTERS, np.uint64)
        lib().oracle_trace_rays_f32(self.h, rays.ctypes.data, hits.ctypes.data, len(rays), threads, int(bruteforce),
                                    cnt.ctypes.data)
        return hits, parse_counters(cnt)

    def intersect_candidates(self, rays: np.ndarray, inst: np.ndarray, tri: np.ndarray):
        """The reference's triangle test of one given (instance, triangle
        slot) per ray and the world t of its hit (t = 1e30: not hit)."""
        rays = np.ascontiguousarray(rays, dtype=N.RAY_DTYPE)
        ids = np.ascontiguousarray(np.stack([inst, tri], 1), dtype=np.uint32)
        hits = np.zeros(len(rays), N.HIT_DTYPE)
        assert lib().oracle_intersect_candidates_f32(self.h, rays.ctypes.data, ids.ctypes.data, hits.ctypes.data,
                                                     len(rays)) == 0
        return hits

    def trace_rays_f64(self, rays: np.ndarray, threads=8):
        rays = np.ascontiguousarray(rays, dtype=N.RAY_DTYPE)
        tuv = np.zeros((len(rays), 3), np.float64)
        ids = np.zeros((len(rays), 2), np.uint32)
        lib().oracle_trace_rays_f64(self.h, rays.ctypes.data, tuv.ctypes.data, ids.ctypes.data, len(rays), threads)
        return tuv, ids

    def shadow_rays(self, rays: np.ndarray, threads=8):
        rays = np.ascontiguousarray(rays, dtype=N.RAY_DTYPE)
        occ = np.zeros(len(rays), np.uint32)
        lib().oracle_shadow_rays_f32(self.h, rays.ctypes.data, occ.ctypes.data, len(rays), threads)
        return occ

    def render(self, camera, params, accum=None, threads=8, rows=(0, 0)):
        if accum is None:
            accum = np.zeros((camera.height, camera.width, 4), np.float32)
        cnt = np.zeros(N_COUNTERS, np.uint64)
        ms = lib().oracle_render_f32(self.h, C.byref(camera), C.byref(params), accum.ctypes.data, threads,
                                     cnt.ctypes.data, rows[0], rows[1])
        return accum, parse_counters(cnt), ms

    def render_gi(self, gbuffer, params, threads=8):
        """The realtime GI pass over a G-buffer: dict of (H, W, 4) float32
        arrays position / normal / base_color / emissive."""
        arrs = [np.ascontiguousarray(gbuffer[k], np.float32) for k in ("position", "normal", "base_color", "emissive")]
        h, w = arrs[0].shape[:2]
        out = np.zeros((h, w, 4), np.float32)
        lib().oracle_render_gi_f32(self.h, C.byref(params), *[a.ctypes.data for a in arrs], w, h, out.ctypes.data,
                                   threads, None)
        return out

    def grad_fd(self, camera, params, target, mat_params, ids, h=1e-6, threads=8, frozen=False, rows=None,
                eval32=False, per_pixel=False):
        """dL/dparam[ids] by f64 central differences along the f32 decisions;
        frozen=True also holds the sampled specular half vectors fixed (the
        detached roughness derivative); frozen="free" replays nothing (f64
        renders with free decisions: the derivative of the estimator itself).
        rows=(r0, r1): the loss over those rows only; pixels whose target is
        NaN are left out.  eval32: perturbed renders in f32 (use h ~ 1e-3).
        per_pixel: also return each pixel's share, (len(ids), rows, width)."""
        target = np.ascontiguousarray(target, np.float32)
        mp = np.ascontiguousarray(mat_params, np.float64)
        ids = np.ascontiguousarray(ids, np.uint32)
        out = np.zeros(len(ids), np.float64)
        loss = C.c_double()
        r0, r1 = (0, camera.height) if rows is None else rows
        px = np.zeros((len(ids), r1 - r0, camera.width), np.float64) if per_pixel else None
        rc = lib().oracle_grad_fd_rows(self.h, C.byref(camera), C.byref(params), target.ctypes.data, mp.ctypes.data,
                                       len(mp) // 7, ids.ctypes.data, len(ids), h, out.ctypes.data, C.byref(loss),
                                       threads, 2 if frozen == "free" else int(bool(frozen)), r0, r1, int(eval32),
                                       px.ctypes.data if per_pixel else None)
        assert rc == 0
        return (out, loss.value, px) if per_pixel else (out, loss.value)

    def render_pair(self, camera, params, replay=True, threads=8):
        a32 = np.zeros((camera.height, camera.width, 4), np.float32)
        a64 = np.zeros((camera.height, camera.width, 4), np.float64)
        flips = np.zeros(2, np.uint64)
        lib().oracle_render_pair(self.h, C.byref(camera), C.byref(params), a32.ctypes.data, a64.ctypes.data, threads,
                                 int(replay), flips.ctypes.data)
        return a32, a64, flips


def random_rays(n: int, rng: np.random.Generator, center, spread: float, target_spread: float | None = None):
    """Rays from points in a box around `center` towards other such points."""
    target_spread = spread if target_spread is None else target_spread
    c = np.asarray(center, np.float64)
    o = c + rng.uniform(-spread, spread, (n, 3))
    t = c + rng.uniform(-target_spread, target_spread, (n, 3))
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(n, N.RAY_DTYPE)
    rays["origin"] = o.astype(np.float32)
    rays["dir"] = d.astype(np.float32)
    rays["t_max"] = np.float32(1e30)
    return rays


def same_bits(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Element-wise bitwise equality of float arrays, any NaN equal to any NaN."""
    a = np.asarray(a)
    b = np.asarray(b)
    return (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
