"""Host mirror of the reference's plugin API on top of libpulse_hip.

Reference interface (src/lib.rs:13-19, src/path_tracer/mod.rs:35-118):
  * ``PulsePlugin`` / ``PulsePathTracerPlugin`` -- registered on the app;
  * ``PulsePathTracerCamera { resolution: Option<UVec2>, accumulation_count,
    previous_transform }`` -- the per-camera component;
  * ``reset_accumulation_on_movement`` -- zeroes the count when the camera
    moves, increments it otherwise;
  * ``PulsePathTracerNode::run`` -- dispatches ``path_trace`` into the rgba32f
    accumulation target (src/path_tracer/node.rs:40-119).
Here the node's dispatch is ``pulse_trace`` on a device context; the target
is a device buffer owned by the node (read back with ``readback``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .scene import SceneBuffers


@dataclass
class PulsePathTracerCamera:
    """src/path_tracer/mod.rs:100-105."""
    resolution: tuple | None = None
    accumulation_count: int = 0
    previous_transform: np.ndarray = field(default_factory=lambda: np.eye(4))


def reset_accumulation_on_movement(cam: PulsePathTracerCamera, current_transform: np.ndarray) -> None:
    """src/path_tracer/mod.rs:107-118."""
    if not np.array_equal(current_transform, cam.previous_transform):
        cam.accumulation_count = 0
    else:
        cam.accumulation_count += 1
    cam.previous_transform = np.array(current_transform, copy=True)


class Device:
    """A pulse_ctx on one HIP device (one per device; calls serialized)."""

    def __init__(self, device: int = 0):
        self._ctx = C.c_void_p()
        N.check(N.lib().pulse_create(device, C.byref(self._ctx)))

    @property
    def handle(self):
        return self._ctx

    def close(self):
        if self._ctx and self._ctx.value:
            N.lib().pulse_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        self.close()

    # --- memory helpers ---
    def alloc(self, nbytes: int) -> "DeviceArray":
        p = C.c_void_p()
        N.check(N.lib().pulse_device_alloc(self._ctx, nbytes, C.byref(p)))
        return DeviceArray(self, p, nbytes)

    def to_device(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = self.alloc(max(a.nbytes, 4))
        N.check(N.lib().pulse_copy_h2d(self._ctx, d.ptr, a.ctypes.data, a.nbytes))
        return d

    def sync(self) -> None:
        N.check(N.lib().pulse_sync(self._ctx))

    # --- scene and tracing ---
    def upload_scene(self, scene: SceneBuffers) -> None:
        d = scene.desc()
        N.check(N.lib().pulse_scene_upload(self._ctx, C.byref(d)))

    def trace_rays(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=N.RAY_DTYPE)
        n = len(rays)
        hits = np.zeros(n, N.HIT_DTYPE)
        if n == 0:
            return hits
        dr = self.to_device(rays)
        dh = self.alloc(hits.nbytes)
        N.check(N.lib().pulse_trace_rays(self._ctx, dr.ptr, dh.ptr, n))
        self.sync()
        dh.copy_to(hits)
        return hits

    def trace_shadow_rays(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=N.RAY_DTYPE)
        n = len(rays)
        occ = np.zeros(n, np.uint32)
        if n == 0:
            return occ
        dr = self.to_device(rays)
        do = self.alloc(occ.nbytes)
        N.check(N.lib().pulse_trace_shadow_rays(self._ctx, dr.ptr, do.ptr, n))
        self.sync()
        do.copy_to(occ)
        return occ

    def set_option(self, option: int, value: int) -> None:
        N.check(N.lib().pulse_set_option(self._ctx, option, value))

    def get_option(self, option: int) -> int:
        v = C.c_int64()
        N.check(N.lib().pulse_get_option(self._ctx, option, C.byref(v)))
        return v.value

    def stats(self) -> N.Stats:
        s = N.Stats()
        N.check(N.lib().pulse_get_stats(self._ctx, C.byref(s)))
        return s


class DeviceArray:
    def __init__(self, dev: Device, p: C.c_void_p, nbytes: int):
        self.dev, self.p, self.nbytes = dev, p, nbytes

    @property
    def ptr(self):
        return self.p

    def copy_to(self, out: np.ndarray) -> np.ndarray:
        assert out.nbytes <= self.nbytes and out.flags["C_CONTIGUOUS"]
        N.check(N.lib().pulse_copy_d2h(self.dev.handle, out.ctypes.data, self.p, out.nbytes))
        return out

    def copy_from(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        N.check(N.lib().pulse_copy_h2d(self.dev.handle, self.p, a.ctypes.data, a.nbytes))

    def free(self):
        if self.p is not None and self.p.value:
            N.lib().pulse_device_free(self.dev.handle, self.p)
            self.p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def make_params(accumulation_count=0, spp=1, max_depth=5, mode=N.MODE_REF_GGX, nee_eval=N.NEE_LAMBERT,
                rr_start_depth=0, flags=0) -> N.Params:
    p = N.Params()
    p.accumulation_count = accumulation_count
    p.spp = spp
    p.max_depth = max_depth
    p.mode = mode
    p.nee_eval = nee_eval
    p.rr_start_depth = rr_start_depth
    p.flags = flags
    return p


class PulsePathTracerNode:
    """PulsePathTracerNode::run (src/path_tracer/node.rs:40-119) on a Device:
    owns the rgba32f accumulation target of one camera."""

    def __init__(self, device: Device, width: int, height: int):
        self.device, self.width, self.height = device, width, height
        self.target = device.alloc(width * height * 16)
        N.check(N.lib().pulse_memset_d(device.handle, self.target.ptr, 0, width * height * 16))

    def run(self, camera: N.Camera, params: N.Params) -> None:
        assert camera.width == self.width and camera.height == self.height
        N.check(N.lib().pulse_trace(self.device.handle, C.byref(camera), C.byref(params), self.target.ptr))

    def readback(self) -> np.ndarray:
        self.device.sync()
        out = np.zeros((self.height, self.width, 4), np.float32)
        return self.target.copy_to(out)


class PulsePlugin:
    """src/lib.rs:13-19: owns the scene side (device context + scene upload)."""

    def __init__(self, device: int = 0):
        self.device = Device(device)
        self._generation = None

    def prepare_scene(self, scene: SceneBuffers) -> None:
        self.device.upload_scene(scene)


class PulsePathTracerPlugin:
    """src/path_tracer/mod.rs:35-91: one accumulation node per camera."""

    def __init__(self, plugin: PulsePlugin):
        self.plugin = plugin
        self.nodes: dict = {}

    def render(self, camera_id, cam: PulsePathTracerCamera, camera: N.Camera, params: N.Params | None = None):
        w, h = cam.resolution or (camera.width, camera.height)
        node = self.nodes.get(camera_id)
        if node is None or (node.width, node.height) != (w, h):
            node = self.nodes[camera_id] = PulsePathTracerNode(self.plugin.device, w, h)
        p = params or make_params()
        p.accumulation_count = cam.accumulation_count
        node.run(camera, p)
        return node
