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
import weakref
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
        self._children = []  # weak references to scenes and batches of this context
        N.check(N.lib().pulse_create(device, C.byref(self._ctx)))

    @property
    def handle(self):
        return self._ctx

    def close(self):
        if self._ctx and self._ctx.value:
            for ref in self._children:
                obj = ref()
                if obj is not None:
                    obj.close()
            self._children = []
            N.lib().pulse_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def _adopt(self, obj) -> None:
        self._children = [r for r in self._children if r() is not None]
        self._children.append(weakref.ref(obj))

    # --- several scenes and batched views ---
    def create_scene(self, scene: SceneBuffers) -> "DeviceScene":
        """A device-resident scene beside the uploaded one (pulse_scene_create)."""
        return DeviceScene(self, scene)

    def update_instances(self, scene: SceneBuffers, target: "DeviceScene" = None) -> None:
        """Per-frame instance update (pulse_scene_update_instances) of the
        uploaded scene, or of `target`: only instances, TLAS and instance
        indices travel to the device."""
        d = scene.desc()
        N.check(N.lib().pulse_scene_update_instances(self._ctx, target.handle if target is not None else None,
                                                     C.byref(d)))

    def refit_instances(self, object_world: np.ndarray, target: "DeviceScene" = None) -> None:
        """Instance motion on the device (pulse_scene_refit_instances): new
        object-to-world matrices, (n, 4, 4) column-major as the instance
        records store them (or (n, 16)); the 4-wide TLAS is refit, after
        which only traversal loop 2 runs on the scene."""
        m = np.ascontiguousarray(object_world, dtype=np.float32).reshape(-1, 16)
        N.check(N.lib().pulse_scene_refit_instances(self._ctx, target.handle if target is not None else None,
                                                    m.ctypes.data, len(m)))

    def bind_scene(self, scene: "DeviceScene" = None) -> None:
        """Makes `scene` (None: the uploaded scene) the one pulse_trace uses."""
        N.check(N.lib().pulse_scene_bind(self._ctx, scene.handle if scene is not None else None))

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

    def selftest_reciprocal(self) -> int:
        """Values of |a| in [2^-100, 2^126) where the kernels' fast reciprocal
        differs from 1.0f / a (expected 0)."""
        n = C.c_uint64(0)
        N.check(N.lib().pulse_selftest_reciprocal(self._ctx, C.byref(n)))
        return n.value

    # --- scene and tracing ---
    def rebuild_instances(self, object_world: np.ndarray, target: "DeviceScene" = None) -> None:
        """As refit_instances, with a new 4-wide TLAS built on the device
        (pulse_scene_rebuild_instances) instead of a refit of the last one."""
        m = np.ascontiguousarray(object_world, dtype=np.float32).reshape(-1, 16)
        N.check(N.lib().pulse_scene_rebuild_instances(self._ctx, target.handle if target is not None else None,
                                                      m.ctypes.data, len(m)))

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


class DeviceScene:
    """pulse_scene: a validated scene in device memory, owned by a Device."""

    def __init__(self, dev: Device, scene: SceneBuffers):
        self.dev = dev
        self.buffers = scene
        self._h = C.c_void_p()
        d = scene.desc()
        N.check(N.lib().pulse_scene_create(dev.handle, C.byref(d), C.byref(self._h)))
        self.n_materials = len(scene.buffers["materials"])
        dev._adopt(self)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h and self._h.value:
            N.lib().pulse_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """pulse_batch: many views (scene, camera, accumulation buffer) rendered
    by one captured launch sequence per frame."""

    def __init__(self, dev: Device, views, params: N.Params):
        """views: list of (DeviceScene, Camera); accumulation buffers are
        allocated here (width*height rgba32f each)."""
        self.dev = dev
        self.params = params
        self.scenes = [sc for sc, _ in views]
        self.cameras = [cam for _, cam in views]
        self.accum = [dev.alloc(cam.width * cam.height * 16) for cam in self.cameras]
        arr = (N.View * len(views))()
        for i, (sc, cam) in enumerate(views):
            arr[i].scene = sc.handle.value
            arr[i].camera = cam
            arr[i].accum_rgba = self.accum[i].ptr.value
        self._h = C.c_void_p()
        N.check(N.lib().pulse_batch_create(dev.handle, arr, len(views), C.byref(params), C.byref(self._h)))
        dev._adopt(self)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h and self._h.value:
            N.lib().pulse_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, accumulation_count: int = 0) -> None:
        N.check(N.lib().pulse_batch_render(self._h, accumulation_count))

    def images(self):
        self.dev.sync()
        return [a.copy_to(np.zeros((c.height, c.width, 4), np.float32)) for a, c in zip(self.accum, self.cameras)]

    def config(self) -> N.BatchConfig:
        c = N.BatchConfig()
        N.check(N.lib().pulse_batch_get_config(self._h, C.byref(c)))
        return c

    def set_config(self, **fields) -> None:
        """Changes the named fields of the current launch shape (a new
        traversal loop picks its own stack size unless one is given)."""
        c = self.config()
        if "traversal_loop" in fields and "stack_entries" not in fields:
            c.stack_entries = 0
        for k, v in fields.items():
            setattr(c, k, v)
        N.check(N.lib().pulse_batch_set_config(self._h, C.byref(c)))

    def set_camera(self, view: int, cam: N.Camera) -> None:
        N.check(N.lib().pulse_batch_set_camera(self._h, view, C.byref(cam)))
        self.cameras[view] = cam

    def autotune(self, reps: int = 1):
        c = N.BatchConfig()
        ms = C.c_double()
        N.check(N.lib().pulse_batch_autotune(self._h, reps, C.byref(c), C.byref(ms)))
        return c, ms.value

    def stats(self) -> N.Stats:
        s = N.Stats()
        N.check(N.lib().pulse_batch_get_stats(self._h, C.byref(s)))
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

    def present(self, width: int, height: int, fmt: int = N.PRESENT_RGBA16F) -> np.ndarray:
        """The upscaling node's draw of the accumulation buffer into a
        width x height view target (pulse_present), read back to the host."""
        item = {N.PRESENT_RGBA32F: 16, N.PRESENT_RGBA16F: 8}.get(fmt, 4)
        dt = {N.PRESENT_RGBA32F: np.float32, N.PRESENT_RGBA16F: np.float16}.get(fmt, np.uint8)
        dst = self.device.alloc(width * height * item)
        N.check(N.lib().pulse_present(self.device.handle, self.target.ptr, self.width, self.height, dst.ptr, width,
                                      height, fmt))
        self.device.sync()
        return dst.copy_to(np.zeros((height, width, 4), dt))


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


# ---------------------------------------------------------------------------
# Differentiable form
# ---------------------------------------------------------------------------
def material_params(materials: np.ndarray) -> np.ndarray:
    """(n, 7) f64 parameter table {base.rgb, roughness, emissive.rgb}."""
    p = np.zeros((len(materials), N.GRAD_PER_MATERIAL), np.float64)
    p[:, 0:3] = materials["base_color"][:, :3]
    p[:, 3] = materials["perceptual_roughness"]
    p[:, 4:7] = materials["emissive"][:, :3]
    return p


def apply_material_params(materials: np.ndarray, params: np.ndarray) -> np.ndarray:
    m = materials.copy()
    m["base_color"][:, :3] = params[:, 0:3]
    m["perceptual_roughness"] = params[:, 3]
    m["emissive"][:, :3] = params[:, 4:7]
    return m


class PulseAdjoint:
    """Forward/backward pair of the differentiable form: renders an image
    with pulse_trace and returns dL/d(material parameters) of
    L = mean((image - target)^2) from the seed-replay adjoint."""

    def __init__(self, device: Device, width: int, height: int, n_materials: int):
        self.device, self.width, self.height, self.n_materials = device, width, height, n_materials
        self.image = device.alloc(width * height * 16)
        self.target = device.alloc(width * height * 16)
        self.grad = device.alloc(max(n_materials * N.GRAD_PER_MATERIAL * 4, 4))
        self.replay = None
        self.image_in = None

    def set_target(self, target: np.ndarray) -> None:
        self.target.copy_from(np.ascontiguousarray(target, np.float32))

    def set_materials(self, materials: np.ndarray) -> None:
        m = np.ascontiguousarray(materials, N.MATERIAL_DTYPE)
        N.check(N.lib().pulse_scene_update_materials(self.device.handle, m.ctypes.data, len(m)))

    def forward(self, camera: N.Camera, params: N.Params) -> np.ndarray:
        N.check(N.lib().pulse_trace(self.device.handle, C.byref(camera), C.byref(params), self.image.ptr))
        self.device.sync()
        return self.image.copy_to(np.zeros((self.height, self.width, 4), np.float32))

    def loss(self) -> float:
        out = C.c_double()
        N.check(N.lib().pulse_image_loss(self.device.handle, self.image.ptr, self.target.ptr,
                                         self.width * self.height, C.byref(out)))
        return out.value

    def backward(self, camera: N.Camera, params: N.Params, want_replay: bool = False, image: np.ndarray = None):
        """dL/d(material parameters).  dL/dI = 2 (image - target) / (3 N) is
        taken from the last forward image, or from `image` when given (an
        independently sampled image decorrelates dL/dI from the replayed
        paths; any array turns L into a fixed linear functional of I)."""
        if want_replay and self.replay is None:
            self.replay = self.device.alloc(self.width * self.height * 16)
        src = self.image
        if image is not None:
            if self.image_in is None:
                self.image_in = self.device.alloc(self.width * self.height * 16)
            self.image_in.copy_from(np.ascontiguousarray(image, np.float32))
            src = self.image_in
        N.check(N.lib().pulse_adjoint(self.device.handle, C.byref(camera), C.byref(params), src.ptr,
                                      self.target.ptr, self.grad.ptr, self.replay.ptr if want_replay else None))
        self.device.sync()
        g = self.grad.copy_to(np.zeros(self.n_materials * N.GRAD_PER_MATERIAL, np.float32))
        g = g.reshape(self.n_materials, N.GRAD_PER_MATERIAL)
        if want_replay:
            return g, self.replay.copy_to(np.zeros((self.height, self.width, 4), np.float32))
        return g

    def forward_backward(self, camera: N.Camera, params: N.Params):
        loss = C.c_double()
        N.check(N.lib().pulse_forward_adjoint(self.device.handle, C.byref(camera), C.byref(params), self.target.ptr,
                                              self.image.ptr, self.grad.ptr, C.byref(loss)))
        self.device.sync()
        g = self.grad.copy_to(np.zeros(self.n_materials * N.GRAD_PER_MATERIAL, np.float32))
        return loss.value, g.reshape(self.n_materials, N.GRAD_PER_MATERIAL)


class MultiViewAdjoint:
    """The batched differentiable form: several views of the bound scene
    (cameras of any resolution), each rendered by pulse_trace, and the
    gradient of the sum of their losses from one pulse_adjoint_views call
    (one launch per N.ADJOINT_MAX_VIEWS views)."""

    def __init__(self, device: Device, cameras, n_materials: int):
        self.device, self.n_materials = device, n_materials
        self.cameras = list(cameras)
        self.images = [device.alloc(c.width * c.height * 16) for c in self.cameras]
        self.targets = [device.alloc(c.width * c.height * 16) for c in self.cameras]
        self.replays = None
        self.grad = device.alloc(max(n_materials * N.GRAD_PER_MATERIAL * 4, 4))
        self.views = (N.AdjointView * len(self.cameras))()
        for v, (cam, img, tgt) in enumerate(zip(self.cameras, self.images, self.targets)):
            self.views[v].camera = cam
            self.views[v].image = img.ptr
            self.views[v].target = tgt.ptr

    def _shape(self, v: int):
        return (self.cameras[v].height, self.cameras[v].width, 4)

    def set_targets(self, targets) -> None:
        for buf, t in zip(self.targets, targets):
            buf.copy_from(np.ascontiguousarray(t, np.float32))

    def forward(self, params: N.Params):
        """Renders every view; returns the images."""
        lib = N.lib()
        for cam, img in zip(self.cameras, self.images):
            N.check(lib.pulse_trace(self.device.handle, C.byref(cam), C.byref(params), img.ptr))
        self.device.sync()
        return [img.copy_to(np.zeros(self._shape(v), np.float32)) for v, img in enumerate(self.images)]

    def losses(self):
        out, res = C.c_double(), []
        for cam, img, tgt in zip(self.cameras, self.images, self.targets):
            N.check(N.lib().pulse_image_loss(self.device.handle, img.ptr, tgt.ptr, cam.width * cam.height,
                                             C.byref(out)))
            res.append(out.value)
        return res

    def backward(self, params: N.Params, want_replay: bool = False):
        """dL/d(material parameters) of L = the sum of the views' losses."""
        if want_replay and self.replays is None:
            self.replays = [self.device.alloc(c.width * c.height * 16) for c in self.cameras]
        for v in range(len(self.cameras)):
            self.views[v].replay_rgba = self.replays[v].ptr if want_replay else None
        N.check(N.lib().pulse_adjoint_views(self.device.handle, self.views, len(self.cameras), C.byref(params),
                                            self.grad.ptr))
        self.device.sync()
        g = self.grad.copy_to(np.zeros(self.n_materials * N.GRAD_PER_MATERIAL, np.float32))
        g = g.reshape(self.n_materials, N.GRAD_PER_MATERIAL)
        if want_replay:
            return g, [r.copy_to(np.zeros(self._shape(v), np.float32)) for v, r in enumerate(self.replays)]
        return g
