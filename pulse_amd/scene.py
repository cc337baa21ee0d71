"""Scene side of the drop-in: meshes/materials/instances -> the 13 buffers.

Mirrors what the reference's scene plugin produces each frame
(src/scene/mod.rs: prepare_extracted_mesh_assets 372-448, prepare_mesh_data
471-510, prepare_mesh_instances 561-697, queue_scene_bind_group 911-1090).
The heavy lifting (BLAS/TLAS builders, light CDFs) runs in the C++ builder of
libpulse_hip (include/pulse_scene.h); this module owns the numpy arrays and
the camera helpers.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N


@dataclass
class Mesh:
    positions: np.ndarray  # (V, 3) f32
    normals: np.ndarray    # (V, 3) f32
    indices: np.ndarray    # (T*3,) u32
    uvs: np.ndarray | None = None  # (V, 2) f32

    @property
    def triangle_count(self) -> int:
        return len(self.indices) // 3


@dataclass
class MaterialDesc:
    base_color: tuple = (0.8, 0.8, 0.8, 1.0)
    emissive: tuple = (0.0, 0.0, 0.0, 1.0)
    perceptual_roughness: float = 0.5
    reflectance: float = 0.5
    metallic: float = 0.0

    def to_c(self) -> N.Material:
        m = N.Material()
        for i in range(4):
            m.base_color[i] = self.base_color[i]
            m.emissive[i] = self.emissive[i]
        m.perceptual_roughness = self.perceptual_roughness
        m.reflectance = self.reflectance
        m.metallic = self.metallic
        return m


@dataclass
class SceneBuffers:
    """The 13 bindings of bind group 0 as numpy arrays (host memory)."""
    instance_count: int
    light_count: int
    buffers: dict
    info: dict = field(default_factory=dict)
    builder: object = None  # the SceneBuilder that made them, when kept for instance updates

    def desc(self) -> N.SceneDesc:
        d = N.SceneDesc()
        d.uniform.instance_count = self.instance_count
        d.uniform.light_count = self.light_count
        for name in N.BUFFER_NAMES:
            a = self.buffers[name]
            v = getattr(d, name)
            v.data = a.ctypes.data if a.size else None
            v.bytes = a.nbytes
        return d


class SceneBuilder:
    """Collects meshes, materials and instances and runs the host builder."""

    def __init__(self):
        self._h = C.c_void_p()
        N.check(N.lib().pulse_sb_create(C.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            N.lib().pulse_sb_destroy(self._h)
            self._h = C.c_void_p()

    def add_mesh(self, mesh: Mesh) -> int:
        pos = np.ascontiguousarray(mesh.positions, dtype=np.float32)
        nrm = np.ascontiguousarray(mesh.normals, dtype=np.float32)
        idx = np.ascontiguousarray(mesh.indices, dtype=np.uint32)
        uvs = None if mesh.uvs is None else np.ascontiguousarray(mesh.uvs, dtype=np.float32)
        out = C.c_uint32()
        N.check(N.lib().pulse_sb_add_mesh(self._h, pos.ctypes.data, nrm.ctypes.data,
                                          None if uvs is None else uvs.ctypes.data, len(pos), idx.ctypes.data,
                                          len(idx), C.byref(out)))
        return out.value

    def add_material(self, m: MaterialDesc) -> int:
        out = C.c_uint32()
        cm = m.to_c()
        N.check(N.lib().pulse_sb_add_material(self._h, C.byref(cm), C.byref(out)))
        return out.value

    def add_instance(self, mesh_id: int, material_id: int, transform: np.ndarray) -> None:
        """transform: 4x4 object->world matrix (row-major numpy, math convention)."""
        cm = np.ascontiguousarray(np.asarray(transform, dtype=np.float32).T.reshape(16))  # column-major
        N.check(N.lib().pulse_sb_add_instance(self._h, mesh_id, material_id, cm.ctypes.data))

    def build(self, threads: int = 8, keep_builder: bool = False) -> SceneBuffers:
        N.check(N.lib().pulse_sb_build(self._h, threads))
        return self._read(N.BUFFER_NAMES, None, keep_builder)

    def set_instance_transform(self, instance_id: int, transform: np.ndarray) -> None:
        """New object->world matrix (row-major numpy) for an instance."""
        cm = np.ascontiguousarray(np.asarray(transform, dtype=np.float32).T.reshape(16))
        N.check(N.lib().pulse_sb_set_instance_transform(self._h, instance_id, cm.ctypes.data))

    def update_instances(self, prev: SceneBuffers) -> SceneBuffers:
        """Instance records, light tables and TLAS for the current transforms
        (pulse_sb_update_instances); the mesh buffers are shared with `prev`."""
        N.check(N.lib().pulse_sb_update_instances(self._h))
        return self._read(INSTANCE_BUFFERS, prev, prev.builder is not None)

    def _read(self, names, prev, keep_builder) -> SceneBuffers:
        d = N.SceneDesc()
        N.check(N.lib().pulse_sb_get_desc(self._h, C.byref(d)))
        info = N.BuildInfo()
        N.check(N.lib().pulse_sb_get_info(self._h, C.byref(info)))
        bufs = dict(prev.buffers) if prev is not None else {}
        for name in names:
            v = getattr(d, name)
            dt = N.BUFFER_DTYPES[name]
            n = v.bytes // dt.itemsize
            if n:
                raw = (C.c_char * v.bytes).from_address(v.data)
                bufs[name] = np.frombuffer(bytes(raw), dtype=dt).copy()
            else:
                bufs[name] = np.zeros(0, dtype=dt)
        return SceneBuffers(d.uniform.instance_count, d.uniform.light_count, bufs,
                            {f: getattr(info, f) for f, _ in N.BuildInfo._fields_},
                            self if keep_builder else None)


# The buffers an instance update rewrites (plus the light tables, which it
# recomputes identically when only transforms change).
INSTANCE_BUFFERS = ("instances", "tlas_nodes", "instance_indices", "light_emission_strength_cdf",
                    "light_triangle_area_cdfs", "light_mesh_areas", "light_indices")


# ---------------------------------------------------------------------------
# Camera (Bevy View.inverse_view_proj / world_position)
# ---------------------------------------------------------------------------
def look_at(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """Camera-to-world transform (right-handed, camera looks down -z)."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = r, u, -f, eye
    return m


def perspective_infinite_reverse_rh(fov_y: float, aspect: float, near: float) -> np.ndarray:
    """Bevy's default 3D projection (glam Mat4::perspective_infinite_reverse_rh)."""
    f = np.cos(0.5 * fov_y) / np.sin(0.5 * fov_y)
    p = np.zeros((4, 4))
    p[0, 0] = f / aspect
    p[1, 1] = f
    p[3, 2] = -1.0
    p[2, 3] = near
    return p


def make_camera(width: int, height: int, eye, target, fov_y: float = np.pi / 4, near: float = 0.1) -> N.Camera:
    cam_to_world = look_at(eye, target)
    view_proj = perspective_infinite_reverse_rh(fov_y, width / height, near) @ np.linalg.inv(cam_to_world)
    ivp = np.linalg.inv(view_proj).astype(np.float32)
    c = N.Camera()
    cm = ivp.T.reshape(16)  # column-major
    for i in range(16):
        c.inverse_view_proj[i] = float(cm[i])
    for i in range(3):
        c.world_position[i] = float(np.float32(eye[i]))
    c.width = width
    c.height = height
    return c
