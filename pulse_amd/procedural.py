"""Procedural scenes for the benchmark workloads (synthetic input data).

* ``cornell``: a Cornell box -- 5 axis-aligned walls (10 triangles, one mesh),
  two boxes rotated about y (one 12-triangle box mesh, two instances) and a
  ceiling quad emitter (2 triangles): 36 triangles, 4 instances, 3 materials
  with albedo U[0.2, 0.8]^3 and roughness U[0.1, 1.0], emitter radiance 15.
* ``icosphere_field``: N seeded icospheres (subdivision s: 20*4^s triangles)
  with per-vertex displacement N(0, 0.05 r) along the normal, instanced with
  uniformly random rotations and translations in a cube, a subset of the
  instances emissive.

Random numbers come from a counter-based PCG hash (the reference's rand_u
mixing function applied to (seed, stream, index) counters), so every draw is
reproducible and vectorizable.
"""
from __future__ import annotations

import numpy as np

from .scene import MaterialDesc, Mesh, SceneBuilder, make_camera


def _pcg_hash(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64)
    state = (x * np.uint64(747796405) + np.uint64(2891336453)) & np.uint64(0xFFFFFFFF)
    word = (((state >> ((state >> np.uint64(28)) + np.uint64(4))) ^ state) * np.uint64(277803737)) & np.uint64(0xFFFFFFFF)
    return ((word >> np.uint64(22)) ^ word).astype(np.uint32)


class CounterRng:
    """uniform(stream, n): n floats in [0, 1) from hash(seed, stream, i)."""

    def __init__(self, seed: int):
        self.seed = np.uint32(seed & 0xFFFFFFFF)

    def bits(self, stream: int, n: int) -> np.ndarray:
        i = np.arange(n, dtype=np.uint64)
        s = _pcg_hash(np.full(1, (int(self.seed) * 0x9E3779B9 + stream * 0x85EBCA6B) & 0xFFFFFFFF, np.uint64))[0]
        return _pcg_hash(_pcg_hash(i ^ np.uint64(s)).astype(np.uint64) + i * np.uint64(0x27D4EB2F))

    def uniform(self, stream: int, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.bits(stream, n) >> np.uint32(8)).astype(np.float64) * (1.0 / (1 << 24))
        return lo + (hi - lo) * u

    def normal(self, stream: int, n: int) -> np.ndarray:
        u1 = self.uniform(2 * stream + 1000, n)
        u2 = self.uniform(2 * stream + 1001, n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
def _quad(p0, p1, p2, p3, normal):
    pos = np.array([p0, p1, p2, p3], np.float32)
    nrm = np.tile(np.asarray(normal, np.float32), (4, 1))
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    return pos, nrm, idx


def _merge(parts):
    pos, nrm, idx, base = [], [], [], 0
    for p, n, i in parts:
        pos.append(p)
        nrm.append(n)
        idx.append(i + base)
        base += len(p)
    return Mesh(np.concatenate(pos), np.concatenate(nrm), np.concatenate(idx).astype(np.uint32))


def box_mesh() -> Mesh:
    """Unit cube [-0.5, 0.5]^3, outward flat normals, 12 triangles."""
    h = 0.5
    faces = []
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            n = np.zeros(3)
            n[axis] = sgn
            u = np.zeros(3)
            u[(axis + 1) % 3] = 1.0
            v = np.cross(n, u)
            c = n * h
            pts = [c - u * h - v * h, c + u * h - v * h, c + u * h + v * h, c - u * h + v * h]
            faces.append(_quad(*pts, n))
    return _merge(faces)


def cornell_walls_mesh() -> Mesh:
    """Floor, ceiling, back, left and right walls of [-1, 1]^3, inward normals."""
    a = 1.0
    return _merge([
        _quad((-a, -a, a), (a, -a, a), (a, -a, -a), (-a, -a, -a), (0, 1, 0)),    # floor
        _quad((-a, a, -a), (a, a, -a), (a, a, a), (-a, a, a), (0, -1, 0)),      # ceiling
        _quad((-a, -a, -a), (a, -a, -a), (a, a, -a), (-a, a, -a), (0, 0, 1)),   # back
        _quad((-a, -a, a), (-a, -a, -a), (-a, a, -a), (-a, a, a), (1, 0, 0)),   # left
        _quad((a, -a, -a), (a, -a, a), (a, a, a), (a, a, -a), (-1, 0, 0)),      # right
    ])


def quad_mesh(size: float = 0.5) -> Mesh:
    """Horizontal quad at y = 0 facing down (-y), side `size`."""
    h = size / 2
    p, n, i = _quad((-h, 0, -h), (h, 0, -h), (h, 0, h), (-h, 0, h), (0, -1, 0))
    return Mesh(p, n, i)


_ICO_CACHE: dict = {}


def unit_icosphere(subdiv: int):
    """(vertices (V,3) on the unit sphere, faces (F,3)) with F = 20 * 4^subdiv."""
    if subdiv in _ICO_CACHE:
        return _ICO_CACHE[subdiv]
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                  (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                  (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], np.int64)
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
        m01, m12, m20 = inv[:nf] + base, inv[nf:2 * nf] + base, inv[2 * nf:] + base
        f = np.concatenate([np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
                            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    _ICO_CACHE[subdiv] = (v, f)
    return v, f


def displaced_icosphere(subdiv: int, radius: float, rng: CounterRng, stream: int) -> Mesh:
    v, f = unit_icosphere(subdiv)
    disp = rng.normal(stream, len(v)) * (0.05 * radius)
    pos = v * (radius + disp)[:, None]
    fn = np.cross(pos[f[:, 1]] - pos[f[:, 0]], pos[f[:, 2]] - pos[f[:, 0]])  # area-weighted
    vn = np.zeros_like(pos)
    for k in range(3):
        np.add.at(vn, f[:, k], fn)
    vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-30)
    uv = np.stack([0.5 + np.arctan2(v[:, 2], v[:, 0]) / (2 * np.pi), np.arccos(np.clip(v[:, 1], -1, 1)) / np.pi], 1)
    return Mesh(pos.astype(np.float32), vn.astype(np.float32), f.reshape(-1).astype(np.uint32), uv.astype(np.float32))


# ---------------------------------------------------------------------------
# transforms
# ---------------------------------------------------------------------------
def trs(translation=(0, 0, 0), rotation_y: float = 0.0, scale=(1, 1, 1)) -> np.ndarray:
    c, s = np.cos(rotation_y), np.sin(rotation_y)
    r = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    m = np.eye(4)
    m[:3, :3] = r @ np.diag(scale)
    m[:3, 3] = translation
    return m


def random_rotation(u1: float, u2: float, u3: float) -> np.ndarray:
    """Uniform random rotation from three uniforms (Shoemake 1992)."""
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    x, y, z, w = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2), b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


# ---------------------------------------------------------------------------
# scenes
# ---------------------------------------------------------------------------
def cornell(seed: int = 0, threads: int = 8):
    """Cornell box (walls, two boxes, ceiling light); returns (SceneBuffers,
    camera factory)."""
    rng = CounterRng(seed)
    u = rng.uniform(1, 16)
    sb = SceneBuilder()
    walls = sb.add_mesh(cornell_walls_mesh())
    box = sb.add_mesh(box_mesh())
    quad = sb.add_mesh(quad_mesh(0.5))
    m_walls = sb.add_material(MaterialDesc(base_color=(*(0.2 + 0.6 * u[0:3]), 1.0),
                                           perceptual_roughness=float(0.1 + 0.9 * u[3])))
    m_boxes = sb.add_material(MaterialDesc(base_color=(*(0.2 + 0.6 * u[4:7]), 1.0),
                                           perceptual_roughness=float(0.1 + 0.9 * u[7])))
    m_light = sb.add_material(MaterialDesc(base_color=(*(0.2 + 0.6 * u[8:11]), 1.0), emissive=(15.0, 15.0, 15.0, 1.0),
                                           perceptual_roughness=float(0.1 + 0.9 * u[11])))
    sb.add_instance(walls, m_walls, np.eye(4))
    sb.add_instance(box, m_boxes, trs((-0.35, -0.4, -0.3), 0.3, (0.6, 1.2, 0.6)))
    sb.add_instance(box, m_boxes, trs((0.4, -0.7, 0.3), -0.3, (0.6, 0.6, 0.6)))
    sb.add_instance(quad, m_light, trs((0.0, 0.99, 0.0)))
    scene = sb.build(threads)

    def camera(width=256, height=256):
        return make_camera(width, height, eye=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0))

    return scene, camera


def cornell_lights(strengths=(2.0, 40.0, 8.0), seed: int = 0, threads: int = 8):
    """The Cornell box with several ceiling quad emitters of different
    radiance (one emissive material each), for light selection by emission
    strength; returns (SceneBuffers, camera factory)."""
    rng = CounterRng(seed)
    u = rng.uniform(1, 8)
    sb = SceneBuilder()
    walls = sb.add_mesh(cornell_walls_mesh())
    box = sb.add_mesh(box_mesh())
    quad = sb.add_mesh(quad_mesh(0.3))
    m_walls = sb.add_material(MaterialDesc(base_color=(*(0.2 + 0.6 * u[0:3]), 1.0),
                                           perceptual_roughness=float(0.1 + 0.9 * u[3])))
    m_boxes = sb.add_material(MaterialDesc(base_color=(*(0.2 + 0.6 * u[4:7]), 1.0),
                                           perceptual_roughness=float(0.1 + 0.9 * u[7])))
    sb.add_instance(walls, m_walls, np.eye(4))
    sb.add_instance(box, m_boxes, trs((-0.35, -0.4, -0.3), 0.3, (0.6, 1.2, 0.6)))
    n = len(strengths)
    for k, e in enumerate(strengths):
        m = sb.add_material(MaterialDesc(base_color=(0.8, 0.8, 0.8, 1.0), emissive=(e, e, e, 1.0),
                                         perceptual_roughness=0.5))
        x = -0.6 + 1.2 * (k + 0.5) / n
        sb.add_instance(quad, m, trs((x, 0.99, 0.1 * k)))
    scene = sb.build(threads)

    def camera(width=256, height=256):
        return make_camera(width, height, eye=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0))

    return scene, camera


def icosphere_field(n_meshes: int = 64, subdiv: int = 6, n_instances: int = 800, n_emissive: int = 32,
                    box: float = 100.0, radius=(1.5, 4.0), roughness=(0.05, 1.0), emission: float = 8.0,
                    seed: int = 2, threads: int = 8, n_materials: int | None = None, keep_builder: bool = False,
                    per_instance_materials: bool = False):
    """Field of instanced displaced icospheres in a `box`-sized cube; returns
    (SceneBuffers, camera factory).

    Mesh k gets material k % n_materials (default: one material per mesh);
    instance i uses mesh i % n_meshes; `n_emissive` instances (chosen by the
    seeded hash) use one extra emissive material.  per_instance_materials:
    instance i gets material i instead (n_instances materials: per-instance
    albedo / roughness parameters for the adjoint).
    """
    rng = CounterRng(seed)
    if per_instance_materials:
        n_materials = n_instances
    n_materials = n_meshes if n_materials is None else n_materials
    sb = SceneBuilder()
    radii = rng.uniform(2, n_meshes, *radius)
    meshes = [sb.add_mesh(displaced_icosphere(subdiv, float(radii[k]), rng, 100 + k)) for k in range(n_meshes)]
    alb = rng.uniform(3, 3 * n_materials, 0.2, 0.8).reshape(n_materials, 3)
    rough = rng.uniform(4, n_materials, *roughness)
    mats = [sb.add_material(MaterialDesc(base_color=(*alb[k], 1.0), perceptual_roughness=float(rough[k])))
            for k in range(n_materials)]
    m_emit = sb.add_material(MaterialDesc(base_color=(0.8, 0.8, 0.8, 1.0), emissive=(emission, emission, emission, 1.0),
                                          perceptual_roughness=0.5))
    order = np.argsort(rng.bits(5, n_instances), kind="stable")
    emissive = set(order[:n_emissive].tolist())
    rot = rng.uniform(6, 3 * n_instances).reshape(n_instances, 3)
    tr = rng.uniform(7, 3 * n_instances, -box / 2, box / 2).reshape(n_instances, 3)
    for i in range(n_instances):
        m = np.eye(4)
        m[:3, :3] = random_rotation(*rot[i])
        m[:3, 3] = tr[i]
        mesh = meshes[i % n_meshes]
        mat = mats[i] if per_instance_materials else mats[(i % n_meshes) % n_materials]
        sb.add_instance(mesh, m_emit if i in emissive else mat, m)
    scene = sb.build(threads, keep_builder=keep_builder)

    def camera(width=1920, height=1080):
        return make_camera(width, height, eye=(0.0, 0.0, 1.3 * box), target=(0.0, 0.0, 0.0))

    return scene, camera


def batch_workload(n_views: int = 32, seed: int = 3, tri_range=(10_000, 2_000_000),
                   resolutions=((512, 512), (1280, 720), (1920, 1080)), threads: int = 8):
    """Views of the batched form: per view a resolution drawn from
    `resolutions` and an icosphere field whose unique triangle count is drawn
    log-uniformly from `tri_range` (subdivision and mesh count chosen to
    match; 12.5 instances per mesh, 4 % emissive instances, box scaled for
    the instance density of the 800-instance field).  Returns a list of
    (SceneBuffers, Camera, info dict)."""
    rng = CounterRng(seed)
    lo, hi = np.log(tri_range[0]), np.log(tri_range[1])
    u_tris = rng.uniform(1, n_views, lo, hi)
    u_res = rng.bits(2, n_views)
    views = []
    for v in range(n_views):
        target = float(np.exp(u_tris[v]))
        subdiv = 2
        for s in range(2, 7):
            if 20 * 4 ** s * 4 <= target:
                subdiv = s
        n_meshes = int(min(64, max(1, round(target / (20 * 4 ** subdiv)))))
        n_inst = max(1, int(round(12.5 * n_meshes)))
        n_emit = max(1, n_inst // 25)
        box = 100.0 * (n_inst / 800.0) ** (1.0 / 3.0)
        scene, camf = icosphere_field(n_meshes=n_meshes, subdiv=subdiv, n_instances=n_inst, n_emissive=n_emit,
                                      box=box, seed=100 + v, threads=threads)
        w, h = resolutions[int(u_res[v]) % len(resolutions)]
        views.append((scene, camf(w, h), {"unique_triangles": n_meshes * 20 * 4 ** subdiv, "subdiv": subdiv,
                                          "meshes": n_meshes, "instances": n_inst, "width": w, "height": h}))
    return views
