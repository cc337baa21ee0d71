"""Per-frame instance updates: pulse_sb_update_instances gives the buffers a
fresh pulse_sb_build with the new transforms gives (the reference rebuilds
instance records and TLAS every frame, src/scene/mod.rs:561-697), and the
device update refuses changes it cannot apply in place."""
import numpy as np
import pytest

from pulse_amd import _native as N
from pulse_amd import procedural as P
from pulse_amd.scene import MaterialDesc, SceneBuilder


def _make(transforms, emissive=(3,), keep=False):
    rng = P.CounterRng(11)
    sb = SceneBuilder()
    meshes = [sb.add_mesh(P.displaced_icosphere(2, 1.0 + k, rng, 50 + k)) for k in range(3)]
    mat = sb.add_material(MaterialDesc(base_color=(0.5, 0.6, 0.7, 1.0), perceptual_roughness=0.4))
    emit = sb.add_material(MaterialDesc(base_color=(0.8, 0.8, 0.8, 1.0), emissive=(6.0, 6.0, 6.0, 1.0)))
    for i, m in enumerate(transforms):
        sb.add_instance(meshes[i % 3], emit if i in emissive else mat, m)
    return sb.build(2, keep_builder=keep)


def _transforms(n, seed):
    u = P.CounterRng(seed).uniform(1, 6 * n).reshape(n, 6)
    out = []
    for i in range(n):
        m = np.eye(4)
        m[:3, :3] = P.random_rotation(*u[i, :3])
        m[:3, 3] = (u[i, 3:] - 0.5) * 30.0
        out.append(m)
    return out


def test_update_equals_fresh_build():
    t0 = _transforms(24, 1)
    a = _make(t0, keep=True)
    t1 = list(t0)
    moved = _transforms(24, 2)
    for i in (0, 3, 5, 17, 23):
        t1[i] = moved[i]
        a.builder.set_instance_transform(i, moved[i])
    b = a.builder.update_instances(a)
    fresh = _make(t1)
    for name in N.BUFFER_NAMES:
        assert b.buffers[name].tobytes() == fresh.buffers[name].tobytes(), name
    assert b.instance_count == fresh.instance_count and b.light_count == fresh.light_count
    # the mesh buffers are shared, not rebuilt
    assert b.buffers["blas_nodes"] is a.buffers["blas_nodes"]
    assert b.buffers["tlas_nodes"].tobytes() != a.buffers["tlas_nodes"].tobytes()


def test_builder_rejects_bad_instance_id():
    a = _make(_transforms(4, 3), keep=True)
    with pytest.raises(N.PulseError, match="INVALID_ARG"):
        a.builder.set_instance_transform(4, np.eye(4))
