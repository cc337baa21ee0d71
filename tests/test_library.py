"""libpulse_hip.so loads, exports every function declared in include/*.h,
and reports errors through status codes (no GPU needed)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from pulse_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(pulse_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 25
    lib = N.lib()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    bound = {s[0] for s in N.SIGNATURES}
    assert names <= bound, sorted(names - bound)


def test_version_and_error_reporting_without_gpu():
    lib = N.lib()
    assert b"gfx950" in lib.pulse_version()
    rc = lib.pulse_build_blas(None, 0, None, 0, None, None)
    assert rc == -1 and b"null" in lib.pulse_last_error()
    with pytest.raises(N.PulseError):
        N.check(lib.pulse_sb_get_desc(None, None))


def test_builder_rejects_bad_input():
    from pulse_amd.scene import Mesh, SceneBuilder
    import numpy as np

    sb = SceneBuilder()
    with pytest.raises(N.PulseError, match="INVALID_ARG"):
        sb.add_mesh(Mesh(np.zeros((3, 3), np.float32), np.zeros((3, 3), np.float32), np.array([0, 1, 5], np.uint32)))
    with pytest.raises(N.PulseError):
        sb.add_instance(0, 0, np.eye(4))
