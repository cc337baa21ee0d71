"""The Rust FFI declarations documented for maintainers agree with the C ABI.

Every `#[repr(C)]` struct in the Rust blocks of INTEGRATION.md is laid out by
C rules from its documented field types and compared, field by field (name,
order, offset) and in total size, with

* a probe compiled by gcc from include/pulse_hip.h that prints sizeof and
  offsetof of the C struct, and
* the ctypes mirror in pulse_amd/_native.py that the tests call through.

Every `pub fn` the Rust blocks declare must be declared in include/*.h and
exported by the library.
"""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from pulse_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]

# Rust struct -> (C type, ctypes mirror)
STRUCTS = {
    "PulseBufferView": ("pulse_buffer_view", N.BufferView),
    "PulseSceneUniform": ("pulse_scene_uniform", N.SceneUniform),
    "PulseSceneDesc": ("pulse_scene_desc", N.SceneDesc),
    "PulseCamera": ("pulse_camera", N.Camera),
    "PulseParams": ("pulse_params", N.Params),
    "PulseMaterial": ("pulse_material", N.Material),
    "PulseView": ("pulse_view", N.View),
    "PulseBatchConfig": ("pulse_batch_config", N.BatchConfig),
    "PulseGBuffer": ("pulse_gbuffer", N.GBuffer),
    "PulseAdjointView": ("pulse_adjoint_view", N.AdjointView),
}
SCALARS = {"u32": (4, 4), "i32": (4, 4), "f32": (4, 4), "u64": (8, 8), "f64": (8, 8), "usize": (8, 8)}


def _rust_blocks() -> str:
    text = (ROOT / "INTEGRATION.md").read_text()
    return "\n".join(re.findall(r"```rust\n(.*?)```", text, re.S))


def _rust_structs(src: str) -> dict:
    out = {}
    for name, body in re.findall(r"#\[repr\(C\)\][^\n]*?pub struct (\w+)\s*\{(.*?)\}", src, re.S):
        fields = re.findall(r"pub (\w+):\s*([^,]+?)\s*(?:,|$)", body.strip())
        out[name] = [(f, t.strip()) for f, t in fields]
    return out


def _layout(structs: dict, name: str):
    """(size, align, [(field, offset)]) by C layout rules."""
    off, align, fields = 0, 1, []
    for f, t in structs[name]:
        if t.startswith("*"):
            sz, al = 8, 8
        elif t.startswith("["):
            m = re.match(r"\[(\w+);\s*(\d+)\]", t)
            esz, al = SCALARS[m.group(1)]
            sz = esz * int(m.group(2))
        elif t in SCALARS:
            sz, al = SCALARS[t]
        else:
            sz, al, _ = _layout(structs, t)
        off = (off + al - 1) // al * al
        fields.append((f, off))
        off += sz
        align = max(align, al)
    return (off + align - 1) // align * align, align, fields


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    """{c type: (sizeof, {field: offsetof})} from a probe built against the header."""
    structs = _rust_structs(_rust_blocks())
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pulse_hip.h"', "int main(void) {"]
    for rname, (cname, _) in STRUCTS.items():
        lines.append(f'printf("S {cname} %zu\\n", sizeof({cname}));')
        for f, _ in structs.get(rname, []):
            lines.append(f'printf("F {cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    d = tmp_path_factory.mktemp("probe")
    (d / "probe.c").write_text("\n".join(lines))
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), "-o", str(d / "probe"), str(d / "probe.c")],
                   check=True, capture_output=True)
    out = subprocess.run([str(d / "probe")], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.splitlines():
        p = line.split()
        if p[0] == "S":
            res.setdefault(p[1], [0, {}])[0] = int(p[2])
        else:
            res.setdefault(p[1], [0, {}])[1][p[2]] = int(p[3])
    return res


def test_every_documented_struct_is_known():
    structs = _rust_structs(_rust_blocks())
    assert set(structs) == set(STRUCTS), set(structs) ^ set(STRUCTS)


@pytest.mark.parametrize("rname", sorted(STRUCTS))
def test_documented_struct_matches_header_and_ctypes(c_layout, rname):
    structs = _rust_structs(_rust_blocks())
    cname, mirror = STRUCTS[rname]
    size, _, fields = _layout(structs, rname)
    csize, coffs = c_layout[cname]
    assert size == csize, f"{rname}: Rust layout {size} B, C sizeof {csize} B"
    assert size == C.sizeof(mirror), f"{rname}: ctypes mirror is {C.sizeof(mirror)} B"
    assert [f for f, _ in fields] == [f for f, _ in mirror._fields_], f"{rname}: field order differs from ctypes"
    for f, off in fields:
        assert coffs[f] == off, f"{rname}.{f}: Rust offset {off}, C offset {coffs[f]}"
        assert getattr(mirror, f).offset == off


def test_documented_functions_are_declared_and_exported():
    src = _rust_blocks()
    names = set(re.findall(r"pub fn (\w+)\s*\(", src))
    assert len(names) >= 20
    header = "\n".join(p.read_text() for p in (ROOT / "include").glob("*.h"))
    lib = N.lib()
    for n in sorted(names):
        assert re.search(rf"\b{n}\s*\(", header), f"{n} is not declared in include/*.h"
        assert hasattr(lib, n), f"{n} is not exported by the library"
