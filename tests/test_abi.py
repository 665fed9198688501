"""The C ABI: the built library exports every entry point the header declares, the ctypes
mirrors have the header's layout, and without a GPU the product path fails loudly."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from crashcat_amd import _abi
from crashcat_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "crashcat_amd.h")


def _declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cc_[a-z_0-9]+)\s*\(", src)))


def _built():
    return os.path.exists(_lib.LIB_PATH)


@pytest.mark.skipif(not _built(), reason="HIP library not built")
def test_library_exports_every_declared_symbol():
    names = _declared_functions()
    assert len(names) >= 14
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert set(_lib.SYMBOLS) == set(names)


STRUCTS = {
    "cc_settings": _abi.Settings, "cc_shape_desc": _abi.ShapeDesc, "cc_hull_data": _abi.HullData,
    "cc_body_desc": _abi.BodyDesc, "cc_body_state": _abi.BodyState,
    "cc_point_constraint_desc": _abi.PointConstraintDesc, "cc_contact_out": _abi.ContactOut,
    "cc_step_counters": _abi.StepCounters, "cc_totals": _abi.Totals, "cc_config": _abi.Config,
}


def test_ctypes_layout_matches_header(tmp_path):
    """Compile a probe against the header with the host C compiler and compare every struct's
    size and every field's offset with the ctypes mirrors."""
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            a, b, c = ln.split()
            got[(a, b)] = int(c)
    for cname, cls in STRUCTS.items():
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


def test_constants_match_header():
    src = open(HEADER).read()

    def val(name):
        m = re.search(r"#define\s+" + name + r"\s+\(?(-?\d+)", src) or re.search(name + r"\s*=\s*(-?\d+)", src)
        assert m, name
        return int(m.group(1))

    for name in ["CC_SHAPE_SPHERE", "CC_SHAPE_BOX", "CC_SHAPE_CAPSULE", "CC_SHAPE_CONVEX_HULL", "CC_MOTION_STATIC",
                 "CC_MOTION_KINEMATIC", "CC_MOTION_DYNAMIC", "CC_PRECISION_FP32", "CC_PRECISION_FP64",
                 "CC_MAX_OBJECT_LAYERS"]:
        assert val(name) == getattr(_abi, name), name


@pytest.mark.skipif(not _built() or os.path.exists("/dev/kfd"), reason="needs the built library and no GPU")
def test_no_gpu_fails_loudly():
    """Without a device the context cannot be created: the product path raises rather than
    falling back to anything on the CPU."""
    with pytest.raises(_lib.CrashcatError):
        _lib.Context("fp32")


def test_missing_library_raises(monkeypatch):
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libcrashcat_amd.so")
    with pytest.raises(_lib.CrashcatError):
        _lib.lib()
