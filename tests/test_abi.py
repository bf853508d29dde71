"""The C ABI library loads without a GPU and exports every function that
include/*.h declares; host-only helpers work without a device."""
import ctypes
import glob
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\**\s+\**(ando_[a-z_0-9]+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_functions():
    names = declared_functions()
    assert "ando_step" in names and "ando_pcg" in names and len(names) >= 18


def test_library_exports_every_declared_symbol():
    from andosim_amd import native

    lib = ctypes.CDLL(native.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(native.ABI_SYMBOLS) == set(declared_functions())


def test_host_helper_lumped_masses_matches_closed_form():
    """Sum of lumped masses = rho * t * total area (state.cpp:19-33)."""
    from andosim_amd import native, scenes

    sc = scenes.drop21()
    m = native.lumped_masses(sc.verts, sc.tris, sc.material)
    total = sc.material["density"] * sc.material["thickness"] * 1.0  # 1 m x 1 m sheet
    assert abs(float(np.sum(m, dtype=np.float64)) - total) < 1e-6 * total * 10


def test_errors_are_reported_not_raised():
    from andosim_amd import native

    L = native.load()
    rc = L.ando_step(None, None, None, 0)
    assert rc != 0
    assert len(L.ando_last_error()) > 0


def _header_struct_fields(name):
    text = open(os.path.join(ROOT, "include", "andosim_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    m = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + r"\s*;", text)
    assert m, name
    return [f.group(1) for f in re.finditer(r"([a-z_0-9]+)\s*(?:\[[A-Z_0-9]+\])?\s*;", m.group(1))]


@pytest.mark.parametrize("cname,pyname", [("ando_stats", "Stats"), ("ando_params", "Params"),
                                          ("ando_material", "Material"), ("ando_contact", "Contact"),
                                          ("ando_rigid_body", "RigidBody"), ("ando_energy", "Energy"),
                                          ("ando_collision_metrics", "CollisionMetrics")])
def test_ctypes_structs_follow_the_header(cname, pyname):
    """The ctypes mirrors in native.py list the header's struct fields in the
    header's order, so the layouts agree (a field added to the header but not
    to the mirror would shift every later one)."""
    native = pytest.importorskip("andosim_amd.native")
    cls = getattr(native, pyname, None)
    if cls is None:
        pytest.skip(f"native.py has no {pyname} mirror")
    assert [f[0] for f in cls._fields_] == _header_struct_fields(cname)


STRUCTS = [("ando_stats", "Stats"), ("ando_params", "Params"), ("ando_material", "Material"),
           ("ando_contact", "Contact"), ("ando_rigid_body", "RigidBody"), ("ando_energy", "Energy"),
           ("ando_collision_metrics", "CollisionMetrics")]


def _c_layout():
    """sizeof and (offsetof, sizeof) of every field of the header's structs,
    from a probe compiled by the host C compiler."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "andosim_hip.h"', "int main(void) {"]
    for cname, _ in STRUCTS:
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for f in _header_struct_fields(cname):
            lines.append(f'  printf("{cname}.{f} %zu %zu\\n", offsetof({cname}, {f}), sizeof((({cname}*)0)->{f}));')
    lines += ["  return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "probe.c"), os.path.join(d, "probe")
        open(src, "w").write("\n".join(lines) + "\n")
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    layout = {}
    for ln in out.splitlines():
        key, *vals = ln.split()
        layout[key] = tuple(int(v) for v in vals)
    return layout


@pytest.mark.parametrize("cname,pyname", STRUCTS)
def test_ctypes_struct_offsets_and_sizes_match_the_c_compiler(cname, pyname):
    """Names alone would miss a type drift (a float for a double, int32 for
    int64): every field's offset and size and the struct's size must equal
    what the C compiler lays out for the header."""
    native = pytest.importorskip("andosim_amd.native")
    cls = getattr(native, pyname)
    layout = _c_layout()
    assert ctypes.sizeof(cls) == layout[cname][0], (cname, ctypes.sizeof(cls), layout[cname])
    for f, _ in cls._fields_:
        d = getattr(cls, f)
        assert (d.offset, d.size) == layout[f"{cname}.{f}"], (cname, f)
