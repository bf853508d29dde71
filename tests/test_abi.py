"""The C ABI library loads without a GPU and exports every function that
include/*.h declares; host-only helpers work without a device."""
import ctypes
import glob
import os
import re

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
