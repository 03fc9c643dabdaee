"""CPU tests of the C ABI: libsimpeg_dc.so loads, exports every function include/simpeg_dc.h
declares, and rejects bad arguments before touching the GPU."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from simpeg_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "simpeg_dc.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sdc_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_are_exported():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    # the Python binding types every declared function
    assert set(names) == set(_lib.EXPORTED)


def test_version_and_default_params():
    lib = _lib.load()
    assert b"gfx950" in lib.sdc_version()
    p = _lib.Params()
    assert lib.sdc_default_params(C.byref(p)) == _lib.SDC_OK
    assert p.nu_pre >= 1 and p.nu_post >= 1 and p.cheb_lmax == 2.0
    assert 0 < p.reliable_delta < 1


def _desc(nx=2, ny=2, nz=2, bc=_lib.SDC_BC_NEUMANN):
    h = [np.ones(max(n, 1)) for n in (nx, ny, nz)]
    d = _lib.MeshDesc()
    d.nx, d.ny, d.nz = nx, ny, nz
    d.hx, d.hy, d.hz = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in h]
    d.bc_type = bc
    return d, h


@pytest.mark.parametrize("case", ["zero_cells", "bad_bc", "robin_without_beta", "negative_width"])
def test_create_rejects_bad_arguments(case):
    lib = _lib.load()
    if case == "zero_cells":
        d, keep = _desc(nx=0)
    elif case == "bad_bc":
        d, keep = _desc(bc=7)
    elif case == "robin_without_beta":
        d, keep = _desc(bc=_lib.SDC_BC_ROBIN)
    else:
        d, keep = _desc()
        keep[0][0] = -1.0
    h = C.c_void_p()
    rc = lib.sdc_create(C.byref(d), C.byref(h))
    assert rc == _lib.SDC_ERR_ARG
    assert not h.value
    assert len(lib.sdc_last_error()) > 0
    with pytest.raises(ValueError):
        _lib.check(rc, "sdc_create")


def test_null_handle_calls():
    lib = _lib.load()
    assert lib.sdc_destroy(None) == _lib.SDC_OK
    assert lib.sdc_ld(None) == -1
    assert lib.sdc_set_sigma(None, None, None) == _lib.SDC_ERR_ARG
    assert lib.sdc_set_coefficients(None, None, None, None, None, None) == _lib.SDC_ERR_ARG
    assert lib.sdc_solve(None, None, None, 1, None, None, None) == _lib.SDC_ERR_ARG


def _header_struct_fields(name):
    """(field, C type) of `typedef struct { ... } name;` in declaration order."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct \{([^}]*)\}\s*" + name + r"\s*;", text).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"((?:const\s+)?[A-Za-z_0-9]+\s*\**)\s*(.*)$", decl)
        ctype = m.group(1).replace(" ", "")
        for var in m.group(2).split(","):
            var = var.strip()
            stars = len(var) - len(var.lstrip("*"))
            fields.append((var.lstrip("*").strip(), ctype + "*" * stars))
    return fields


_CTYPE = {"int32_t": C.c_int32, "double": C.c_double, "uint32_t": C.c_uint32}


@pytest.mark.parametrize("cname,pyname", [("sdc_mesh_desc", "MeshDesc"), ("sdc_params", "Params"),
                                          ("sdc_solve_opts", "SolveOpts"),
                                          ("sdc_solve_stats", "SolveStats")])
def test_ctypes_structs_mirror_the_header(cname, pyname):
    """Every struct the C ABI passes has the same fields, in the same order and of the same
    scalar types, in its ctypes mirror (a field added on one side only shifts the others)."""
    hdr = _header_struct_fields(cname)
    py = getattr(_lib, pyname)._fields_
    assert [f for f, _ in hdr] == [f for f, _ in py]
    for (f, ct), (_, pt) in zip(hdr, py):
        if ct in _CTYPE:
            assert pt is _CTYPE[ct], (f, ct, pt)
        else:
            assert ct.endswith("*"), (f, ct)
            assert issubclass(pt, (C._Pointer, C.c_void_p)) or pt is C.c_void_p, (f, ct, pt)
