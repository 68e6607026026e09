"""The C-ABI library loads and exports every entry point of include/pts.h."""
import ctypes
import os
import re

import pytest

import ptstokes
from ptstokes import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pts.h")


def header_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|double|const char\s*\*)\s*(pts_\w+)\s*\(", txt, re.M)))


def test_header_declares_the_api():
    syms = header_symbols()
    for s in ("pts_create", "pts_destroy", "pts_compute_P", "pts_compute_tau", "pts_compute_dV",
              "pts_compute_V", "pts_compute_Pt_tau", "pts_bc_x", "pts_bc_y", "pts_bc_z", "pts_pt_solve"):
        assert s in syms


def test_library_exports_every_header_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_header_symbol():
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert bound == set(header_symbols())


def test_field_names_match_header_enum():
    txt = open(HEADER).read()
    body = txt[txt.index("typedef enum pts_field"):txt.index("} pts_field;")]
    names = re.findall(r"PTS_([A-Z0-9_]+)", body)
    names = [n for n in names if n != "NFIELDS"]
    assert len(names) == len(_lib.FIELDS)


def test_package_loads_native_library():
    lib = ptstokes.load()
    assert lib.pts_last_error(None) is not None


def test_missing_library_raises():
    with pytest.raises(RuntimeError, match="not built"):
        _lib.load(os.path.join(ROOT, "no_such_dir", "libpts.so"))


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present: creation succeeds")
def test_no_gpu_means_error_not_fallback():
    # Without a device the library reports a HIP error; nothing computes on the host.
    with pytest.raises(ptstokes.PtsError) as e:
        ptstokes.Context(2, 31, 31)
    assert e.value.code == _lib.PTS_ERR_HIP


def test_julia_binding_ccalls_every_header_symbol():
    # Julia is absent here, so this checks the binding's text: each entry point is ccall'ed.
    jl = open(os.path.join(ROOT, "pseudotransientstokes.jl_amd", "julia", "PTStokesAMD.jl")).read()
    called = set(re.findall(r"ccall\(\(:(pts_\w+),", jl))
    assert called == set(header_symbols())


def _zero_args(argtypes):
    vals = []
    for a in argtypes:
        if a is ctypes.c_double:
            vals.append(0.0)
        elif a in (ctypes.c_int, ctypes.c_int32, ctypes.c_int64):
            vals.append(0)
        else:
            vals.append(None)
    return vals


def test_every_entry_point_rejects_a_null_context():
    # error codes, never a crash: checked before any device call
    lib = ptstokes.load()
    for name, _, args in _lib.SIGNATURES:
        if name in ("pts_last_error", "pts_err_from_norms"):
            continue
        rc = getattr(lib, name)(*_zero_args(args))
        want = _lib.PTS_OK if name == "pts_destroy" else _lib.PTS_ERR_ARG
        assert rc == want, (name, rc)


@pytest.mark.parametrize("grid,msg", [
    ((4, 31, 31, 1), "ndim"),
    ((2, 2, 31, 1), ">= 3"),
    ((3, 31, 31, 2), ">= 3"),
    ((2, 65001, 31, 1), "too large"),
    ((2, 65000, 65000, 1), "32-bit"),
    ((3, 1400, 1400, 1400), "32-bit"),
])
def test_create_rejects_bad_grids(grid, msg):
    # argument checks come before device initialisation, so they run without a GPU
    lib = ptstokes.load()
    ndim, nx, ny, nz = grid
    g = _lib.Grid(ndim, nx, ny, nz, 1, _lib.PTS_FD_MUL, 0, 0)
    h = ctypes.c_void_p()
    assert lib.pts_create(ctypes.byref(g), ctypes.byref(h)) == _lib.PTS_ERR_ARG
    assert msg in lib.pts_last_error(None).decode()
    assert not h.value


def test_create_rejects_bad_fd_mode():
    lib = ptstokes.load()
    g = _lib.Grid(2, 31, 31, 1, 1, 7, 0, 0)
    h = ctypes.c_void_p()
    assert lib.pts_create(ctypes.byref(g), ctypes.byref(h)) == _lib.PTS_ERR_ARG
    assert "fd_mode" in lib.pts_last_error(None).decode()
