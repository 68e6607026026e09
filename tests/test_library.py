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


def test_julia_script_entry_points_use_the_binding():
    # julia/scripts/*.jl replay the reference scripts' entry points (ENV knobs
    # NX, NY(, NZ), USE_RETURN, DO_SAVE, DO_SAVE_VIZ; README.md:71-79) on
    # PTStokesAMD; Julia is absent here, so check them as text: every
    # PTStokesAMD function they call exists in the binding
    jdir = os.path.join(ROOT, "pseudotransientstokes.jl_amd", "julia")
    binding = open(os.path.join(jdir, "PTStokesAMD.jl")).read()
    defined = set(re.findall(r"^(?:function\s+)?([A-Za-z_][\w!]*)\s*\(", binding, re.M))
    defined |= set(re.findall(r"^const\s+(\w+)", binding, re.M)) | set(re.findall(r"^(?:mutable\s+)?struct\s+(\w+)", binding, re.M))
    sdir = os.path.join(jdir, "scripts")
    common = open(os.path.join(sdir, "common.jl")).read()
    scripts = {f: open(os.path.join(sdir, f)).read() for f in sorted(os.listdir(sdir)) if f != "common.jl"}
    assert {"Stokes2D.jl", "Stokes2D_ve.jl", "Stokes2D_ve_perf.jl", "Stokes3D_ve_perf.jl",
            "Stokes3D_ve_multixpu.jl"} <= set(scripts)
    for name, txt in list(scripts.items()) + [("common.jl", common)]:
        used = set(re.findall(r"\bPT\.([A-Za-z_][\w!]*)", txt))
        assert used <= defined, (name, used - defined)
    for name, txt in scripts.items():
        assert 'include(joinpath(@__DIR__, "common.jl"))' in txt, name
        for k in ("NX", "NY", "USE_RETURN", "DO_SAVE", "DO_SAVE_VIZ"):
            assert f'knob("{k}"' in txt, (name, k)
        if "3D" in name:
            assert 'knob("NZ"' in txt, name
