"""ctypes binding of libpts (include/pts.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (hipcc for
gfx950) next to this package.  There is no fallback: if the library is
missing or cannot be loaded, every entry point raises.
"""
import ctypes
import os

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libpts.so")

PTS_OK = 0
PTS_ERR_ARG = -1
PTS_ERR_HIP = -2
PTS_ERR_NAN = -3
PTS_ERR_NOCONV = -4
PTS_ERR_OOM = -5
PTS_ERR_STATE = -6

PTS_FD_DIV = 0
PTS_FD_MUL = 1

FIELDS = [
    "Pt", "Vx", "Vy", "Vz",
    "txx", "tyy", "tzz", "txy", "txz", "tyz",
    "txx_o", "tyy_o", "tzz_o", "txy_o", "txz_o", "tyz_o",
    "Mus", "Mus2", "Mustau", "Gdtau", "dtau_Rho",
    "divV", "Rx", "Ry", "Rz", "dVx", "dVy", "dVz", "rhog",
]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
# Unicode spellings used by the reference scripts
ALIASES = {"τxx": "txx", "τyy": "tyy", "τzz": "tzz", "τxy": "txy", "τxz": "txz", "τyz": "tyz",
           "τxx_o": "txx_o", "τyy_o": "tyy_o", "τzz_o": "tzz_o", "τxy_o": "txy_o", "τxz_o": "txz_o",
           "τyz_o": "tyz_o", "Musτ": "Mustau", "Gdτ": "Gdtau", "dτ_Rho": "dtau_Rho", "∇V": "divV",
           "ρg": "rhog", "RhoG": "rhog"}


class PtsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libpts error {code}: {msg}")
        self.code = code


class Grid(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("nz", ctypes.c_int32), ("ve", ctypes.c_int32), ("fd_mode", ctypes.c_int32),
                ("body_force", ctypes.c_int32), ("device", ctypes.c_int32)]


class Norms(ctypes.Structure):
    _fields_ = [("Vmin", ctypes.c_double), ("Vmax", ctypes.c_double), ("Pmin", ctypes.c_double),
                ("Pmax", ctypes.c_double), ("sumsq", ctypes.c_double * 4), ("maxabs", ctypes.c_double * 4),
                ("len", ctypes.c_int64 * 4)]


class PtParams(ctypes.Structure):
    _fields_ = [("r", ctypes.c_double), ("G", ctypes.c_double), ("dt", ctypes.c_double),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("lx", ctypes.c_double)]


class SolveParams(ctypes.Structure):
    _fields_ = [("eps", ctypes.c_double), ("nout", ctypes.c_int64), ("iter_max", ctypes.c_int64),
                ("norm_kind", ctypes.c_int32), ("use_graph", ctypes.c_int32)]


class SolveResult(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int64), ("err", ctypes.c_double), ("comp", ctypes.c_double * 4),
                ("converged", ctypes.c_int32), ("nchecks", ctypes.c_int32)]


class DcParams(ctypes.Structure):
    _fields_ = [("eps", ctypes.c_double), ("norm_kind", ctypes.c_int32), ("max_outer", ctypes.c_int32),
                ("rtol_inner", ctypes.c_double), ("nout_inner", ctypes.c_int64), ("iter_max_inner", ctypes.c_int64)]


class DcResult(ctypes.Structure):
    _fields_ = [("outer", ctypes.c_int32), ("converged", ctypes.c_int32), ("inner_iters", ctypes.c_int64),
                ("err", ctypes.c_double), ("err_outer", ctypes.c_double * 32)]


# (name, restype, argtypes) of every symbol in include/pts.h
_vp = ctypes.c_void_p
_d = ctypes.c_double
_i = ctypes.c_int
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)
SIGNATURES = [
    ("pts_create", _i, [ctypes.POINTER(Grid), ctypes.POINTER(_vp)]),
    ("pts_destroy", _i, [_vp]),
    ("pts_last_error", ctypes.c_char_p, [_vp]),
    ("pts_field_shape", _i, [_vp, _i, ctypes.POINTER(_i64)]),
    ("pts_upload", _i, [_vp, _i, _dp, _i64]),
    ("pts_download", _i, [_vp, _i, _dp, _i64]),
    ("pts_fill", _i, [_vp, _i, _d]),
    ("pts_copy", _i, [_vp, _i, _i]),
    ("pts_swap", _i, [_vp, _i, _i]),
    ("pts_sync", _i, [_vp]),
    ("pts_get_stream", _i, [_vp, ctypes.POINTER(_vp)]),
    ("pts_device_bytes", _i, [_vp, ctypes.POINTER(_i64)]),
    ("pts_placement_times", _i, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int32]),
    ("pts_set_knob", _i, [_vp, ctypes.c_char_p, ctypes.c_int64]),
    ("pts_norms_dense_noalloc", _i, [_i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("pts_init_inclusion", _i, [_vp, _d, _d, _d, _d, _d, _d, _d, _d, _d, _i32]),
    ("pts_init_grf_viscosity", _i, [_vp, _i32, _dp, _d, _d, _d, _d, _d, _d, _d, _d]),
    ("pts_smooth", _i, [_vp, _i, _i, _d]),
    ("pts_maxloc", _i, [_vp, _i, _i]),
    ("pts_bc_x", _i, [_vp, _i]),
    ("pts_bc_y", _i, [_vp, _i]),
    ("pts_bc_z", _i, [_vp, _i]),
    ("pts_iter_params", _i, [_vp, _d, _d, _d, _d, _d, _d]),
    ("pts_assign_tau", _i, [_vp]),
    ("pts_compute_P", _i, [_vp, _d, _d, _d, _d]),
    ("pts_compute_tau", _i, [_vp, _d, _d, _d, _d, _d]),
    ("pts_compute_dV", _i, [_vp, _d, _d, _d]),
    ("pts_compute_V", _i, [_vp]),
    ("pts_compute_Res", _i, [_vp, _d, _d, _d]),
    ("pts_compute_Pt_tau", _i, [_vp, _d, _d, _d, _d, _d, _d]),
    ("pts_compute_V_perf", _i, [_vp, _d, _d, _d]),
    ("pts_compute_Res_perf", _i, [_vp, _d, _d, _d]),
    ("pts_norms_stored", _i, [_vp, ctypes.POINTER(Norms)]),
    ("pts_err_from_norms", _d, [ctypes.POINTER(Norms), _i32, _d, _i32, _dp]),
    ("pts_pt_iterate", _i, [_vp, ctypes.POINTER(PtParams), _i64]),
    ("pts_pt_solve", _i, [_vp, ctypes.POINTER(PtParams), ctypes.POINTER(SolveParams),
                          ctypes.POINTER(SolveResult)]),
    ("pts_dc_solve", _i, [_vp, ctypes.POINTER(PtParams), ctypes.POINTER(DcParams), ctypes.POINTER(DcResult)]),
    ("pts_adjoint_set_source", _i, [_vp, _dp, _i64, _dp, _i64]),
    ("pts_adjoint_upload", _i, [_vp, _i32, _dp, _i64]),
    ("pts_adjoint_download", _i, [_vp, _i32, _dp, _i64]),
    ("pts_adjoint_iterate", _i, [_vp, ctypes.POINTER(PtParams), _i64]),
    ("pts_adjoint_solve", _i, [_vp, ctypes.POINTER(PtParams), ctypes.POINTER(SolveParams),
                               ctypes.POINTER(SolveResult)]),
    ("pts_grad_logmu", _i, [_vp, ctypes.POINTER(PtParams), _dp, _i64]),
    ("pts_checkpoint", _i, [_vp, _i32]),
    ("pts_adjoint_ve_solve", _i, [_vp, ctypes.POINTER(PtParams), _i32, ctypes.POINTER(SolveParams),
                                  ctypes.POINTER(SolveResult), _dp, _i64]),
]

_lib = None


def load(path=None):
    """Load libpts.so (raises if it is absent: there is no CPU fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libpts.so not built ({p}); run __graft_entry__.build()")
    lib = ctypes.CDLL(p)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
