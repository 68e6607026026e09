"""pts_norms_dense_noalloc (the norm reductions of pts_norms_stored over
caller-owned dense device arrays, with its partial results in the library's
file-scope device array) against pts_norms_stored on the same data, through
the C ABI: min/max and max|R| bit-identical, Σ R² to 1e-13 relative (the two
reductions add in different orders), on consecutive calls with shrinking
inputs, and without allocating device memory."""
import ctypes

import numpy as np
import pytest

from ptstokes import Context, _lib

pytestmark = pytest.mark.gpu

NAMES = ["Vx", "Pt", "Rx", "Ry", "Rz", "divV"]


class Hip:
    """The caller's own device memory and stream, through the HIP runtime."""

    def __init__(self):
        h = ctypes.CDLL("libamdhip64.so")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipStreamSynchronize.argtypes = [vp]
        h.hipStreamDestroy.argtypes = [vp]
        h.hipMemGetInfo.argtypes = [ctypes.POINTER(sz), ctypes.POINTER(sz)]
        self.h = h

    def chk(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        self.chk(self.h.hipMalloc(ctypes.byref(p), max(8, a.nbytes)))
        self.chk(self.h.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1))  # host to device
        return p

    def download(self, p, n):
        out = np.empty(n, dtype=np.float64)
        self.chk(self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2))  # device to host
        return out

    def free_bytes(self):
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        self.chk(self.h.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)))
        return f.value

    def stream(self):
        s = ctypes.c_void_p()
        self.chk(self.h.hipStreamCreate(ctypes.byref(s)))
        return s


def _one_call(lib, hip, stream, ndim, nx, ny, nz, seed, nan_in=None):
    ctx = Context(ndim, nx, ny, nz, ve=False, fd_mode="mul")
    rng = np.random.default_rng(seed)
    host = {}
    for n in NAMES:
        if ndim == 2 and n == "Rz":
            continue
        a = rng.uniform(-2.0, 3.0, size=ctx.shape(n))
        if n == nan_in:
            a.reshape(-1)[a.size // 3] = np.nan
        host[n] = a
        ctx.upload(n, a)
    want = ctx.norms_stored()
    ctx.close()
    # caller-owned device arrays in the reference's dense column-major shapes
    dev = {n: hip.upload(a.reshape(-1, order="F")) for n, a in host.items()}
    out = hip.upload(np.full(12, -7.0))
    free0 = hip.free_bytes()
    print(f"\nno-alloc case ndim={ndim} {nx}x{ny}x{nz} seed={seed} nan_in={nan_in}", flush=True)
    rc = lib.pts_norms_dense_noalloc(ndim, nx, ny, nz, dev["Vx"], dev["Pt"], dev["Rx"], dev["Ry"], dev.get("Rz"),
                                     dev["divV"], stream, out)
    hip.chk(hip.h.hipStreamSynchronize(stream))
    free1 = hip.free_bytes()
    assert rc == 0, lib.pts_last_error(None)
    got = hip.download(out, 12)
    for p in list(dev.values()) + [out]:
        hip.chk(hip.h.hipFree(p))
    ref = np.array([want.Vmin, want.Vmax, want.Pmin, want.Pmax] + list(want.sumsq) + list(want.maxabs))
    return got, ref, host, free0, free1


def _check(got, ref, host, ndim):
    exact = [0, 1, 2, 3, 8, 9, 10, 11]
    assert np.array_equal(got[exact], ref[exact], equal_nan=True), (got, ref)
    for q in range(4, 8):
        if np.isnan(ref[q]):
            assert np.isnan(got[q])
        else:
            assert got[q] == pytest.approx(ref[q], rel=1e-13, abs=0.0), (q, got[q], ref[q])
    # and against numpy on the host arrays
    assert got[0] == np.min(host["Vx"]) and got[1] == np.max(host["Vx"])
    assert got[2] == np.min(host["Pt"]) and got[3] == np.max(host["Pt"])
    comps = ["Rx", "Ry", "Rz", "divV"]
    for c, n in enumerate(comps):
        if n not in host:
            assert got[4 + c] == 0.0 and got[8 + c] == 0.0
            continue
        a = host[n]
        if np.isnan(a).any():
            assert np.isnan(got[4 + c]) and np.isnan(got[8 + c])
        else:
            assert got[8 + c] == np.max(np.abs(a))
            assert got[4 + c] == pytest.approx(float(np.sum(a * a)), rel=1e-12)


def test_noalloc_norms_equal_stored_norms_on_consecutive_calls():
    lib = _lib.load()
    hip = Hip()
    stream = hip.stream()
    cases = [(3, 45, 38, 27, 1, None), (3, 20, 17, 13, 2, None), (2, 37, 29, 1, 3, None), (2, 21, 9, 1, 4, "Ry")]
    for i, (ndim, nx, ny, nz, seed, nan_in) in enumerate(cases):
        got, ref, host, free0, free1 = _one_call(lib, hip, stream, ndim, nx, ny, nz, seed, nan_in)
        _check(got, ref, host, ndim)
        # no device memory taken by the call itself (the first call in a
        # process may load the code object that holds the scratch array)
        if i > 0:
            assert free1 == free0, (free0, free1)
    hip.chk(hip.h.hipStreamDestroy(stream))


def test_noalloc_rejects_bad_arguments():
    lib = _lib.load()
    hip = Hip()
    p = hip.upload(np.zeros(64))
    assert lib.pts_norms_dense_noalloc(4, 8, 8, 8, p, p, p, p, p, p, None, p) == _lib.PTS_ERR_ARG
    assert lib.pts_norms_dense_noalloc(3, 8, 2, 8, p, p, p, p, p, p, None, p) == _lib.PTS_ERR_ARG
    assert lib.pts_norms_dense_noalloc(3, 8, 8, 8, p, p, p, p, None, p, None, p) == _lib.PTS_ERR_ARG
    assert lib.pts_norms_dense_noalloc(3, 2000, 2000, 2000, p, p, p, p, p, p, None, p) == _lib.PTS_ERR_ARG
    assert b"2^31" in lib.pts_last_error(None)
    hip.chk(hip.h.hipFree(p))
