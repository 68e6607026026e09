#!/usr/bin/env python3
"""Decode the reference's BSON golden files into JSON fixtures (data only).

The reference's ReferenceTests golden files (test/reftest-files/*.bson, written
by test/test_2D.jl:35-39 and test/test_3D_{1,2,3}.jl) each hold two Julia
arrays: ``X`` (sampled cell-centre x coordinates) and ``P`` (the pressure field
``Pt`` sampled at ``ceil.(LinRange(1, n, 12))`` in every dimension,
test_2D.jl:19-21).  BSON.jl stores an array as a sub-document
``{tag:"array", type:{...}, size:[...], data:<binary little-endian Float64,
column-major>}``.

This script reads the BSON bytes with a small stdlib-only parser (no code from
the file is executed) and writes ``<name>.json`` next to itself with the values
as hex floats (exact round trip), the shape, the sample indices and the grid
size.  Run it where the reference is mounted:

    python tests/golden/decode_bson.py /root/reference/test/reftest-files
"""
import json
import math
import os
import struct
import sys

FILES = {
    # file stem: (script that produced it, n, ndim)
    "test_Stokes_2D_v": ("scripts/Stokes2D.jl", 63, 2),
    "test_Stokes_2D_ve": ("scripts/Stokes2D_ve.jl", 63, 2),
    "test_Stokes_2D_ve_perf": ("scripts/Stokes2D_ve_perf.jl", 63, 2),
    "test_Stokes3D_ve_multixpu": ("scripts/stokes_3D/Stokes3D_ve_multixpu.jl", 31, 3),
    "test_Stokes3D_ve_multixpu_perf": ("scripts/stokes_3D/Stokes3D_ve_multixpu_perf.jl", 31, 3),
    "test_Stokes3D_ve_perf": ("scripts/stokes_3D/Stokes3D_ve_perf.jl", 31, 3),
}


def _read_doc(buf, off):
    size = struct.unpack_from("<i", buf, off)[0]
    end = off + size
    off += 4
    out = {}
    while off < end - 1:
        typ = buf[off]
        off += 1
        zero = buf.index(b"\0", off)
        name = buf[off:zero].decode()
        off = zero + 1
        if typ == 0x01:
            out[name] = struct.unpack_from("<d", buf, off)[0]
            off += 8
        elif typ in (0x03, 0x04):
            out[name], off = _read_doc(buf, off)
        elif typ == 0x02:
            ln = struct.unpack_from("<i", buf, off)[0]
            out[name] = buf[off + 4:off + 4 + ln - 1].decode()
            off += 4 + ln
        elif typ == 0x05:
            ln = struct.unpack_from("<i", buf, off)[0]
            out[name] = bytes(buf[off + 5:off + 5 + ln])
            off += 5 + ln
        elif typ == 0x10:
            out[name] = struct.unpack_from("<i", buf, off)[0]
            off += 4
        elif typ == 0x12:
            out[name] = struct.unpack_from("<q", buf, off)[0]
            off += 8
        elif typ == 0x08:
            out[name] = bool(buf[off])
            off += 1
        else:
            raise ValueError(f"unsupported BSON type 0x{typ:02x} at {off}")
    return out, end


def _array(doc):
    assert doc["tag"] == "array"
    names = [doc["type"]["name"][k] for k in sorted(doc["type"]["name"], key=int)]
    assert names == ["Core", "Float64"], names
    shape = [doc["size"][k] for k in sorted(doc["size"], key=int)]
    data = doc["data"]
    count = math.prod(shape)
    assert len(data) == 8 * count
    vals = list(struct.unpack("<%dd" % count, data))
    return shape, vals


def sample_indices(n, k=12):
    """Julia ``Int.(ceil.(LinRange(1, n, k)))`` as 1-based indices."""
    out = []
    for i in range(k):
        # LinRange(1,n,k)[i+1] = 1 + i*(n-1)/(k-1), computed like Julia's lerp
        t = i / (k - 1)
        v = (1 - t) * 1 + t * n
        out.append(int(math.ceil(v)))
    return out


def main(src_dir, dst_dir):
    for stem, (script, n, ndim) in FILES.items():
        with open(os.path.join(src_dir, stem + ".bson"), "rb") as f:
            buf = f.read()
        doc, _ = _read_doc(buf, 0)
        pshape, pvals = _array(doc["P"])
        xshape, xvals = _array(doc["X"])
        assert pshape == [12] * ndim and xshape == [12]
        fix = {
            "source": "test/reftest-files/%s.bson" % stem,
            "script": script,
            "ndim": ndim,
            "n": n,
            "sample_indices_1based": sample_indices(n),
            "P_shape": pshape,
            "P_colmajor_hex": [v.hex() for v in pvals],
            "X_hex": [v.hex() for v in xvals],
        }
        with open(os.path.join(dst_dir, stem + ".json"), "w") as f:
            json.dump(fix, f, indent=0)
        print(stem, pshape, "P range", min(pvals), max(pvals))


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test/reftest-files"
    main(src, os.path.dirname(os.path.abspath(__file__)))
