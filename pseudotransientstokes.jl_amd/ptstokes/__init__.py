"""ptstokes -- host side of the MI355X pseudo-transient Stokes library.

Mirrors the reference's per-kernel call surface (``Context``) and its script
drivers (``StokesModel``) over the C ABI of libpts (include/pts.h).  Every
computation runs in the HIP kernels of libpts on the GPU; nothing here
computes on the CPU.
"""
from ._lib import FIELDS, PtsError, load, LIB_PATH  # noqa: F401
from .context import Context  # noqa: F401
from .models import PRESETS, RE_OPT, R_OPT, StokesModel, config  # noqa: F401
