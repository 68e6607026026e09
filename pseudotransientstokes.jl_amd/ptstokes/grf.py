"""Seeded Gaussian random field for log10 viscosity (no counterpart in the
reference; used by the 3D mixed-precision and 2D adjoint configurations).

Random Fourier modes: f(x) = sigma*sqrt(2/N) * sum_m cos(k_m . x + phi_m)
with k_m ~ Normal(0, 2/corr_len^2) per component and phi_m ~ U[0, 2*pi),
so that E[f(x) f(x')] -> sigma^2 exp(-|x-x'|^2 / corr_len^2) as N grows.
The modes are drawn on the host from numpy's PCG64 generator with the given
seed and the field is evaluated on the device (pts_init_grf_viscosity).
"""
import numpy as np


def modes(seed, ndim, corr_len=0.5, nmodes=1024):
    rng = np.random.default_rng(seed)
    k = rng.normal(0.0, np.sqrt(2.0) / corr_len, size=(nmodes, 3))
    if ndim == 2:
        k[:, 2] = 0.0
    phi = rng.uniform(0.0, 2.0 * np.pi, size=nmodes)
    return np.column_stack([k, phi])


def evaluate(md, x, y, z=None, sigma=1.0):
    """Host evaluation of the field at points (for checks on small grids)."""
    ph = np.multiply.outer(x, md[:, 0]) + np.multiply.outer(y, md[:, 1]) + md[:, 3]
    if z is not None:
        ph = ph + np.multiply.outer(z, md[:, 2])
    return sigma * np.sqrt(2.0 / md.shape[0]) * np.cos(ph).sum(axis=-1)
