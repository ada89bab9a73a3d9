"""Synthetic input fields (host side, numpy).

``grf`` draws a Gaussian random field the way the reference's random initial
condition does in spirit (src/shear_sim_3d.cpp:2793-3078): Box-Muller noise
convolved with exp(-d^2 / ell) (the reference's kernel uses llinv = 1/ell,
src/shear_sim_3d.cpp:57 and :3032, not 1/(2 ell^2)).  The convolution runs
through an FFT on a periodic box, so a seeded field is reproducible here and
on any host; the reference's own generator is unseeded glibc rand()
(src/shear_sim_3d.cpp:3060-3061) and cannot be matched draw for draw.
"""
from __future__ import annotations

import numpy as np


def grf(shape, ell, seed, unit_variance=True):
    """Periodic Gaussian random field on ``shape`` = (nz, ny, nx) (C order,
    x fastest), kernel exp(-d^2/ell) in cell units, zero mean, unit variance
    when ``unit_variance``."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    spec = np.fft.rfftn(noise)
    axes = [np.fft.fftfreq(n) * n for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1]) * shape[-1]]
    # Fourier transform of exp(-d^2/ell) per axis (periodised): a Gaussian
    # in wavenumber, exp(-pi^2 ell k^2 / n^2).
    filt = np.ones(spec.shape)
    for ax, (k, n) in enumerate(zip(axes, shape)):
        g = np.exp(-(np.pi ** 2) * ell * (k / n) ** 2)
        sh = [1] * len(shape)
        sh[ax] = len(k)
        filt = filt * g.reshape(sh)
    f = np.fft.irfftn(spec * filt, s=shape, axes=list(range(len(shape))))
    f -= f.mean()
    if unit_variance:
        f /= f.std()
    return f


def grf_vec3(shape, ell, seed):
    """Three independent unit-variance GRF components, AoS (N, 3)."""
    comps = [grf(shape, ell, seed * 3 + c) for c in range(3)]
    return np.stack([c.ravel() for c in comps], axis=1)
