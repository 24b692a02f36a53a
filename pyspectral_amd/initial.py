"""Initial conditions (physical-space float64 arrays, or spectral states where noted).

The closed-form and shaped-noise fields follow src/pyspectral/initial.py:10-87 so that runs
started from the reference's cases are identical; ``random_phase_2d`` is the freely-decaying
turbulence field with energy spectrum E(k) ~ k^4 exp(-(k/k0)^2), uniform random phases and a
prescribed total energy, built directly in spectral space (it has no reference counterpart).
"""

from __future__ import annotations

import numpy as np

from pyspectral_amd.grid import (dealias_mask, hermitian_weights, k_squared, physical_grid,
                                 retained, wavenumbers)


def taylor_green_2d(n: int) -> np.ndarray:
    """omega = 2 sin x sin y: steady Euler solution, decays as exp(-2 nu t)."""
    x, y = physical_grid(n, 2)
    return 2.0 * np.sin(x) * np.sin(y)


def smooth_2d(n: int) -> np.ndarray:
    x, y = physical_grid(n, 2)
    w = np.exp(0.8 * np.cos(x) + 0.6 * np.sin(2.0 * y + 0.4 * np.cos(x)))
    w = w + 0.5 * np.sin(x + y) * np.cos(y)
    return w - w.mean()


def _shaped_noise(n: int, dim: int, k0: float, seed: int, ncomp: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((ncomp,) + (n,) * dim)
    spec = np.fft.rfftn(noise, axes=tuple(range(-dim, 0)))
    kk = np.sqrt(k_squared(n, dim))
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.sqrt(kk**4 * np.exp(-2.0 * (kk / k0) ** 2) /
                      np.where(kk > 0, kk ** (dim - 1), 1.0))
    amp[kk == 0] = 0.0
    return spec * amp * dealias_mask(n, dim)


def random_2d(n: int, k0: float = 8.0, seed: int = 0, energy: float = 0.5) -> np.ndarray:
    spec = _shaped_noise(n, 2, k0, seed, 1)[0]
    k2 = k_squared(n, 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        e = 0.5 * np.sum(np.where(k2 > 0, hermitian_weights(n, 2) * np.abs(spec) ** 2 / k2,
                                  0.0)) / float(n) ** 4
    return np.fft.irfftn(spec, s=(n, n), axes=(0, 1)) * np.sqrt(energy / e)


def taylor_green_3d(n: int) -> np.ndarray:
    x, y, z = physical_grid(n, 3)
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u * v)
    return np.stack(np.broadcast_arrays(u, v, w))


def random_3d(n: int, k0: float = 4.0, seed: int = 0, energy: float = 0.5) -> np.ndarray:
    uh = _shaped_noise(n, 3, k0, seed, 3)
    ks = wavenumbers(n, 3)
    k2 = sum(k * k for k in ks)
    inv = np.where(k2 > 0, 1.0 / np.where(k2 > 0, k2, 1.0), 0.0)
    div = sum(k * c for k, c in zip(ks, uh))
    uh = np.stack([c - k * div * inv for k, c in zip(ks, uh)])
    u = np.fft.irfftn(uh, s=(n, n, n), axes=(1, 2, 3))
    e = 0.5 * np.mean(np.sum(u * u, axis=0))
    return u * np.sqrt(energy / e)


def to_spectral(fld: np.ndarray, dim: int, dealias: str = "2/3") -> np.ndarray:
    """Physical field -> dealiased r2c state (complex128)."""
    n = fld.shape[-1]
    return np.fft.rfftn(fld, axes=tuple(range(-dim, 0))) * dealias_mask(n, dim, dealias)


def random_phase_2d(n: int, k0: float = 10.0, seed: int = 0, energy: float = 0.5,
                    dtype: type = np.complex128) -> np.ndarray:
    """Spectral vorticity with E(k) ~ k^4 exp(-(k/k0)^2) and uniform random phases.

    Built on the retained 2/3-rule modes only: |omega_hat|^2 = c k^2 E(k) / (2 pi k) (continuous
    shell area), phase theta_k ~ U[0, 2 pi) from ``default_rng(seed)`` drawn over the retained
    block row by row, the ky = 0 column made Hermitian (omega_hat(-kx, 0) = conj(omega_hat(kx, 0)),
    omega_hat(0, 0) = 0), then c set so that the kinetic energy (diagnostics.py:61-67) equals
    ``energy``.  Returns the NumPy r2c state (n, n/2+1) in ``dtype``.
    """
    jm = retained(n)
    nh = n // 2 + 1
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(n, 1.0 / n)
    keep_x = np.nonzero(3.0 * np.abs(kx) < n)[0]
    out = np.zeros((n, nh), dtype=dtype)
    ky = np.arange(jm, dtype=np.float64)
    w_h = np.where((ky == 0) | (ky == n // 2), 1.0, 2.0)
    e_tot = 0.0
    rows = {}
    for i in keep_x:
        k2 = kx[i] ** 2 + ky**2
        k = np.sqrt(k2)
        with np.errstate(divide="ignore", invalid="ignore"):
            amp2 = np.where(k > 0, k2 * k**4 * np.exp(-((k / k0) ** 2)) / (2 * np.pi * k), 0.0)
        theta = rng.uniform(0.0, 2.0 * np.pi, size=jm)
        rows[i] = np.sqrt(amp2) * np.exp(1j * theta)
    for i in keep_x:  # Hermitian ky = 0 column: take the kx > 0 half, mirror it
        kxi = kx[i]
        if kxi < 0:
            rows[i][0] = np.conj(rows[int(np.rint(-kxi))][0])
        elif kxi == 0:
            rows[i][0] = 0.0
    for i in keep_x:
        k2 = kx[i] ** 2 + ky**2
        with np.errstate(divide="ignore", invalid="ignore"):
            e_tot += 0.5 * np.sum(np.where(k2 > 0, w_h * np.abs(rows[i]) ** 2 / k2, 0.0))
    e_tot /= float(n) ** 4
    scale = np.sqrt(energy / e_tot) if e_tot > 0 else 0.0
    for i in keep_x:
        out[i, :jm] = rows[i] * scale
    return out
