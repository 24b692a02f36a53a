"""Spectral grid conventions (src/pyspectral/grid.py:1-67): physical arrays ``(n,)*d``, r2c
spectra over the last axis with shape ``(n,)*(d-1) + (n//2+1,)``, integer wavenumbers."""

from __future__ import annotations

import numpy as np


def spectral_shape(n: int, dim: int) -> tuple[int, ...]:
    return (n,) * (dim - 1) + (n // 2 + 1,)


def wavenumbers(n: int, dim: int) -> tuple[np.ndarray, ...]:
    """Broadcastable float64 wavenumber arrays, full axes in FFT order, half axis 0..n/2."""
    full = np.fft.fftfreq(n, 1.0 / n)
    half = np.arange(n // 2 + 1, dtype=np.float64)
    shapes = [(n, 1), (1, n // 2 + 1)] if dim == 2 else [(n, 1, 1), (1, n, 1), (1, 1, n // 2 + 1)]
    axes = [full] * (dim - 1) + [half]
    return tuple(a.reshape(s) for a, s in zip(axes, shapes))


def k_squared(n: int, dim: int) -> np.ndarray:
    out = np.zeros(spectral_shape(n, dim))
    for k in wavenumbers(n, dim):
        out = out + k * k
    return out


def dealias_mask(n: int, dim: int, kind: str = "2/3") -> np.ndarray:
    """``"2/3"``: keep 3|k_i| < n on every axis; ``"phase-shift"``: |k|^2 < 2 (n/3)^2."""
    if kind == "2/3":
        keep = np.ones(spectral_shape(n, dim), dtype=bool)
        for k in wavenumbers(n, dim):
            keep = keep & (3.0 * np.abs(k) < n)
        return keep
    if kind == "phase-shift":
        return k_squared(n, dim) < 2.0 * (n / 3.0) ** 2
    raise ValueError(kind)


def retained(n: int) -> int:
    """Number of retained non-negative wavenumbers per axis under the 2/3 rule (3 k < n)."""
    return (n - 1) // 3 + 1


def hermitian_weights(n: int, dim: int) -> np.ndarray:
    w = np.full(n // 2 + 1, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    return np.broadcast_to(w.reshape((1,) * (dim - 1) + (-1,)), spectral_shape(n, dim))


def physical_grid(n: int, dim: int) -> tuple[np.ndarray, ...]:
    x = 2.0 * np.pi * np.arange(n) / n
    if dim == 2:
        return x[:, None], x[None, :]
    return x[:, None, None], x[None, :, None], x[None, None, :]
