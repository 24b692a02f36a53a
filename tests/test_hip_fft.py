"""The FFT engine against numpy.fft and closed forms (GPU only), through psx_fft_lines: the
plans, work-group shapes and twiddle tables of the step kernels' line transforms.

Bounds: max|a - b| / max|b| below 2e-6 against numpy.fft on random rows (fp32 butterflies,
errors of a few ulp times log2 L); the inverse of the forward transform returns L x; Parseval
within 1e-6; a single Fourier mode and an impulse give their closed forms.  The row counts are
not multiples of the rows per work-group, so the ragged last work-group is covered too."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from pyspectral_amd import _lib

pytestmark = pytest.mark.gpu
LENGTHS = [64, 96, 128, 192, 256, 384, 512, 1024, 2048, 4096, 8192]


def _fft(x: np.ndarray, inverse: bool) -> np.ndarray:
    lib = _lib.load()
    t = torch.from_numpy(x.astype(np.complex64)).cuda()
    _lib.check(lib.psx_fft_lines(t.data_ptr(), x.shape[-1], x.shape[0], int(inverse),
                                 _lib.stream_ptr()))
    return t.cpu().numpy().astype(np.complex128)


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("L", LENGTHS)
def test_fft_lines_match_numpy(L):
    rng = np.random.default_rng(L)
    rows = 37
    x = (rng.normal(size=(rows, L)) + 1j * rng.normal(size=(rows, L))).astype(np.complex64)
    xd = x.astype(np.complex128)
    fwd = _fft(x, False)
    inv = _fft(x, True)
    e_f = _rel(fwd, np.fft.fft(xd, axis=1))
    e_i = _rel(inv, np.fft.ifft(xd, axis=1) * L)
    back = _fft(fwd.astype(np.complex64), True) / L
    e_r = _rel(back, xd)
    pars = abs(np.sum(np.abs(fwd) ** 2) / (L * np.sum(np.abs(xd) ** 2)) - 1)
    print(f"  L={L}: forward {e_f:.2e} inverse {e_i:.2e} round trip {e_r:.2e} "
          f"Parseval {pars:.2e}", flush=True)
    assert e_f < 2e-6 and e_i < 2e-6 and e_r < 2e-6 and pars < 1e-6


@pytest.mark.parametrize("L", [96, 8192])
def test_fft_lines_closed_forms(L):
    """x_n = exp(2 pi i k0 n / L) -> X_k = L [k = k0]; x_n = [n = n0] -> X_k = exp(-2 pi i k n0 / L)
    (the sign convention of numpy.fft.fft and of the reference's cuFFT forward transforms)."""
    n = np.arange(L)
    k0, n0 = 5, 7
    mode = np.exp(2j * np.pi * k0 * n / L)[None, :]
    X = _fft(mode.astype(np.complex64), False)[0]
    want = np.zeros(L, complex)
    want[k0] = L
    assert np.abs(X - want).max() < 1e-5 * L
    imp = np.zeros((1, L), complex)
    imp[0, n0] = 1
    X = _fft(imp.astype(np.complex64), False)[0]
    assert np.abs(X - np.exp(-2j * np.pi * n * n0 / L)).max() < 1e-6
