"""The FFT engine's plans and twiddle tables (host code; no GPU needed).

Every line plan's radices multiply to its length, and every twiddle entry -- generated from
long double sin/cos and rounded to fp32 -- is within half an fp32 ulp of exp(-2 pi i m k / (Ns R))
computed in FP64, in both pass orders (psx_fft_twiddles)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from pyspectral_amd import _lib

LENGTHS = [64, 96, 128, 192, 256, 384, 512, 1024, 2048, 4096, 8192]


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="libpsx.so not built")
@pytest.mark.parametrize("L", LENGTHS)
@pytest.mark.parametrize("inverse", [0, 1])
def test_twiddles_are_rounded_fp64_values(L, inverse):
    lib = _lib.load()
    cap = 4 * L
    out = np.zeros(2 * cap, dtype=np.float32)
    rad = np.zeros(8, dtype=np.int32)
    npass = ctypes.c_int()
    rc = lib.psx_fft_twiddles(L, inverse, out.ctypes.data_as(ctypes.c_void_p), cap,
                              rad.ctypes.data_as(ctypes.c_void_p), ctypes.byref(npass))
    assert rc == 0, lib.psx_last_error()
    radices = [int(r) for r in rad[: npass.value]]
    assert int(np.prod(radices)) == L and all(r in (2, 3, 4, 5, 8, 12, 16, 32) for r in radices)
    want = []
    ns = radices[0]
    for R in radices[1:]:
        m, k = np.meshgrid(np.arange(1, R), np.arange(ns), indexing="ij")
        want.append(np.exp(-2j * np.pi * (m * k).ravel() / (ns * R)))
        ns *= R
    want = np.concatenate(want) if want else np.zeros(0, complex)
    got = out[: 2 * len(want)].reshape(-1, 2)
    for comp, w in ((got[:, 0], want.real), (got[:, 1], want.imag)):
        ulp = np.spacing(np.abs(w).astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(comp.astype(np.float64) - w) <= 0.5 * ulp + 1e-15)
