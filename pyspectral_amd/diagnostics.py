"""Device diagnostics for the ``hip`` variants, FP64 whatever the working precision.

Same methods and normalisations as the reference's ``Diagnostics``
(src/pyspectral/diagnostics.py:20-109): energy <|u|^2>/2, enstrophy <|omega|^2>/2,
palinstrophy, band energy, the shell spectrum with sum_k E(k) = E, and sum_i max|u_i| for the
CFL condition.  The reductions run in libpsx (``psx_reduce``, ``psx_spectrum``); the velocity
maxima come from the line kernel run on the solver's transform buffers (``psx_velocity_max``),
so no extra inverse FFT is needed for the solver's own state.
"""

from __future__ import annotations

import ctypes
import math
from typing import Any

import numpy as np

from pyspectral_amd import _lib


class DeviceDiagnostics:
    def __init__(self, solver: Any) -> None:
        self.s = solver
        self.n, self.dim = solver.n, solver.dim
        self.nshell = int(math.floor(math.sqrt(self.dim) * (self.n // 2) + 0.5)) + 1
        self._scratch: Any = None

    # ------------------------------------------------------------------ helpers
    def _ptr(self, state: Any) -> tuple[ctypes.c_void_p, int]:
        """Device pointer and row pitch of a state (the solver's padded rows are used in place;
        anything else is converted to a contiguous tensor of the working precision)."""
        s = self.s
        if not hasattr(state, "data_ptr"):
            state = s.ops.asarray(np.asarray(state), dtype=s.cdt)
        if tuple(state.shape) != s.state_shape():
            raise ValueError(f"state shape {tuple(state.shape)} != {s.state_shape()}")
        own = s.state_native()
        if state.data_ptr() == own.data_ptr() and state.stride() == own.stride():
            return ctypes.c_void_p(state.data_ptr()), s.buf.pitch
        if state.dtype != s.cdt or not state.is_contiguous():
            state = state.to(s.cdt).contiguous()
        self._keep = state
        return ctypes.c_void_p(state.data_ptr()), state.shape[-1]

    def _reduce(self, state: Any, klo2: float = 0.0, khi2: float = -1.0) -> np.ndarray:
        out = (ctypes.c_double * 4)()
        ptr, pitch = self._ptr(state)
        _lib.check(self.s.lib.psx_reduce(self.s.ctx, ptr, pitch, klo2, khi2, out,
                                         _lib.stream_ptr()))
        return np.array(out[:])

    # ------------------------------------------------------------------ reference API
    def energy(self, state: Any) -> float:
        return float(self._reduce(state)[0])

    def enstrophy(self, state: Any) -> float:
        return float(self._reduce(state)[1])

    def palinstrophy(self, state: Any) -> float:
        return float(self._reduce(state)[2])

    def band_energy(self, state: Any, klo: float, khi: float) -> float:
        return float(self._reduce(state, klo * klo, khi * khi)[3])

    def all(self, state: Any, klo: float = 0.0, khi: float = -1.0) -> dict[str, float]:
        """Energy, enstrophy, palinstrophy and band energy from one FP64 reduction."""
        e, z, p, b = self._reduce(state, klo * klo, khi * khi if khi >= 0 else -1.0)
        return {"energy": float(e), "enstrophy": float(z), "palinstrophy": float(p),
                "band_energy": float(b)}

    def spectrum(self, state: Any) -> tuple[np.ndarray, np.ndarray]:
        out = (ctypes.c_double * self.nshell)()
        ptr, pitch = self._ptr(state)
        _lib.check(self.s.lib.psx_spectrum(self.s.ctx, ptr, pitch, out, self.nshell,
                                           _lib.stream_ptr()))
        return np.arange(self.nshell), np.array(out[:])

    def velocity_max(self, state: Any) -> np.ndarray:
        """max|u_i| per component (physical space) of ``state``."""
        s = self.s
        if hasattr(state, "data_ptr") and state.data_ptr() == s.state_native().data_ptr():
            solver = s
        else:
            solver = self._scratch_solver()
            solver.set_state_device(state if hasattr(state, "data_ptr")
                                    else s.ops.asarray(np.asarray(state)))
        out = (ctypes.c_double * 3)()
        _lib.check(solver.lib.psx_velocity_max(solver.ctx, out, _lib.stream_ptr()))
        return np.array(out[: self.dim])

    def max_velocity_sum(self, state: Any) -> float:
        return float(self.velocity_max(state).sum())

    def _scratch_solver(self) -> Any:
        if self._scratch is None:
            self._scratch = type(self.s)(self.s.cfg)
        return self._scratch
