"""The solver interface of the reference (src/pyspectral/backends/base.py:15-144): capability
class variables, the state/step API, the linear operator and the host-side state checks.
The ``hip`` variants subclass it, so code written against the reference's ``SpectralSolver``
drives them unchanged."""

from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Any, ClassVar

import numpy as np

from pyspectral_amd.config import Config, ffts_per_stage
from pyspectral_amd.grid import dealias_mask, spectral_shape


class BackendUnavailable(RuntimeError):
    """The backend cannot run here (no device, missing library, unsupported option)."""


class SpectralSolver(ABC):
    """Pseudo-spectral Navier-Stokes on [0, 2 pi)^d.  State: 2D vorticity (n, n/2+1), 3D
    velocity (3, n, n, n/2+1), unnormalised ``numpy.fft.rfftn`` coefficients."""

    backend: ClassVar[str] = "base"
    variant: ClassVar[str] = "default"
    default_form: ClassVar[str] = "basdevant"
    dims: ClassVar[tuple[int, ...]] = (2, 3)
    dtypes: ClassVar[tuple[str, ...]] = ("float32", "float64")
    phase_shift: ClassVar[bool] = False
    forms: ClassVar[tuple[str, ...] | None] = None

    def __init__(self, cfg: Config) -> None:
        if cfg.dim not in self.dims:
            raise BackendUnavailable(f"{self.backend}: dim={cfg.dim} not supported")
        if cfg.dtype not in self.dtypes:
            raise BackendUnavailable(f"{self.backend}: dtype={cfg.dtype} not supported")
        if cfg.dealias != "2/3" and not self.phase_shift:
            raise BackendUnavailable(f"{self.backend}: only 2/3-rule dealiasing")
        self.cfg = cfg
        self.n = cfg.n
        self.dim = cfg.dim
        self.form = cfg.resolved_form(self.default_form)
        if self.forms is not None and self.form not in self.forms:
            raise BackendUnavailable(f"{self.backend}: form={self.form} not supported")
        self.force_amp = cfg.force_amp
        self.t = 0.0
        self.nstep = 0
        self._diag: Any = None

    @classmethod
    def available(cls) -> tuple[bool, str]:
        return True, ""

    @abstractmethod
    def set_state(self, state_hat: np.ndarray) -> None:
        """Load a spectral state (host complex array, NumPy convention)."""

    @abstractmethod
    def state_native(self) -> Any:
        """Current state as a device array in NumPy layout."""

    @abstractmethod
    def _step(self, dt: float) -> None:
        """Advance one IF-RK4 step."""

    @property
    @abstractmethod
    def ops(self) -> Any:
        """Array shim (:mod:`pyspectral_amd.ops`)."""

    def get_state(self) -> np.ndarray:
        return np.asarray(self.ops.to_numpy(self.state_native()), dtype=np.complex128)

    def step(self, dt: float) -> None:
        self._step(dt)
        self.t += dt
        self.nstep += 1

    def advance(self, nsteps: int, dt: float) -> None:
        for _ in range(nsteps):
            self.step(dt)

    def synchronize(self) -> None:
        self.ops.synchronize()

    def set_forcing_amplitude(self, amp: float) -> None:
        self.force_amp = float(amp)

    @property
    def diag(self) -> Any:
        if self._diag is None:
            self._diag = self._make_diag()
        return self._diag

    def _make_diag(self) -> Any:
        raise NotImplementedError

    @property
    def fft_per_step(self) -> int:
        return 4 * ffts_per_stage(self.dim, self.form)

    def linear_operator(self, k2: Any) -> Any:
        """L(k) on a k^2 array (base.py:114-123)."""
        c = self.cfg
        L = -(c.nu * k2)
        if c.alpha != 0.0:
            L = L - (k2 * 0.0 + c.alpha) * (k2 <= c.drag_k2)
        if self.force_amp != 0.0:
            band = (k2 >= c.force_klo**2) * (k2 <= c.force_khi**2)
            L = L + (k2 * 0.0 + self.force_amp) * band
        return L

    def mask_np(self) -> np.ndarray:
        return dealias_mask(self.n, self.dim, self.cfg.dealias)

    def state_shape(self) -> tuple[int, ...]:
        s = spectral_shape(self.n, self.dim)
        return s if self.dim == 2 else (3,) + s

    def check_state(self, state_hat: np.ndarray) -> np.ndarray:
        """Shape check and dealias mask on the host, as complex128 (base.py:134-139)."""
        a = np.asarray(state_hat, dtype=np.complex128)
        if a.shape != self.state_shape():
            raise ValueError(f"state shape {a.shape} != {self.state_shape()}")
        return a * self.mask_np()

    def describe(self) -> str:
        return (f"{self.backend}/{self.variant} dim={self.dim} n={self.n} {self.cfg.dtype} "
                f"form={self.form} ({self.fft_per_step} FFTs/step)")
