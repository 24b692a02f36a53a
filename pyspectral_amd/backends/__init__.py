"""Variant registry with the reference's interface (src/pyspectral/backends/__init__.py:17-121).

Entries are loaded lazily; a bare backend name resolves to its ``default`` variant.  This
package registers the MI355X variants; INTEGRATION.md shows the lines that add them to the
reference's own ``REGISTRY``.
"""

from __future__ import annotations

import importlib
from dataclasses import dataclass, field
from typing import Any

from pyspectral_amd.backends.base import BackendUnavailable, SpectralSolver
from pyspectral_amd.config import Config


@dataclass(frozen=True)
class Variant:
    backend: str
    variant: str
    module: str
    cls: str
    color: str
    gpu: bool
    description: str
    kwargs: dict[str, Any] = field(default_factory=dict)
    default: bool = False

    @property
    def name(self) -> str:
        return f"{self.backend}/{self.variant}"

    def load(self) -> type[SpectralSolver]:
        mod = importlib.import_module(f"pyspectral_amd.backends.{self.module}")
        return getattr(mod, self.cls)

    def available(self) -> tuple[bool, str]:
        try:
            return self.load().available()
        except Exception as e:  # noqa: BLE001
            return False, f"{type(e).__name__}: {e}"

    def make(self, cfg: Config) -> SpectralSolver:
        cls = self.load()
        ok, why = cls.available()
        if not ok:
            raise BackendUnavailable(f"{self.name}: {why}")
        return cls(cfg, **self.kwargs)


COLORS = {"hip": "#ed1c24"}

REGISTRY: list[Variant] = [
    Variant("hip", "fused", "hip_fused", "HipSolver", COLORS["hip"], True,
            "hand-written HIP (gfx950) Stockham FFTs with the whole IF-RK4 stage fused in",
            default=True),
    Variant("hip", "fused-graph", "hip_fused", "HipGraphSolver", COLORS["hip"], True,
            "as above, one time step replayed as a hipGraph (parameters in device memory)"),
]


def names() -> list[str]:
    return [v.name for v in REGISTRY]


def get(name: str) -> Variant:
    for v in REGISTRY:
        if v.name == name:
            return v
    matches = [v for v in REGISTRY if v.backend == name]
    if matches:
        return next((v for v in matches if v.default), matches[-1])
    raise KeyError(f"unknown backend {name!r}; choose from {names()}")


def make_solver(name: str, cfg: Config) -> SpectralSolver:
    return get(name).make(cfg)


def available() -> list[tuple[Variant, bool, str]]:
    return [(v, *v.available()) for v in REGISTRY]


__all__ = ["REGISTRY", "Variant", "BackendUnavailable", "SpectralSolver", "get", "make_solver",
           "names", "available", "COLORS"]
