"""Run configuration, identical in fields and validation to the reference's ``Config``
(src/pyspectral/config.py:12-92), so either object drives the ``hip`` variants."""

from __future__ import annotations

from dataclasses import dataclass, replace

FORMS_2D = ("basdevant", "advective")
FORMS_3D = ("basdevant", "rotational")
DEALIAS = ("2/3", "phase-shift")

_SPLIT = {(2, "basdevant"): (2, 2), (2, "advective"): (4, 1),
          (3, "basdevant"): (3, 5), (3, "rotational"): (6, 3)}


@dataclass(frozen=True)
class Config:
    """Periodic box of side 2 pi.  Linear operator integrated exactly:
    ``L(k) = -(nu k^2 + alpha [|k| <= alpha_kmax]) + force_amp [force_klo <= |k| <= force_khi]``
    (config.py:14-26)."""

    n: int
    dim: int = 2
    nu: float = 1e-3
    alpha: float = 0.0
    alpha_kmax: float = 0.0
    force_amp: float = 0.0
    force_klo: float = 0.0
    force_khi: float = 0.0
    dtype: str = "float32"
    form: str = "auto"
    dealias: str = "2/3"

    def __post_init__(self) -> None:
        if self.dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        if self.n < 8 or self.n % 2:
            raise ValueError("n must be even and >= 8")
        if self.dtype not in ("float32", "float64"):
            raise ValueError("dtype must be float32 or float64")
        forms = FORMS_2D if self.dim == 2 else FORMS_3D
        if self.form != "auto" and self.form not in forms:
            raise ValueError(f"form must be one of {forms} for dim={self.dim}")
        if self.dealias not in DEALIAS:
            raise ValueError(f"dealias must be one of {DEALIAS}")

    def resolved_form(self, default: str | None = None) -> str:
        if self.form != "auto":
            return self.form
        return default or "basdevant"

    @property
    def drag_k2(self) -> float:
        """Upper k^2 of the drag band (1e30 when the drag is uniform)."""
        return self.alpha_kmax**2 if self.alpha_kmax > 0 else 1e30

    def with_(self, **kw: object) -> Config:
        return replace(self, **kw)  # type: ignore[arg-type]

    @property
    def is_fp64(self) -> bool:
        return self.dtype == "float64"


def ffts_per_stage(dim: int, form: str) -> int:
    """Real FFTs per right-hand-side evaluation (config.py:73-81)."""
    return sum(_SPLIT[(dim, form)])


def fft_split(dim: int, form: str) -> tuple[int, int]:
    """(inverse, forward) FFT counts per stage (config.py:84-92)."""
    return _SPLIT[(dim, form)]
