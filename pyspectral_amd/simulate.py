"""Time loop with the semantics of src/pyspectral/simulate.py:14-96.

Rules kept from the reference: fixed dt, or CFL-adaptive dt re-evaluated every ``cfl_every``
steps (dt = min(dt_max, cfl * dx / sum_i max|u_i|), dt_max if that sum is not a positive
finite number); the step is shortened to land exactly on ``t_end`` and on every diagnostics
or callback time; with ``target_power`` the band growth rate is reset every ``cfl_every``
steps to A = P / (2 E_band); (t, E, Z) is recorded at the start, at every diagnostics time
and at the end.

Two drivers implement them.  Solvers with a device controller (``device_loop`` true: the
hip variants) run the whole loop on the GPU in captured chunks, and the host only collects the
history; that is the default when no callback or wall-clock limit asks for host control
between steps.  Every other case runs the host loop below, one ``solver.step`` at a time.
"""

from __future__ import annotations

import math
import time
from collections.abc import Callable
from dataclasses import dataclass, field
from typing import Any


@dataclass
class RunResult:
    """History of a run (src/pyspectral/simulate.py:14-29); ``amp`` holds the band growth
    rate of each step."""

    t: list[float] = field(default_factory=list)
    energy: list[float] = field(default_factory=list)
    enstrophy: list[float] = field(default_factory=list)
    dt: list[float] = field(default_factory=list)
    amp: list[float] = field(default_factory=list)
    steps: int = 0
    wall: float = 0.0
    finished: bool = True

    def as_dict(self) -> dict[str, Any]:
        return {"t": self.t, "energy": self.energy, "enstrophy": self.enstrophy, "dt": self.dt,
                "steps": self.steps, "wall_s": self.wall}


def cfl_dt(solver: Any, cfl: float, dt_max: float) -> float:
    """CFL step from the current state's velocity maxima, capped at ``dt_max``."""
    umax = solver.diag.max_velocity_sum(solver.state_native())
    if not (umax > 0.0 and math.isfinite(umax)):
        return dt_max
    return min(dt_max, cfl * (2.0 * math.pi / solver.n) / umax)


class _Landing:
    """Upcoming event times (diagnostics, callback) that steps must land on."""

    def __init__(self, t: float, every: float | None) -> None:
        self.every = every
        self.next = t + every if every else math.inf

    def reached(self, t: float, eps: float) -> bool:
        if t < self.next - eps:
            return False
        self.next += self.every  # type: ignore[operator]
        return True


def run(solver: Any, t_end: float, *, dt: float | None = None, cfl: float = 0.5,
        dt_max: float = 0.05, cfl_every: int = 10, diag_dt: float | None = None,
        target_power: float | None = None, callback: Callable[[Any], None] | None = None,
        callback_dt: float | None = None, wall_limit: float | None = None,
        result: RunResult | None = None, device: bool | None = None) -> RunResult:
    """Integrate ``solver`` to ``t_end``; ``device`` forces (True) or forbids (False) the
    device controller, None picks it whenever the run allows it."""
    res = result if result is not None else RunResult()
    can_device = (getattr(solver, "device_loop", False) and callback is None
                  and wall_limit is None)
    if device and not can_device:
        raise ValueError("the device controller needs a hip solver, no callback and no "
                         "wall-clock limit")
    if device is None:
        device = can_device
    if device:
        return _run_on_device(solver, t_end, res, dt=dt, cfl=cfl, dt_max=dt_max,
                              cfl_every=cfl_every, diag_dt=diag_dt, target_power=target_power)
    return _run_on_host(solver, t_end, res, dt=dt, cfl=cfl, dt_max=dt_max,
                        cfl_every=cfl_every, diag_dt=diag_dt, target_power=target_power,
                        callback=callback, callback_dt=callback_dt, wall_limit=wall_limit)


def _run_on_device(solver: Any, t_end: float, res: RunResult, **kw: Any) -> RunResult:
    t0 = time.perf_counter()
    diag_dt = kw.pop("diag_dt")
    # a continued run keeps the record schedule of its earlier part
    first = res.t[-1] + diag_dt if (res.t and diag_dt) else None
    out = solver.run_device(t_end, diag_dt=diag_dt, first_diag=first, record_start=not res.t,
                            last_record=res.t[-1] if res.t else None, **kw)
    res.t += out["t"]
    res.energy += out["energy"]
    res.enstrophy += out["enstrophy"]
    res.dt += out["dt"]
    res.amp += out["amp"]
    res.steps += out["steps"]
    res.finished = out["finished"]
    res.wall += time.perf_counter() - t0
    return res


def _run_on_host(solver: Any, t_end: float, res: RunResult, *, dt: float | None, cfl: float,
                 dt_max: float, cfl_every: int, diag_dt: float | None,
                 target_power: float | None, callback: Callable[[Any], None] | None,
                 callback_dt: float | None, wall_limit: float | None) -> RunResult:
    d, cfg = solver.diag, solver.cfg

    def record() -> None:
        st = solver.state_native()
        res.t.append(solver.t)
        res.energy.append(d.energy(st))
        res.enstrophy.append(d.enstrophy(st))

    if not res.t:
        record()
    diag_at = _Landing(res.t[-1], diag_dt)
    call_at = _Landing(solver.t, callback_dt)
    eps = 1e-12 * max(1.0, t_end)
    step_dt = dt if dt is not None else cfl_dt(solver, cfl, dt_max)
    start = time.perf_counter()
    k = 0
    while solver.t < t_end - eps:
        if k % cfl_every == 0:
            if dt is None and k:
                step_dt = cfl_dt(solver, cfl, dt_max)
            if target_power is not None:
                eb = d.band_energy(solver.state_native(), cfg.force_klo, cfg.force_khi)
                solver.set_forcing_amplitude(target_power / (2.0 * eb) if eb > 0 else 0.0)
        h = min(step_dt, t_end - solver.t, diag_at.next - solver.t, call_at.next - solver.t)
        solver.step(h)
        res.dt.append(h)
        res.amp.append(solver.force_amp)
        k += 1
        if diag_at.reached(solver.t, eps):
            record()
        if callback is not None and call_at.reached(solver.t, eps):
            callback(solver)
        if wall_limit is not None and time.perf_counter() - start > wall_limit:
            break
    solver.synchronize()
    res.finished = solver.t >= t_end - eps
    res.wall += time.perf_counter() - start
    res.steps += k
    if res.finished and res.t[-1] < solver.t - eps:
        record()
    return res
