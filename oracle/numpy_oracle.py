"""CPU oracle: a NumPy/SciPy restatement of pyspectral's reference solver path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this module; only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker / CPU baseline, never as the thing measured or shipped.

What it restates (reference paths relative to the pyspectral repository):

* ``OracleSolver``      -- ``src/pyspectral/backends/numpy_ref.py:16-153`` (IF-RK4 / RK4, all
  nonlinear forms, phase-shift dealiasing) with ``base.py:83-139`` (step bookkeeping, L(k),
  mask/shape checks).
* ``OracleDiagnostics`` -- ``src/pyspectral/diagnostics.py:20-109``.
* ``run``               -- ``src/pyspectral/simulate.py:31-96`` (CFL, landing on t_end and on
  diagnostic/callback times, constant-power forcing).
* grid helpers          -- ``src/pyspectral/grid.py:13-67``; FFT counts ``config.py:73-92``.
* initial conditions    -- ``src/pyspectral/initial.py:10-87``.
* ``pad_spectrum``/``rel_err`` -- ``src/pyspectral/validation.py:25-40``.
* ``traffic_model``     -- ``src/pyspectral/benchmark.py:43-55``.

How each part is pinned (tests/test_oracle.py): the reference's own closed forms and invariants
(exact Taylor-Green decay, inviscid energy conservation, divergence-free projection,
dE/dt = -2 nu Z, forms agreeing, TGV3D E=0.125 / Z=0.375, band-forcing growth e^{A t}),
brute-force DFT checks, and golden vectors written by ``tests/golden/make_golden.py`` from the
reference itself (trajectories, nonlinear terms, diagnostics, run() sequences, ICs,
pad_spectrum).
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Any, Callable

import numpy as np
import scipy.fft as sfft

FFT_SPLIT = {(2, "basdevant"): (2, 2), (2, "advective"): (4, 1),
             (3, "basdevant"): (3, 5), (3, "rotational"): (6, 3)}  # config.py:84-92


# --------------------------------------------------------------------------- grid (grid.py)
def spectral_shape(n: int, dim: int) -> tuple[int, ...]:
    """grid.py:13-15 -- r2c over the last axis."""
    return (n,) * (dim - 1) + (n // 2 + 1,)


def wavenumbers(n: int, dim: int) -> tuple[np.ndarray, ...]:
    """grid.py:18-24 -- integer wavenumbers as broadcastable float64 arrays."""
    full = np.fft.fftfreq(n, 1.0 / n)
    half = np.arange(n // 2 + 1, dtype=np.float64)
    axes = [full] * (dim - 1) + [half]
    out = []
    for i, a in enumerate(axes):
        shape = [1] * dim
        shape[i] = a.size
        out.append(a.reshape(shape))
    return tuple(out)


def k_squared(n: int, dim: int) -> np.ndarray:
    """grid.py:27-33."""
    acc = np.zeros(spectral_shape(n, dim))
    for k in wavenumbers(n, dim):
        acc = acc + k * k
    return acc


def dealias_mask(n: int, dim: int, kind: str = "2/3") -> np.ndarray:
    """grid.py:36-50 -- cubic 2/3 truncation 3|k_i| < n, or spherical |k|^2 < 2 (n/3)^2."""
    if kind == "2/3":
        keep = np.ones(spectral_shape(n, dim), dtype=bool)
        for k in wavenumbers(n, dim):
            keep = keep & (3.0 * np.abs(k) < n)
        return keep
    if kind == "phase-shift":
        return k_squared(n, dim) < 2.0 * (n / 3.0) ** 2
    raise ValueError(kind)


def hermitian_weights(n: int, dim: int) -> np.ndarray:
    """grid.py:53-59 -- weight 1 on the first and last half-axis plane, 2 elsewhere."""
    w = np.full(n // 2 + 1, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    return np.broadcast_to(w.reshape((1,) * (dim - 1) + (-1,)), spectral_shape(n, dim))


def physical_grid(n: int, dim: int) -> tuple[np.ndarray, ...]:
    """grid.py:62-67."""
    x = 2.0 * np.pi * np.arange(n) / n
    out = []
    for i in range(dim):
        shape = [1] * dim
        shape[i] = n
        out.append(x.reshape(shape))
    return tuple(out)


def _cfg_fields(cfg: Any) -> dict[str, Any]:
    keys = ("n", "dim", "nu", "alpha", "alpha_kmax", "force_amp", "force_klo", "force_khi",
            "dtype", "form", "dealias")
    return {k: getattr(cfg, k) for k in keys}


# --------------------------------------------------------------------------- solver
class OracleSolver:
    """Restatement of ``NumpySolver`` (numpy_ref.py:16-153) and the shared ``SpectralSolver``
    bookkeeping it inherits (base.py:36-139).

    ``cfg`` is any object with the ``Config`` fields (config.py:28-38).
    """

    def __init__(self, cfg: Any, integrator: str = "ifrk4", workers: int | None = None,
                 seed: int = 1234) -> None:
        c = _cfg_fields(cfg)
        if integrator not in ("ifrk4", "rk4"):
            raise ValueError(integrator)
        self.cfg = cfg
        self.n, self.dim = int(c["n"]), int(c["dim"])
        self.form = c["form"] if c["form"] != "auto" else "basdevant"   # config.py:53-57
        self.nu, self.alpha = float(c["nu"]), float(c["alpha"])
        self.ka2 = c["alpha_kmax"] ** 2 if c["alpha_kmax"] > 0 else 1e30  # config.py:59-62
        self.klo, self.khi = float(c["force_klo"]), float(c["force_khi"])
        self.force_amp = float(c["force_amp"])
        self.dealias = c["dealias"]
        self.integrator = integrator
        n, d = self.n, self.dim
        # numpy_ref.py:35 -- threads only on big grids
        self.workers = workers if workers is not None else (-1 if n**d >= 2**16 else 1)
        fp64 = c["dtype"] == "float64"
        self.rdt = np.float64 if fp64 else np.float32
        self.cdt = np.complex128 if fp64 else np.complex64
        self.k = tuple(k.astype(self.rdt) for k in wavenumbers(n, d))
        self.k2 = k_squared(n, d).astype(self.rdt)
        safe = np.where(self.k2 > 0, self.k2, 1)
        self.inv_k2 = np.where(self.k2 > 0, 1.0 / safe, 0).astype(self.rdt)
        self.mask = dealias_mask(n, d, self.dealias)
        self.s = (n,) * d
        self.axes = tuple(range(-d, 0))
        self.rng = np.random.default_rng(seed)
        self.shift: np.ndarray | None = None
        self.state = np.zeros(self.state_shape(), self.cdt)
        self.t = 0.0
        self.nstep = 0
        # 3D only, no reference counterpart: a source held fixed over a step and added to the
        # nonlinear term of every stage (the LES subgrid-stress divergence); state layout
        self.source: np.ndarray | None = None

    # -- shared API (base.py) --------------------------------------------------------
    def state_shape(self) -> tuple[int, ...]:
        s = spectral_shape(self.n, self.dim)
        return s if self.dim == 2 else (3,) + s

    def check_state(self, state_hat: np.ndarray) -> np.ndarray:
        a = np.asarray(state_hat, dtype=np.complex128)
        if a.shape != self.state_shape():
            raise ValueError(f"state shape {a.shape} != {self.state_shape()}")
        return a * self.mask

    def set_state(self, state_hat: np.ndarray) -> None:
        self.state = self.check_state(state_hat).astype(self.cdt)

    def get_state(self) -> np.ndarray:
        return np.asarray(self.state, dtype=np.complex128)

    def set_forcing_amplitude(self, amp: float) -> None:
        self.force_amp = float(amp)

    def step(self, dt: float) -> None:
        self._step(dt)
        self.t += dt
        self.nstep += 1

    def advance(self, nsteps: int, dt: float) -> None:
        for _ in range(nsteps):
            self.step(dt)

    def linear_operator(self, k2: np.ndarray) -> np.ndarray:
        """base.py:114-123: L = -nu k^2 - alpha [k^2 <= ka^2] + A [klo^2 <= k^2 <= khi^2]."""
        L = -(self.nu * k2)
        if self.alpha != 0.0:
            L = L - (k2 * 0.0 + self.alpha) * (k2 <= self.ka2)
        if self.force_amp != 0.0:
            inband = (k2 >= self.klo**2) * (k2 <= self.khi**2)
            L = L + (k2 * 0.0 + self.force_amp) * inband
        return L

    # -- transforms ----------------------------------------------------------------
    def _irfft(self, a: np.ndarray) -> np.ndarray:
        return sfft.irfftn(a, s=self.s, axes=self.axes, workers=self.workers)

    def _rfft(self, a: np.ndarray) -> np.ndarray:
        return sfft.rfftn(a, axes=self.axes, workers=self.workers)

    def _phase(self, sign: float) -> Any:
        if self.shift is None:
            return 1.0
        arg = sum(k * float(dx) for k, dx in zip(self.k, self.shift))
        return np.exp(sign * 1j * arg).astype(self.cdt)

    # -- physics (numpy_ref.py:75-123) ---------------------------------------------
    def nonlinear(self, st: np.ndarray) -> np.ndarray:
        """Dealiased nonlinear term; in 3D also Leray-projected."""
        fwd_shift, back_shift = self._phase(+1.0), self._phase(-1.0)
        if self.dim == 2:
            kx, ky = self.k
            w = st * fwd_shift
            psi = w * self.inv_k2
            u = self._irfft(1j * ky * psi)
            v = self._irfft(-1j * kx * psi)
            if self.form == "basdevant":
                a = self._rfft(v * v - u * u)
                b = self._rfft(u * v)
                nl = kx * ky * a + (kx * kx - ky * ky) * b
            else:  # advective: -(u w_x + v w_y)
                wx = self._irfft(1j * kx * w)
                wy = self._irfft(1j * ky * w)
                nl = -self._rfft(u * wx + v * wy)
            return (nl * back_shift * self.mask).astype(self.cdt)
        k = self.k
        uh = st * fwd_shift
        u = self._irfft(uh)
        if self.form == "rotational":
            wh = np.stack([1j * (k[1] * uh[2] - k[2] * uh[1]),
                           1j * (k[2] * uh[0] - k[0] * uh[2]),
                           1j * (k[0] * uh[1] - k[1] * uh[0])])
            om = self._irfft(wh)
            cross = np.stack([u[1] * om[2] - u[2] * om[1],
                              u[2] * om[0] - u[0] * om[2],
                              u[0] * om[1] - u[1] * om[0]])
            nl = self._rfft(cross)
        else:  # basdevant: N_i = -i k_j T_ij with T the traceless product tensor
            q = (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / 3.0
            t11, t22, t12, t13, t23 = self._rfft(np.stack(
                [u[0] * u[0] - q, u[1] * u[1] - q, u[0] * u[1], u[0] * u[2], u[1] * u[2]]))
            t33 = -(t11 + t22)
            nl = -1j * np.stack([k[0] * t11 + k[1] * t12 + k[2] * t13,
                                 k[0] * t12 + k[1] * t22 + k[2] * t23,
                                 k[0] * t13 + k[1] * t23 + k[2] * t33])
        if self.source is not None:
            nl = nl + self.source
        nl = nl * back_shift * self.mask
        div = (k[0] * nl[0] + k[1] * nl[1] + k[2] * nl[2]) * self.inv_k2
        nl = np.stack([nl[i] - k[i] * div for i in range(3)])
        return nl.astype(self.cdt)

    def rhs(self, st: np.ndarray) -> np.ndarray:
        return self.nonlinear(st) + (self.linear_operator(self.k2) * st).astype(self.cdt)

    def _new_shift(self) -> None:
        if self.dealias == "phase-shift":
            self.shift = self.rng.uniform(0.0, 2.0 * np.pi / self.n, size=self.dim)

    def _half_shift(self) -> None:
        if self.shift is not None:
            self.shift = self.shift + np.pi / self.n

    def _step(self, dt: float) -> None:
        """numpy_ref.py:134-153 -- Lawson IF-RK4 (or explicit RK4)."""
        u0 = self.state
        self._new_shift()
        if self.integrator == "rk4":
            k1 = self.rhs(u0)
            k2 = self.rhs(u0 + 0.5 * dt * k1)
            self._half_shift()
            k3 = self.rhs(u0 + 0.5 * dt * k2)
            k4 = self.rhs(u0 + dt * k3)
            self.state = (u0 + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)).astype(self.cdt)
            return
        L = self.linear_operator(self.k2)
        E = np.exp(L * dt).astype(self.rdt)
        Eh = np.exp(L * (0.5 * dt)).astype(self.rdt)
        k1 = self.nonlinear(u0)
        k2 = self.nonlinear(Eh * (u0 + 0.5 * dt * k1))
        self._half_shift()
        k3 = self.nonlinear(Eh * u0 + 0.5 * dt * k2)
        k4 = self.nonlinear(E * u0 + dt * Eh * k3)
        self.state = (E * u0 + dt / 6.0 * (E * k1 + 2.0 * Eh * (k2 + k3) + k4)).astype(self.cdt)


# --------------------------------------------------------------------------- diagnostics
class OracleDiagnostics:
    """diagnostics.py:20-109 on host arrays, in float64."""

    def __init__(self, n: int, dim: int) -> None:
        self.n, self.dim = n, dim
        self.norm = float(n) ** (2 * dim)
        self.kvec = wavenumbers(n, dim)
        w = np.full(n // 2 + 1, 2.0)
        w[0] = w[-1] = 1.0
        self.w1 = w.reshape((1,) * (dim - 1) + (-1,)) / self.norm
        self.nshell = int(np.floor(np.sqrt(dim) * (n // 2) + 0.5)) + 1

    @property
    def k2(self) -> np.ndarray:
        return sum(k * k for k in self.kvec)

    @staticmethod
    def _inv_k2(k2: np.ndarray) -> np.ndarray:
        return (k2 > 0) / (k2 + (k2 == 0))

    def _density(self, st: np.ndarray) -> np.ndarray:
        st = np.asarray(st)
        a = (st.real ** 2 + st.imag ** 2)
        if self.dim == 3:
            a = a.sum(axis=0)
        return a * self.w1

    def _energy_density(self, st: np.ndarray) -> np.ndarray:
        d = self._density(st)
        return d * self._inv_k2(self.k2) if self.dim == 2 else d

    def energy(self, st: np.ndarray) -> float:
        return 0.5 * float(self._energy_density(st).sum())

    def enstrophy(self, st: np.ndarray) -> float:
        d = self._density(st)
        if self.dim == 3:
            d = d * self.k2
        return 0.5 * float(d.sum())

    def palinstrophy(self, st: np.ndarray) -> float:
        k2 = self.k2
        d = self._density(st) * k2
        if self.dim == 3:
            d = d * k2
        return 0.5 * float(d.sum())

    def band_energy(self, st: np.ndarray, klo: float, khi: float) -> float:
        k2 = self.k2
        d = self._energy_density(st)
        return 0.5 * float((d * ((k2 >= klo * klo) * (k2 <= khi * khi))).sum())

    def spectrum(self, st: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        d = self._energy_density(st)
        shell = (np.sqrt(self.k2) + 0.5).astype(np.int64).reshape(-1)
        e = np.bincount(shell, 0.5 * d.reshape(-1), self.nshell)
        return np.arange(len(e)), np.asarray(e, dtype=np.float64)

    def max_velocity_sum(self, st: np.ndarray) -> float:
        s = (self.n,) * self.dim
        axes = tuple(range(-self.dim, 0))
        if self.dim == 2:
            kx, ky = self.kvec
            psi = st * self._inv_k2(self.k2)
            u = sfft.irfftn(psi * ky * 1j, s=s, axes=axes, workers=-1)
            v = sfft.irfftn(psi * kx * (-1j), s=s, axes=axes, workers=-1)
            return float(np.abs(u).max() + np.abs(v).max())
        return float(sum(np.abs(sfft.irfftn(st[i], s=s, axes=axes, workers=-1)).max()
                         for i in range(3)))


# --------------------------------------------------------------------------- driver
@dataclass
class RunRecord:
    """simulate.py:14-28."""

    t: list[float] = field(default_factory=list)
    energy: list[float] = field(default_factory=list)
    enstrophy: list[float] = field(default_factory=list)
    dt: list[float] = field(default_factory=list)
    amp: list[float] = field(default_factory=list)
    steps: int = 0
    wall: float = 0.0
    finished: bool = True


def cfl_dt(umax: float, n: int, cfl: float, dt_max: float) -> float:
    """simulate.py:31-37 (umax = sum_i max|u_i|)."""
    dx = 2.0 * math.pi / n
    if umax <= 0.0 or not math.isfinite(umax):
        return dt_max
    return min(dt_max, cfl * dx / umax)


def run(solver: Any, t_end: float, diag: Any, *, dt: float | None = None, cfl: float = 0.5,
        dt_max: float = 0.05, cfl_every: int = 10, diag_dt: float | None = None,
        target_power: float | None = None, klo: float = 0.0, khi: float = 0.0,
        callback: Callable[[Any], None] | None = None, callback_dt: float | None = None,
        dt_sequence: list[float] | None = None,
        amp_sequence: list[float] | None = None) -> RunRecord:
    """simulate.py:40-96.  ``diag`` provides energy/enstrophy/band_energy/max_velocity_sum on
    the solver's state.  ``dt_sequence`` replays a recorded dt sequence instead of the CFL
    rule (used to give the oracle exactly the steps another solver took); ``amp_sequence``
    likewise replays the band growth rate of every step."""
    res = RunRecord()

    def record() -> None:
        st = solver.get_state()
        res.t.append(solver.t)
        res.energy.append(diag.energy(st))
        res.enstrophy.append(diag.enstrophy(st))

    record()
    next_diag = (res.t[-1] + diag_dt) if diag_dt else math.inf
    next_cb = solver.t + callback_dt if callback_dt else math.inf
    if dt_sequence is not None:
        seq = iter(dt_sequence)
    if amp_sequence is not None:
        aseq = iter(amp_sequence)
    cur_dt = dt if dt is not None else cfl_dt(diag.max_velocity_sum(solver.get_state()),
                                              solver.n, cfl, dt_max)
    t0 = time.perf_counter()
    k = 0
    eps = 1e-12 * max(1.0, t_end)
    while solver.t < t_end - eps:
        if k % cfl_every == 0:
            if dt is None and k > 0:
                cur_dt = cfl_dt(diag.max_velocity_sum(solver.get_state()), solver.n, cfl,
                                dt_max)
            if target_power is not None:
                eb = diag.band_energy(solver.get_state(), klo, khi)
                solver.set_forcing_amplitude(target_power / (2.0 * eb) if eb > 0 else 0.0)
        h = min(cur_dt, t_end - solver.t, next_diag - solver.t, next_cb - solver.t)
        if dt_sequence is not None:
            h = next(seq)
        if amp_sequence is not None:
            solver.set_forcing_amplitude(next(aseq))
        solver.step(h)
        res.dt.append(h)
        res.amp.append(solver.force_amp)
        k += 1
        if solver.t >= next_diag - eps:
            record()
            next_diag += diag_dt  # type: ignore[operator]
        if callback is not None and solver.t >= next_cb - eps:
            callback(solver)
            next_cb += callback_dt  # type: ignore[operator]
    res.finished = solver.t >= t_end - eps
    res.wall = time.perf_counter() - t0
    res.steps = k
    if res.finished and res.t[-1] < solver.t - eps:
        record()
    return res


# --------------------------------------------------------------------------- initial fields
def taylor_green_2d(n: int) -> np.ndarray:
    """initial.py:10-17 -- omega = 2 sin x sin y."""
    x, y = physical_grid(n, 2)
    return 2.0 * np.sin(x) * np.sin(y)


def smooth_2d(n: int) -> np.ndarray:
    """initial.py:20-25."""
    x, y = physical_grid(n, 2)
    w = np.exp(0.8 * np.cos(x) + 0.6 * np.sin(2.0 * y + 0.4 * np.cos(x)))
    w = w + 0.5 * np.sin(x + y) * np.cos(y)
    return w - w.mean()


def _shaped_noise(n: int, dim: int, k0: float, seed: int, ncomp: int) -> np.ndarray:
    """initial.py:28-41 -- Gaussian noise shaped to shell spectrum k^4 exp(-2 (k/k0)^2)."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((ncomp,) + (n,) * dim)
    spec = np.fft.rfftn(noise, axes=tuple(range(-dim, 0)))
    kk = np.sqrt(sum(k * k for k in wavenumbers(n, dim)))
    area = kk ** (dim - 1)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.sqrt(kk**4 * np.exp(-2.0 * (kk / k0) ** 2) / np.where(kk > 0, area, 1.0))
    amp[kk == 0] = 0.0
    return spec * amp * dealias_mask(n, dim)


def random_2d(n: int, k0: float = 8.0, seed: int = 0, energy: float = 0.5) -> np.ndarray:
    """initial.py:44-51."""
    spec = _shaped_noise(n, 2, k0, seed, 1)[0]
    k2 = k_squared(n, 2)
    w8 = hermitian_weights(n, 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        e = 0.5 * np.sum(np.where(k2 > 0, w8 * np.abs(spec) ** 2 / k2, 0.0)) / float(n) ** 4
    return np.fft.irfftn(spec, s=(n, n), axes=(0, 1)) * np.sqrt(energy / e)


def taylor_green_3d(n: int) -> np.ndarray:
    """initial.py:54-63."""
    x, y, z = physical_grid(n, 3)
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u * v)
    return np.stack(np.broadcast_arrays(u, v, w))


def random_3d(n: int, k0: float = 4.0, seed: int = 0, energy: float = 0.5) -> np.ndarray:
    """initial.py:66-76 -- solenoidal projection of shaped noise."""
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
    """initial.py:79-87."""
    n = fld.shape[-1]
    return np.fft.rfftn(fld, axes=tuple(range(-dim, 0))) * dealias_mask(n, dim, dealias)


# --------------------------------------------------------------------------- validation helpers
def pad_spectrum(a: np.ndarray, n_to: int, dim: int) -> np.ndarray:
    """validation.py:25-35 -- zero-pad or truncate an r2c spectrum, keeping amplitudes."""
    n = a.shape[-2]
    out = np.zeros(a.shape[: a.ndim - dim] + spectral_shape(n_to, dim), dtype=a.dtype)
    m = min(n, n_to) // 2
    kept = list(range(0, m)) + list(range(-m + 1, 0))
    src = [np.array([k % n for k in kept])] * (dim - 1) + [np.arange(m)]
    dst = [np.array([k % n_to for k in kept])] * (dim - 1) + [np.arange(m)]
    out[(...,) + np.ix_(*dst)] = a[(...,) + np.ix_(*src)] * (n_to / n) ** dim
    return out


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """validation.py:38-40."""
    return float(np.abs(a - b).max() / np.abs(b).max())


def traffic_model(n: int, dim: int, form: str, dtype: str) -> dict[str, float]:
    """benchmark.py:43-55 -- minimal bytes per step of the reference's fused schedule."""
    sr = 8 if dtype == "float64" else 4
    sc = 2 * sr
    npts = float(n) ** dim
    nmodes = float(np.prod(spectral_shape(n, dim)))
    ni, nf = FFT_SPLIT[(dim, form)]
    ncomp = 1 if dim == 2 else 3
    fft = 4 * (ni + nf) * (nmodes * sc + npts * sr)
    phys = 4 * (ni + nf) * npts * sr
    post = (4 * (nf + ncomp + ni) + 7 * ncomp) * nmodes * sc
    return {"fft_bytes": fft, "pointwise_bytes": phys + post, "total_bytes": fft + phys + post,
            "points": npts, "ffts_per_step": 4 * (ni + nf)}
