"""MI355X variants ``hip/fused`` and ``hip/fused-graph``: the IF-RK4 step on libpsx.

Host side of the path the reference implements in src/pyspectral/backends/cupy_cuda.py:32-213.
Torch owns the device buffers (state, RK accumulator, work fields); the C context borrows
their pointers and owns its plans, FP64 parameter block and captured graph.  Every call is
enqueued on torch's current stream.
"""

from __future__ import annotations

import ctypes
import math
from typing import Any

import numpy as np

from pyspectral_amd import _lib
from pyspectral_amd.backends.base import BackendUnavailable, SpectralSolver
from pyspectral_amd.config import Config
from pyspectral_amd.grid import retained
from pyspectral_amd.ops import TorchOps

_FORM_CODE = {"basdevant": 0, "rotational": 1}


def flow_params(cfg: Config, form: str) -> _lib.FlowParams:
    return _lib.FlowParams(n=cfg.n, dim=cfg.dim, form=_FORM_CODE[form], fp64=int(cfg.is_fp64),
                           nu=cfg.nu, alpha=cfg.alpha, ka2=cfg.drag_k2,
                           klo2=cfg.force_klo**2, khi2=cfg.force_khi**2,
                           dealias=1 if cfg.dealias == "phase-shift" else 0)


def check_ring_bounds(chunk: int, depth: int, cap_step: int, cap_rec: int) -> None:
    """The device loop's rings are drained before half of either can be overwritten: every
    step writes one step entry and at most one record, and up to ``depth + 1`` chunks run
    between two drains (plus the initial record)."""
    if chunk < 1 or depth < 1:
        raise ValueError("chunk and depth must be >= 1")
    if chunk * (depth + 1) + 1 > min(cap_step, cap_rec) // 2:
        raise ValueError(f"chunk * (depth + 1) = {chunk * (depth + 1)} exceeds half the "
                         f"step ring ({cap_step}) or the record ring ({cap_rec})")


class _Buffers:
    """Device buffers plus the C context that borrows them."""

    def __init__(self, params: _lib.FlowParams, state_shape: tuple[int, ...], cdt: Any) -> None:
        import torch

        lib = _lib.load()
        lay = _lib.Layout()
        if lib.psx_query_layout(ctypes.byref(params), ctypes.byref(lay)) != 0:
            raise BackendUnavailable(lib.psx_last_error().decode())
        self.layout = lay
        # state rows padded to whole 128-byte lines; the solver exposes the NumPy-layout view
        padded = state_shape[:-1] + (lay.state_pitch,)
        self.u0_rows = torch.zeros(padded, dtype=cdt, device="cuda")
        self.acc = torch.zeros(padded, dtype=cdt, device="cuda")
        self.u0 = self.u0_rows[..., : state_shape[-1]]
        self.pitch = lay.state_pitch
        self.work = torch.zeros(lay.work_elems, dtype=cdt, device="cuda")
        ctx = ctypes.c_void_p()
        _lib.check(lib.psx_create(ctypes.byref(params), ctypes.c_void_p(self.u0_rows.data_ptr()),
                                  ctypes.c_void_p(self.acc.data_ptr()),
                                  ctypes.c_void_p(self.work.data_ptr()), ctypes.byref(ctx)))
        self.ctx = ctx
        self.lib = lib

    def close(self) -> None:
        if self.ctx:
            import torch

            torch.cuda.synchronize()
            self.lib.psx_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass


class HipSolver(SpectralSolver):
    """Fused HIP kernels and the project's own FFTs (2D Basdevant; 3D Basdevant/rotational)."""

    backend = "hip"
    variant = "fused"
    default_form = "basdevant"
    forms = ("basdevant", "rotational")
    phase_shift = True
    device_loop = True  # run_device: the time loop of simulate.run on the GPU

    def __init__(self, cfg: Config, graph: bool = False, seed: int = 1234) -> None:
        super().__init__(cfg)
        if self.dim == 2 and self.form != "basdevant":
            raise BackendUnavailable("hip: 2D supports the basdevant form only")
        # phase-shift dealiasing: the grid shifts come from the oracle's generator and seed
        # (numpy_ref.py:46,125-128), drawn one step ahead on the host; the device-resident
        # loop has no generator of its own, so such solvers step from the host
        self.ps = cfg.dealias == "phase-shift"
        self.rng = np.random.default_rng(seed)
        self._shift: np.ndarray | None = None  # the next step's shift, once drawn
        if self.ps:
            self.device_loop = False
        ok, why = self.available()
        if not ok:
            raise BackendUnavailable(f"hip: {why}")
        import torch

        self.torch = torch
        self.cdt = torch.complex128 if cfg.is_fp64 else torch.complex64
        self.params = flow_params(cfg, self.form)
        self.buf = _Buffers(self.params, self.state_shape(), self.cdt)
        self.lib = self.buf.lib
        self.graph = graph
        self._key: tuple[float, float] | None = None
        self._ops = TorchOps()

    @classmethod
    def available(cls) -> tuple[bool, str]:
        try:
            import torch

            if not torch.cuda.is_available():
                return False, "no HIP device visible to torch"
            _lib.load()
            return _lib.device_info()
        except Exception as e:  # noqa: BLE001
            return False, f"{type(e).__name__}: {e}"

    @property
    def ops(self) -> Any:
        return self._ops

    @property
    def ctx(self) -> ctypes.c_void_p:
        return self.buf.ctx

    # ------------------------------------------------------------------ state
    def set_state(self, state_hat: np.ndarray) -> None:
        """cupy_cuda.py:134-145: mask on the host, copy in, compute the first stage inputs."""
        a = self.check_state(state_hat)
        host = self.torch.from_numpy(a.astype(np.complex128 if self.cfg.is_fp64
                                              else np.complex64))
        self.buf.u0.copy_(host.to(self.buf.u0.device))
        self.buf.acc.zero_()
        self._set_shifts(self._next_shift(), None)
        _lib.check(self.lib.psx_prepare(self.ctx, _lib.stream_ptr()))

    def set_state_device(self, state: Any) -> None:
        """Load a device tensor already in NumPy layout and masked (no host round trip)."""
        if tuple(state.shape) != self.state_shape():
            raise ValueError(f"state shape {tuple(state.shape)} != {self.state_shape()}")
        self.buf.u0.copy_(state.to(self.cdt))
        self._mask_device(self.buf.u0)
        self._set_shifts(self._next_shift(), None)
        _lib.check(self.lib.psx_prepare(self.ctx, _lib.stream_ptr()))

    def _next_shift(self) -> np.ndarray | None:
        """The shift of the next step (drawn now if not yet drawn), as numpy_ref.py:125-128."""
        if not self.ps:
            return None
        if self._shift is None:
            self._shift = self.rng.uniform(0.0, 2.0 * np.pi / self.n, size=self.dim)
        return self._shift

    def _set_shifts(self, cur: np.ndarray | None, nxt: np.ndarray | None) -> None:
        if cur is None:
            return
        nxt = cur if nxt is None else nxt
        a = (ctypes.c_double * 3)(*cur)
        b = (ctypes.c_double * 3)(*nxt)
        _lib.check(self.lib.psx_set_phase_shift(self.ctx, a, b, _lib.stream_ptr()))

    def _mask_device(self, t: Any) -> None:
        if self.ps:
            m = self.torch.from_numpy(self.mask_np()).to(t.device)
            t.mul_(m if self.dim == 2 else m.unsqueeze(0))
            return
        n, jm = self.n, retained(self.n)
        lo, hi = jm, n - jm + 1  # full-axis indices jm .. n-jm are outside 3|k| < n
        t[..., jm:] = 0
        if self.dim == 2:
            t[lo:hi, :] = 0
        else:
            t[:, lo:hi, :, :] = 0
            t[:, :, lo:hi, :] = 0

    def state_native(self) -> Any:
        return self.buf.u0

    # ------------------------------------------------------------------ stepping
    def _step(self, dt: float) -> None:
        st = _lib.stream_ptr()
        key = (float(dt), float(self.force_amp))
        if key != self._key:
            _lib.check(self.lib.psx_set_step(self.ctx, key[0], key[1], st))
            self._key = key
        if self.ps:
            cur = self._next_shift()
            self._shift = None
            nxt = self._next_shift()  # drawn now: the last stage prepares its inputs
            self._set_shifts(cur, nxt)
        if self.graph:
            _lib.check(self.lib.psx_step_graph(self.ctx, st))
        else:
            _lib.check(self.lib.psx_step(self.ctx, st))

    # ------------------------------------------------------------------ device time loop
    def run_device(self, t_end: float, *, dt: float | None = None, cfl: float = 0.5,
                   dt_max: float = 0.05, cfl_every: int = 10, diag_dt: float | None = None,
                   target_power: float | None = None, kmax: int = 0, chunk: int | None = None,
                   use_graph: bool = True, depth: int = 3, record_start: bool = True,
                   first_diag: float | None = None,
                   last_record: float | None = None) -> dict[str, Any]:
        """Integrate to ``t_end`` (or ``kmax`` steps) with the controller on the device
        (src/pyspectral/simulate.py:40-96 semantics, no callback).  Chunks of ``chunk`` steps
        are enqueued ``depth`` deep; the host waits only on the oldest chunk's state copy.
        ``last_record``: time of the last record of an earlier part of the same history (a
        continued run records its end state only when that one is older).
        Returns the records (t, E, Z) and the step history (h, A)."""
        if self.ps:
            raise BackendUnavailable("hip: the device-resident loop runs the 2/3 rule only "
                                     "(phase-shift grids are drawn on the host); use step()")
        lib, st = self.lib, _lib.stream_ptr()
        chunk = int(chunk or cfl_every)
        # a step writes one step entry and at most one record: both rings take the same bound
        cap_step = cap_rec = 1 << 14
        check_ring_bounds(chunk, depth, cap_step, cap_rec)
        rp = _lib.RunParams(t0=float(self.t), t_end=float(t_end),
                            dt_fixed=float(dt) if dt is not None else 0.0, cfl=float(cfl),
                            dt_max=float(dt_max), diag_dt=float(diag_dt or 0.0),
                            power=-1.0 if target_power is None else float(target_power),
                            amp0=float(self.force_amp), kmax=int(kmax),
                            cfl_every=int(cfl_every), cap_rec=cap_rec, cap_step=cap_step,
                            record_start=int(record_start),
                            next_diag=float(first_diag) if first_diag is not None else math.nan,
                            last_rec_t=float(last_record) if last_record is not None
                            else math.nan)
        _lib.check(lib.psx_run_begin(self.ctx, ctypes.byref(rp), st))
        self._key = None  # the controller rewrites the parameter block
        rec = np.zeros(3 * cap_rec)
        steps = np.zeros(2 * cap_step)
        out = {"t": [], "energy": [], "enstrophy": [], "dt": [], "amp": []}
        seen_k = seen_rec = 0
        state = _lib.RunState()

        def drain() -> None:
            nonlocal seen_k, seen_rec
            _lib.check(lib.psx_run_poll(self.ctx, ctypes.byref(state),
                                        rec.ctypes.data_as(ctypes.c_void_p),
                                        steps.ctypes.data_as(ctypes.c_void_p), st))
            for i in range(seen_rec, state.nrec):
                j = 3 * (i % cap_rec)
                out["t"].append(float(rec[j]))
                out["energy"].append(float(rec[j + 1]))
                out["enstrophy"].append(float(rec[j + 2]))
            for i in range(seen_k, state.k):
                j = 2 * (i % cap_step)
                out["dt"].append(float(steps[j]))
                out["amp"].append(float(steps[j + 1]))
            seen_k, seen_rec = state.k, state.nrec

        try:
            queued = 0
            peek = _lib.RunState()
            while True:
                _lib.check(lib.psx_run_chunk(self.ctx, chunk, int(use_graph), st))
                queued += 1
                if queued < depth:
                    continue
                _lib.check(lib.psx_run_peek(self.ctx, depth - 1, ctypes.byref(peek)))
                if peek.done:
                    break
                if (peek.k - seen_k) + chunk * depth > cap_step // 2 or \
                        (peek.nrec - seen_rec) + chunk * depth > cap_rec // 2:
                    drain()
            drain()
        finally:
            _lib.check(lib.psx_run_end(self.ctx, st))
        self.nstep += state.k
        self.t = float(state.t)
        self.force_amp = float(state.amp)
        out["steps"] = int(state.k)
        out["finished"] = bool(state.t >= t_end - 1e-12 * max(1.0, t_end))
        return out

    def _make_diag(self) -> Any:
        from pyspectral_amd.diagnostics import DeviceDiagnostics

        return DeviceDiagnostics(self)

    def close(self) -> None:
        self.buf.close()


class HipGraphSolver(HipSolver):
    """Same kernels, one step replayed from a hipGraph captured once."""

    variant = "fused-graph"

    def __init__(self, cfg: Config, seed: int = 1234) -> None:
        super().__init__(cfg, graph=True, seed=seed)
