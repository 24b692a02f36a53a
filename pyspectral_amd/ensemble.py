"""Ensembles of independent 2D flows of uneven grid sizes on one GPU (libpsx ``psx_ens_*``).

Every member is a 2D fp32 Basdevant flow with its own ``Config`` (size, viscosity, drag,
forcing band), state, step and forcing amplitude.  All members advance together: each phase
of an IF-RK4 stage is one persistent launch over the tiles of every member, and the device
time loop (CFL dt, constant-power forcing, records) runs per member inside the same captured
chunks.  The reference has no ensembles; per member the semantics are those of the
single-flow solver (src/pyspectral/backends/cupy_cuda.py:175-195) and of
src/pyspectral/simulate.py:40-96.

``autotune`` picks each member size's spectral-pass variant (column-tile width, one pass or
the four-step split) and the persistent grid by timing the compiled candidates on the
device; every candidate computes the same transforms, so the choice changes run time only
(results agree to fp32 round-off).
"""

from __future__ import annotations

import ctypes
import gc
import hashlib
import json
import math
import os
import time
from collections.abc import Sequence
from pathlib import Path
from typing import Any

import numpy as np

from pyspectral_amd import _lib
from pyspectral_amd.backends.base import BackendUnavailable
from pyspectral_amd.backends.hip_fused import check_ring_bounds, flow_params
from pyspectral_amd.config import Config
from pyspectral_amd.grid import dealias_mask

VARIANT_NAMES = {0: "one pass, 16-column tiles", 1: "one pass, 8-column tiles",
                 2: "one pass, 4-column tiles", 3: "four-step, inner 256",
                 4: "four-step, inner 128"}


def variants_for(n: int) -> list[int]:
    """Compiled spectral-pass variants for member size n (first = default)."""
    lib = _lib.load()
    buf = (ctypes.c_int * 8)()
    k = lib.psx_ens_variants(n, buf, 8)
    return [buf[i] for i in range(min(k, 8))]


class EnsembleSolver:
    """Device buffers and the libpsx ensemble context for a list of member configs."""

    def __init__(self, cfgs: Sequence[Config], variants: Sequence[int] | None = None,
                 wg_per_cu: int | None = None) -> None:
        import torch

        if not cfgs:
            raise ValueError("an ensemble needs at least one member")
        for c in cfgs:
            if c.dim != 2 or c.dtype != "float32" or c.resolved_form() != "basdevant":
                raise BackendUnavailable("ensemble members are 2D basdevant fp32 flows")
            if c.dealias != "2/3":
                raise BackendUnavailable("ensemble members use 2/3-rule dealiasing")
        self.torch = torch
        self.lib = _lib.load()
        self.cfgs = list(cfgs)
        self.count = len(cfgs)
        self.variants = [int(v) for v in variants] if variants is not None else \
            [variants_for(c.n)[0] if variants_for(c.n) else -1 for c in cfgs]
        self.params = (_lib.FlowParams * self.count)(*[flow_params(c, "basdevant")
                                                       for c in cfgs])
        lays = []
        for m in range(self.count):
            lay = _lib.Layout()
            if self.lib.psx_ens_query_layout(ctypes.byref(self.params[m]), self.variants[m],
                                             ctypes.byref(lay)) != 0:
                raise BackendUnavailable(self.lib.psx_last_error().decode())
            lays.append(lay)
        # one arena per kind of buffer (state rows, accumulator, work), each member's slice at a
        # 64 KiB-aligned offset: the members' placement in HBM is then the same whatever the
        # process allocated and freed before (separate tensors from the caching allocator left
        # the line phase 5-15 % slower after an autotuning pass than in a fresh process)
        align = 8192  # complex64 elements in 64 KiB

        def arena(sizes: list[int]) -> list[Any]:
            offs, total = [], 0
            for sz in sizes:
                offs.append(total)
                total += -(-sz // align) * align
            buf = torch.zeros(total, dtype=torch.complex64, device="cuda")
            return [buf[o:o + sz] for o, sz in zip(offs, sizes)]

        row_sizes = [c.n * lay.state_pitch for c, lay in zip(cfgs, lays)]
        self.u0_rows = [r.view(c.n, lay.state_pitch)
                        for r, c, lay in zip(arena(row_sizes), cfgs, lays)]
        self.u0 = [r[:, :c.n // 2 + 1] for r, c in zip(self.u0_rows, cfgs)]
        self.acc = [a.view(c.n, lay.state_pitch) for a, c, lay in zip(arena(row_sizes), cfgs, lays)]
        self.work = arena([lay.work_elems for lay in lays])
        ptrs = lambda ts: (ctypes.c_void_p * self.count)(*[t.data_ptr() for t in ts])  # noqa: E731
        var = (ctypes.c_int * self.count)(*self.variants)
        ctx = ctypes.c_void_p()
        _lib.check(self.lib.psx_ens_create(self.count, self.params, var, ptrs(self.u0_rows),
                                           ptrs(self.acc), ptrs(self.work), ctypes.byref(ctx)))
        self.ctx = ctx
        if wg_per_cu is not None:
            _lib.check(self.lib.psx_ens_set_grid(self.ctx, int(wg_per_cu)))
        self.wg_per_cu = wg_per_cu or 2
        self.t = [0.0] * self.count
        self.nstep = [0] * self.count
        self.force_amp = [float(c.force_amp) for c in cfgs]
        self._key: tuple | None = None

    # ------------------------------------------------------------------ state
    def set_states(self, states: Sequence[np.ndarray]) -> None:
        """Load every member's spectral vorticity (NumPy rfft2 layout; masked here)."""
        if len(states) != self.count:
            raise ValueError("one state per member")
        for m, (c, s) in enumerate(zip(self.cfgs, states)):
            shape = (c.n, c.n // 2 + 1)
            if tuple(s.shape) != shape:
                raise ValueError(f"member {m}: state shape {s.shape} != {shape}")
            a = (np.asarray(s) * dealias_mask(c.n, 2)).astype(np.complex64)
            self.u0[m].copy_(self.torch.from_numpy(a).to("cuda"))
            self.acc[m].zero_()
        _lib.check(self.lib.psx_ens_prepare(self.ctx, _lib.stream_ptr()))

    def get_state(self, m: int) -> np.ndarray:
        return self.u0[m].cpu().numpy().astype(np.complex128)

    def state_native(self, m: int) -> Any:
        return self.u0[m]

    def synchronize(self) -> None:
        self.torch.cuda.synchronize()

    # ------------------------------------------------------------------ stepping
    def step(self, dt: float | Sequence[float], graph: bool = False) -> None:
        """One IF-RK4 step of every member (a scalar dt applies to all)."""
        dts = [float(dt)] * self.count if np.isscalar(dt) else [float(x) for x in dt]
        key = (tuple(dts), tuple(self.force_amp))
        st = _lib.stream_ptr()
        if key != self._key:
            d = (ctypes.c_double * self.count)(*dts)
            a = (ctypes.c_double * self.count)(*self.force_amp)
            _lib.check(self.lib.psx_ens_set_step(self.ctx, d, a, st))
            self._key = key
        if graph:
            _lib.check(self.lib.psx_ens_step_graph(self.ctx, st))
        else:
            _lib.check(self.lib.psx_ens_step(self.ctx, st))
        for m in range(self.count):
            self.t[m] += dts[m]
            self.nstep[m] += 1

    def set_forcing_amplitude(self, amps: Sequence[float]) -> None:
        self.force_amp = [float(a) for a in amps]

    # ------------------------------------------------------------------ device time loop
    def run_device(self, t_end: float | Sequence[float], *, dt: float | None = None,
                   cfl: float = 0.5, dt_max: float = 0.05, cfl_every: int = 10,
                   diag_dt: float | None = None, target_power: float | None = None,
                   kmax: int = 0, chunk: int | None = None, use_graph: bool = True,
                   depth: int = 3, record_start: bool = True) -> list[dict[str, Any]]:
        """Every member integrated with its own controller (simulate.py:40-96 rules) until all
        reach their end time or ``kmax`` steps; returns per member the records (t, E, Z)
        and the step history (h, A)."""
        lib, st, c = self.lib, _lib.stream_ptr(), self.count
        chunk = int(chunk or cfl_every)
        # a step writes one step entry and at most one record: both rings take the same bound
        cap_step = cap_rec = 1 << 12
        check_ring_bounds(chunk, depth, cap_step, cap_rec)
        ends = [float(t_end)] * c if np.isscalar(t_end) else [float(x) for x in t_end]
        rp = (_lib.RunParams * c)()
        for m in range(c):
            rp[m] = _lib.RunParams(
                t0=self.t[m], t_end=ends[m], dt_fixed=float(dt) if dt is not None else 0.0,
                cfl=float(cfl), dt_max=float(dt_max), diag_dt=float(diag_dt or 0.0),
                power=-1.0 if target_power is None else float(target_power),
                amp0=self.force_amp[m], kmax=int(kmax), cfl_every=int(cfl_every),
                cap_rec=cap_rec, cap_step=cap_step, record_start=int(record_start),
                next_diag=math.nan, last_rec_t=math.nan)
        _lib.check(lib.psx_ens_run_begin(self.ctx, rp, st))
        self._key = None
        rec = np.zeros(c * 3 * cap_rec)
        steps = np.zeros(c * 2 * cap_step)
        out = [{"t": [], "energy": [], "enstrophy": [], "dt": [], "amp": []} for _ in range(c)]
        seen_k = [0] * c
        seen_r = [0] * c
        state = (_lib.RunState * c)()
        peek = (_lib.RunState * c)()

        def drain() -> None:
            _lib.check(lib.psx_ens_run_poll(self.ctx, state,
                                            rec.ctypes.data_as(ctypes.c_void_p),
                                            steps.ctypes.data_as(ctypes.c_void_p), st))
            for m in range(c):
                base_r, base_s = m * 3 * cap_rec, m * 2 * cap_step
                for i in range(seen_r[m], state[m].nrec):
                    j = base_r + 3 * (i % cap_rec)
                    out[m]["t"].append(float(rec[j]))
                    out[m]["energy"].append(float(rec[j + 1]))
                    out[m]["enstrophy"].append(float(rec[j + 2]))
                for i in range(seen_k[m], state[m].k):
                    j = base_s + 2 * (i % cap_step)
                    out[m]["dt"].append(float(steps[j]))
                    out[m]["amp"].append(float(steps[j + 1]))
                seen_k[m], seen_r[m] = state[m].k, state[m].nrec

        try:
            queued = 0
            while True:
                _lib.check(lib.psx_ens_run_chunk(self.ctx, chunk, int(use_graph), st))
                queued += 1
                if queued < depth:
                    continue
                _lib.check(lib.psx_ens_run_peek(self.ctx, depth - 1, peek))
                if all(peek[m].done for m in range(c)):
                    break
                if any((peek[m].k - seen_k[m]) + chunk * depth > cap_step // 2 or
                       (peek[m].nrec - seen_r[m]) + chunk * depth > cap_rec // 2
                       for m in range(c)):
                    drain()
            drain()
        finally:
            _lib.check(lib.psx_ens_run_end(self.ctx, st))
        for m in range(c):
            self.t[m] = float(state[m].t)
            self.nstep[m] += int(state[m].k)
            self.force_amp[m] = float(state[m].amp)
            out[m]["steps"] = int(state[m].k)
        return out

    def timed_steps(self, nsteps: int, dt: float) -> dict[str, dict[str, float]]:
        """Per-phase HIP-event times of ``nsteps`` eager steps (ms total, launches)."""
        lib, st = self.lib, _lib.stream_ptr()
        _lib.check(lib.psx_ens_timing_begin(self.ctx))
        for _ in range(nsteps):
            self.step(dt)
        nk = len(_lib.KINDS)
        ms = (ctypes.c_double * nk)()
        cnt = (ctypes.c_int * nk)()
        _lib.check(lib.psx_ens_timing_end(self.ctx, st, ms, cnt))
        return {_lib.KINDS[i]: {"ms_total": ms[i], "launches": cnt[i]}
                for i in range(nk) if cnt[i]}

    def close(self) -> None:
        if self.ctx:
            self.torch.cuda.synchronize()
            self.lib.psx_ens_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass


# --------------------------------------------------------------------------- autotuner
def _time_steps(ens: EnsembleSolver, dt: float, nsteps: int, warm: int = 2) -> float:
    torch = ens.torch
    for _ in range(warm):
        ens.step(dt, graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        ens.step(dt, graph=True)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / nsteps


def _time_candidates(solvers: dict[int, EnsembleSolver], dt: Any, nsteps: int,
                     rounds: int = 3) -> dict[int, float]:
    """Seconds per step of each candidate: the best of ``rounds`` interleaved timings, so that
    a slow moment of the device (clocks, another process) does not pick the candidate timed
    during it."""
    best = {k: math.inf for k in solvers}
    for _ in range(rounds):
        for k, ens in solvers.items():
            best[k] = min(best[k], _time_steps(ens, dt, nsteps))
    return best


def _cache_path() -> Path:
    env = os.environ.get("PSX_AUTOTUNE_CACHE")
    return Path(env) if env else Path.home() / ".cache" / "pyspectral_amd" / "autotune.json"


def _cache_key(cfgs: Sequence[Config], nsteps: int, grids: Sequence[int]) -> str:
    """Device, library build and the ensemble's member sizes and counts: a tuning is reused
    only for the same ensemble shape on the same kind of device with the same kernels."""
    import torch

    props = torch.cuda.get_device_properties(0)
    lib = hashlib.sha256(_lib.LIB_PATH.read_bytes()).hexdigest()[:16]
    counts: dict[int, int] = {}
    for c in cfgs:
        counts[c.n] = counts.get(c.n, 0) + 1
    key = {"device": props.name, "arch": getattr(props, "gcnArchName", ""),
           "cus": props.multi_processor_count, "lib": lib,
           "members": sorted(counts.items()), "nsteps": nsteps, "grids": list(grids)}
    return hashlib.sha256(json.dumps(key, sort_keys=True).encode()).hexdigest()[:24]


def _cache_load(key: str) -> dict[str, Any] | None:
    try:
        return json.loads(_cache_path().read_text()).get(key)
    except (OSError, ValueError):
        return None


def _cache_store(key: str, result: dict[str, Any]) -> None:
    path = _cache_path()
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        try:
            data = json.loads(path.read_text())
        except (OSError, ValueError):
            data = {}
        data[key] = result
        tmp = path.with_suffix(f".tmp{os.getpid()}")
        tmp.write_text(json.dumps(data))
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs a re-tune next time


def autotune(cfgs: Sequence[Config], *, nsteps: int = 5, grids: Sequence[int] = (1, 2, 4),
             seed: int = 0, cache: bool = True) -> dict[str, Any]:
    """Choose each member size's spectral variant, then the persistent grid, by timing
    graph-replayed steps of the compiled candidates on this device (the best of three
    interleaved timings of ``nsteps`` steps each).

    Variants are timed on a sub-ensemble holding only that size's members (same count), on a
    random state; the grid is timed on the whole ensemble with the chosen variants.  Returns
    ``{"variants": [...] (per member), "wg_per_cu": g, "timings": {...}, "cached": bool}``.
    With ``cache`` the result is kept on disk (``$PSX_AUTOTUNE_CACHE``, default
    ``~/.cache/pyspectral_amd/autotune.json``) under a key of the device, the library build
    and the member sizes and counts, and reused instead of re-timing."""
    key = _cache_key(cfgs, nsteps, grids) if cache else ""
    if cache:
        hit = _cache_load(key)
        if hit is not None:
            by_n = {int(k): v for k, v in hit["variant_by_size"].items()}
            return {"variants": [by_n[c.n] for c in cfgs], "wg_per_cu": hit["wg_per_cu"],
                    "timings": hit["timings"], "cached": True}
    res = _autotune(cfgs, nsteps=nsteps, grids=grids, seed=seed)
    if cache:
        _cache_store(key, {"variant_by_size": {str(c.n): v for c, v in zip(cfgs,
                                                                            res["variants"])},
                           "wg_per_cu": res["wg_per_cu"], "timings": res["timings"]})
    res["cached"] = False
    return res


def _autotune(cfgs: Sequence[Config], *, nsteps: int, grids: Sequence[int],
              seed: int) -> dict[str, Any]:
    import torch

    from pyspectral_amd import initial

    by_size: dict[int, list[int]] = {}
    for m, c in enumerate(cfgs):
        by_size.setdefault(c.n, []).append(m)
    best_var: dict[int, int] = {}
    timings: dict[str, Any] = {"variants": {}, "grid": {}}
    for n, members in sorted(by_size.items()):
        sub = [cfgs[m] for m in members]
        states = [initial.random_phase_2d(n, k0=min(10.0, n / 12), seed=seed + i)
                  for i in range(len(sub))]
        dt = 0.2 * (2 * math.pi / n)
        solvers = {}
        for v in variants_for(n):
            solvers[v] = EnsembleSolver(sub, [v] * len(sub))
            solvers[v].set_states(states)
        res = _time_candidates(solvers, dt, nsteps)
        for ens in solvers.values():
            ens.close()
        best_var[n] = min(res, key=res.get)
        timings["variants"][n] = {VARIANT_NAMES[v]: t * 1e3 for v, t in res.items()}
    variants = [best_var[c.n] for c in cfgs]
    states = [initial.random_phase_2d(c.n, k0=min(10.0, c.n / 12), seed=seed + m)
              for m, c in enumerate(cfgs)]
    solvers = {}
    for g in grids:
        solvers[g] = EnsembleSolver(cfgs, variants, wg_per_cu=g)
        solvers[g].set_states(states)
    gres = _time_candidates(solvers, [0.2 * (2 * math.pi / c.n) for c in cfgs], nsteps)
    for ens in solvers.values():
        ens.close()
    # hand the candidates' buffers back to the device, so that the tuned ensemble's buffers
    # are allocated afresh rather than carved out of the cached blocks of the candidates
    ens = None
    solvers.clear()
    gc.collect()
    torch.cuda.empty_cache()
    best_grid = min(gres, key=gres.get)
    timings["grid"] = {g: t * 1e3 for g, t in gres.items()}
    return {"variants": variants, "wg_per_cu": best_grid, "timings": timings}


# ------------------------------------------------------------------ the forced ensemble
def forced_ensemble_configs(members: Sequence[tuple[int, int]] = ((256, 64), (512, 32),
                                                                  (1024, 16), (2048, 4)),
                            power: float = 1e-3, k_f: float = 5.0, drag: float = 0.1,
                            klo: float = 4.0, khi: float = 6.0) -> list[Config]:
    """Members of the batched forced 2D ensemble: constant-power band forcing on
    klo <= |k| <= khi, uniform linear drag, and nu chosen per size so that the enstrophy
    dissipation scale eta = (nu^3 / beta)^(1/6) equals 1 / k_max (k_max = n/3, the 2/3-rule
    cutoff), with beta = P k_f^2 the enstrophy injection rate of power P at wavenumber k_f:
    nu = (3 / n)^2 beta^(1/3)."""
    beta = power * k_f * k_f
    out = []
    for n, count in members:
        nu = (3.0 / n) ** 2 * beta ** (1.0 / 3.0)
        out += [Config(n=n, dim=2, nu=nu, alpha=drag, force_klo=klo, force_khi=khi,
                       dtype="float32")] * count
    return out
