"""Benchmark: IF-RK4 step throughput of the hip/fused variant on one MI355X.

Headline workload (one "step" = one Lawson IF-RK4 step, 4 stages, 16 real FFTs): 2D freely
decaying turbulence on an 8192^2 grid in fp32 -- initial vorticity with E(k) ~ k^4
exp(-(k/k0)^2), k0 = 10, uniform random phases, total energy 0.5; nu = 1e-5, 2/3 dealiasing,
CFL 0.5 with dt re-evaluated every 10 steps by the device-resident controller.  Inputs are
resident in HBM before the timed region starts.

Printed (one JSON line): Mpts/s per RK4 step; % of the FFT roofline
T_roof = 4 t_stage + B_pointwise / B_copy (src/pyspectral/benchmark.py:437-443) with t_stage
the fastest standalone rocFFT leg on the same batched C2R(2) + R2C(2) transforms (hipFFT
in-place and unnormalised on the padded layout, eagerly and replayed from a graph, and
torch.fft; benchmark.py:145-214) and B_copy the device copy bandwidth measured in the same run;
the reference's benchmark row (benchmark.py:413-444); a roofline object for the dominant
kernel from HIP events around every launch of the timed region; the CPU oracle (scipy.fft,
fp32 and fp64) timed on a bounded sample on this host; and, unless --no-extras, the 3D
Taylor-Green vortex on 384^3 and the forced ensemble timed the same way (``extras``).

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-extras]
                    [--workload single|tgv3d|ensemble|fft-yardstick|les-conv]
"""

from __future__ import annotations

import argparse
import ctypes
import ctypes.util
import hashlib
import json
import math
import os
import sys
import time
from pathlib import Path
from typing import Any, Callable

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
CSRC = ROOT / "pyspectral_amd" / "csrc"

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
REF_MPTS = 547.0             # reference 2D 8192^2 fp32 (cupy-cuda-c/fused-inplace, RTX 3070 L)
REF_MPTS_3D = 310.0          # reference 3D 384^3 fp32 (cupy-cuda-c/fused-inplace-lean, same GPU)
FFT_SPLIT = {(2, "basdevant"): (2, 2), (2, "advective"): (4, 1),
             (3, "basdevant"): (3, 5), (3, "rotational"): (6, 3)}  # config.py:78-92

# Sources each size-specialised kernel family is compiled from: a committed traffic
# measurement is used only while these files are unchanged.
KERNEL_SOURCES = {
    "2d": ["psx_k2d_ct.hip", "psx_k2d_body.h", "psx_fft.h", "psx_fft_ct.h", "psx_geo.h",
           "psx_stage.h", "psx_flow.h", "psx_kernels.h"],
    "3d": ["psx_k3d_ct.hip", "psx_fft.h", "psx_fft_ct.h", "psx_geo.h", "psx_stage.h",
           "psx_flow.h", "psx_kernels.h"],
    "ens": ["psx_kens.hip", "psx_ens.h", "psx_k2d_body.h", "psx_fft.h", "psx_fft_ct.h",
            "psx_geo.h", "psx_stage.h", "psx_flow.h", "psx_kernels.h"],
}


# ----------------------------------------------------------------------------- models
def traffic_model(n: int, dim: int, form: str, dtype: str) -> dict[str, float]:
    """The reference's minimal bytes per step (src/pyspectral/benchmark.py:43-55): every FFT
    reads and writes its data once, plus the pointwise passes."""
    sr = 8 if dtype == "float64" else 4
    sc = 2 * sr
    npts = float(n) ** dim
    nmodes = float(n) ** (dim - 1) * (n // 2 + 1)
    ni, nf = FFT_SPLIT[(dim, form)]
    ncomp = 1 if dim == 2 else 3
    fft = 4 * (ni + nf) * (nmodes * sc + npts * sr)
    phys = 4 * (ni + nf) * npts * sr
    post = (4 * (nf + ncomp + ni) + 7 * ncomp) * nmodes * sc
    return {"fft_bytes": fft, "pointwise_bytes": phys + post, "total_bytes": fft + phys + post,
            "points": npts, "ffts_per_step": 4 * (ni + nf)}


def kernel_bytes(n: int, jm: int, csz: int) -> dict[str, float]:
    """Compulsory HBM bytes per launch of each 2D kernel kind (fp32: csz = 8).

    Work fields hold the n x jm retained half-axis columns; the spectral kernel also reads and
    writes the state and the RK accumulator on the M = (2 jm - 1) jm retained modes:
    stage 0 reads u0 and writes acc, stages 1-2 read u0, acc and write acc, stage 3 reads u0,
    acc and writes u0 -- 2.75 state accesses per mode on average over a step.
    """
    field = float(n) * jm * csz
    m = float(2 * jm - 1) * jm * csz
    return {"line": 4 * field, "outer_fwd": 4 * field, "outer_inv": 4 * field,
            "spec": 4 * field + 2.75 * m}


def kernel_bytes_3d(n: int, jm: int, ldw: int, rot: bool, csz: int) -> dict[str, float]:
    """Compulsory HBM bytes per launch of each 3D kernel kind.

    A work field holds the n^2 (x, y) rows of ldw >= jm retained kz columns.  Per stage:
    the z-line kernel reads the 3 (rotational: 6) velocity/vorticity fields and writes the 5
    (3) product fields; the forward y pass reads and writes the 5 (3) product fields; the
    spectral (x) kernel reads them, writes the 3 (6) next-stage input fields and reads/writes
    the 3-component state and accumulator on the (2 jm - 1)^2 jm retained modes (2.75
    accesses per mode on average over a step, as in 2D); the inverse y pass reads and writes
    the 3 (6) input fields.
    """
    field = float(n) * n * ldw * csz
    m = 3.0 * (2 * jm - 1) ** 2 * jm * csz
    nin, nout = (6, 3) if rot else (3, 5)
    return {"line": (nin + nout) * field, "mid_fwd": 2 * nout * field,
            "spec": (nout + nin) * field + 2.75 * m, "mid_inv": 2 * nin * field}


def source_sha(family: str) -> str:
    h = hashlib.sha256()
    for name in KERNEL_SOURCES[family]:
        h.update(name.encode())
        h.update((CSRC / name).read_bytes())
    return h.hexdigest()[:16]


def measured_traffic(kernel: str, workload: str, family: str,
                     variants: dict[str, int] | None = None) -> tuple[float | None, str]:
    """HBM bytes per launch from a committed rocprofv3 counter summary
    (profiles/traffic_<kernel>_<workload>.json), if it was measured on the current sources
    (and, where the file names the ensemble variants it was measured with, on these)."""
    f = ROOT / "profiles" / f"traffic_{kernel}_{workload}.json"
    if not f.exists():
        return None, f"no counter measurement ({f.name})"
    d = json.loads(f.read_text())
    sha = source_sha(family)
    if d.get("source_sha") != sha:
        return None, (f"{f.name} was measured on kernel sources {d.get('source_sha')}, "
                      f"current {sha}: not used")
    if "variants" in d and d["variants"] != variants:
        return None, (f"{f.name} was measured with variants {d['variants']}, "
                      f"this run's are {variants}: not used")
    return float(d["bytes_per_launch"]), f"{f.name} (sources {sha})"


# ----------------------------------------------------------------------------- machine
def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or platform.machine()


def cpu_cores() -> int:
    """This job's CPU share: OMP_NUM_THREADS when set (the GPU host sets it), else all."""
    n = os.cpu_count() or 1
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(n, env) if env > 0 else n


def copy_bandwidth(torch, nbytes: int = 1 << 30, repeats: int = 10) -> float:
    """Device copy bandwidth, read + write counted (benchmark.py:99-116), GB/s."""
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    ts = []
    for _ in range(repeats):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_time(e1) / 1e3)
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes / float(np.median(ts)) / 1e9


# ----------------------------------------------------------------------------- FFT yardstick
HIPFFT_R2C, HIPFFT_C2R, HIPFFT_D2Z, HIPFFT_Z2D = 0x2A, 0x2C, 0x6A, 0x6C
_HIPFFT: ctypes.CDLL | None = None


def hipfft() -> ctypes.CDLL:
    """hipFFT (rocFFT underneath) through ctypes; only the calls the yardstick needs."""
    global _HIPFFT
    if _HIPFFT is None:
        cands = [os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libhipfft.so"),
                 ctypes.util.find_library("hipfft")]
        path = next((c for c in cands if c and os.path.exists(c)), cands[-1])
        lib = ctypes.CDLL(path)
        c_int, c_void_p, P = ctypes.c_int, ctypes.c_void_p, ctypes.POINTER
        lib.hipfftPlanMany.argtypes = [P(c_void_p), c_int, P(c_int), P(c_int), c_int, c_int,
                                       P(c_int), c_int, c_int, c_int, c_int]
        for f in ("hipfftExecC2R", "hipfftExecR2C", "hipfftExecZ2D", "hipfftExecD2Z"):
            getattr(lib, f).argtypes = [c_void_p, c_void_p, c_void_p]
        lib.hipfftSetStream.argtypes = [c_void_p, c_void_p]
        lib.hipfftDestroy.argtypes = [c_void_p]
        for f in ("hipfftPlanMany", "hipfftExecC2R", "hipfftExecR2C", "hipfftExecZ2D",
                  "hipfftExecD2Z", "hipfftSetStream", "hipfftDestroy"):
            getattr(lib, f).restype = c_int
        _HIPFFT = lib
    return _HIPFFT


def _check_fft(r: int, what: str) -> None:
    if r != 0:
        raise RuntimeError(f"{what} returned hipfftResult {r}")


def _time_events(torch, fn: Callable[[], None], repeats: int, inner: int = 1) -> list[float]:
    """Per-call seconds of `inner` back-to-back calls, `repeats` times, HIP events on the
    current stream (what every leg launches on)."""
    fn()
    torch.cuda.synchronize()
    out = []
    for _ in range(repeats):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(inner):
            fn()
        e1.record()
        e1.synchronize()
        out.append(e0.elapsed_time(e1) / 1e3 / inner)
    return out


def fft_ceiling(torch, n: int, dim: int, form: str, fp64: bool = False,
                repeats: int = 10, flows: int = 1) -> dict[str, Any]:
    """Standalone rocFFT time of one RK stage's batched transforms, the reference's
    fft_ceiling (benchmark.py:145-214) on ROCm: in-place unnormalised hipFFT plans on the
    padded layout (complex n^(d-1) x (n/2+1) <-> real n^(d-1) x (n+2)), eagerly and replayed
    from a graph of 10 stages, and torch.fft (out of place, default normalisation); the
    ceiling is the fastest leg.  ``flows`` > 1 batches the transforms of that many flows."""
    ni, nf = FFT_SPLIT[(dim, form)]
    ni, nf = ni * flows, nf * flows
    nw = max(ni, nf)
    nh = n // 2 + 1
    fsc = n ** (dim - 1) * nh
    out: dict[str, Any] = {"batch_c2r": ni, "batch_r2c": nf}
    cdt = torch.complex128 if fp64 else torch.complex64
    try:
        lib = hipfft()
        W = torch.zeros(nw * fsc, dtype=cdt, device="cuda")
        shape = (ctypes.c_int * dim)(*([n] * dim))
        cemb = (ctypes.c_int * dim)(*([n] * (dim - 1) + [nh]))
        remb = (ctypes.c_int * dim)(*([n] * (dim - 1) + [2 * nh]))
        inv, fwd = ctypes.c_void_p(), ctypes.c_void_p()
        _check_fft(lib.hipfftPlanMany(ctypes.byref(inv), dim, shape, cemb, 1, fsc, remb, 1,
                                      2 * fsc, HIPFFT_Z2D if fp64 else HIPFFT_C2R, ni),
                   "hipfftPlanMany C2R")
        _check_fft(lib.hipfftPlanMany(ctypes.byref(fwd), dim, shape, remb, 1, 2 * fsc, cemb, 1,
                                      fsc, HIPFFT_D2Z if fp64 else HIPFFT_R2C, nf),
                   "hipfftPlanMany R2C")
        ex_inv = lib.hipfftExecZ2D if fp64 else lib.hipfftExecC2R
        ex_fwd = lib.hipfftExecD2Z if fp64 else lib.hipfftExecR2C
        ptr = ctypes.c_void_p(W.data_ptr())

        def bind(stream) -> None:
            for p in (inv, fwd):
                _check_fft(lib.hipfftSetStream(p, ctypes.c_void_p(stream.cuda_stream)),
                           "hipfftSetStream")

        def c2r() -> None:
            _check_fft(ex_inv(inv, ptr, ptr), "hipfftExec C2R")

        def r2c() -> None:
            _check_fft(ex_fwd(fwd, ptr, ptr), "hipfftExec R2C")

        def stage() -> None:
            c2r()
            r2c()

        try:
            bind(torch.cuda.current_stream())
            out["hipfft_inplace_stage_s"] = float(np.median(_time_events(torch, stage,
                                                                         repeats)))
            out["hipfft_c2r_batch_s"] = float(np.median(_time_events(torch, c2r, repeats)))
            out["hipfft_r2c_batch_s"] = float(np.median(_time_events(torch, r2c, repeats)))
            try:
                g = torch.cuda.CUDAGraph()
                side = torch.cuda.Stream()
                side.wait_stream(torch.cuda.current_stream())
                with torch.cuda.stream(side):
                    bind(side)
                    stage()
                torch.cuda.current_stream().wait_stream(side)
                torch.cuda.synchronize()
                with torch.cuda.graph(g):
                    bind(torch.cuda.current_stream())
                    for _ in range(10):
                        stage()
                bind(torch.cuda.current_stream())
                ts = _time_events(torch, g.replay, repeats)
                out["hipfft_inplace_graph_stage_s"] = float(np.median(ts)) / 10
            except Exception as e:  # noqa: BLE001
                out["hipfft_graph_error"] = f"{type(e).__name__}: {e}"
        finally:
            torch.cuda.synchronize()
            lib.hipfftDestroy(inv)
            lib.hipfftDestroy(fwd)
            del W
    except Exception as e:  # noqa: BLE001
        out["hipfft_error"] = f"{type(e).__name__}: {e}"
    try:
        full = (n,) * dim
        half = (n,) * (dim - 1) + (nh,)
        axes = tuple(range(-dim, 0))
        Q = torch.zeros((nw,) + half, dtype=cdt, device="cuda")
        R = torch.zeros((nw,) + full, dtype=torch.float64 if fp64 else torch.float32,
                        device="cuda")

        def stage_t() -> None:
            torch.fft.irfftn(Q[:ni], s=full, dim=axes, out=R[:ni])
            torch.fft.rfftn(R[:nf], dim=axes, out=Q[:nf])

        out["torch_stage_s"] = float(np.median(_time_events(torch, stage_t, repeats)))
        del Q, R
    except Exception as e:  # noqa: BLE001
        out["torch_error"] = f"{type(e).__name__}: {e}"
    torch.cuda.empty_cache()
    legs = ("hipfft_inplace_stage_s", "hipfft_inplace_graph_stage_s", "torch_stage_s")
    cands = {k: out[k] for k in legs if k in out}
    out["best_stage_s"] = min(cands.values()) if cands else None
    out["best_leg"] = min(cands, key=cands.get) if cands else None
    return out


# ----------------------------------------------------------------------------- rows
def bench_row(variant: str, n: int, dim: int, form: str, dtype: str, times_s: list[float],
              inner_steps: int, finite: bool, warmup_s: float, peak_mem_mb: float,
              ceil: dict[str, Any] | None, bcopy_gbs: float | None, notes: str) -> dict:
    """The reference's benchmark row (src/pyspectral/benchmark.py:413-444).  Its fixed
    bandwidths (a measured 361 GB/s copy rate and the 448 GB/s datasheet figure of its GPU)
    become the copy bandwidth measured in this run and the MI355X's 8 TB/s."""
    backend, var = variant.split("/")
    ts = np.array(times_s)
    tm = float(np.median(ts))
    model = traffic_model(n, dim, form, dtype)
    pts = model["points"]
    rate = pts / ts / 1e6
    eff = model["total_bytes"] / tm / 1e9
    row: dict[str, Any] = {
        "backend": backend, "variant": var, "size": [n] * dim, "dim": dim, "dtype": dtype,
        "form": form, "metric": "Mpts/s", "value": float(np.median(rate)),
        "repeats": int(len(ts)), "median": float(np.median(rate)), "min": float(rate.min()),
        "max": float(rate.max()), "step_ms": tm * 1e3, "step_ms_min": float(ts.min() * 1e3),
        "step_ms_max": float(ts.max() * 1e3), "inner_steps": inner_steps,
        "bytes_per_unit": model["total_bytes"] / pts, "effective_GBs": eff,
        "pct_copy_bw": 100 * eff / bcopy_gbs if bcopy_gbs else None,
        "pct_spec_bw": 100 * eff / HBM_PEAK_GBS, "flops_per_unit": None,
        "pct_peak_flops": None, "ffts_per_step": model["ffts_per_step"],
        "fft_Gpts_per_s": model["ffts_per_step"] * pts / tm / 1e9, "finite": finite,
        "warmup_s": warmup_s, "peak_mem_MB": peak_mem_mb, "notes": notes,
    }
    if ceil and ceil.get("best_stage_s") and bcopy_gbs:
        t_fft = 4 * ceil["best_stage_s"]
        t_pw = model["pointwise_bytes"] / (bcopy_gbs * 1e9)
        row.update({"fft_ceiling_step_ms": t_fft * 1e3,
                    "roofline_step_ms": (t_fft + t_pw) * 1e3,
                    "pct_fft_roofline": 100 * (t_fft + t_pw) / tm,
                    "fft_share_of_step": t_fft / tm})
    return row


# ----------------------------------------------------------------------------- CPU baseline
def cpu_baseline(n: int, dim: int, state: np.ndarray, dt: float, nu: float,
                 steps: int = 3) -> dict:
    """The CPU oracle (oracle/numpy_oracle.py: scipy.fft, workers = this job's CPU share) on a
    bounded sample of the same workload: one untimed and `steps` timed IF-RK4 steps of the same
    state and dt, in fp32 (the workload's precision) and in fp64 (the oracle's); median step."""
    from oracle import numpy_oracle as O
    from pyspectral_amd.config import Config

    cores = cpu_cores()
    res: dict[str, Any] = {}
    for dtype in ("float32", "float64"):
        s = O.OracleSolver(Config(n=n, dim=dim, nu=nu, dtype=dtype), workers=cores)
        s.set_state(state)
        s.step(dt)
        ts = []
        for _ in range(steps):
            t0 = time.perf_counter()
            s.step(dt)
            ts.append(time.perf_counter() - t0)
        res[dtype] = float(np.median(ts))
        del s
    t32, t64 = res["float32"], res["float64"]
    return {"value": n ** dim / t32 / 1e6, "unit": "Mpts/s", "cores": cores, "kind": "port",
            "sample": f"1 untimed + {steps} timed IF-RK4 steps of the same {n}^{dim} state and "
                      f"dt with the oracle (scipy.fft, workers={cores}), median step: fp32 "
                      f"{t32:.2f} s, fp64 {t64:.2f} s; value is fp32",
            "s_per_step": t32, "fp64_value": n ** dim / t64 / 1e6, "fp64_s_per_step": t64,
            "cpu": cpu_model()}


# ----------------------------------------------------------------------------- timing
def timed_device_loop(torch, s: Any, state: np.ndarray, steps: int, warmup: int,
                      repeats: int, **run_kw: Any) -> dict[str, Any]:
    """Warm up, then time `steps` steps of the device-resident loop (eager chunks, a HIP event
    around every launch) bracketed by synchronisations: the contract's timed region.  Then
    `repeats - 1` more runs of the same steps from the same state for the row's spread, and the
    same loop replayed from captured chunk graphs."""
    from pyspectral_amd import _lib

    s.run_device(math.inf, kmax=warmup, use_graph=False, **run_kw)
    torch.cuda.synchronize()
    s.set_state(state)
    s.t = 0.0
    lib, ctx, st = s.lib, s.ctx, _lib.stream_ptr()
    _lib.check(lib.psx_timing_begin(ctx))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run = s.run_device(math.inf, kmax=steps, use_graph=False, **run_kw)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nk = len(_lib.KINDS)
    ms = (ctypes.c_double * nk)()
    cnt = (ctypes.c_int * nk)()
    _lib.check(lib.psx_timing_end(ctx, st, ms, cnt))
    per_kernel = {_lib.KINDS[i]: {"ms_total": ms[i], "launches": cnt[i],
                                  "us_per_launch": 1e3 * ms[i] / cnt[i]}
                  for i in range(nk) if cnt[i]}
    times = [wall / steps]
    for _ in range(repeats - 1):
        s.set_state(state)
        s.t = 0.0
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        s.run_device(math.inf, kmax=steps, use_graph=False, **run_kw)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t1) / steps)
    s.set_state(state)
    s.t = 0.0
    s.run_device(math.inf, kmax=min(steps, 20), use_graph=True, **run_kw)
    s.set_state(state)
    s.t = 0.0
    torch.cuda.synchronize()
    tg = time.perf_counter()
    s.run_device(math.inf, kmax=steps, use_graph=True, **run_kw)
    torch.cuda.synchronize()
    return {"t_step": wall / steps, "times": times, "per_kernel": per_kernel,
            "gpu_busy_ms_per_step": sum(ms[i] for i in range(nk)) / steps, "run": run,
            "graph_ms_per_step": 1e3 * (time.perf_counter() - tg) / steps}


def roofline_entry(per_kernel: dict, kb: dict[str, float], traffic: float | None,
                   note: str) -> dict[str, Any]:
    dom = max((k for k in per_kernel if k in kb), key=lambda k: per_kernel[k]["ms_total"])
    dom_s = per_kernel[dom]["ms_total"] / 1e3 / per_kernel[dom]["launches"]
    achieved = kb[dom] / dom_s / 1e9
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_source": note, "algorithmic_bytes_per_launch": kb[dom],
            "avg_launch_us": dom_s * 1e6}


# ----------------------------------------------------------------------------- workloads
# depth 1: the host enqueues exactly the chunks that run (no no-op chunks past the end, whose
# launches would dilute the per-launch averages)
RUN_KW = dict(cfl=0.5, dt_max=0.05, cfl_every=10, chunk=10, record_start=False, depth=1)


def single_bench(a: argparse.Namespace) -> dict[str, Any]:
    import torch

    from pyspectral_amd import initial
    from pyspectral_amd.backends import get
    from pyspectral_amd.config import Config

    n = a.n
    cfg = Config(n=n, dim=2, nu=1e-5, dtype="float32")
    t_setup = time.perf_counter()
    w0 = initial.random_phase_2d(n, k0=10.0, seed=0, energy=0.5, dtype=np.complex64)
    s = get("hip/fused").make(cfg)
    s.set_state(w0)
    lay = s.buf.layout
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t_setup
    torch.cuda.reset_peak_memory_stats()
    tw = time.perf_counter()
    r = timed_device_loop(torch, s, w0, a.steps, a.warmup, a.repeats, **RUN_KW)
    warm_s = time.perf_counter() - tw
    peak_mb = torch.cuda.max_memory_allocated() / 2**20
    t_step, per_kernel, dts = r["t_step"], r["per_kernel"], r["run"]["dt"]
    finite = bool(np.isfinite(s.get_state()).all())
    e_final = s.diag.energy(s.state_native())
    s.close()
    del s
    torch.cuda.empty_cache()

    kb = kernel_bytes(n, lay.jm, 8)
    traffic, tnote = measured_traffic(roof_kernel(per_kernel, kb), f"2d_{n}", "2d")
    roof = roofline_entry(per_kernel, kb, traffic, tnote)
    step_bytes = sum(kb[k] * per_kernel[k]["launches"] for k in per_kernel if k in kb) / a.steps
    tm = traffic_model(n, 2, "basdevant", "float32")
    value = n * n / t_step / 1e6
    out = {
        "metric": "Mpts/s per RK4 step", "value": value, "unit": "Mpts/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": t_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": value / REF_MPTS,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"2D freely decaying turbulence {n}^2 fp32, IF-RK4 Basdevant, "
                               "2/3 dealias, CFL 0.5 (dt every 10 steps, device-resident loop)",
                   "n": n, "nu": 1e-5, "k0": 10, "energy0": 0.5, "cfl": 0.5,
                   "variant": "hip/fused"},
        "roofline": roof, "per_kernel": per_kernel,
        "gpu_busy_ms_per_step": r["gpu_busy_ms_per_step"],
        "step_hbm_bytes_algorithmic": step_bytes, "step_hbm_GBs": step_bytes / t_step / 1e9,
        "reference_model_bytes_per_step": tm["total_bytes"],
        "reference_model_effective_GBs": tm["total_bytes"] / t_step / 1e9,
        "graph_ms_per_step": r["graph_ms_per_step"], "repeat_ms_per_step":
            [1e3 * t for t in r["times"]],
        "dt_first": dts[0], "dt_last": dts[-1], "energy_final": e_final, "finite": finite,
        "setup_s": setup_s, "device": torch.cuda.get_device_name(0),
        "vs_baseline_note": "reference 547 Mpts/s: cupy-cuda-c/fused-inplace, RTX 3070 Laptop",
    }
    ceil, bcopy = None, None
    if not a.no_roofline:
        bcopy = copy_bandwidth(torch)
        ceil = fft_ceiling(torch, n, 2, "basdevant", repeats=20)
        out["fft_ceiling"] = ceil
        out["copy_bandwidth_GBs"] = bcopy
    out["row"] = bench_row("hip/fused", n, 2, "basdevant", "float32", r["times"], a.steps,
                           finite, warm_s, peak_mb, ceil, bcopy,
                           "fused HIP kernels, own Stockham FFTs, device-resident loop")
    for k in ("pct_fft_roofline", "roofline_step_ms", "fft_ceiling_step_ms",
              "fft_share_of_step"):
        if k in out["row"]:
            out[k] = out["row"][k]
    if ceil:
        out["rocfft_stage_ms"] = 1e3 * ceil["best_stage_s"] if ceil["best_stage_s"] else None
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, 2, w0, dts[0], 1e-5)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    return out


def roof_kernel(per_kernel: dict, kb: dict[str, float]) -> str:
    return max((k for k in per_kernel if k in kb), key=lambda k: per_kernel[k]["ms_total"])


def tgv3d_bench(a: argparse.Namespace, n: int, cpu: bool) -> dict[str, Any]:
    """The 3D Taylor-Green vortex at Re = 1600 on n^3 (default 384^3), fp32, Basdevant form
    (8 FFTs per stage), CFL 0.5 with dt every 10 steps on the device."""
    import torch

    from pyspectral_amd import initial
    from pyspectral_amd.backends import get
    from pyspectral_amd.config import Config

    nu = 1.0 / 1600.0
    cfg = Config(n=n, dim=3, nu=nu, dtype="float32")
    w0 = initial.to_spectral(initial.taylor_green_3d(n), 3).astype(np.complex64)
    s = get("hip/fused").make(cfg)
    s.set_state(w0)
    lay = s.buf.layout
    torch.cuda.reset_peak_memory_stats()
    tw = time.perf_counter()
    r = timed_device_loop(torch, s, w0, a.steps, a.warmup, a.repeats, **RUN_KW)
    warm_s = time.perf_counter() - tw
    peak_mb = torch.cuda.max_memory_allocated() / 2**20
    t_step, per_kernel = r["t_step"], r["per_kernel"]
    finite = bool(torch.isfinite(torch.view_as_real(s.state_native())).all().item())
    s.close()
    del s
    torch.cuda.empty_cache()
    kb = kernel_bytes_3d(n, lay.jm, lay.ldw, False, 8)
    traffic, tnote = measured_traffic(roof_kernel(per_kernel, kb), f"3d_{n}", "3d")
    roof = roofline_entry(per_kernel, kb, traffic, tnote)
    step_bytes = sum(kb[k] * per_kernel[k]["launches"] for k in per_kernel if k in kb) / a.steps
    tm = traffic_model(n, 3, "basdevant", "float32")
    value = n ** 3 / t_step / 1e6
    out = {
        "metric": "Mpts/s per RK4 step", "value": value, "unit": "Mpts/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": t_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": value / REF_MPTS_3D,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"3D Taylor-Green vortex Re=1600 {n}^3 fp32, IF-RK4 Basdevant "
                               "(8 FFTs per stage), 2/3 dealias, Leray projection, CFL 0.5 "
                               "(dt every 10 steps, device-resident loop)",
                   "n": n, "nu": nu, "cfl": 0.5, "variant": "hip/fused"},
        "roofline": roof, "per_kernel": per_kernel,
        "step_hbm_bytes_algorithmic": step_bytes, "step_hbm_GBs": step_bytes / t_step / 1e9,
        "reference_model_bytes_per_step": tm["total_bytes"],
        "reference_model_effective_GBs": tm["total_bytes"] / t_step / 1e9,
        "graph_ms_per_step": r["graph_ms_per_step"],
        "repeat_ms_per_step": [1e3 * t for t in r["times"]],
        "dt_first": r["run"]["dt"][0], "finite": finite,
        "vs_baseline_note": "reference 310 Mpts/s: cupy-cuda-c/fused-inplace-lean 384^3 fp32, "
                            "RTX 3070 Laptop",
    }
    ceil, bcopy = None, None
    if not a.no_roofline:
        bcopy = copy_bandwidth(torch)
        ceil = fft_ceiling(torch, n, 3, "basdevant", repeats=10)
        out["fft_ceiling"] = ceil
        out["copy_bandwidth_GBs"] = bcopy
    out["row"] = bench_row("hip/fused", n, 3, "basdevant", "float32", r["times"], a.steps,
                           finite, warm_s, peak_mb, ceil, bcopy,
                           "fused HIP kernels, own Stockham FFTs, device-resident loop")
    for k in ("pct_fft_roofline", "roofline_step_ms", "fft_ceiling_step_ms",
              "fft_share_of_step"):
        if k in out["row"]:
            out[k] = out["row"][k]
    if cpu:
        out["cpu_baseline"] = cpu_baseline(n, 3, w0, r["run"]["dt"][0], nu, steps=2)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    return out


def ensemble_bench(a: argparse.Namespace, cpu: bool) -> dict[str, Any]:
    """The batched forced ensemble (64 x 256^2 + 32 x 512^2 + 16 x 1024^2 + 4 x 2048^2, fp32,
    constant-power band forcing 4 <= |k| <= 6 with P = 1e-3, drag 0.1, nu per size for
    k_max eta = 1, CFL 0.5): autotuned plans, device-resident loop, one launch per phase."""
    import torch

    from pyspectral_amd import _lib, initial
    from pyspectral_amd.ensemble import EnsembleSolver, autotune, forced_ensemble_configs

    cfgs = forced_ensemble_configs()
    npts = sum(c.n * c.n for c in cfgs)
    states = [initial.random_phase_2d(c.n, k0=6.0, seed=1000 + m, dtype=np.complex64)
              for m, c in enumerate(cfgs)]
    t0 = time.perf_counter()
    tuned = autotune(cfgs, nsteps=5)
    tune_s = time.perf_counter() - t0
    ens = EnsembleSolver(cfgs, tuned["variants"], wg_per_cu=tuned["wg_per_cu"])
    ens.set_states(states)
    run_kw = dict(RUN_KW, target_power=1e-3)
    ens.run_device(math.inf, kmax=a.warmup, use_graph=False, **run_kw)
    torch.cuda.synchronize()
    lib, st = ens.lib, _lib.stream_ptr()
    _lib.check(lib.psx_ens_timing_begin(ens.ctx))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = ens.run_device(math.inf, kmax=a.steps, use_graph=False, **run_kw)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / a.steps
    nk = len(_lib.KINDS)
    ms = (ctypes.c_double * nk)()
    cnt = (ctypes.c_int * nk)()
    _lib.check(lib.psx_ens_timing_end(ens.ctx, st, ms, cnt))
    per_kernel = {_lib.KINDS[i]: {"ms_total": ms[i], "launches": cnt[i],
                                  "us_per_launch": 1e3 * ms[i] / cnt[i]}
                  for i in range(nk) if cnt[i]}
    ens.set_states(states)
    ens.t = [0.0] * len(cfgs)
    ens.run_device(math.inf, kmax=20, use_graph=True, **run_kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run_device(math.inf, kmax=a.steps, use_graph=True, **run_kw)
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / a.steps
    finite = all(math.isfinite(e) for o in outs for e in o["energy"] + o["dt"])
    ens.close()
    del ens
    torch.cuda.empty_cache()

    # algorithmic bytes of one launch of each phase: the sum over members
    pb: dict[str, float] = {"line": 0.0, "spec": 0.0, "outer_fwd": 0.0, "outer_inv": 0.0}
    for c, v in zip(cfgs, tuned["variants"]):
        kb = kernel_bytes(c.n, (c.n - 1) // 3 + 1, 8)
        pb["line"] += kb["line"]
        pb["spec"] += kb["spec"]
        if v in (3, 4):  # four-step variants have outer passes
            pb["outer_fwd"] += kb["outer_fwd"]
            pb["outer_inv"] += kb["outer_inv"]
    # the line phase's bytes do not depend on the spectral variants or the grid the autotuner
    # picks, so one counter measurement of the forced ensemble serves every tuning; the
    # spectral phase's figure is used only with the variants it was measured with
    rk = roof_kernel(per_kernel, pb)
    by_size = {str(c.n): int(v) for c, v in zip(cfgs, tuned["variants"])}
    traffic, traffic_src = (measured_traffic(rk, "ens_forced", "ens",
                                             by_size if rk == "spec" else None)
                            if rk in ("line", "spec")
                            else (None, f"not measured for the {rk} phase"))
    roof = roofline_entry(per_kernel, pb, traffic, traffic_src)
    step_bytes = sum(pb[k] * per_kernel[k]["launches"] for k in per_kernel if k in pb) / a.steps
    model_bytes = sum(traffic_model(c.n, 2, "basdevant", "float32")["total_bytes"]
                      for c in cfgs)
    value = npts / t_step / 1e6
    fft_roof: dict[str, Any] = {}
    if not a.no_roofline:
        # the FFT roofline of the whole ensemble: per member size, rocFFT on the batched
        # transforms of all members of that size, plus the reference model's pointwise bytes
        # at the measured copy bandwidth
        bcopy = copy_bandwidth(torch)
        counts: dict[int, int] = {}
        for c in cfgs:
            counts[c.n] = counts.get(c.n, 0) + 1
        t_fft, legs = 0.0, {}
        for nn, cnt in sorted(counts.items()):
            ce = fft_ceiling(torch, nn, 2, "basdevant", repeats=10, flows=cnt)
            legs[str(nn)] = {k: v for k, v in ce.items() if k.endswith("_s") or k == "best_leg"}
            t_fft += 4 * (ce["best_stage_s"] or math.nan)
        t_pw = sum(traffic_model(c.n, 2, "basdevant", "float32")["pointwise_bytes"]
                   for c in cfgs) / (bcopy * 1e9)
        fft_roof = {"pct_fft_roofline": 100 * (t_fft + t_pw) / t_step,
                    "roofline_step_ms": 1e3 * (t_fft + t_pw), "fft_ceiling_step_ms": 1e3 * t_fft,
                    "copy_bandwidth_GBs": bcopy, "fft_ceiling": legs}
    out = {
        "metric": "Mpts/s per RK4 step", "value": value, "unit": "Mpts/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": t_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "forced 2D ensemble: 64 x 256^2 + 32 x 512^2 + 16 x 1024^2 + "
                               "4 x 2048^2 fp32, band 4-6 constant power 1e-3, drag 0.1, "
                               "CFL 0.5, device-resident loop, one launch per phase",
                   "members": len(cfgs), "points": npts,
                   "variants": {str(c.n): v for c, v in zip(cfgs, tuned["variants"])},
                   "wg_per_cu": tuned["wg_per_cu"]},
        "roofline": roof, "per_kernel": per_kernel, "step_hbm_bytes_algorithmic": step_bytes,
        "step_hbm_GBs": step_bytes / t_step / 1e9,
        "reference_model_bytes_per_step": model_bytes,
        "reference_model_effective_GBs": model_bytes / t_step / 1e9,
        "graph_ms_per_step": t_graph * 1e3, "autotune_s": tune_s,
        "autotune": tuned["timings"], "autotune_cached": tuned.get("cached", False),
        "finite": finite, **fft_roof,
    }
    if cpu:
        out["cpu_baseline"] = ensemble_cpu_baseline(cfgs, states, outs)
    return out


def ensemble_cpu_baseline(cfgs, states, outs) -> dict:
    """The CPU oracle on one IF-RK4 step of the first member of each size, scaled by the member
    counts to one ensemble step."""
    from oracle import numpy_oracle as O

    cores = cpu_cores()
    seen: dict[int, int] = {}
    for m, c in enumerate(cfgs):
        seen.setdefault(c.n, m)
    total = 0.0
    per = {}
    for n, m in seen.items():
        s = O.OracleSolver(cfgs[m], workers=cores)
        s.set_state(states[m])
        s.step(outs[m]["dt"][0])
        t0 = time.perf_counter()
        s.step(outs[m]["dt"][0])
        dt_s = time.perf_counter() - t0
        count = sum(1 for c in cfgs if c.n == n)
        per[str(n)] = dt_s
        total += dt_s * count
    npts = sum(c.n * c.n for c in cfgs)
    return {"value": npts / total / 1e6, "unit": "Mpts/s", "cores": cores, "kind": "port",
            "sample": "one timed IF-RK4 step (after one untimed) of the first member of each "
                      "size with the oracle (scipy.fft fp32, workers=%d), times the member "
                      "counts: %s s per member step" % (cores, per), "cpu": cpu_model()}


def les_conv_bench(a: argparse.Namespace) -> dict[str, Any]:
    """The learned closure's kernels on 64^3 (bf16 MFMA convolutions with 32 channels, their
    input and weight gradients, the loss, Adam, a training step at batch 8), and GPU time per
    step of the 64^3 LES (no model, Smagorinsky, the network) and of the 256^3 DNS it is
    compared with, each on the device controller from the same kind of forced state.  Network
    weights are random (the timing does not depend on their values)."""
    import torch

    from pyspectral_amd import _lib, initial
    from pyspectral_amd.backends import get
    from pyspectral_amd.config import Config
    from pyspectral_amd.les import C, NLAYER, TAU_C, LESSolver, SgsNet, conv

    lib, n = _lib.load(), 64
    rng = np.random.default_rng(3)
    flop_pt = 2.0 * 27 * C * C

    def events(fn: Callable[[], None], reps: int) -> float:
        # 10 calls between the events: a single call of a 20-us kernel would also time the
        # host's launch latency
        return float(np.median(_time_events(torch, fn, reps, inner=10)))

    out: dict[str, Any] = {"unit": "us per call", "flop_per_point": flop_pt}
    net = SgsNet(n=n, seed=1)
    res: dict[str, Any] = {}
    for nb in (1, 8):
        P = nb * n ** 3
        x = torch.from_numpy(rng.normal(size=(P, C)).astype(np.float32)).to(torch.bfloat16).cuda()
        y = torch.zeros_like(x)
        msk = torch.ones_like(x)
        t_f = events(lambda: conv(x, net.wb[1], n=n, nb=nb, bias=net.b[1], y=y, relu=True), 20)
        t_b = events(lambda: conv(x, net.wt[1], n=n, nb=nb, mask=msk, y=y), 20)
        res[f"conv_fwd_nb{nb}"] = {"us": 1e6 * t_f, "TFLOPs": flop_pt * P / t_f / 1e12}
        res[f"conv_input_grad_nb{nb}"] = {"us": 1e6 * t_b, "TFLOPs": flop_pt * P / t_b / 1e12}
        if nb == 8:
            dw = torch.zeros(27 * C * C, dtype=torch.float32, device="cuda")
            db = torch.zeros(C, dtype=torch.float32, device="cuda")
            st = _lib.stream_ptr()
            for name, fn in (("wgrad_nb8", lib.psx_cnn_wgrad),
                             ("wgrad_noalloc_nb8", lib.psx_cnn_wgrad_noalloc)):
                t_w = events(lambda: _lib.check(fn(x.data_ptr(), y.data_ptr(), dw.data_ptr(),
                                                   db.data_ptr(), n, nb, st)), 10)
                res[name] = {"us": 1e6 * t_w, "TFLOPs": flop_pt * P / t_w / 1e12}
        del x, y, msk
    X = torch.from_numpy(rng.normal(size=(8 * n ** 3, C)).astype(np.float32)).to(
        torch.bfloat16).cuda()
    T = torch.from_numpy(rng.normal(size=(8 * n ** 3, TAU_C)).astype(np.float32)).cuda()
    res["network_forward_nb1"] = {"us": 1e6 * events(lambda: net.forward(X[:n ** 3], 1), 20)}
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    t_tr = events(lambda: net.train_step(X, T, 8, loss), 5)
    # forward 6 layers, input gradients of 5, weight gradients of 6
    res["train_step_batch8"] = {"us": 1e6 * t_tr,
                                "TFLOPs": 17 * flop_pt * 8 * n ** 3 / t_tr / 1e12}
    out["kernels"] = res
    del X, T

    def device_step_us(s: Any, steps: int) -> tuple[float, float]:
        kw = dict(cfl=0.5, dt_max=0.05, cfl_every=10, chunk=10, record_start=False, depth=1,
                  target_power=0.1)
        s.run_device(math.inf, kmax=10, use_graph=True, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = s.run_device(math.inf, kmax=steps, use_graph=True, **kw)
        torch.cuda.synchronize()
        return 1e6 * (time.perf_counter() - t0) / steps, float(np.mean(r["dt"]))

    nu = 1.6e-3
    steps: dict[str, Any] = {}
    for N, model in ((256, "dns"), (64, "none"), (64, "smagorinsky"), (64, "net")):
        cfg = Config(n=N, dim=3, nu=nu, dtype="float32", force_klo=2.0, force_khi=3.0)
        s = (get("hip/fused").make(cfg) if model == "dns"
             else LESSolver(cfg, model, net=net if model == "net" else None))
        s.set_state(initial.to_spectral(initial.random_3d(N, k0=2.5, seed=7, energy=1.0), 3))
        us, dt = device_step_us(s, 50 if N == 256 else 200)
        steps[f"{model}_{N}"] = {"us_per_step": us, "dt_mean": dt}
        s.close()
        del s
        torch.cuda.empty_cache()
    # GPU time per unit simulated time: each grid at the CFL step of its model-free flow (the
    # random network's stress distorts its own flow; a trained one keeps the LES's dt)
    dns = steps["dns_256"]
    dns["gpu_s_per_unit_time"] = 1e-6 * dns["us_per_step"] / dns["dt_mean"]
    dt_les = steps["none_64"]["dt_mean"]
    for k, v in steps.items():
        if k != "dns_256":
            v["gpu_s_per_unit_time"] = 1e-6 * v["us_per_step"] / dt_les
            v["dns_over_les_gpu_time"] = dns["gpu_s_per_unit_time"] / v["gpu_s_per_unit_time"]
    out["steps"] = steps
    out["note"] = ("GPU time per unit simulated time from one forced random state per grid "
                   "(energy 1, band 2-3, constant power 0.1, CFL 0.5), the LES models at the "
                   "model-free LES's mean step; the closure study (python -m "
                   "pyspectral_amd.closure) measures it on its own flows with the trained "
                   "network")
    out["device"] = torch.cuda.get_device_name(0)
    return out


def yardstick(a: argparse.Namespace) -> dict[str, Any]:
    """Only the standalone rocFFT legs of the 8192^2 (C2R 2 + R2C 2) and 384^3 (C2R 3 + R2C 5)
    stages: the run a kernel trace of the yardstick is taken from."""
    import torch

    return {"fft_ceiling_2d_8192": fft_ceiling(torch, 8192, 2, "basdevant", repeats=5),
            "fft_ceiling_3d_384": fft_ceiling(torch, 384, 3, "basdevant", repeats=5),
            "device": torch.cuda.get_device_name(0)}


def _extra(d: dict[str, Any]) -> dict[str, Any]:
    keys = ("value", "unit", "ms_per_step", "graph_ms_per_step", "repeat_ms_per_step",
            "pct_fft_roofline", "roofline_step_ms", "rocfft_stage_ms", "roofline", "per_kernel",
            "step_hbm_GBs", "config", "finite", "vs_baseline", "autotune_s", "autotune_cached",
            "fft_ceiling_step_ms", "cpu_baseline")
    out = {k: d[k] for k in keys if k in d}
    if "fft_ceiling" in d:
        out["fft_ceiling"] = {k: v for k, v in d["fft_ceiling"].items()
                              if k.endswith("_s") or k == "best_leg" or isinstance(v, dict)}
    return out


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--repeats", type=int, default=5,
                    help="runs of the timed steps for the row's median/min/max (the first is "
                         "the timed region)")
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-roofline", action="store_true", help="skip rocFFT/copy ceilings")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline only: skip the 384^3 and ensemble timings")
    ap.add_argument("--n3", type=int, default=384, help="grid of the tgv3d workload")
    ap.add_argument("--workload", choices=("single", "ensemble", "tgv3d", "fft-yardstick",
                                           "les-conv"),
                    default="single",
                    help="single: one 8192^2 flow (default, with the other two as extras); "
                         "ensemble: the forced ensemble; tgv3d: the 3D Taylor-Green vortex "
                         "on n3^3; fft-yardstick: the rocFFT legs only; les-conv: the "
                         "closure's network kernels and the LES/DNS step times")
    a = ap.parse_args(argv)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the path advances a single flow on one "
              "MI355X (multi-GPU sharding is a design note, not a run); use --gpus 1",
              file=sys.stderr)
        return 2
    if a.workload == "fft-yardstick":
        print(json.dumps(yardstick(a)), flush=True)
        return 0
    if a.workload == "les-conv":
        print(json.dumps(les_conv_bench(a)), flush=True)
        return 0
    if a.workload == "ensemble":
        out = ensemble_bench(a, not a.no_cpu)
    elif a.workload == "tgv3d":
        out = tgv3d_bench(a, a.n3, not a.no_cpu)
    else:
        out = single_bench(a)
        if not a.no_extras:
            out["extras"] = {"tgv3d_384": _extra(tgv3d_bench(a, 384, False)),
                             "ensemble": _extra(ensemble_bench(a, False))}
    print(json.dumps(out), flush=True)
    ok = out["finite"] and all(e.get("finite", True) for e in out.get("extras", {}).values())
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
