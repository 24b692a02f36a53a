"""The reference's validation protocols on the HIP path (GPU only).

- The 2D Taylor-Green vortex on 256^2 in fp64 (nu = 1e-2, dt = 1e-3, IF-RK4 to t = 1):
  against the exact decay exp(-2 nu t) and the FP64 oracle run on the same grid.
- The reference's own TGV case (validation.py:59-68: n = 32, nu = 0.05, dt = 0.01, t = 1) in
  fp64 and fp32, against the published errors (results/validation_2d.json, "tgv2d").
- IF-RK4 temporal self-convergence on a smooth nonlinear flow (validation.py:86-100): order
  near 4 and the published error at every dt (results/validation_2d.json, "temporal").
- Spectral convergence in n (validation.py:103-119) against the published errors at the grid
  sizes of the protocol (8 ... 64, 20 and 40 included; results/validation_2d.json, "spatial").

Maximum relative errors are max|a - b| / max|b| (validation.py:36-38).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import numpy_oracle as O
from pyspectral_amd import initial
from pyspectral_amd.backends import BackendUnavailable, get
from pyspectral_amd.config import Config

pytestmark = pytest.mark.gpu

# results/validation_2d.json of the reference (numpy/scipy-fft FP64 runs)
PUB_TGV = {"float64": 3.06746797486421e-15, "float32": 7.777842007764804e-07}
PUB_TEMPORAL = {0.04: 4.3819985044200083e-07, 0.02: 2.777248871034911e-08,
                0.01: 1.7426140327364963e-09, 0.005: 1.0904312121842099e-10,
                0.0025: 6.8145811136031484e-12}
PUB_TEMPORAL_ORDER = 3.993782286607557
PUB_SPATIAL = {8: 0.09457572679027483, 12: 0.06184973471137069, 16: 0.005421877805026815,
               20: 0.000707740167908263, 24: 0.0006398754099795953, 32: 7.81320123782286e-06,
               40: 7.900165112528533e-07, 48: 7.803592630563717e-08, 64: 6.137047484046012e-11}


def _make(cfg, name="hip/fused"):
    try:
        return get(name).make(cfg)
    except BackendUnavailable as e:  # a GPU machine without the library must fail, not skip
        raise AssertionError(f"{name} unavailable with a GPU present: {e}") from e


def _slope(x, y):
    return float(np.polyfit(np.log(x), np.log(y), 1)[0])


def test_tgv2d_256_fp64_to_t1():
    """The Taylor-Green vortex, 256^2 fp64, 1000 IF-RK4 steps of 1e-3."""
    n, nu, dt, steps = 256, 1e-2, 1e-3, 1000
    cfg = Config(n=n, dim=2, nu=nu, dtype="float64")
    w0 = initial.to_spectral(initial.taylor_green_2d(n), 2)
    s = _make(cfg, "hip/fused-graph")
    s.set_state(w0)
    s.advance(steps, dt)
    got = s.get_state()
    exact = w0 * math.exp(-2 * nu * s.t)
    ref = O.OracleSolver(cfg)
    ref.set_state(w0)
    ref.advance(steps, dt)
    e_exact = O.rel_err(got, exact)
    e_oracle = O.rel_err(got, ref.get_state())
    print(f"  TGV 256^2 fp64 t={s.t:.6f}: vs exact {e_exact:.2e}, vs oracle {e_oracle:.2e}, "
          f"oracle vs exact {O.rel_err(ref.get_state(), exact):.2e}", flush=True)
    assert abs(s.t - 1.0) < 1e-12
    assert e_exact < 1e-13 and e_oracle < 1e-13


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_tgv2d_reference_case(dtype):
    """validation.py:59-68 with its defaults; the FP64 IF factors keep fp32 near the oracle's
    own fp32 error."""
    n, nu, dt = 32, 0.05, 0.01
    s = _make(Config(n=n, dim=2, nu=nu, dtype=dtype))
    w0 = initial.to_spectral(initial.taylor_green_2d(n), 2)
    s.set_state(w0)
    s.advance(100, dt)
    err = O.rel_err(s.get_state(), w0 * math.exp(-2 * nu * s.t))
    print(f"  TGV n=32 {dtype}: {err:.3e} (published numpy: {PUB_TGV[dtype]:.3e})", flush=True)
    assert err < (1e-13 if dtype == "float64" else 5e-6)


def test_temporal_convergence_order():
    """validation.py:86-100: smooth_2d, n = 64, nu = 0.01, t = 1, reference run at dt_min/8."""
    n, nu, t_end = 64, 0.01, 1.0
    w0 = initial.to_spectral(initial.smooth_2d(n), 2)

    def solve(dt):
        s = _make(Config(n=n, dim=2, nu=nu, dtype="float64"), "hip/fused-graph")
        s.set_state(w0)
        s.advance(int(round(t_end / dt)), dt)
        return s.get_state()

    dts = sorted(PUB_TEMPORAL, reverse=True)
    ref = solve(min(dts) / 8)
    errs = [O.rel_err(solve(dt), ref) for dt in dts]
    order = _slope(np.array(dts), np.array(errs))
    dev = [abs(e - PUB_TEMPORAL[dt]) / PUB_TEMPORAL[dt] for dt, e in zip(dts, errs)]
    print(f"  errors {['%.4e' % e for e in errs]}; order {order:.4f} (published "
          f"{PUB_TEMPORAL_ORDER:.4f}); max deviation from published {max(dev):.2e}", flush=True)
    assert abs(order - PUB_TEMPORAL_ORDER) < 0.01
    assert max(dev) < 1e-2


def test_spatial_convergence():
    """validation.py:103-119: smooth_2d, nu = 0.02, t = 0.5, dt = 2.5e-3, reference n = 192;
    the error of each n, after zero-padding its spectrum to 192, against the published one."""
    nu, t_end, dt, n_ref = 0.02, 0.5, 2.5e-3, 192

    def solve(n):
        s = _make(Config(n=n, dim=2, nu=nu, dtype="float64"), "hip/fused-graph")
        s.set_state(initial.to_spectral(initial.smooth_2d(n), 2))
        s.advance(int(round(t_end / dt)), dt)
        return s.get_state()

    ref = solve(n_ref)
    w_ref = np.fft.irfftn(ref, s=(n_ref, n_ref), axes=(0, 1))
    errs = {}
    for n in PUB_SPATIAL:
        w = np.fft.irfftn(O.pad_spectrum(solve(n), n_ref, 2), s=(n_ref, n_ref), axes=(0, 1))
        errs[n] = float(np.abs(w - w_ref).max() / np.abs(w_ref).max())
    dev = {n: abs(errs[n] - PUB_SPATIAL[n]) / PUB_SPATIAL[n] for n in errs}
    print("  " + ", ".join(f"n={n}: {errs[n]:.4e} ({dev[n]:.1e})" for n in errs), flush=True)
    assert max(dev.values()) < 1e-3
    assert errs[64] < 1e-10
