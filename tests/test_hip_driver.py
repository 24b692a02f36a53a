"""Device-resident time loop (psx_run_*) against the oracle's host loop (GPU only).

The oracle runs src/pyspectral/simulate.py's rules on the host in FP64 (oracle.run).  In
fp64 the device controller must take the same steps (count and sizes to round-off), set the
same band growth rates and record the same (t, E, Z); in fp32 the steps agree to the working
precision, and the state is compared with the oracle replaying the device's own step and
amplitude sequence.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import numpy_oracle as O
from pyspectral_amd import initial, simulate
from pyspectral_amd.backends import get
from pyspectral_amd.config import Config

pytestmark = pytest.mark.gpu


def _ic(n, dim, seed=0):
    if dim == 2:
        return initial.to_spectral(initial.random_2d(n, k0=4, seed=seed), 2)
    return initial.to_spectral(initial.random_3d(n, k0=2, seed=seed), 3)


def _cfg(n, dim, dtype, form="basdevant"):
    return Config(n=n, dim=dim, nu=0.02 if dim == 2 else 0.05, alpha=0.1, force_amp=0.0,
                  force_klo=1.0, force_khi=3.0, dtype=dtype, form=form)


@pytest.mark.parametrize("dim,n,dtype,form", [(2, 64, "float64", "basdevant"),
                                              (2, 64, "float32", "basdevant"),
                                              (2, 96, "float64", "basdevant"),
                                              (3, 16, "float64", "basdevant"),
                                              (3, 16, "float64", "rotational"),
                                              (3, 64, "float32", "basdevant")])
def test_device_loop_matches_oracle_controller(dim, n, dtype, form):
    print(f"case driver {dim}D n={n} {dtype} {form}", flush=True)
    cfg = _cfg(n, dim, dtype, form)
    s = get("hip/fused").make(cfg)
    ic = _ic(n, dim)
    s.set_state(ic)
    t_end = 0.25 if dim == 2 else 0.2
    kw = dict(cfl=0.5, dt_max=0.02, cfl_every=3, diag_dt=0.06, target_power=0.05)
    res = simulate.run(s, t_end, device=True, **kw)
    ref = O.OracleSolver(cfg.with_(dtype="float64"))
    ref.set_state(ic)
    diag = O.OracleDiagnostics(n, dim)
    oref = O.run(ref, t_end, diag, klo=cfg.force_klo, khi=cfg.force_khi, **kw)
    print(f"  steps {res.steps} / oracle {oref.steps}; records {len(res.t)}", flush=True)
    assert res.finished and res.steps == oref.steps and len(res.t) == len(oref.t)
    tol_dt = 1e-11 if dtype == "float64" else 2e-5
    np.testing.assert_allclose(res.dt, oref.dt, rtol=tol_dt, atol=0)
    np.testing.assert_allclose(res.amp, oref.amp, rtol=10 * tol_dt, atol=0)
    np.testing.assert_allclose(res.t, oref.t, rtol=tol_dt, atol=0)
    assert res.t[-1] == pytest.approx(t_end, abs=1e-12)
    tol_rec = 1e-11 if dtype == "float64" else 2e-5
    np.testing.assert_allclose(res.energy, oref.energy, rtol=tol_rec)
    np.testing.assert_allclose(res.enstrophy, oref.enstrophy, rtol=tol_rec)
    if dtype == "float64":
        err = O.rel_err(s.get_state(), ref.get_state())
    else:
        # same steps and amplitudes as the device took
        rep = O.OracleSolver(cfg.with_(dtype="float64"))
        rep.set_state(ic)
        O.run(rep, t_end, diag, klo=cfg.force_klo, khi=cfg.force_khi, dt_sequence=res.dt,
              amp_sequence=res.amp, **kw)
        err = O.rel_err(s.get_state(), rep.get_state())
    print(f"  state rel err {err:.3e}", flush=True)
    assert err < (1e-11 if dtype == "float64" else 2e-5), err


def test_fixed_dt_landing_is_exact():
    """Fixed dt: every step size and record time equals the host rule's bit for bit."""
    cfg = _cfg(32, 2, "float64")
    s = get("hip/fused").make(cfg)
    s.set_state(_ic(32, 2))
    res = simulate.run(s, 0.1, dt=0.03, diag_dt=0.025, device=True)
    ref = O.OracleSolver(cfg)
    ref.set_state(_ic(32, 2))
    oref = O.run(ref, 0.1, O.OracleDiagnostics(32, 2), dt=0.03, diag_dt=0.025)
    assert res.dt == oref.dt and res.t == oref.t
    np.testing.assert_allclose(res.energy, oref.energy, rtol=1e-12)
    assert O.rel_err(s.get_state(), ref.get_state()) < 1e-12


def test_graph_chunks_match_eager_chunks():
    cfg = _cfg(128, 2, "float32")
    ic = _ic(128, 2, seed=3)
    out = []
    for use_graph in (False, True):
        s = get("hip/fused").make(cfg)
        s.set_state(ic)
        r = s.run_device(0.2, cfl=0.5, dt_max=0.02, cfl_every=4, diag_dt=0.05,
                         target_power=0.05, use_graph=use_graph, chunk=5)
        out.append((r, s.get_state()))
    (a, sa), (b, sb) = out
    assert a["dt"] == b["dt"] and a["amp"] == b["amp"] and a["energy"] == b["energy"]
    assert np.array_equal(sa, sb)


def test_step_budget_and_eager_steps_after_a_run():
    """kmax stops a run with t_end = inf; eager steps afterwards use fresh parameters."""
    cfg = _cfg(64, 2, "float64")
    s = get("hip/fused").make(cfg)
    ref = O.OracleSolver(cfg)
    ic = _ic(64, 2, seed=5)
    s.set_state(ic)
    ref.set_state(ic)
    r = s.run_device(math.inf, dt=0.004, kmax=23, chunk=10)
    assert r["steps"] == 23 and len(r["dt"]) == 23 and not r["finished"]
    assert s.nstep == 23 and s.t == pytest.approx(23 * 0.004)
    ref.advance(23, 0.004)
    s.advance(4, 0.01)
    ref.advance(4, 0.01)
    assert O.rel_err(s.get_state(), ref.get_state()) < 1e-12


def test_full_size_controller_run_8192():
    """The 8192^2 fp32 workload under the device loop: CFL steps track the host rule
    applied to the device's own maxima, and the state stays finite."""
    n = 8192
    cfg = Config(n=n, dim=2, nu=1e-5, dtype="float32")
    s = get("hip/fused").make(cfg)
    s.set_state(initial.random_phase_2d(n, k0=10.0, seed=0, dtype=np.complex64))
    r = s.run_device(math.inf, cfl=0.5, dt_max=0.05, cfl_every=10, kmax=20)
    assert r["steps"] == 20
    dts = np.array(r["dt"])
    assert np.all(np.isfinite(dts)) and np.all(dts > 0) and np.all(dts <= 0.05)
    # dt is constant between controller updates
    for k in range(20):
        assert dts[k] == dts[k - k % 10]
    umax = s.diag.max_velocity_sum(s.state_native())
    assert math.isfinite(umax)
