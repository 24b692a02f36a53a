"""Pin the CPU oracle (oracle/numpy_oracle.py) before trusting it.

Three kinds of checks:
* golden vectors written from the reference itself (tests/golden/make_golden.py),
* the reference's own closed forms and invariants (tests/test_reference.py, test_grid.py of
  the reference, restated),
* brute-force DFT evaluation of the nonlinear term on tiny grids, independent of scipy.fft,
plus planted mistakes (in the oracle, on the CPU) that these checks must catch.
"""

from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest

from oracle import numpy_oracle as O
from pyspectral_amd.config import Config

GOLD = Path(__file__).resolve().parent / "golden"


def _npz(name: str) -> dict:
    with np.load(GOLD / name, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


# ----------------------------------------------------------------------------- golden
@pytest.mark.parametrize("dim,form", [(2, "basdevant"), (2, "advective"), (3, "basdevant"),
                                      (3, "rotational")])
@pytest.mark.parametrize("akm", [0, 2])
def test_trajectory_matches_reference(dim, form, akm):
    g = _npz("trajectories.npz")
    key = f"{dim}d_{form}_akm{akm}"
    n = 32 if dim == 2 else 16
    cfg = Config(n=n, dim=dim, nu=0.01, alpha=0.5, alpha_kmax=float(akm), force_amp=0.3,
                 force_klo=1, force_khi=2.5, dtype="float64", form=form)
    s = O.OracleSolver(cfg)
    s.set_state(g[f"ic_{key}"])
    s.advance(5, 0.01)
    s.set_forcing_amplitude(0.1)
    s.advance(3, 0.007)
    assert O.rel_err(s.get_state(), g[f"out_{key}"]) < 1e-13


def test_phase_shift_and_rk4_trajectories():
    g = _npz("trajectories.npz")
    s = O.OracleSolver(Config(n=16, dim=3, nu=0.02, dtype="float64", dealias="phase-shift"))
    s.set_state(g["ic_3d_phase"])
    s.advance(3, 0.01)
    assert O.rel_err(s.get_state(), g["out_3d_phase"]) < 1e-13
    s = O.OracleSolver(Config(n=32, dim=2, nu=0.01, dtype="float64"), integrator="rk4")
    s.set_state(g["ic_2d_rk4"])
    s.advance(4, 0.01)
    assert O.rel_err(s.get_state(), g["out_2d_rk4"]) < 1e-13


@pytest.mark.parametrize("dim,form", [(2, "basdevant"), (2, "advective"), (3, "basdevant"),
                                      (3, "rotational")])
def test_nonlinear_matches_reference(dim, form):
    g = _npz("nonlinear.npz")
    n = 32 if dim == 2 else 16
    s = O.OracleSolver(Config(n=n, dim=dim, nu=0.01, dtype="float64", form=form))
    out = s.nonlinear(s.check_state(g[f"in_{dim}d_{form}"]))
    assert O.rel_err(out, g[f"out_{dim}d_{form}"]) < 1e-13


@pytest.mark.parametrize("dim,n", [(2, 32), (3, 16)])
def test_diagnostics_match_reference(dim, n):
    g = _npz("diagnostics.npz")
    st = g[f"state_{dim}d"]
    d = O.OracleDiagnostics(n, dim)
    got = np.array([d.energy(st), d.enstrophy(st), d.palinstrophy(st),
                    d.band_energy(st, 1.0, 2.5), d.max_velocity_sum(st)])
    np.testing.assert_allclose(got, g[f"scalars_{dim}d"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(d.spectrum(st)[1], g[f"spectrum_{dim}d"], rtol=1e-12,
                               atol=1e-300)


def test_run_driver_matches_reference():
    g = _npz("run_cfl_power.npz")
    cfg = Config(n=32, dim=2, nu=0.005, alpha=0.1, force_amp=0.0, force_klo=2.0, force_khi=4.0,
                 dtype="float64")
    s = O.OracleSolver(cfg)
    s.set_forcing_amplitude(1.0)
    s.set_state(g["ic"])
    res = O.run(s, 0.37, O.OracleDiagnostics(32, 2), cfl=0.5, dt_max=0.05, diag_dt=0.1,
                target_power=0.05, klo=2.0, khi=4.0)
    assert res.steps == int(g["steps"])
    np.testing.assert_allclose(res.dt, g["dt"], rtol=1e-12)
    np.testing.assert_allclose(res.t, g["t"], rtol=1e-12)
    np.testing.assert_allclose(res.energy, g["energy"], rtol=1e-11)
    np.testing.assert_allclose(res.enstrophy, g["enstrophy"], rtol=1e-11)
    assert O.rel_err(s.get_state(), g["final"]) < 1e-11
    assert s.force_amp == pytest.approx(float(g["final_amp"]), rel=1e-11)


def test_initial_conditions_match_reference():
    g = _npz("ic.npz")
    np.testing.assert_allclose(O.random_2d(32, k0=4), g["random_2d_32_k4"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(O.random_2d(64, k0=6), g["random_2d_64_k6"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(O.random_3d(16, k0=2), g["random_3d_16_k2"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(O.random_2d(16, k0=3), g["random_2d_16_k3"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(O.smooth_2d(32), g["smooth_2d_32"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(O.random_2d(48, k0=5.0, seed=5, energy=0.2),
                               g["random_2d_48_seed5"], rtol=0, atol=1e-13)


def test_pad_spectrum_matches_reference():
    g = _npz("pad.npz")
    np.testing.assert_array_equal(O.pad_spectrum(g["w16"], 24, 2), g["w16_to24"])
    np.testing.assert_array_equal(O.pad_spectrum(g["w16"], 10, 2), g["w16_to10"])
    np.testing.assert_array_equal(O.pad_spectrum(g["u12"], 16, 3), g["u12_to16"])


# ----------------------------------------------------------------------------- closed forms
@pytest.mark.parametrize("form", ["basdevant", "advective"])
@pytest.mark.parametrize("dtype,tol", [("float64", 1e-13), ("float32", 2e-6)])
def test_tgv2d_exact_decay(form, dtype, tol):
    s = O.OracleSolver(Config(n=16, dim=2, nu=0.1, dtype=dtype, form=form))
    w0 = O.to_spectral(O.taylor_green_2d(16), 2)
    s.set_state(w0)
    s.advance(20, 0.05)
    assert O.rel_err(s.get_state(), w0 * math.exp(-0.2)) < tol


def test_rk4_is_fourth_order_and_ifrk4_exact():
    errs, errs_if = [], []
    dts = (0.2, 0.1, 0.05, 0.025)
    for dt in dts:
        for integ, store in (("rk4", errs), ("ifrk4", errs_if)):
            s = O.OracleSolver(Config(n=16, dim=2, nu=0.5, dtype="float64"), integrator=integ)
            w0 = O.to_spectral(O.taylor_green_2d(16), 2)
            s.set_state(w0)
            s.advance(int(round(1.0 / dt)), dt)
            store.append(O.rel_err(s.get_state(), w0 * math.exp(-2 * 0.5 * s.t)))
    order = np.polyfit(np.log(dts), np.log(errs), 1)[0]
    assert 3.8 < order < 4.3
    assert max(errs_if) < 1e-13


def test_forms_agree():
    w0 = O.to_spectral(O.smooth_2d(32), 2)
    a = O.OracleSolver(Config(n=32, dim=2, nu=0.01, dtype="float64", form="basdevant"))
    b = O.OracleSolver(Config(n=32, dim=2, nu=0.01, dtype="float64", form="advective"))
    for s in (a, b):
        s.set_state(w0)
        s.advance(5, 0.01)
    assert O.rel_err(a.get_state(), b.get_state()) < 1e-12
    u0 = O.to_spectral(O.random_3d(16, k0=2), 3)
    a = O.OracleSolver(Config(n=16, dim=3, nu=0.01, dtype="float64", form="basdevant"))
    b = O.OracleSolver(Config(n=16, dim=3, nu=0.01, dtype="float64", form="rotational"))
    for s in (a, b):
        s.set_state(u0)
        s.advance(5, 0.01)
    assert O.rel_err(a.get_state(), b.get_state()) < 1e-12


def test_inviscid_energy_and_divergence():
    n = 16
    u0 = O.to_spectral(O.random_3d(n, k0=2), 3)
    s = O.OracleSolver(Config(n=n, dim=3, nu=0.0, dtype="float64"))
    s.set_state(u0)
    d = O.OracleDiagnostics(n, 3)
    e0 = d.energy(u0)
    s.advance(10, 0.002)
    u = s.get_state()
    assert abs(d.energy(u) - e0) / e0 < 1e-9
    kx, ky, kz = O.wavenumbers(n, 3)
    assert np.abs(kx * u[0] + ky * u[1] + kz * u[2]).max() < 1e-10 * np.abs(u).max()


def test_energy_budget_2d():
    n, nu, dt = 32, 0.01, 1e-3
    s = O.OracleSolver(Config(n=n, dim=2, nu=nu, dtype="float64"))
    s.set_state(O.to_spectral(O.smooth_2d(n), 2))
    d = O.OracleDiagnostics(n, 2)
    e, z = [], []
    for _ in range(3):
        e.append(d.energy(s.state))
        z.append(d.enstrophy(s.state))
        s.step(dt)
    dedt = (e[2] - e[0]) / (2 * dt)
    assert abs(dedt + 2 * nu * z[1]) < 1e-6 * abs(dedt)


def test_tgv3d_diagnostics_closed_form():
    d = O.OracleDiagnostics(16, 3)
    u = O.to_spectral(O.taylor_green_3d(16), 3)
    assert d.energy(u) == pytest.approx(0.125, rel=1e-14)
    assert d.enstrophy(u) == pytest.approx(0.375, rel=1e-14)
    assert d.spectrum(u)[1].sum() == pytest.approx(0.125, rel=1e-14)
    assert d.max_velocity_sum(u) == pytest.approx(2.0, rel=1e-14)  # max|u| = max|v| = 1


def test_forcing_band_and_drag_closed_forms():
    n = 32
    cfg = Config(n=n, dim=2, nu=0.0, dtype="float64", force_amp=1.0, force_klo=1, force_khi=1.5)
    s = O.OracleSolver(cfg)
    w0 = O.to_spectral(O.taylor_green_2d(n), 2)
    s.set_state(w0)
    s.advance(10, 0.01)
    np.testing.assert_allclose(s.get_state(), w0 * math.exp(0.1), rtol=1e-12, atol=1e-9)
    s = O.OracleSolver(Config(n=16, dim=2, nu=0.0, alpha=1.0, alpha_kmax=1.5, dtype="float64"))
    L = s.linear_operator(s.k2)
    assert L[1, 1] == -1.0 and L[2, 0] == 0.0


def test_phase_shift_close_to_two_thirds():
    n = 32
    u0 = O.to_spectral(O.random_3d(n, k0=2), 3)
    a = O.OracleSolver(Config(n=n, dim=3, nu=0.02, dtype="float64"))
    b = O.OracleSolver(Config(n=n, dim=3, nu=0.02, dtype="float64", dealias="phase-shift"))
    for s in (a, b):
        s.set_state(u0)
        s.advance(10, 0.01)
    d = O.OracleDiagnostics(n, 3)
    assert d.energy(b.get_state()) == pytest.approx(d.energy(a.get_state()), rel=1e-3)


def test_grid_conventions():
    m = O.dealias_mask(12, 2)
    assert m[3, 3] and not m[4, 0] and not m[0, 4] and m.sum() == 28
    assert O.k_squared(8, 3)[1, 2, 3] == 14
    rng = np.random.default_rng(0)
    f = rng.standard_normal((16, 16))
    fh = np.fft.rfftn(f)
    assert np.isclose(np.sum(O.hermitian_weights(16, 2) * np.abs(fh) ** 2) / 16**4,
                      np.mean(f**2))


def test_traffic_model_matches_published_bytes():
    # benchmark.py:43-55 at 8192^2 fp32: 20.136 GB per step (300 B/pt)
    tm = O.traffic_model(8192, 2, "basdevant", "float32")
    assert tm["total_bytes"] / 1e9 == pytest.approx(20.136, rel=1e-3)
    assert tm["total_bytes"] / tm["points"] == pytest.approx(300.0, rel=2e-3)


# ----------------------------------------------------------------------------- brute force
def _dft2(a: np.ndarray, sign: float) -> np.ndarray:
    n = a.shape[0]
    j = np.arange(n)
    w = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)
    return w @ a @ w.T


def test_nonlinear_2d_brute_force():
    """N = -(u w_x + v w_y) by direct O(n^4) DFTs on an 8x8 grid (no scipy, no r2c)."""
    n = 8
    rng = np.random.default_rng(11)
    x = np.arange(n) * 2 * np.pi / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    w = np.zeros((n, n))
    for (a, b) in ((1, 0), (0, 1), (1, 1), (1, -2), (2, 1)):  # retained modes, 3|k| < 8
        w += rng.standard_normal() * np.cos(a * X + b * Y + rng.uniform(0, 6))
    W = _dft2(w, -1.0)  # full complex spectrum, numpy.fft sign convention
    k = np.fft.fftfreq(n, 1.0 / n)
    KX, KY = np.meshgrid(k, k, indexing="ij")
    K2 = KX**2 + KY**2
    psi = np.where(K2 > 0, W / np.where(K2 > 0, K2, 1), 0)
    u = _dft2(1j * KY * psi, 1.0).real / n**2
    v = _dft2(-1j * KX * psi, 1.0).real / n**2
    wx = _dft2(1j * KX * W, 1.0).real / n**2
    wy = _dft2(1j * KY * W, 1.0).real / n**2
    N = -_dft2(u * wx + v * wy, -1.0)
    mask = (3 * np.abs(KX) < n) & (3 * np.abs(KY) < n)
    expect = (N * mask)[:, : n // 2 + 1]
    for form in ("basdevant", "advective"):
        s = O.OracleSolver(Config(n=n, dim=2, nu=0.0, dtype="float64", form=form))
        got = s.nonlinear(s.check_state(np.fft.rfftn(w)))
        assert np.abs(got - expect).max() < 1e-12 * np.abs(expect).max()


def test_nonlinear_3d_single_mode_pair_closed_form():
    """Two Fourier modes of a solenoidal field: the projected nonlinear term equals
    -P[(u.grad) u], with the gradients taken analytically, on an 8^3 grid."""
    n = 8
    x = np.arange(n) * 2 * np.pi / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = np.stack([np.sin(Y) + 0.3 * np.cos(Z), np.sin(Z) * 0.5, np.cos(X) * 0.7])
    du = [[np.zeros_like(X) for _ in range(3)] for _ in range(3)]  # du[i][j] = d u_i / d x_j
    du[0][1] = np.cos(Y)
    du[0][2] = -0.3 * np.sin(Z)
    du[1][2] = 0.5 * np.cos(Z)
    du[2][0] = -0.7 * np.sin(X)
    adv = np.stack([sum(u[j] * du[i][j] for j in range(3)) for i in range(3)])
    A = np.stack([np.fft.fftn(-adv[i]) for i in range(3)])
    k = np.fft.fftfreq(n, 1.0 / n)
    KX, KY, KZ = np.meshgrid(k, k, k, indexing="ij")
    K2 = KX**2 + KY**2 + KZ**2
    div = (KX * A[0] + KY * A[1] + KZ * A[2]) / np.where(K2 > 0, K2, 1)
    P = np.stack([A[0] - KX * div, A[1] - KY * div, A[2] - KZ * div])
    mask = (3 * np.abs(KX) < n) & (3 * np.abs(KY) < n) & (3 * np.abs(KZ) < n)
    expect = (P * mask)[..., : n // 2 + 1]
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    for form in ("basdevant", "rotational"):
        s = O.OracleSolver(Config(n=n, dim=3, nu=0.0, dtype="float64", form=form))
        got = s.nonlinear(s.check_state(uh))
        assert np.abs(got - expect).max() < 1e-12 * np.abs(expect).max()


# ----------------------------------------------------------------------------- planted mistakes
def test_planted_mistakes_are_caught(monkeypatch):
    """A wrong sign in the nonlinear term or a wrong stage weight must fail the golden check."""
    g = _npz("trajectories.npz")
    cfg = Config(n=32, dim=2, nu=0.01, alpha=0.5, alpha_kmax=0.0, force_amp=0.3, force_klo=1,
                 force_khi=2.5, dtype="float64")
    orig = O.OracleSolver.nonlinear

    def run_once() -> float:
        s = O.OracleSolver(cfg)
        s.set_state(g["ic_2d_basdevant_akm0"])
        s.advance(5, 0.01)
        s.set_forcing_amplitude(0.1)
        s.advance(3, 0.007)
        return O.rel_err(s.get_state(), g["out_2d_basdevant_akm0"])

    assert run_once() < 1e-13
    monkeypatch.setattr(O.OracleSolver, "nonlinear", lambda self, st: -orig(self, st))
    assert run_once() > 1e-6
    monkeypatch.setattr(O.OracleSolver, "nonlinear", orig)
    orig_lin = O.OracleSolver.linear_operator
    monkeypatch.setattr(O.OracleSolver, "linear_operator",
                        lambda self, k2: orig_lin(self, k2) * 1.01)
    assert run_once() > 1e-6


def test_bench_models_match_oracle_and_reference_row():
    """bench.py's own traffic model (the product's measurement code may not import the oracle)
    equals the oracle's restatement of benchmark.py:43-55 for every form, and its benchmark row
    carries every key of the reference's row (benchmark.py:413-444)."""
    import bench

    for dim, form in ((2, "basdevant"), (2, "advective"), (3, "basdevant"), (3, "rotational")):
        for dtype in ("float32", "float64"):
            assert bench.traffic_model(96, dim, form, dtype) == O.traffic_model(96, dim, form,
                                                                                dtype)
    assert bench.traffic_model(8192, 2, "basdevant", "float32")["total_bytes"] == 20135477248.0
    ceil = {"best_stage_s": 3e-3}
    row = bench.bench_row("hip/fused", 64, 2, "basdevant", "float32", [1e-3, 1.1e-3, 0.9e-3],
                          10, True, 0.1, 10.0, ceil, 5000.0, "x")
    ref_keys = {"backend", "variant", "size", "dim", "dtype", "form", "metric", "value",
                "repeats", "median", "min", "max", "step_ms", "step_ms_min", "step_ms_max",
                "inner_steps", "bytes_per_unit", "effective_GBs", "pct_copy_bw", "pct_spec_bw",
                "flops_per_unit", "pct_peak_flops", "ffts_per_step", "fft_Gpts_per_s",
                "finite", "warmup_s", "peak_mem_MB", "notes", "fft_ceiling_step_ms",
                "roofline_step_ms", "pct_fft_roofline", "fft_share_of_step"}
    assert ref_keys <= set(row)
    assert row["step_ms"] == pytest.approx(1.0) and row["repeats"] == 3
    t_roof = 4 * 3e-3 + bench.traffic_model(64, 2, "basdevant", "float32")[
        "pointwise_bytes"] / 5000e9
    assert row["pct_fft_roofline"] == pytest.approx(100 * t_roof / 1e-3)
