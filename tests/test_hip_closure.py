"""What the trained subgrid-stress network gains (GPU only): a short version of the closure
study (256^3 forced DNS, 40 filtered snapshots, 200 Adam steps, 6 time units of 64^3 LES),
against the time-averaged filtered-DNS spectrum.  The network must beat the coarse run
without a model on the E(k) error up to the filter cutoff, and train to a validation loss
well below that of predicting zero.  Before that, the a-priori energy-transfer machinery: the
stress source of the LES kernels against the oracle's, shell by shell, and its sum against
the physical-space mean of tau_ij du_i/dx_j; the source kernel's per-shell factor; and the
per-shell factor fitted a posteriori on a window before the reported one."""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from oracle import les_oracle as LO
from pyspectral_amd import initial
from pyspectral_amd.closure import Shells, closure_study, target_source
from pyspectral_amd.config import Config
from pyspectral_amd.les import TAU_C, LESSolver

pytestmark = pytest.mark.gpu

PAIRS = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]


def test_transfer_of_a_stress_matches_oracle_and_parseval():
    print("case a-priori transfer of a stress, L=64", flush=True)
    L = 64
    cfg = Config(n=L, dim=3, nu=1.6e-3, dtype="float32")
    probe = LESSolver(cfg, "none")
    probe.set_state(initial.to_spectral(initial.random_3d(L, k0=6, seed=5), 3))
    uh = probe.get_state()  # masked, fp32-rounded
    g = LO.gradient(uh, L)
    S = 0.5 * (g + np.swapaxes(g, 0, 1))
    rng = np.random.default_rng(1)
    nut = 1.0 + rng.uniform(size=(L, L, L))
    # an eddy-viscosity stress plus noise: mostly draining, not orthogonal to the strain
    tau = np.stack([-2.0 * nut * S[i, j] + 0.3 * rng.normal(size=(L, L, L)) * S.std()
                    for i, j in PAIRS])
    tau_rms = 0.7
    T = np.zeros((L ** 3, TAU_C), dtype=np.float32)
    T[:, :6] = (tau / tau_rms).reshape(6, -1).T
    src = target_source(probe, torch.from_numpy(T).cuda(), tau_rms)
    dev = src[..., : L // 2 + 1].cpu().numpy()
    sh = Shells(L, probe.diag.nshell)
    t_dev = sh.transfer(uh, dev)
    t_ref = sh.transfer(uh, LO.stress_source(tau, L))
    tS = float(np.mean(sum(tau[c] * (g[i, j] + (g[j, i] if i != j else 0.0))
                           for c, (i, j) in enumerate(PAIRS))))
    e_shell = float(np.abs(t_dev - t_ref).max() / np.abs(t_ref).max())
    e_sum = abs(t_dev.sum() - tS) / abs(tS)
    print(f"  per shell {e_shell:.2e}; sum {t_dev.sum():.6e} vs <tau_ij du_i/dx_j> {tS:.6e} "
          f"({e_sum:.2e})", flush=True)
    assert tS < 0
    assert e_shell < 1e-4 and e_sum < 1e-4


def test_source_shell_gain_scales_each_shell():
    print("case per-shell factor on the Smagorinsky source, L=64", flush=True)
    L = 64
    cfg = Config(n=L, dim=3, nu=1.6e-3, dtype="float32")
    st = initial.to_spectral(initial.random_3d(L, k0=6, seed=2), 3)
    plain = LESSolver(cfg, "smagorinsky")
    plain.set_state(st)
    plain.update_source()
    gain = np.random.default_rng(3).uniform(0.3, 2.0, size=40)  # shells past 39 keep 1
    scaled = LESSolver(cfg, "smagorinsky", shell_gain=gain)
    scaled.set_state(st)
    scaled.update_source()
    a = plain.src[..., : L // 2 + 1].cpu().numpy()
    b = scaled.src[..., : L // 2 + 1].cpu().numpy()
    sh = Shells(L, plain.diag.nshell)
    g = np.ones(sh.nshell)
    g[:40] = gain
    expect = a * g[sh.idx].reshape(a.shape[1:])[None].astype(np.float32)
    err = float(np.abs(b - expect).max() / np.abs(a).max())
    print(f"  max deviation {err:.2e}", flush=True)
    assert err < 1e-6


def test_trained_network_beats_no_model():
    print("case closure study (short)", flush=True)
    r = closure_study(t_spin=3.0, n_snap=40, snap_every=0.1, n_val=8, train_steps=200,
                      t_les=6.0, t_avg=3.0, t_cost=1.0, backscatter=(1.0,), mixed=((1.0, 0.04),),
                      fit_models=("net_bs1",), fit_iters=2, fit_t_les=3.0, fit_t_avg=2.0,
                      log=lambda m: print("  " + m, flush=True))
    les = r["les"]
    assert all(les[m]["finite"] for m in les)
    assert r["training"]["val_loss_after"] < 0.5 * r["training"]["val_loss_before"]
    assert les["net_bs1"]["Ek_rel_err_mean"] < les["none"]["Ek_rel_err_mean"]
    assert les["net_bs1"]["Ek_rel_err_max"] < les["none"]["Ek_rel_err_max"]
    speedup = les["net_bs1"]["speedup_vs_dns"]
    assert math.isfinite(speedup) and speedup > 1
    # a priori: the filtered-DNS stress drains the resolved scales on both windows; the
    # per-shell factor fitted on the calibration window, applied by the source kernel, equals
    # the factor times the network's transfer, and lowers the held-out error
    ap = r["apriori"]
    for w in ("calibration", "held_out"):
        assert sum(ap["T"][w]["target"][1:]) < 0
    assert ap["kernel_check"]["net_bs1_shellgain"] < 1e-4
    held = ap["errors"]["held_out"]
    assert held["net_bs1_shellgain"]["rel_err_mean"] < held["net_bs1"]["rel_err_mean"]
    assert les["net_bs1_shellgain"]["finite"]
    # a posteriori: the per-shell factor is fitted on [1, 3] and reported on [3, 6]; the
    # returned iterate is the one with the smallest maximum error on the fit window
    fit = r["aposteriori_fit"]["net_bs1"]
    assert fit["disjoint_from_report"] and len(fit["history"]) == 3
    assert fit["err_max"] == min(h["err_max"] for h in fit["history"])
    assert les["net_bs1_fit"]["finite"]
