"""What the trained subgrid-stress network gains (GPU only): a short version of the closure
study (256^3 forced DNS, 40 filtered snapshots, 200 Adam steps, 6 time units of 64^3 LES),
against the time-averaged filtered-DNS spectrum.  The network must beat the coarse run
without a model on the E(k) error up to the filter cutoff, and train to a validation loss
well below that of predicting zero."""

from __future__ import annotations

import math

import pytest

from pyspectral_amd.closure import closure_study

pytestmark = pytest.mark.gpu


def test_trained_network_beats_no_model():
    print("case closure study (short)", flush=True)
    r = closure_study(t_spin=3.0, n_snap=40, snap_every=0.1, n_val=8, train_steps=200,
                      t_les=6.0, t_avg=3.0, t_cost=1.0, backscatter=(1.0,), mixed=((1.0, 0.04),),
                      log=lambda m: print("  " + m, flush=True))
    les = r["les"]
    assert all(les[m]["finite"] for m in les)
    assert r["training"]["val_loss_after"] < 0.5 * r["training"]["val_loss_before"]
    assert les["net_bs1"]["Ek_rel_err_mean"] < les["none"]["Ek_rel_err_mean"]
    assert les["net_bs1"]["Ek_rel_err_max"] < les["none"]["Ek_rel_err_max"]
    speedup = les["net_bs1"]["speedup_vs_dns"]
    assert math.isfinite(speedup) and speedup > 1
