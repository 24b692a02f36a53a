"""Ensembles of uneven 2D flows (psx_ens_*) against the oracle and single-flow runs (GPU only).

Every compiled (size, spectral variant) body runs as a member of one mixed ensemble and is
checked against the FP64 oracle on its own inputs (per-member dt and forcing amplitude);
members whose plans equal the single-flow kernels' must agree with them bit for bit; graph
replays equal eager steps; the device time loop of each member takes the same steps as the
single-flow device loop; the autotuner only changes run time.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import numpy_oracle as O
from pyspectral_amd import initial
from pyspectral_amd.backends import get
from pyspectral_amd.config import Config
from pyspectral_amd.ensemble import (EnsembleSolver, autotune, forced_ensemble_configs,
                                     variants_for)

pytestmark = pytest.mark.gpu
SIZES = (64, 128, 256, 512, 1024, 2048)


def _cfg(n, i=0):
    return Config(n=n, dim=2, nu=1e-3 * (1 + 0.1 * i), alpha=0.02, force_amp=0.2,
                  force_klo=2, force_khi=4, dtype="float32")


def _state(n, seed):
    return initial.random_phase_2d(n, k0=min(10.0, n / 12), seed=seed)


def test_every_variant_matches_oracle():
    members = [(n, v) for n in SIZES for v in variants_for(n)]
    print(f"case mixed ensemble of {len(members)} members: {members}", flush=True)
    cfgs = [_cfg(n, i) for i, (n, _) in enumerate(members)]
    ens = EnsembleSolver(cfgs, [v for _, v in members])
    states = [_state(n, 10 + i) for i, (n, _) in enumerate(members)]
    ens.set_states(states)
    dts = [0.2 * (2 * math.pi / n) * (1 + 0.05 * i) for i, (n, _) in enumerate(members)]
    ens.step(dts)
    ens.set_forcing_amplitude([0.1 * (i % 3) for i in range(len(members))])
    ens.step(dts)
    for i, ((n, v), c) in enumerate(zip(members, cfgs)):
        ref = O.OracleSolver(c.with_(dtype="float64"))
        ref.set_state(states[i])
        ref.step(dts[i])
        ref.set_forcing_amplitude(0.1 * (i % 3))
        ref.step(dts[i])
        err = O.rel_err(ens.get_state(i), ref.get_state())
        print(f"  n={n} variant={v} rel err {err:.3e}", flush=True)
        assert err < 5e-6, (n, v, err)
    ens.close()


def test_member_equals_single_flow():
    """A 256^2 member (default variant: one pass, 8-column tiles, 8 elements per thread) and
    the single-flow kernels (16-column tiles, 16 elements per thread) run different radix
    plans, so the states agree to fp32 round-off of the transforms."""
    c = _cfg(256)
    ens = EnsembleSolver([_cfg(64), c, _cfg(512)],
                         [0, variants_for(256)[0], variants_for(512)[0]])
    st = [_state(64, 1), _state(256, 2), _state(512, 3)]
    ens.set_states(st)
    s = get("hip/fused").make(c)
    s.set_state(st[1])
    for k in range(3):
        ens.step([0.01, 0.004, 0.002])
        s.step(0.004)
    err = O.rel_err(ens.get_state(1), s.get_state())
    print(f"  ensemble vs single flow rel err {err:.3e}", flush=True)
    assert err < 2e-6
    ens.close()


def test_graph_replays_match_eager():
    cfgs = [_cfg(128), _cfg(256, 1), _cfg(1024, 2)]
    st = [_state(c.n, 20 + i) for i, c in enumerate(cfgs)]
    a = EnsembleSolver(cfgs, [variants_for(c.n)[-1] for c in cfgs])
    b = EnsembleSolver(cfgs, [variants_for(c.n)[-1] for c in cfgs])
    a.set_states(st)
    b.set_states(st)
    for k in range(4):
        dts = [0.003 * (1 + k % 2)] * 3
        for e in (a, b):
            e.set_forcing_amplitude([0.1 * (k % 2)] * 3)
        a.step(dts)
        b.step(dts, graph=True)
        a.synchronize()
        for m in range(3):
            assert np.array_equal(a.get_state(m), b.get_state(m)), (k, m)
    a.close()
    b.close()


def test_device_loop_matches_single_flow_loops():
    """Each member's device loop takes the steps of its single-flow device loop (same count,
    step sizes, amplitudes and records to fp32 round-off)."""
    sizes = (64, 256, 512)
    cfgs = [Config(n=n, dim=2, nu=2e-3 * (1 + 0.2 * i), alpha=0.1, force_klo=1, force_khi=3,
                   dtype="float32") for i, n in enumerate(sizes)]
    st = [initial.to_spectral(initial.random_2d(c.n, k0=4, seed=30 + i), 2)
          for i, c in enumerate(cfgs)]
    ens = EnsembleSolver(cfgs)
    ens.set_states(st)
    ends = [0.2, 0.15, 0.2]
    kw = dict(cfl=0.5, dt_max=0.02, cfl_every=3, diag_dt=0.05, target_power=0.05)
    outs = ens.run_device(ends, **kw)
    for i, c in enumerate(cfgs):
        s = get("hip/fused").make(c)
        s.set_state(st[i])
        r = s.run_device(ends[i], **kw)
        o = outs[i]
        err = O.rel_err(ens.get_state(i), s.get_state())
        print(f"  member n={c.n}: steps {o['steps']} / {r['steps']}, state rel err {err:.3e}",
              flush=True)
        assert o["steps"] == r["steps"] and len(o["t"]) == len(r["t"])
        # the landing step h = t_end - t inherits the absolute round-off of t
        np.testing.assert_allclose(o["dt"], r["dt"], rtol=1e-5, atol=1e-8)
        np.testing.assert_allclose(o["amp"], r["amp"], rtol=1e-4)
        np.testing.assert_allclose(o["t"], r["t"], rtol=1e-5, atol=1e-8)
        np.testing.assert_allclose(o["energy"], r["energy"], rtol=1e-5)
        assert err < 1e-5
    ens.close()


def test_autotuner_changes_time_only(tmp_path, monkeypatch):
    monkeypatch.setenv("PSX_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    cfgs = [_cfg(256), _cfg(256, 1), _cfg(512, 2), _cfg(1024, 3)]
    tuned = autotune(cfgs, nsteps=2, grids=(1, 2))
    assert not tuned["cached"]
    # the second call reads the on-disk cache: same choices, no timing runs
    again = autotune(cfgs, nsteps=2, grids=(1, 2))
    assert again["cached"] and again["variants"] == tuned["variants"]
    assert again["wg_per_cu"] == tuned["wg_per_cu"]
    # another ensemble shape is tuned afresh
    assert not autotune(cfgs[:2], nsteps=2, grids=(1, 2))["cached"]
    print(f"  autotune: {tuned}", flush=True)
    assert all(v in variants_for(c.n) for v, c in zip(tuned["variants"], cfgs))
    assert tuned["wg_per_cu"] in (1, 2)
    st = [_state(c.n, 40 + i) for i, c in enumerate(cfgs)]
    a = EnsembleSolver(cfgs)
    b = EnsembleSolver(cfgs, tuned["variants"], wg_per_cu=tuned["wg_per_cu"])
    for e in (a, b):
        e.set_states(st)
        for _ in range(3):
            e.step(0.004)
    for m in range(len(cfgs)):
        assert O.rel_err(b.get_state(m), a.get_state(m)) < 2e-6
    a.close()
    b.close()


def test_forced_ensemble_full_size_runs():
    """The 116-member forced ensemble (64 x 256^2, 32 x 512^2, 16 x 1024^2, 4 x 2048^2) under
    the device loop with constant-power forcing: finite, power control engaged."""
    cfgs = forced_ensemble_configs()
    assert len(cfgs) == 116 and sum(c.n * c.n for c in cfgs) == 46137344
    ens = EnsembleSolver(cfgs)
    ens.set_states([initial.random_phase_2d(c.n, k0=6.0, seed=100 + m, dtype=np.complex64)
                    for m, c in enumerate(cfgs)])
    outs = ens.run_device(math.inf, cfl=0.5, dt_max=0.05, cfl_every=10, target_power=1e-3,
                          kmax=20)
    for m, o in enumerate(outs):
        assert o["steps"] == 20
        assert all(0 < h <= 0.05 for h in o["dt"])
        assert all(a > 0 for a in o["amp"])
        assert all(math.isfinite(e) and e > 0 for e in o["energy"])
    ens.close()
