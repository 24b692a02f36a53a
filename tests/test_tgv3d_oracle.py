"""The oracle's 3D Taylor-Green vortex (Re = 1600) against data the reference ships (CPU).

tests/golden/tgv3d.npz holds the reference's published fp32 run on 64^3
(results/validation_tgv3d_n64.json) and the HiOCFD 512^3 spectral history
(data/reference/tgv_re1600_reference.json), written by tests/golden/make_tgv_fixtures.py.
Up to t = 1 the flow is resolved on 64^3, so the oracle's FP64 run must agree with the
published 2 nu Z and E to the fp32 noise of that run, and with the 512^3 dissipation to the
accuracy of its -dE/dt column (about 3e-5 relative, as the data file's notes state).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from oracle import numpy_oracle as O
from pyspectral_amd.config import Config

G = np.load(Path(__file__).resolve().parent / "golden" / "tgv3d.npz")
NU = 1.0 / 1600.0


def test_fixture_closed_forms():
    """E(0) = 1/8 and 2 nu Z(0) = 2 nu 3/8 for this initial condition (both data sets)."""
    assert abs(G["hiocfd_E"][0] - 0.125) < 1e-9
    assert abs(G["pub64_E"][0] - 0.125) < 1e-12
    assert abs(G["pub64_eps"][0] - 2 * NU * 0.375) < 1e-12
    assert abs(float(G["hiocfd_nu"]) - NU) < 1e-15


def test_oracle_tgv3d_64_matches_published_history():
    n, dt = 64, 0.02
    s = O.OracleSolver(Config(n=n, dim=3, nu=NU, dtype="float64"))
    s.set_state(O.to_spectral(O.taylor_green_3d(n), 3))
    d = O.OracleDiagnostics(n, 3)
    ts, eps, E = [0.0], [2 * NU * d.enstrophy(s.state)], [d.energy(s.state)]
    for k in range(1, 51):
        s.step(dt)
        if k % 5 == 0:
            ts.append(k * dt)
            eps.append(2 * NU * d.enstrophy(s.state))
            E.append(d.energy(s.state))
    ts, eps, E = np.array(ts), np.array(eps), np.array(E)
    pub_eps = np.interp(ts, G["pub64_t"], G["pub64_eps"])
    pub_E = np.interp(ts, G["pub64_t"], G["pub64_E"])
    hi = np.interp(ts, G["hiocfd_t"], G["hiocfd_eps"])
    assert np.abs(eps - pub_eps).max() / pub_eps.max() < 2e-6
    assert np.abs(E - pub_E).max() / pub_E.max() < 2e-6
    assert np.abs(eps - hi).max() / hi.max() < 5e-5
    # a planted mistake fails the check: 1 % more viscosity
    s2 = O.OracleSolver(Config(n=n, dim=3, nu=1.01 * NU, dtype="float64"))
    s2.set_state(O.to_spectral(O.taylor_green_3d(n), 3))
    s2.advance(50, dt)
    assert abs(2 * NU * d.enstrophy(s2.state) - pub_eps[-1]) / pub_eps.max() > 2e-6
