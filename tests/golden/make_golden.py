"""Write golden input/output vectors from the pyspectral reference (run in the build container).

The reference source is imported from ``/root/reference/src`` only here; the vectors are
stored as small ``.npz`` fixtures next to this script so that nothing else ever needs the
reference at run time.  Every fixture holds inputs and the reference's outputs only.

Functions covered (the ones no closed form fully pins):

* ``traj_*``      NumpySolver 8-step trajectories under the protocol of
                  ``tests/test_backends.py:24-47`` (5 steps dt=0.01, forcing amplitude change to
                  0.1, 3 steps dt=0.007), every (dim, form) x alpha_kmax in {0, 2}, fp64; plus
                  one phase-shift run.
* ``nonlinear_*`` NumpySolver.nonlinear on a fixed random state (numpy_ref.py:75-119).
* ``diag_*``      Diagnostics outputs (diagnostics.py:65-109) on a fixed state.
* ``run_cfl_power`` simulate.run with CFL + constant power (simulate.py:40-96).
* ``ic_*``        initial.random_2d / random_3d / smooth_2d (initial.py:20-76).
* ``pad_*``       validation.pad_spectrum up/down (validation.py:25-35).

Usage: ``python tests/golden/make_golden.py`` (writes ``tests/golden/*.npz``).
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference/src")
OUT = Path(__file__).resolve().parent


def main() -> None:
    sys.path.insert(0, str(REF))
    from pyspectral import initial
    from pyspectral.backends.numpy_ref import NumpySolver
    from pyspectral.config import Config
    from pyspectral.diagnostics import Diagnostics
    from pyspectral.simulate import run
    from pyspectral.validation import pad_spectrum

    # ---- initial conditions
    np.savez(OUT / "ic.npz",
             random_2d_32_k4=initial.random_2d(32, k0=4),
             random_2d_64_k6=initial.random_2d(64, k0=6),
             random_3d_16_k2=initial.random_3d(16, k0=2),
             random_2d_16_k3=initial.random_2d(16, k0=3),
             smooth_2d_32=initial.smooth_2d(32),
             random_2d_48_seed5=initial.random_2d(48, k0=5.0, seed=5, energy=0.2))

    # ---- trajectories (test_backends.py protocol)
    traj = {}
    for dim, form in ((2, "basdevant"), (2, "advective"), (3, "basdevant"), (3, "rotational")):
        for akm in (0.0, 2.0):
            n = 32 if dim == 2 else 16
            cfg = Config(n=n, dim=dim, nu=0.01, alpha=0.5, alpha_kmax=akm, force_amp=0.3,
                         force_klo=1, force_khi=2.5, dtype="float64", form=form)
            ic = (initial.to_spectral(initial.random_2d(n, k0=4), 2) if dim == 2
                  else initial.to_spectral(initial.random_3d(n, k0=2), 3))
            s = NumpySolver(cfg)
            s.set_state(ic)
            s.advance(5, 0.01)
            s.set_forcing_amplitude(0.1)
            s.advance(3, 0.007)
            key = f"{dim}d_{form}_akm{int(akm)}"
            traj[f"ic_{key}"] = ic
            traj[f"out_{key}"] = s.get_state()
    # phase-shift dealiasing (seeded shifts, numpy_ref.py:125-132)
    cfg = Config(n=16, dim=3, nu=0.02, dtype="float64", dealias="phase-shift")
    ic = initial.to_spectral(initial.random_3d(16, k0=2), 3)
    s = NumpySolver(cfg)
    s.set_state(ic)
    s.advance(3, 0.01)
    traj["ic_3d_phase"] = ic
    traj["out_3d_phase"] = s.get_state()
    # explicit RK4 integrator on a nonlinear 2D flow
    cfg = Config(n=32, dim=2, nu=0.01, dtype="float64")
    ic = initial.to_spectral(initial.smooth_2d(32), 2)
    s = NumpySolver(cfg, integrator="rk4")
    s.set_state(ic)
    s.advance(4, 0.01)
    traj["ic_2d_rk4"] = ic
    traj["out_2d_rk4"] = s.get_state()
    np.savez(OUT / "trajectories.npz", **traj)

    # ---- nonlinear term on a fixed state
    nl = {}
    for dim, form in ((2, "basdevant"), (2, "advective"), (3, "basdevant"), (3, "rotational")):
        n = 32 if dim == 2 else 16
        cfg = Config(n=n, dim=dim, nu=0.01, dtype="float64", form=form)
        s = NumpySolver(cfg)
        ic = (initial.to_spectral(initial.random_2d(n, k0=5, seed=3), 2) if dim == 2
              else initial.to_spectral(initial.random_3d(n, k0=3, seed=3), 3))
        nl[f"in_{dim}d_{form}"] = ic
        nl[f"out_{dim}d_{form}"] = s.nonlinear(s.check_state(ic))
    np.savez(OUT / "nonlinear.npz", **nl)

    # ---- diagnostics
    dg = {}
    for dim, n in ((2, 32), (3, 16)):
        st = (initial.to_spectral(initial.random_2d(n, k0=4, seed=7), 2) if dim == 2
              else initial.to_spectral(initial.random_3d(n, k0=2, seed=7), 3))
        d = Diagnostics(n, dim)
        k, e = d.spectrum(st)
        dg[f"state_{dim}d"] = st
        dg[f"scalars_{dim}d"] = np.array([d.energy(st), d.enstrophy(st), d.palinstrophy(st),
                                          d.band_energy(st, 1.0, 2.5), d.max_velocity_sum(st)])
        dg[f"spectrum_{dim}d"] = e
    np.savez(OUT / "diagnostics.npz", **dg)

    # ---- run(): CFL-adaptive dt + constant power, landing on t_end and diag times
    cfg = Config(n=32, dim=2, nu=0.005, alpha=0.1, force_amp=0.0, force_klo=2.0, force_khi=4.0,
                 dtype="float64")
    s = NumpySolver(cfg)
    s.set_forcing_amplitude(1.0)
    ic = initial.to_spectral(initial.random_2d(32, k0=3, seed=2, energy=0.1), 2)
    s.set_state(ic)
    res = run(s, 0.37, cfl=0.5, dt_max=0.05, diag_dt=0.1, target_power=0.05)
    np.savez(OUT / "run_cfl_power.npz", ic=ic, t=np.array(res.t), energy=np.array(res.energy),
             enstrophy=np.array(res.enstrophy), dt=np.array(res.dt), steps=res.steps,
             final=s.get_state(), final_amp=s.force_amp)

    # ---- pad_spectrum
    w = initial.to_spectral(initial.random_2d(16, k0=3, seed=4), 2)
    u = initial.to_spectral(initial.random_3d(12, k0=2, seed=4), 3)
    np.savez(OUT / "pad.npz", w16=w, w16_to24=pad_spectrum(w, 24, 2),
             w16_to10=pad_spectrum(w, 10, 2), u12=u, u12_to16=pad_spectrum(u, 16, 3))
    print("golden vectors written to", OUT)


if __name__ == "__main__":
    main()
