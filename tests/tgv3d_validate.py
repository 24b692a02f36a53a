"""3D Taylor-Green vortex at Re = 1600 on the HIP path: dissipation histories, the FP64
oracle at 64^3, self-convergence in n and the 512^3 spectral reference.

    python tests/tgv3d_validate.py [--sizes 64,128,192,256,384] [--t-end 10] [--out FILE]

Protocol of the reference's ``validation.tgv3d`` (validation.py:152-200): u = sin x cos y cos z,
v = -cos x sin y cos z, w = 0, nu = 1/1600, Basdevant form, CFL 0.5 (dt every 10 steps,
dt_max 0.05), E and Z recorded every 0.05; dissipation 2 nu Z.  Here the time loop runs on
the device (run_device).  Reported per grid (fp32):

- the peak of 2 nu Z and its time; max |eps - eps_512| / peak_512 over t <= t_end against the
  HiOCFD 512^3 history (tests/golden/tgv3d.npz);
- the same against the reference's published fp32 run on the same grid (n <= 256);
- self-convergence: max |eps_n - eps_nmax| / peak over t <= t_end.

At 64^3 the fp64 run is replayed by the FP64 oracle (oracle/numpy_oracle.py) with the
device's own dt sequence, and the E and 2 nu Z histories and the final state are compared.
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

NU = 1.0 / 1600.0
GOLDEN = ROOT / "tests" / "golden" / "tgv3d.npz"


def gpu_run(n: int, dtype: str, t_end: float, form: str = "basdevant") -> dict:
    import torch

    from pyspectral_amd import initial
    from pyspectral_amd.backends import get
    from pyspectral_amd.config import Config

    s = get("hip/fused").make(Config(n=n, dim=3, nu=NU, dtype=dtype, form=form))
    s.set_state(initial.to_spectral(initial.taylor_green_3d(n), 3))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = s.run_device(t_end, cfl=0.5, dt_max=0.05, cfl_every=10, diag_dt=0.05)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = {"n": n, "dtype": dtype, "form": form, "steps": r["steps"], "wall_s": wall,
           "ms_per_step": 1e3 * wall / max(1, r["steps"]), "t": r["t"], "E": r["energy"],
           "eps": [2 * NU * z for z in r["enstrophy"]], "dt": r["dt"],
           "finished": r["finished"]}
    if n == 64:
        out["state"] = s.get_state()
    s.close()
    return out


def peak(t: list[float], eps: list[float]) -> tuple[float, float]:
    i = int(np.argmax(eps))
    return float(eps[i]), float(t[i])


def oracle_replay(run: dict) -> dict:
    """The FP64 oracle on the same initial state with the device's dt sequence."""
    from oracle import numpy_oracle as O
    from pyspectral_amd import initial
    from pyspectral_amd.config import Config

    n = run["n"]
    s = O.OracleSolver(Config(n=n, dim=3, nu=NU, dtype="float64"))
    s.set_state(initial.to_spectral(initial.taylor_green_3d(n), 3))
    d = O.OracleDiagnostics(n, 3)
    rec_t = np.array(run["t"])
    E, eps = [d.energy(s.state)], [2 * NU * d.enstrophy(s.state)]
    t = 0.0
    j = 1
    for h in run["dt"]:
        s.step(h)
        t += h
        if j < len(rec_t) and abs(t - rec_t[j]) < 1e-9 * max(1.0, t):
            E.append(d.energy(s.state))
            eps.append(2 * NU * d.enstrophy(s.state))
            j += 1
    if j != len(rec_t):
        raise RuntimeError(f"oracle replay met {j} of {len(rec_t)} record times")
    E, eps = np.array(E), np.array(eps)
    return {"E_rel_err": float(np.abs(np.array(run["E"]) - E).max() / np.abs(E).max()),
            "eps_rel_err": float(np.abs(np.array(run["eps"]) - eps).max() / np.abs(eps).max()),
            "state_rel_err": O.rel_err(run["state"], s.get_state()),
            "oracle_peak": peak(list(rec_t), list(eps))}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default="64,128,192,256,384")
    ap.add_argument("--t-end", type=float, default=10.0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--no-oracle", action="store_true")
    a = ap.parse_args()
    sizes = [int(x) for x in a.sizes.split(",")]
    g = np.load(GOLDEN)
    tr, er = g["hiocfd_t"], g["hiocfd_eps"]
    p512 = float(er.max())
    res: dict = {"t_end": a.t_end, "nu": NU, "hiocfd_peak": [p512, float(tr[np.argmax(er)])],
                 "runs": {}}
    runs = {}
    for n in sizes:
        print(f"run {n}^3 fp32 to t={a.t_end}", flush=True)
        r = gpu_run(n, "float32", a.t_end)
        runs[n] = r
        t, eps = np.array(r["t"]), np.array(r["eps"])
        sel = t <= tr.max()
        row = {"steps": r["steps"], "wall_s": r["wall_s"], "ms_per_step": r["ms_per_step"],
               "finished": r["finished"], "peak": peak(r["t"], r["eps"]),
               "err_vs_512_over_peak": float(np.abs(eps[sel] - np.interp(t[sel], tr, er)).max()
                                             / p512)}
        if f"pub{n}_t" in g:
            tp, ep = g[f"pub{n}_t"], g[f"pub{n}_eps"]
            sp = tp <= a.t_end + 1e-9
            row["err_vs_published_over_peak"] = float(
                np.abs(np.interp(tp[sp], t, eps) - ep[sp]).max() / ep[sp].max())
            row["published_peak"] = peak(list(tp[sp]), list(ep[sp]))
        res["runs"][str(n)] = row
        print(f"  {json.dumps(row)}", flush=True)
    nmax = max(sizes)
    tm, em = np.array(runs[nmax]["t"]), np.array(runs[nmax]["eps"])
    for n in sizes:
        if n != nmax:
            e = np.interp(tm, np.array(runs[n]["t"]), np.array(runs[n]["eps"]))
            res["runs"][str(n)]["err_vs_nmax_over_peak"] = float(np.abs(e - em).max() / em.max())
    if 64 in sizes and not a.no_oracle:
        print("run 64^3 fp64 and the FP64 oracle replay", flush=True)
        r64 = gpu_run(64, "float64", a.t_end)
        t0 = time.perf_counter()
        cmp = oracle_replay(r64)
        cmp["oracle_s"] = time.perf_counter() - t0
        cmp["steps"] = r64["steps"]
        cmp["gpu_peak"] = peak(r64["t"], r64["eps"])
        e32 = np.array(runs[64]["eps"])
        cmp["fp32_vs_fp64_eps_rel_err"] = float(np.abs(e32 - np.array(r64["eps"])).max() /
                                                np.abs(np.array(r64["eps"])).max()) \
            if len(e32) == len(r64["eps"]) else math.nan
        res["fp64_64_vs_oracle"] = cmp
        print(f"  {json.dumps(cmp)}", flush=True)
    txt = json.dumps(res)
    if a.out:
        Path(a.out).parent.mkdir(parents=True, exist_ok=True)
        Path(a.out).write_text(json.dumps({"summary": res, "histories": {
            str(n): {"t": runs[n]["t"], "eps": runs[n]["eps"], "E": runs[n]["E"]}
            for n in runs}}))
    print(txt, flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
