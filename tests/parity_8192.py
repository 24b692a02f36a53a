"""Parity of the 8192^2 fp32 workload after 100 IF-RK4 steps against the FP64 oracle.

The workload: 2D freely decaying turbulence on 8192^2, E(k) ~ k^4 exp(-(k/k0)^2), k0 = 10,
uniform random phases (seed 0), E = 0.5; nu = 1e-5, 2/3 dealiasing, CFL 0.5 with dt
re-evaluated every 10 steps by the device controller.  One FP64 oracle step takes seconds to
tens of seconds on a CPU, so the check runs in two stages:

    python tests/parity_8192.py gpu --part 0 --out DIR   # on the GPU: run, save rows
    python tests/parity_8192.py gpu --part 1 --out DIR   # (each part in a run of its own)
    python tests/parity_8192.py gpu --part 2 --out DIR
    python tests/parity_8192.py oracle --inp DIR         # on a CPU: replay dt, compare

The GPU part runs the hip/fused solver with the device-resident controller and saves the dt
sequence it took, the FP64 energy/enstrophy of its final state, a full-state checksum and one
third (``--part``) of the retained spectral rows (complex64, the solver's own values; a
third, 40 MB, fits the size limit of one run's result folder).  The oracle stage integrates
the same masked initial state in FP64 with that dt sequence (the inputs are identical) and
reports, over the retained modes of the rows it has, max|a - b| / max|b| (the reference's
rel_err, validation.py:38-40) and the relative errors of E and Z; it keeps the oracle's final
retained modes beside the GPU files, so parts that arrive later are compared without a rerun.  The GPU run is deterministic: the three parts must
carry the same dt sequence and checksum.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

N, NU, K0, SEED, STEPS, PARTS = 8192, 1e-5, 10.0, 0, 100, 3


def retained_rows(n: int) -> np.ndarray:
    jm = (n - 1) // 3 + 1
    return np.concatenate([np.arange(jm), np.arange(n - jm + 1, n)])


def initial_state() -> np.ndarray:
    from pyspectral_amd import initial

    return initial.random_phase_2d(N, k0=K0, seed=SEED, energy=0.5, dtype=np.complex64)


def gpu_part(part: int, out: Path, steps: int) -> None:
    import torch

    from pyspectral_amd.backends import get
    from pyspectral_amd.config import Config

    out.mkdir(parents=True, exist_ok=True)
    print(f"gpu part {part}: {N}^2 fp32, {steps} steps", flush=True)
    s = get("hip/fused").make(Config(n=N, dim=2, nu=NU, dtype="float32"))
    s.set_state(initial_state())
    r = s.run_device(math.inf, cfl=0.5, dt_max=0.05, cfl_every=10, kmax=steps,
                     record_start=False)
    torch.cuda.synchronize()
    sn = s.state_native()
    e, z = s.diag.energy(sn), s.diag.enstrophy(sn)
    st = s.get_state()  # complex128 copy of the fp32 values
    rows = retained_rows(N)[part::PARTS]
    jm = (N - 1) // 3 + 1
    np.save(out / f"rows_{part}.npy", st[rows, :jm].astype(np.complex64))
    meta = {"dt": r["dt"], "t": s.t, "energy": e, "enstrophy": z,
            "checksum": [float(np.sum(np.abs(st) ** 2)),
                         float(np.sum(st.real * np.arange(st.shape[1])))],
            "finite": bool(np.isfinite(st).all())}
    (out / f"meta_{part}.json").write_text(json.dumps(meta))
    print(f"gpu part {part}: t={s.t:.6g}, E={e:.12g}, Z={z:.12g}, rows {len(rows)}", flush=True)


def oracle_part(inp: Path, workers: int) -> dict:
    from oracle import numpy_oracle as O
    from pyspectral_amd.config import Config

    metas = [json.loads((inp / f"meta_{p}.json").read_text())
             for p in range(PARTS) if (inp / f"meta_{p}.json").exists()]
    parts = [p for p in range(PARTS) if (inp / f"rows_{p}.npy").exists()]
    if not metas or not parts:
        raise SystemExit(f"no GPU results in {inp}")
    m0 = metas[0]
    for m in metas[1:]:
        if m["dt"] != m0["dt"] or m["checksum"] != m0["checksum"]:
            raise SystemExit("the GPU parts disagree: the run was not deterministic")
    dts = m0["dt"]
    rows = retained_rows(N)
    jm = (N - 1) // 3 + 1
    cache = inp / "oracle_final.npz"  # the oracle's final retained modes, E, Z and its dts
    t0 = time.perf_counter()
    if cache.exists() and list(np.load(cache)["dt"]) == dts:
        z = np.load(cache)
        ref_rows, e_ref, z_ref = z["rows"], float(z["energy"]), float(z["enstrophy"])
        s_per_step = float(z["s_per_step"])
    else:
        s = O.OracleSolver(Config(n=N, dim=2, nu=NU, dtype="float64"), workers=workers)
        s.set_state(initial_state())
        for k, h in enumerate(dts):
            s.step(h)
            if k % 10 == 0 or k == len(dts) - 1:
                print(f"oracle step {k + 1}/{len(dts)} ({time.perf_counter() - t0:.0f} s)",
                      flush=True)
        s_per_step = (time.perf_counter() - t0) / len(dts)
        ref = s.get_state()
        d = O.OracleDiagnostics(N, 2)
        e_ref, z_ref = d.energy(ref), d.enstrophy(ref)
        ref_rows = ref[rows, :jm]
        np.savez(cache, rows=ref_rows, energy=e_ref, enstrophy=z_ref, dt=np.array(dts),
                 s_per_step=s_per_step)
    bmax = float(np.abs(ref_rows).max())
    emax, nmodes = 0.0, 0
    for p in parts:
        got = np.load(inp / f"rows_{p}.npy").astype(np.complex128)
        b = ref_rows[p::PARTS]
        emax = max(emax, float(np.abs(got - b).max()))
        nmodes += got.size
    res = {"n": N, "steps": len(dts), "t": float(sum(dts)), "dt_first": dts[0],
           "dt_last": dts[-1], "modes_compared": nmodes,
           "modes_retained": int(len(rows) * jm), "rel_err": emax / bmax,
           "energy_rel_err": abs(m0["energy"] - e_ref) / e_ref,
           "enstrophy_rel_err": abs(m0["enstrophy"] - z_ref) / z_ref,
           "energy_oracle": e_ref, "enstrophy_oracle": z_ref,
           "oracle_s_per_step": s_per_step, "oracle_workers": workers}
    print(json.dumps(res), flush=True)
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=("gpu", "oracle"))
    ap.add_argument("--part", type=int, default=0, choices=range(PARTS))
    ap.add_argument("--out", default="parity8192_out", help="where the GPU parts are written")
    ap.add_argument("--inp", default="parity8192_out", help="where the oracle reads them")
    ap.add_argument("--steps", type=int, default=STEPS)
    ap.add_argument("--workers", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    if a.mode == "gpu":
        gpu_part(a.part, Path(a.out), a.steps)
        return 0
    res = oracle_part(Path(a.inp), a.workers)
    return 0 if res["rel_err"] < 1e-5 else 1


if __name__ == "__main__":
    raise SystemExit(main())
