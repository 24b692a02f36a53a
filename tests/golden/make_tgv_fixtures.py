"""Write the 3D Taylor-Green (Re = 1600) comparison data as a small fixture.

Reads two kinds of data files the reference ships (JSON, no code is imported or run):

* ``data/reference/tgv_re1600_reference.json``: the HiOCFD C3.5 512^3 spectral DNS history
  (t, E, dissipation -dE/dt, enstrophy; 2000 samples, t = 0 .. 19.99), the reference's own
  ground truth for ``validation.tgv3d`` (validation.py:47-55, 152-200).
* ``results/validation_tgv3d_n{64,128,192,256}.json``: the reference's published fp32 runs of
  that protocol (cupy-cuda-c variant, CFL 0.5, records every 0.05): t, E and 2 nu Z.

Usage: ``python tests/golden/make_tgv_fixtures.py`` (writes ``tests/golden/tgv3d.npz``).
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "tgv3d.npz"


def main() -> None:
    d = json.loads((REF / "data/reference/tgv_re1600_reference.json").read_text())
    arrays = {"hiocfd_t": np.array(d["t"]), "hiocfd_E": np.array(d["E"]),
              "hiocfd_eps": np.array(d["dissipation"]),
              "hiocfd_enstrophy": np.array(d["enstrophy"]),
              "hiocfd_nu": np.array(d["nu"])}
    for n in (64, 128, 192, 256):
        r = json.loads((REF / f"results/validation_tgv3d_n{n}.json").read_text())
        arrays[f"pub{n}_t"] = np.array(r["t"])
        arrays[f"pub{n}_E"] = np.array(r["energy"])
        arrays[f"pub{n}_eps"] = np.array(r["dissipation_2nuZ"])
        arrays[f"pub{n}_steps"] = np.array(r["steps"])
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
