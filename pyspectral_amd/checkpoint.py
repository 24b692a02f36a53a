"""Checkpoints of a run in the reference's file format (src/pyspectral/validation.py:204-221).

One ``.npz`` file holds the state (complex64 for fp32 flows, complex128 otherwise, NumPy
rfftn layout), the time ``t``, the band-forcing growth rate ``amp`` and the run history as a
JSON string ``hist`` (RunResult.as_dict); extra arrays (e.g. spectra) may ride along.  A file
written here loads with the reference's ``_load_ck`` and the other way round, so chunked
runs (``wall_limit`` + checkpoint, as scripts/validate.sh drives them) can move between the
reference's variants and these.  Loading uses ``allow_pickle=False``.
"""

from __future__ import annotations

import json
from pathlib import Path
from typing import Any

import numpy as np

from pyspectral_amd.simulate import RunResult


def save(path: str | Path, solver: Any, result: RunResult, **extra: Any) -> None:
    p = Path(path)
    p.parent.mkdir(parents=True, exist_ok=True)
    st = solver.get_state()
    st = st.astype(np.complex64) if solver.cfg.dtype == "float32" else st
    np.savez(p, state=st, t=solver.t, amp=solver.force_amp, hist=json.dumps(result.as_dict()),
             **extra)


def load(path: str | Path, solver: Any) -> tuple[RunResult, Any]:
    """Restore state, time and forcing amplitude into ``solver``; return the history and the
    open archive (for extra arrays)."""
    z = np.load(Path(path), allow_pickle=False)
    solver.set_state(z["state"])
    solver.t = float(z["t"])
    solver.set_forcing_amplitude(float(z["amp"]))
    h = json.loads(str(z["hist"]))
    res = RunResult(t=h["t"], energy=h["energy"], enstrophy=h["enstrophy"], dt=h["dt"],
                    steps=h["steps"], wall=h["wall_s"])
    return res, z
