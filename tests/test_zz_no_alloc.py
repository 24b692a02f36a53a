"""psx_cnn_wgrad_noalloc (static per-GPU partial-sum array, no allocation) against
psx_cnn_wgrad (partial sums in a stream-ordered allocation of its own), through the C ABI.

Both launch the same two kernels with the same 512 partial rows and the same fixed-order
reduction, so their results must be equal bit for bit: two calls in a row, the second on a
smaller input (the static array keeps the first call's partial sums, which the second must
overwrite completely), and a third with both entry points accumulating onto non-zero dw, db.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from pyspectral_amd import _lib

pytestmark = pytest.mark.gpu
C, N = 32, 64


def _bf16(rng: np.random.Generator, shape: tuple[int, ...], scale: float = 1.0) -> torch.Tensor:
    return torch.from_numpy((scale * rng.normal(size=shape)).astype(np.float32)).to(
        torch.bfloat16).cuda()


def _wgrad(fn, x: torch.Tensor, dz: torch.Tensor, nb: int, dw: torch.Tensor,
           db: torch.Tensor) -> None:
    _lib.check(fn(x.data_ptr(), dz.data_ptr(), dw.data_ptr(), db.data_ptr(), N, nb,
                  _lib.stream_ptr()))


def test_noalloc_wgrad_matches_allocating_entry_point():
    lib = _lib.load()
    assert lib.psx_cnn_wgrad_scratch_bytes() == 4 * 512 * (27 * C * C + C)
    rng = np.random.default_rng(21)
    for call, nb in enumerate((2, 1)):  # the second call is smaller
        print(f"case wgrad noalloc call {call} nb={nb}", flush=True)
        P = nb * N ** 3
        x, dz = _bf16(rng, (P, C)), _bf16(rng, (P, C), 1e-3)
        out = []
        for fn in (lib.psx_cnn_wgrad, lib.psx_cnn_wgrad_noalloc):
            dw = torch.zeros(27 * C * C, dtype=torch.float32, device="cuda")
            db = torch.zeros(C, dtype=torch.float32, device="cuda")
            _wgrad(fn, x, dz, nb, dw, db)
            torch.cuda.synchronize()
            out.append((dw.cpu(), db.cpu()))
        (dw_a, db_a), (dw_b, db_b) = out
        assert torch.isfinite(dw_a).all() and float(dw_a.abs().max()) > 0
        assert torch.equal(dw_a, dw_b) and torch.equal(db_a, db_b), call
    # both accumulate onto what dw and db hold
    dw0 = torch.from_numpy(rng.normal(size=27 * C * C).astype(np.float32)).cuda()
    db0 = torch.from_numpy(rng.normal(size=C).astype(np.float32)).cuda()
    res = []
    for fn in (lib.psx_cnn_wgrad, lib.psx_cnn_wgrad_noalloc):
        dw, db = dw0.clone(), db0.clone()
        _wgrad(fn, x, dz, 1, dw, db)
        torch.cuda.synchronize()
        res.append((dw.cpu(), db.cpu()))
    assert torch.equal(res[0][0], res[1][0]) and torch.equal(res[0][1], res[1][1])


def test_noalloc_wgrad_rejects_sizes_outside_its_limits():
    lib = _lib.load()
    x = torch.zeros((1, C), dtype=torch.bfloat16, device="cuda")
    dw = torch.zeros(27 * C * C, dtype=torch.float32, device="cuda")
    db = torch.zeros(C, dtype=torch.float32, device="cuda")
    args = (x.data_ptr(), x.data_ptr(), dw.data_ptr(), db.data_ptr())
    st = _lib.stream_ptr()
    assert lib.psx_cnn_wgrad_noalloc(*args, N, 1 << 22, st) != 0
    assert b"outside" in lib.psx_last_error()
    assert lib.psx_cnn_wgrad_noalloc(*args, N, 0, st) != 0
    assert lib.psx_cnn_wgrad_noalloc(*args, 32, 1, st) != 0
    assert b"n must be 64" in lib.psx_last_error()
