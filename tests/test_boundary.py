"""The drop-in boundary on the CPU: the C ABI library, the registry and the host-side mirror of
the reference interface (no compute calls; those need a GPU)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import numpy_oracle as O
from pyspectral_amd import _lib, grid, initial
from pyspectral_amd.backends import REGISTRY, BackendUnavailable, get, names
from pyspectral_amd.config import Config, fft_split, ffts_per_stage

ROOT = Path(__file__).resolve().parents[1]
GOLD = ROOT / "tests" / "golden"


def _header_symbols() -> list[str]:
    text = (ROOT / "include" / "psx.h").read_text()
    decl = r"^\s*(?:int|const char\*|unsigned long)\s+(psx_\w+)\s*\("
    return sorted(set(re.findall(decl, text, re.M)))


def test_header_and_binding_agree():
    assert set(_header_symbols()) == set(_lib.SIGNATURES)


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="libpsx.so not built")
def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in _header_symbols():
        assert hasattr(lib, name), name


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="libpsx.so not built")
@pytest.mark.parametrize("n,dim,form,fp64,ok", [
    (8192, 2, 0, 0, True), (256, 2, 0, 1, True), (32, 2, 0, 0, True), (16, 3, 1, 1, True),
    (384, 3, 0, 0, True), (96, 2, 0, 0, True), (20, 2, 0, 1, True), (40, 3, 0, 0, True),
    (10, 2, 0, 0, False), (60, 2, 0, 0, False), (32, 2, 1, 0, False), (7, 2, 0, 0, False)])
def test_layout_query(n, dim, form, fp64, ok):
    lib = _lib.load()
    p = _lib.FlowParams(n=n, dim=dim, form=form, fp64=fp64, nu=1e-3, alpha=0.0, ka2=1e30,
                        klo2=0.0, khi2=0.0)
    lay = _lib.Layout()
    rc = lib.psx_query_layout(ctypes.byref(p), ctypes.byref(lay))
    assert (rc == 0) == ok, lib.psx_last_error()
    if ok:
        nh = n // 2 + 1
        rows = n ** (dim - 1)
        assert lay.state_pitch >= nh and (lay.state_pitch * (16 if fp64 else 8)) % 128 == 0
        assert lay.state_elems == (1 if dim == 2 else 3) * rows * lay.state_pitch
        assert lay.jm == grid.retained(n) and lay.ldw >= lay.jm
        assert lay.nfields == (2 if dim == 2 else (6 if form else 5))
        assert lay.work_elems == lay.nfields * rows * lay.ldw
    else:
        assert lib.psx_last_error()


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="libpsx.so not built")
@pytest.mark.parametrize("n", [8, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256,
                               320, 384, 512, 640, 768, 1024, 1280, 1536, 2048, 4096, 8192])
@pytest.mark.parametrize("dim,form", [(2, 0), (3, 0), (3, 1)])
@pytest.mark.parametrize("fp64", [0, 1])
def test_planner_picks_compiled_variants(n, dim, form, fp64):
    """Every launch shape the planner picks exists in the compiled kernel set and fits a
    work-group (the launchers reject other shapes with an error, never a bad launch)."""
    if dim == 3 and n > 512:
        pytest.skip("3D grids above 512 are not planned")
    lib = _lib.load()
    p = _lib.FlowParams(n=n, dim=dim, form=form, fp64=fp64, nu=1e-3, alpha=0.0, ka2=1e30,
                        klo2=0.0, khi2=0.0)
    lay = _lib.Layout()
    assert lib.psx_query_layout(ctypes.byref(p), ctypes.byref(lay)) == 0, lib.psx_last_error()
    ok2d = {8, 16, 20, 24, 32} if not fp64 else {8, 12, 16, 20}
    # 3D fp32 sizes with compile-time kernels: any plan the planner makes is compiled (24 for
    # radix-3 lines, 4 for the narrow column tiles of 64^3)
    ok3d = ({4, 8, 12, 16, 24} if (not fp64 and n in (64, 96, 128, 192, 256, 384))
            else {8, 12, 16, 20})
    ok = ok2d if dim == 2 else ok3d
    for e, t in ((lay.e_line, lay.t_line), (lay.e_spec, lay.t_spec),
                 (lay.e_outer, lay.t_outer), (lay.e_mid, lay.t_mid)):
        if e:
            assert e in ok and 1 <= t <= 512
    # the line kernel's work-groups hold whole rows and cover the n^(d-1) rows exactly: a
    # tile that does not divide them would leave the last rows untransformed
    assert (lay.t_line * lay.e_line) % n == 0
    row_tile = lay.t_line * lay.e_line // n
    assert (n ** (dim - 1)) % row_tile == 0, (n, dim, row_tile)


def test_registry_surface():
    assert names() == ["hip/fused", "hip/fused-graph"]
    assert get("hip").variant == "fused"
    with pytest.raises(KeyError):
        get("nope")
    for v in REGISTRY:
        ok, why = v.available()
        assert isinstance(ok, bool) and isinstance(why, str)


def test_unavailable_without_gpu_raises_backend_unavailable():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(BackendUnavailable):
        get("hip/fused").make(Config(n=16, dim=2))


def test_config_mirrors_reference():
    assert ffts_per_stage(2, "basdevant") == 4 and ffts_per_stage(3, "rotational") == 9
    assert fft_split(3, "basdevant") == (3, 5)
    assert Config(n=16, dim=3).resolved_form() == "basdevant"
    assert Config(n=16, alpha_kmax=2.0).drag_k2 == 4.0 and Config(n=16).drag_k2 == 1e30
    for bad in ({"n": 7}, {"n": 16, "dim": 4}, {"n": 16, "dtype": "float16"},
                {"n": 16, "form": "rotational"}, {"n": 16, "dealias": "x"}):
        with pytest.raises(ValueError):
            Config(**bad)


def test_grid_matches_oracle():
    for n, d in ((12, 2), (16, 3), (18, 2)):
        assert grid.spectral_shape(n, d) == O.spectral_shape(n, d)
        np.testing.assert_array_equal(grid.dealias_mask(n, d), O.dealias_mask(n, d))
        np.testing.assert_array_equal(grid.k_squared(n, d), O.k_squared(n, d))
        assert grid.retained(n) == int(O.dealias_mask(n, d).reshape(-1, n // 2 + 1)[0].sum())


def test_initial_conditions_match_reference_vectors():
    with np.load(GOLD / "ic.npz", allow_pickle=False) as g:
        np.testing.assert_allclose(initial.random_2d(32, k0=4), g["random_2d_32_k4"], atol=1e-13)
        np.testing.assert_allclose(initial.random_3d(16, k0=2), g["random_3d_16_k2"], atol=1e-13)
        np.testing.assert_allclose(initial.smooth_2d(32), g["smooth_2d_32"], atol=1e-13)


def test_random_phase_ic_spectrum_and_energy():
    n, k0 = 64, 6.0
    w = initial.random_phase_2d(n, k0=k0, seed=3, energy=0.5)
    d = O.OracleDiagnostics(n, 2)
    assert d.energy(w) == pytest.approx(0.5, rel=1e-12)
    assert np.all(w[~grid.dealias_mask(n, 2)] == 0)
    # Hermitian ky = 0 column -> the physical field is real (irfft round trip is exact)
    f = np.fft.irfftn(w, s=(n, n))
    np.testing.assert_allclose(np.fft.rfftn(f), w, atol=1e-9 * np.abs(w).max())
    k, e = d.spectrum(w)
    sel = (k >= 3) & (k <= 12)
    model = k[sel] ** 4 * np.exp(-((k[sel] / k0) ** 2))
    ratio = e[sel] / model
    assert ratio.std() / ratio.mean() < 0.35  # shell sums follow k^4 exp(-(k/k0)^2)


def test_checkpoint_round_trip(tmp_path):
    """The reference's checkpoint format (validation.py:204-221): state, t, amp and the JSON
    history survive a save/load into a fresh solver; the keys are the reference's."""
    from pyspectral_amd import checkpoint
    from pyspectral_amd.simulate import RunResult

    cfg = Config(n=16, dim=2, nu=0.01, force_amp=0.2, force_klo=1, force_khi=3,
                 dtype="float32")
    a = O.OracleSolver(cfg)
    a.set_state(O.to_spectral(O.random_2d(16, k0=3, seed=2), 2))
    a.advance(3, 0.01)
    a.set_forcing_amplitude(0.35)
    res = RunResult(t=[0.0, 0.03], energy=[0.5, 0.49], enstrophy=[2.0, 1.9],
                    dt=[0.01] * 3, steps=3, wall=0.1)
    f = tmp_path / "ck.npz"
    checkpoint.save(f, a, res, spectrum=np.arange(4.0))
    z = np.load(f, allow_pickle=False)
    assert set(z.files) == {"state", "t", "amp", "hist", "spectrum"}
    assert z["state"].dtype == np.complex64
    b = O.OracleSolver(cfg)
    r2, z2 = checkpoint.load(f, b)
    assert np.array_equal(b.get_state(), a.get_state()) and b.t == a.t
    assert b.force_amp == 0.35 and r2.as_dict() == res.as_dict()
    assert np.array_equal(z2["spectrum"], np.arange(4.0))


def test_device_loop_ring_bounds():
    """The device loop refuses chunk sizes that could overwrite undrained ring entries."""
    import pytest

    from pyspectral_amd.backends.hip_fused import check_ring_bounds

    check_ring_bounds(1500, 3, 1 << 14, 1 << 14)
    with pytest.raises(ValueError):
        check_ring_bounds(2048, 3, 1 << 14, 1 << 14)
    with pytest.raises(ValueError):
        check_ring_bounds(342, 3, 1 << 12, 1 << 10)
    with pytest.raises(ValueError):
        check_ring_bounds(0, 3, 1 << 14, 1 << 14)


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="libpsx.so not built")
@pytest.mark.parametrize("n", [8, 16, 32, 48, 96, 256, 1024, 8192])
@pytest.mark.parametrize("dim,fp64", [(2, 0), (2, 1), (3, 0), (3, 1)])
def test_phase_shift_layout_holds_the_spherical_mask(n, dim, fp64):
    """dealias = 1: the stored half-axis columns are exactly those the spherical mask
    |k|^2 < 2 (n/3)^2 of the reference (grid.py:36-50) can reach."""
    if dim == 3 and n > 512:
        pytest.skip("3D grids above 512 are not planned")
    lib = _lib.load()
    p = _lib.FlowParams(n=n, dim=dim, form=0, fp64=fp64, nu=1e-3, alpha=0.0, ka2=1e30,
                        klo2=0.0, khi2=0.0, dealias=1)
    lay = _lib.Layout()
    assert lib.psx_query_layout(ctypes.byref(p), ctypes.byref(lay)) == 0, lib.psx_last_error()
    m = grid.dealias_mask(n, 2, "phase-shift")
    assert lay.jm == int(np.nonzero(m.any(axis=0))[0].max()) + 1
    assert lay.jm <= n // 2 and lay.ldw >= lay.jm
    p.dealias = 2
    assert lib.psx_query_layout(ctypes.byref(p), ctypes.byref(lay)) != 0
