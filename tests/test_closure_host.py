"""Host side of the a-priori transfer analysis of the closure study (CPU): the shell sums of
Re(conj(u_hat) . s_hat) against the physical-space mean of u . s (Parseval), the shell index of
single modes, and the per-shell factor fit."""

from __future__ import annotations

import numpy as np

from pyspectral_amd.closure import Shells, fit_shell_gain


def test_transfer_sums_to_the_physical_mean():
    L = 16
    rng = np.random.default_rng(0)
    u = rng.normal(size=(3, L, L, L))
    s = rng.normal(size=(3, L, L, L)) + 0.5 * u  # correlated, so the mean is not small
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    sh = np.fft.rfftn(s, axes=(1, 2, 3))
    sh_ = Shells(L, int(np.floor(np.sqrt(3) * L / 2 + 0.5)) + 1)
    t = sh_.transfer(uh, sh)
    ref = float(np.mean(np.sum(u * s, axis=0)))
    assert abs(t.sum() - ref) < 1e-10 * abs(ref)


def test_single_modes_land_in_their_shells():
    L = 16
    sh = Shells(L, int(np.floor(np.sqrt(3) * L / 2 + 0.5)) + 1)
    for (kx, ky, kz), shell in (((3, 4, 0), 5), ((1, 1, 1), 2), ((0, 0, 2), 2), ((-2, 1, 2), 3),
                                ((2, 2, 2), 3), ((8, 8, 8), 14)):
        uh = np.zeros((3, L, L, L // 2 + 1), dtype=complex)
        if kz < 0:
            kx, ky, kz = -kx, -ky, -kz
        uh[0, kx % L, ky % L, kz] = 1.0
        t = sh.transfer(uh, uh)
        w = 1.0 if kz in (0, L // 2) else 2.0
        expect = np.zeros(sh.nshell)
        expect[shell] = w / float(L) ** 6
        np.testing.assert_allclose(t, expect, rtol=1e-14, atol=0)


def test_shell_gain_fit():
    tt = np.array([0.0, 1.0, -2.0, -4.0, -1.0, -3.0, -1.0])
    tm = np.array([0.0, -1.0, -1.0, -0.5, -8.0, -3e-6, -2.0])
    g = fit_shell_gain(tt, tm, kmax=5)
    # shell 0 and shells past kmax keep 1; a sign mismatch keeps 1; a negligible model
    # transfer keeps 1; ratios are clipped to [0.25, 4]
    np.testing.assert_allclose(g, [1.0, 1.0, 2.0, 4.0, 0.25, 1.0, 1.0])
