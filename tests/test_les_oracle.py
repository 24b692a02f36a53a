"""Pins of the LES closure oracle (CPU): finite differences for the network gradients,
brute-force loops for the periodic convolution, Adam's closed form, the Smagorinsky stress and
source of a shear flow in closed form, the power of the stress source (Parseval), the box
filter's transfer function, the subgrid-stress target of two shear modes in closed form and
the divergence-free source."""

from __future__ import annotations

import math

import numpy as np

from oracle import les_oracle as LO


def _net(rng, cin, ch, cout, nl):
    ws, bs = [], []
    for i in range(nl):
        ci = cin if i == 0 else ch
        co = cout if i == nl - 1 else ch
        ws.append(rng.normal(0, 0.4, (27, co, ci)))
        bs.append(rng.normal(0, 0.1, co))
    return ws, bs


def test_conv_matches_brute_force():
    rng = np.random.default_rng(0)
    n, ci, co = 4, 3, 2
    x = rng.normal(size=(1, n, n, n, ci))
    w = rng.normal(size=(27, co, ci))
    b = rng.normal(size=co)
    y = LO.conv_fwd(x, w, b)
    ref = np.zeros((1, n, n, n, co))
    for X in range(n):
        for Y in range(n):
            for Z in range(n):
                for t, (dx, dy, dz) in enumerate(LO.TAPS):
                    ref[0, X, Y, Z] += w[t] @ x[0, (X + dx) % n, (Y + dy) % n, (Z + dz) % n]
                ref[0, X, Y, Z] += b
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_conv_adjoint_identities():
    """<conv(x), z> = <x, conv_bwd_data(z)> and d<conv(x), z>/dw = conv_wgrad(x, z)."""
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 4, 4, 4, 5))
    z = rng.normal(size=(2, 4, 4, 4, 3))
    w = rng.normal(size=(27, 3, 5))
    lhs = (LO.conv_fwd(x, w) * z).sum()
    assert math.isclose(lhs, (x * LO.conv_bwd_data(z, w)).sum(), rel_tol=1e-12)
    dw, db = LO.conv_wgrad(x, z)
    assert math.isclose(lhs, (dw * w).sum(), rel_tol=1e-12)
    np.testing.assert_allclose(db, z.reshape(-1, 3).sum(0))


def test_network_gradient_finite_differences():
    rng = np.random.default_rng(2)
    ws, bs = _net(rng, 3, 4, 2, 3)
    x = rng.normal(size=(1, 4, 4, 4, 3))
    t = rng.normal(size=(1, 4, 4, 4, 2))

    def loss():
        return LO.mse(LO.net_forward(x, ws, bs)[-1], t, 2)[0]

    acts = LO.net_forward(x, ws, bs)
    _, g = LO.mse(acts[-1], t, 2)
    gws, gbs = LO.net_backward(acts, ws, g)
    h = 1e-6
    for layer in range(3):
        for idx in [(0, 0, 0), (13, 1, 1), (26, 0, 2)]:
            idx = tuple(min(i, s - 1) for i, s in zip(idx, ws[layer].shape))
            w0 = ws[layer][idx]
            ws[layer][idx] = w0 + h
            lp = loss()
            ws[layer][idx] = w0 - h
            lm = loss()
            ws[layer][idx] = w0
            assert math.isclose((lp - lm) / (2 * h), gws[layer][idx], rel_tol=1e-5, abs_tol=1e-9)
        b0 = bs[layer][0]
        bs[layer][0] = b0 + h
        lp = loss()
        bs[layer][0] = b0 - h
        lm = loss()
        bs[layer][0] = b0
        assert math.isclose((lp - lm) / (2 * h), gbs[layer][0], rel_tol=1e-5, abs_tol=1e-9)


def test_adam_closed_form():
    """First step moves every weight by lr * sign(g) (m_hat = g, v_hat = g^2)."""
    w, m, v = np.zeros(4), np.zeros(4), np.zeros(4)
    g = np.array([0.5, -2.0, 1e-3, -1e-4])
    w1, m1, v1 = LO.adam(w, m, v, g, 1, 0.01, eps=0.0)
    np.testing.assert_allclose(w1, -0.01 * np.sign(g), rtol=1e-12)
    # constant gradient: every step is again lr * sign(g)
    w2, _, _ = LO.adam(w1, m1, v1, g, 2, 0.01, eps=0.0)
    np.testing.assert_allclose(w2, -0.02 * np.sign(g), rtol=1e-12)


def test_smagorinsky_shear_flow():
    """u = (sin y, 0, 0): S_12 = cos(y)/2, |S| = |cos y|, tau_12 = -(Cs D)^2 |cos y| cos y."""
    L = 16
    y = 2 * np.pi * np.arange(L) / L
    u = np.zeros((3, L, L, L))
    u[0] = np.sin(y)[None, :, None]
    st = np.fft.rfftn(u, axes=(1, 2, 3))
    g = LO.gradient(st, L)
    tau = LO.smagorinsky(g, 0.01)
    cy = np.cos(y)[None, :, None] * np.ones((L, L, L))
    np.testing.assert_allclose(tau[3], -0.01 * np.abs(cy) * cy, atol=1e-12)
    for c in (0, 1, 2, 4, 5):
        np.testing.assert_allclose(tau[c], 0, atol=1e-12)


def test_backscatter_clipping():
    """Smagorinsky stress only drains energy (-tau:S = 2 (Cs D)^2 |S|^3 >= 0) and is kept; its
    negative is pure backscatter and is zeroed; a mixed field keeps exactly the points where
    the transfer is forward.  Shear flow u = (sin y, 0, 0): -tau:S = 2 c |S|^3."""
    L = 16
    rng = np.random.default_rng(3)
    u = rng.normal(size=(3, L, L, L))
    g = LO.gradient(np.fft.rfftn(u, axes=(1, 2, 3)), L)
    smag = LO.smagorinsky(g, 0.02)
    np.testing.assert_array_equal(LO.clip_backscatter(smag, g), smag)
    assert np.all(LO.clip_backscatter(-smag, g) == 0)
    sign = np.where(rng.random((L, L, L)) < 0.5, 1.0, -1.0)
    mixed = smag * sign
    got = LO.clip_backscatter(mixed, g)
    keep = (sign > 0) | np.all(smag == 0, axis=0)
    np.testing.assert_array_equal(got, np.where(keep, mixed, 0.0))
    np.testing.assert_array_equal(LO.clip_backscatter(mixed, g, keep=0.25),
                                  np.where(keep, mixed, 0.25 * mixed))


def test_filter_transfer_and_truncation():
    L, Lc = 32, 16
    x = 2 * np.pi * np.arange(L) / L
    f = np.cos(3 * x)[:, None, None] + np.cos(7 * x)[None, :, None] * np.ones((L, L, L))
    ft = LO.filter_truncate(f, Lc, 5.0)
    # cos(3x): coefficient Lc^3/2 at kx = +-3 times G(3); cos(7x) is outside 3|k| < 16
    assert math.isclose(ft[3, 0, 0].real, Lc ** 3 / 2 * LO.box_transfer(5.0, np.array(3.0)),
                        rel_tol=1e-12)
    assert np.abs(ft[0, 7, 0]) < 1e-9 and np.abs(ft[0, 16 - 7, 0]) < 1e-9
    assert LO.box_transfer(5.0, np.array(0.0)) == 1.0


def test_source_is_divergence_free_and_masked():
    rng = np.random.default_rng(3)
    L = 16
    tau = rng.normal(size=(6, L, L, L))
    s = LO.stress_source(tau, L)
    k = LO.wavenumbers(L)
    kx, ky, kz = k[:, None, None], k[None, :, None], np.arange(L // 2 + 1)[None, None, :]
    div = kx * s[0] + ky * s[1] + kz * s[2]
    assert np.abs(div).max() < 1e-9 * np.abs(s).max()
    assert np.all(s[:, 6:11] == 0) and np.all(s[..., 6:] == 0)


def _random_solenoidal(rng, L, kmax=None):
    """A random divergence-free velocity spectrum (3, L, L, L/2+1) on the 2/3-retained modes,
    rfftn of a real field."""
    u = np.fft.rfftn(rng.normal(size=(3, L, L, L)), axes=(1, 2, 3))
    k = LO.wavenumbers(L)
    kx, ky, kz = k[:, None, None], k[None, :, None], np.arange(L // 2 + 1)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    d = (kx * u[0] + ky * u[1] + kz * u[2]) / np.where(k2 > 0, k2, 1)
    u = np.stack([u[0] - kx * d, u[1] - ky * d, u[2] - kz * d])
    mask = (3 * np.abs(kx) < L) & (3 * np.abs(ky) < L) & (3 * np.abs(kz) < L)
    return u * mask, (kx, ky, kz)


def test_smagorinsky_source_removes_energy_at_the_closed_form_rate():
    """For a divergence-free retained state u, the power of the source s = P(-i k_j tau_ij)
    is sum_k conj(u_i) s_i = L^3 sum_x tau_ij du_i/dx_j (Parseval, exactly), and for the
    Smagorinsky stress tau_ij = -2 (Cs D)^2 |S| S_ij that is -(Cs D)^2 L^3 sum_x |S|^3 < 0.
    A sign or factor error in the stress, the divergence or the projection moves the
    result off this value."""
    rng = np.random.default_rng(5)
    L, csd2 = 16, 0.013
    u, _ = _random_solenoidal(rng, L)
    g = LO.gradient(u, L)
    tau = LO.smagorinsky(g, csd2)
    s = LO.stress_source(tau, L)
    w = np.full(L // 2 + 1, 2.0)
    w[0] = w[L // 2] = 1.0
    power = float(np.sum(w * np.real(np.conj(u) * s)))
    S = 0.5 * (g + np.swapaxes(g, 0, 1))
    mag = np.sqrt(2.0 * np.einsum("ij...,ij...->...", S, S))
    want = -csd2 * L ** 3 * float(np.sum(mag ** 3))
    assert power < 0
    assert math.isclose(power, want, rel_tol=1e-10), (power, want)
    # the same identity for any symmetric stress (here a random one)
    tr = rng.normal(size=(6, L, L, L))
    sr = LO.stress_source(tr, L)
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    tg = sum(tr[c] * (g[i, i] if i == j else g[i, j] + g[j, i])
             for c, (i, j) in enumerate(pairs))
    assert math.isclose(float(np.sum(w * np.real(np.conj(u) * sr))), L ** 3 * float(np.sum(tg)),
                        rel_tol=1e-10)


def test_smagorinsky_source_of_a_shear_flow_closed_form():
    """u = (sin y, 0, 0): |S| = |cos y|, tau_12 = -(Cs D)^2 |cos y| cos y at every grid point,
    and the source is s_x = -d tau_12 / dy = -2 (Cs D)^2 |cos y| sin y.  |cos y| cos y =
    sum over odd m of a_m cos(m y), a_m = (I(m) + (I(m + 2) + I(m - 2)) / 2) / pi with
    I(k) = 2 sin(k pi / 2) / k, so the retained coefficients of s_x are
    s_x(0, m, 0) = i (Cs D)^2 c_m m L^3 / 2 (m > 0), where the grid aliases the harmonics
    beyond it onto c_m = sum_j a_|m + j L| (a_m ~ 8 / (pi m^3), summed to convergence)."""
    L, c = 64, 0.01
    y = 2 * np.pi * np.arange(L) / L
    u = np.zeros((3, L, L, L))
    u[0] = np.sin(y)[None, :, None]
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    g = LO.gradient(uh, L)
    tau = LO.smagorinsky(g, c)
    want_t12 = -c * np.abs(np.cos(y)) * np.cos(y)
    np.testing.assert_allclose(tau[3], np.broadcast_to(want_t12[None, :, None], (L, L, L)),
                               atol=1e-14)
    assert np.abs(tau[[0, 1, 2, 4, 5]]).max() < 1e-14
    s = LO.stress_source(tau, L)

    def I(k):  # noqa: E743
        return math.pi if k == 0 else 2 * math.sin(k * math.pi / 2) / k

    ref = np.zeros((3, L, L, L // 2 + 1), dtype=complex)
    def a(m):
        return (I(m) + 0.5 * (I(m + 2) + I(m - 2))) / math.pi

    for m in range(1, (L - 1) // 3 + 1, 2):
        cm = sum(a(abs(m + j * L)) for j in range(-4000, 4001))
        ref[0, 0, m, 0] = 1j * c * cm * m * L ** 3 / 2
        ref[0, 0, L - m, 0] = -1j * c * cm * m * L ** 3 / 2
    assert math.isclose(a(1), 8 / (3 * math.pi))
    err = np.abs(s - ref).max() / np.abs(ref).max()
    assert err < 1e-9, err


def test_sgs_target_of_two_shear_modes_closed_form():
    """u = (A cos(p y), B cos(q x), 0) on the fine grid: with the separable box filter
    G(k) = g(kx) g(ky) g(kz), filt(u1 u2) = g(p) g(q) u1 u2 = filt(u1) filt(u2), so tau_12 = 0;
    tau_11 = A^2/2 (1 - g(p)^2) + A^2/2 (g(2p) - g(p)^2) cos(2 p y), tau_22 likewise in x;
    every other component is 0; the filtered velocity is (A g(p) cos(p y), B g(q) cos(q x), 0)
    on the coarse grid."""
    L, Lc, kc = 32, 16, 5.0
    A, B, p, q = 0.7, -1.3, 2, 1
    x = 2 * np.pi * np.arange(L) / L
    u = np.zeros((3, L, L, L))
    u[0] = A * np.cos(p * x)[None, :, None]
    u[1] = B * np.cos(q * x)[:, None, None]
    ub, tau = LO.sgs_target(u, Lc, kc)

    def gk(k):
        return float(LO.box_transfer(kc, np.array(float(k))))

    xc = 2 * np.pi * np.arange(Lc) / Lc
    t11 = A * A / 2 * (1 - gk(p) ** 2) + A * A / 2 * (gk(2 * p) - gk(p) ** 2) * np.cos(2 * p * xc)
    t22 = B * B / 2 * (1 - gk(q) ** 2) + B * B / 2 * (gk(2 * q) - gk(q) ** 2) * np.cos(2 * q * xc)
    np.testing.assert_allclose(tau[0], np.broadcast_to(t11[None, :, None], (Lc,) * 3), atol=1e-13)
    np.testing.assert_allclose(tau[1], np.broadcast_to(t22[:, None, None], (Lc,) * 3), atol=1e-13)
    assert np.abs(tau[2:]).max() < 1e-13
    uc = LO.physical(ub, Lc)
    np.testing.assert_allclose(uc[0], np.broadcast_to((A * gk(p) * np.cos(p * xc))[None, :, None],
                                                      (Lc,) * 3), atol=1e-13)
    np.testing.assert_allclose(uc[1], np.broadcast_to((B * gk(q) * np.cos(q * xc))[:, None, None],
                                                      (Lc,) * 3), atol=1e-13)
    assert gk(p) < 1 and gk(2 * p) < gk(p) ** 2 + 1e-3  # the filter acts on both modes
