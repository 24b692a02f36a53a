"""Pins of the LES closure oracle (CPU): finite differences for the network gradients,
brute-force loops for the periodic convolution, Adam's closed form, the Smagorinsky stress of
a shear flow, the box filter's transfer function and the divergence-free source."""

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
