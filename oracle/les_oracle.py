"""CPU (numpy, FP64) restatement of the LES closure path: TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module (it is the checker of the HIP kernels, never the thing
measured or shipped).  The reference has no LES, no subgrid model and no network; what it
holds is the rfftn state convention (src/pyspectral/grid.py:13-67) and the sharp spectral
truncation src/pyspectral/validation.py:25-35, which ``filter_truncate`` follows (times the
box-filter transfer function).  Everything else is pinned by mathematics in
tests/test_les_oracle.py: finite differences for every gradient, brute-force loops for the
convolution, closed forms for Adam and for the Smagorinsky stress of a shear flow, the
divergence-free source, and the filter's transfer function.

Layouts (as the kernels): network tensors (nb, n, n, n, C) point-major; weights
(27, C_out, C_in) with tap = (dx+1)*9 + (dy+1)*3 + (dz+1) and input x[p + (dx, dy, dz)].
"""

from __future__ import annotations

import math

import numpy as np

TAPS = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def shift(x: np.ndarray, s: tuple[int, int, int]) -> np.ndarray:
    """x'[p] = x[p + s] on periodic cubes (axes 1..3 of (nb, n, n, n, C))."""
    return np.roll(x, shift=(-s[0], -s[1], -s[2]), axis=(1, 2, 3))


def conv_fwd(x: np.ndarray, w: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    """y[p][o] = sum_tap sum_i w[tap][o][i] x[p + s(tap)][i] + b[o]."""
    y = np.zeros(x.shape[:-1] + (w.shape[1],))
    for t, s in enumerate(TAPS):
        y += shift(x, s) @ w[t].T
    return y + (0.0 if b is None else b)


def conv_bwd_data(dz: np.ndarray, w: np.ndarray) -> np.ndarray:
    """dx[q][i] = sum_tap sum_o dz[q - s(tap)][o] w[tap][o][i]."""
    dx = np.zeros(dz.shape[:-1] + (w.shape[2],))
    for t, s in enumerate(TAPS):
        dx += shift(dz, (-s[0], -s[1], -s[2])) @ w[t]
    return dx


def conv_wgrad(x: np.ndarray, dz: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """dw[tap][o][i] = sum_p dz[p][o] x[p + s(tap)][i]; db[o] = sum_p dz[p][o]."""
    co, ci = dz.shape[-1], x.shape[-1]
    dzf = dz.reshape(-1, co)
    dw = np.stack([dzf.T @ shift(x, s).reshape(-1, ci) for s in TAPS])
    return dw, dzf.sum(axis=0)


def net_forward(x: np.ndarray, ws: list[np.ndarray], bs: list[np.ndarray]) -> list[np.ndarray]:
    """Activations [x, y1, ..., yL]: ReLU after every layer but the last."""
    acts = [x]
    for i, (w, b) in enumerate(zip(ws, bs)):
        z = conv_fwd(acts[-1], w, b)
        acts.append(np.maximum(z, 0.0) if i < len(ws) - 1 else z)
    return acts


def mse(y: np.ndarray, t: np.ndarray, nch: int) -> tuple[float, np.ndarray]:
    """Mean square over the first nch channels and its gradient (zero elsewhere)."""
    d = y[..., :nch] - t[..., :nch]
    count = d.size
    g = np.zeros_like(y)
    g[..., :nch] = 2.0 * d / count
    return float((d * d).sum() / count), g


def net_backward(acts: list[np.ndarray], ws: list[np.ndarray],
                 g: np.ndarray) -> tuple[list[np.ndarray], list[np.ndarray]]:
    """Gradients of the loss with respect to every weight and bias, given dL/dy_out = g."""
    gws, gbs = [None] * len(ws), [None] * len(ws)
    dz = g
    for i in range(len(ws) - 1, -1, -1):
        gws[i], gbs[i] = conv_wgrad(acts[i], dz)
        if i > 0:
            dz = conv_bwd_data(dz, ws[i]) * (acts[i] > 0)
    return gws, gbs  # type: ignore[return-value]


def adam(w: np.ndarray, m: np.ndarray, v: np.ndarray, g: np.ndarray, t: int, lr: float,
         b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8) -> tuple[np.ndarray, ...]:
    """One Adam step (Kingma & Ba, Algorithm 1) at step number t >= 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


# ----------------------------------------------------------------------- spectral side
def wavenumbers(L: int) -> np.ndarray:
    return np.fft.fftfreq(L, 1.0 / L)


def physical(state: np.ndarray, L: int) -> np.ndarray:
    """Velocity (3, L, L, L) of a state in the rfftn convention."""
    return np.fft.irfftn(state, s=(L, L, L), axes=(1, 2, 3))


def gradient(state: np.ndarray, L: int) -> np.ndarray:
    """g[i][j] = d u_i / d x_j (3, 3, L, L, L), spectral derivative (Nyquist planes zero)."""
    k = wavenumbers(L)
    kk = [k[:, None, None], k[None, :, None], np.arange(L // 2 + 1)[None, None, :]]
    nyq = np.ones((L, L, L // 2 + 1))
    nyq[L // 2, :, :] = 0
    nyq[:, L // 2, :] = 0
    nyq[:, :, L // 2] = 0
    return np.stack([np.stack([np.fft.irfftn(1j * kk[j] * state[i] * nyq, s=(L, L, L),
                                             axes=(0, 1, 2)) for j in range(3)])
                     for i in range(3)])


def box_transfer(kc: float, k: np.ndarray) -> np.ndarray:
    """sin(k D/2) / (k D/2), D = pi / kc (1 at k = 0)."""
    h = 0.5 * math.pi / kc
    return np.where(k == 0, 1.0, np.sin(k * h) / np.where(k == 0, 1.0, k * h))


def filter_truncate(field: np.ndarray, Lc: int, kc: float | None) -> np.ndarray:
    """Real fields (..., Lf, Lf, Lf) -> rfftn spectra on the coarse grid (..., Lc, Lc, Lc/2+1):
    the modes retained by the coarse 2/3 rule, times the box transfer function (kc not None),
    normalised as the coarse grid's own unnormalised DFT."""
    Lf = field.shape[-1]
    F = np.fft.fftn(field, axes=(-3, -2, -1)) * (Lc / Lf) ** 3
    kcg = wavenumbers(Lc).astype(int)
    KX, KY, KZ = np.meshgrid(kcg, kcg, np.arange(Lc // 2 + 1), indexing="ij")
    keep = (3 * np.abs(KX) < Lc) & (3 * np.abs(KY) < Lc) & (3 * np.abs(KZ) < Lc)
    g = np.ones(KX.shape)
    if kc is not None:
        g = box_transfer(kc, KX) * box_transfer(kc, KY) * box_transfer(kc, KZ)
    return F[..., KX % Lf, KY % Lf, KZ % Lf] * (g * keep)


def sgs_target(u: np.ndarray, Lc: int, kc: float) -> tuple[np.ndarray, np.ndarray]:
    """Filtered velocity spectra (3, Lc, Lc, Lc/2+1) and tau_ij = filt(u_i u_j) - filt u_i
    filt u_j (6, Lc, Lc, Lc) in the order 11, 22, 33, 12, 13, 23, from the fine velocity u."""
    ub = filter_truncate(u, Lc, kc)
    uc = physical(ub, Lc)
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    tau = []
    for i, j in pairs:
        fij = np.fft.irfftn(filter_truncate(u[i] * u[j], Lc, kc), s=(Lc,) * 3, axes=(0, 1, 2))
        tau.append(fij - uc[i] * uc[j])
    return ub, np.stack(tau)


def smagorinsky(g: np.ndarray, csd2: float) -> np.ndarray:
    """tau_ij = -2 (Cs D)^2 |S| S_ij, |S| = sqrt(2 S_ij S_ij); order 11, 22, 33, 12, 13, 23."""
    S = 0.5 * (g + np.swapaxes(g, 0, 1))
    mag = np.sqrt(2.0 * np.einsum("ij...,ij...->...", S, S))
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    return np.stack([-2.0 * csd2 * mag * S[i, j] for i, j in pairs])


def clip_backscatter(tau: np.ndarray, g: np.ndarray, keep: float = 0.0) -> np.ndarray:
    """Stress (order 11, 22, 33, 12, 13, 23) times ``keep`` wherever -tau_ij S_ij < 0, S the
    strain rate of the gradient g[i][j] = d u_i / d x_j (no reference counterpart)."""
    S = 0.5 * (g + np.swapaxes(g, 0, 1))
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    ts = sum((1.0 if i == j else 2.0) * tau[c] * S[i, j] for c, (i, j) in enumerate(pairs))
    return np.where(ts > 0, keep * tau, tau)


def stress_source(tau: np.ndarray, L: int) -> np.ndarray:
    """S_i = -i k_j tau_ij (rfftn of the physical stress), Leray-projected, 2/3-masked."""
    th = np.fft.rfftn(tau, axes=(1, 2, 3))
    t11, t22, t33, t12, t13, t23 = th
    k = wavenumbers(L)
    kx, ky, kz = k[:, None, None], k[None, :, None], np.arange(L // 2 + 1)[None, None, :]
    s = -1j * np.stack([kx * t11 + ky * t12 + kz * t13,
                        kx * t12 + ky * t22 + kz * t23,
                        kx * t13 + ky * t23 + kz * t33])
    k2 = kx * kx + ky * ky + kz * kz
    ik2 = np.where(k2 > 0, 1.0 / np.where(k2 > 0, k2, 1), 0.0)
    d = (kx * s[0] + ky * s[1] + kz * s[2]) * ik2
    s = np.stack([s[0] - kx * d, s[1] - ky * d, s[2] - kz * d])
    mask = (3 * np.abs(kx) < L) & (3 * np.abs(ky) < L) & (3 * np.abs(kz) < L)
    return s * mask
