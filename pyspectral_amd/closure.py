"""Learned-closure LES study: forced 3D DNS -> filtered snapshots -> network training -> coarse
LES with no model, Smagorinsky and the network, compared with the filtered DNS.

    python -m pyspectral_amd.closure [--quick] [--out results.json]

Workload (defaults): DNS 256^3 fp32, constant-power band forcing on 2 <= |k| <= 3 with
P = 0.1, nu = 1.6e-3 (Re_lambda about 80); snapshots box-filtered (k_c = 21) onto the 64^3
LES grid; six 3x3x3 convolutions of 32 channels trained with Adam (batch 8, cosine
learning-rate decay) on the GPU; 64^3 LES runs from the filtered DNS field with the same
forcing, averaged over all but its first 3 time units: no model, Smagorinsky
(Cs = 0.17), the network with backscatter clipping and the network without it, and mixed
models (the network plus a Smagorinsky stress of a small constant).  Reported:
the time-averaged E(k) error of each LES against the time-averaged filtered DNS spectrum for
1 <= k <= k_c, the resolved dissipation 2 nu Z, the DNS's Re_lambda, k_max eta and
large-eddy time, and GPU time per unit simulated time of the DNS and each LES.

A priori (before any LES runs), on filtered snapshots: the subgrid energy-transfer spectrum
T(k) = sum over the shell of Re(conj(u_hat) . s_hat), s the projected stress divergence, of the
filtered-DNS stress (the target) and of each model, averaged over a calibration window drawn
from the training snapshots and, separately, over the held-out snapshots.  A per-shell factor
on the network's source, fitted on the calibration window only (target transfer over network
transfer), is reported on the held-out window and then run a posteriori as its own LES.
Every kernel is libpsx's; torch provides device memory and the batch gathers; the shell sums
of the transfer run on the host.
"""

from __future__ import annotations

import argparse
import json
import math
import time
from typing import Any

import numpy as np

from pyspectral_amd import initial
from pyspectral_amd.backends.hip_fused import HipSolver
from pyspectral_amd.config import Config
from pyspectral_amd.les import C, TAU_C, FilterPipeline, LESSolver, SgsNet, fft3, les_op, rms


def _spectrum(solver: Any, half_state: Any) -> np.ndarray:
    return solver.diag.spectrum(half_state.contiguous())[1]


def turbulence_scales(E: float, Z: float, ek: np.ndarray, nu: float, n: int) -> dict:
    """Scales of a homogeneous isotropic flow from its energy E, enstrophy Z (<|omega|^2>/2)
    and shell spectrum: u' = sqrt(2E/3), dissipation eps = 2 nu Z, Taylor microscale
    lambda = sqrt(15 nu u'^2 / eps), Re_lambda = u' lambda / nu, Kolmogorov eta =
    (nu^3 / eps)^(1/4) against the 2/3-rule cutoff k_max = n/3, integral scale
    L = pi / (2 u'^2) sum_k E(k) / k and large-eddy time L / u'."""
    u = math.sqrt(2.0 * E / 3.0)
    eps = 2.0 * nu * Z
    lam = math.sqrt(15.0 * nu * u * u / eps)
    k = np.arange(1, len(ek))
    L = math.pi / (2.0 * u * u) * float(np.sum(ek[1:] / k))
    return {"energy": E, "dissipation": eps, "u_rms": u, "taylor_lambda": lam,
            "re_lambda": u * lam / nu, "kmax_eta": (n / 3.0) * (nu ** 3 / eps) ** 0.25,
            "integral_scale": L, "t_eddy": L / u}


def _dns_and_snapshots(cfg_dns: Config, n_les: int, kc: float, nu: float, t_spin: float,
                       n_snap: int, snap_every: float, t_cost: float, seed: int,
                       run_kw: dict, log: Any) -> dict[str, Any]:
    """Forced DNS: spin-up, filtered snapshots (model inputs X, targets T), the time-averaged
    filtered spectrum and dissipation, the flow's scales, and the GPU time per unit simulated
    time of one continuous run of t_cost."""
    import torch

    n_dns = cfg_dns.n
    dns = HipSolver(cfg_dns)
    dns.set_state(initial.to_spectral(initial.random_3d(n_dns, k0=2.5, seed=seed,
                                                        energy=0.5), 3))
    spin = dns.run_device(t_spin, **run_kw)
    torch.cuda.synchronize()
    log(f"DNS {n_dns}^3 spin-up to t={dns.t:.2f}: {spin['steps']} steps")
    L3 = n_les ** 3
    pipe = FilterPipeline(n_dns, n_les, kc)
    X = torch.zeros((n_snap, L3, C), dtype=torch.bfloat16, device="cuda")
    T = torch.zeros((n_snap, L3, TAU_C), dtype=torch.float32, device="cuda")
    # the filtered velocity of every snapshot (NumPy layout), for the a-priori transfer
    U = torch.zeros((n_snap, 3, n_les, n_les, n_les // 2 + 1), dtype=torch.complex64,
                    device="cuda")
    pipe(dns.buf.u0_rows, dns.buf.pitch, X[0], T[0])  # unscaled pilot for the scales
    g_rms = rms(torch.view_as_real(pipe.grad[2 * L3: 6 * L3]).reshape(-1), 2, 2)
    tau_rms = rms(T[0].reshape(-1), TAU_C, 6)
    log(f"scales: velocity-gradient rms {g_rms:.4g}, stress rms {tau_rms:.4g}")
    les_probe = LESSolver(cfg_dns.with_(n=n_les), "none")
    ek_f = np.zeros(les_probe.diag.nshell)
    ek_half = [np.zeros(les_probe.diag.nshell), np.zeros(les_probe.diag.nshell)]
    ek_dns = np.zeros(dns.diag.nshell)
    zf, e_dns, z_dns = [], [], []
    for i in range(n_snap):
        dns.run_device(dns.t + snap_every, **run_kw)
        sn = dns.state_native()
        e_dns.append(dns.diag.energy(sn))
        z_dns.append(dns.diag.enstrophy(sn))
        ek_dns += _spectrum(dns, sn)
        ubar = pipe(dns.buf.u0_rows, dns.buf.pitch, X[i], T[i], g_scale=1.0 / g_rms,
                    tau_scale=1.0 / tau_rms)
        hs = pipe.half_state(ubar)
        U[i].copy_(hs)
        ek_s = _spectrum(les_probe, hs)
        ek_f += ek_s
        ek_half[0 if 2 * i < n_snap else 1] += ek_s
        zf.append(les_probe.diag.enstrophy(hs.contiguous()))
        if (i + 1) % 500 == 0:
            log(f"  snapshot {i + 1}/{n_snap} at t={dns.t:.2f}")
    ek_f /= n_snap
    ek_dns /= n_snap
    # stationarity of the reference: the two halves of the sampling window
    h0, h1 = ek_half[0] / max(1, (n_snap + 1) // 2), ek_half[1] / max(1, n_snap // 2)
    kk = np.arange(1, int(kc) + 1)
    half_diff = (h1[kk] - h0[kk]) / ek_f[kk]
    filt_state = pipe.half_state(pipe.ubar).cpu().numpy().astype(np.complex128)
    # the cost: one continuous device-controlled run, as the LES runs are timed
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dns.run_device(dns.t + t_cost, **run_kw)
    torch.cuda.synchronize()
    per_time = (time.perf_counter() - t0) / t_cost
    info = {"gpu_s_per_unit_time": per_time, "simulated_time": t_spin + n_snap * snap_every,
            "filtered_Ek": ek_f.tolist(), "filtered_2nuZ": 2 * nu * float(np.mean(zf)),
            "filtered_Ek_halves": [h0.tolist(), h1.tolist()],
            "filtered_Ek_half_diff_max": float(np.abs(half_diff).max()),
            "scales": {"g_rms": g_rms, "tau_rms": tau_rms},
            **turbulence_scales(float(np.mean(e_dns)), float(np.mean(z_dns)), ek_dns, nu,
                                n_dns)}
    log(f"DNS: {n_snap} filtered snapshots; {per_time:.3f} GPU s per unit time; "
        f"Re_lambda {info['re_lambda']:.1f}, k_max eta {info['kmax_eta']:.2f}, "
        f"large-eddy time {info['t_eddy']:.2f}; filtered E(k) second half vs first half of "
        f"the window: max {info['filtered_Ek_half_diff_max']:.3f}")
    return {"info": info, "X": X, "T": T, "U": U, "filt_state": filt_state}


# ------------------------------------------------------------------ a-priori energy transfer
class Shells:
    """Shell index (int)(|k| + 0.5) (the spectrum's binning) and Hermitian weight / L^6 of
    every mode of an L^3 half spectrum (NumPy rfftn layout)."""

    def __init__(self, L: int, nshell: int) -> None:
        k = np.fft.fftfreq(L, 1.0 / L)
        kz = np.arange(L // 2 + 1, dtype=np.float64)
        k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2
        self.idx = (np.sqrt(k2) + 0.5).astype(np.int64).ravel()
        w = np.full(L // 2 + 1, 2.0)
        w[0] = w[-1] = 1.0
        self.w = np.broadcast_to(w / float(L) ** 6, k2.shape).ravel()
        self.L, self.nshell = L, nshell

    def transfer(self, u: np.ndarray, s: np.ndarray) -> np.ndarray:
        """Per shell, sum of w Re(conj(u_hat) . s_hat) / L^6 over the three components: the
        rate of change of E(k) that the source s drives (sum over k = d/dt of <|u|^2>/2)."""
        d = np.zeros(u.shape[1:], dtype=np.float64)
        for c in range(3):
            d += (u[c].real.astype(np.float64) * s[c].real + u[c].imag.astype(np.float64)
                  * s[c].imag)
        return np.bincount(self.idx, weights=d.ravel() * self.w,
                           minlength=self.nshell)[: self.nshell]


def target_source(probe: LESSolver, stress: Any, tau_rms: float) -> Any:
    """Projected divergence source of a point-major fp32 stress [L^3][8] (channels 11, 22, 33,
    12, 13, 23, in units of tau_rms), through the LES's own source kernels; returns the source
    in the state layout, (3, L, L, pitch)."""
    import torch

    L = probe.n
    if getattr(probe, "_ybuf", None) is None:
        probe._ybuf = torch.zeros((L ** 3, C), dtype=torch.float32, device="cuda")
    probe._ybuf[:, :TAU_C].copy_(stress)
    les_op("tau_from_net", None, probe.tau, L=L, scale=tau_rms, in2=probe._ybuf)
    fft3(probe.tau, L, 3, inverse=False)
    les_op("source", probe.tau, probe.src, L=L, pitch=probe.buf.pitch)
    return probe.src


def _half(src: Any, L: int) -> np.ndarray:
    return src[..., : L // 2 + 1].cpu().numpy()


def fit_shell_gain(t_target: np.ndarray, t_model: np.ndarray, kmax: int,
                   lo: float = 0.25, hi: float = 4.0) -> np.ndarray:
    """Per-shell factor target / model on shells 1..kmax where both transfers have the same sign
    and the model's is not negligible (1e-3 of the largest target magnitude), clipped to
    [lo, hi]; 1 elsewhere."""
    g = np.ones_like(t_target)
    k = np.arange(len(g))
    ok = (k >= 1) & (k <= kmax) & (np.sign(t_target) == np.sign(t_model)) & \
        (np.abs(t_model) > 1e-3 * np.abs(t_target).max())
    g[ok] = np.clip(t_target[ok] / t_model[ok], lo, hi)
    return g


def apriori_transfer(net: SgsNet, U: Any, T: Any, n_val: int, n_cal: int, cfg_les: Config,
                     kc: float, g_rms: float, tau_rms: float, models: list[dict],
                     gain_models: tuple[str, ...], log: Any) -> dict[str, Any]:
    """A-priori subgrid energy transfer T(k) of the filtered-DNS stress and of each model on
    two disjoint windows: ``n_cal`` snapshots spread over the training snapshots (calibration)
    and the last ``n_val`` (held out from training).  For each key of ``gain_models`` a
    per-shell factor is fitted on the calibration window and evaluated on the held-out one,
    both as the product with the model's held-out transfer and through the source kernel's
    own per-shell factor."""
    import torch

    n_snap = U.shape[0]
    n_train = n_snap - n_val
    L = cfg_les.n
    cal = np.unique(np.linspace(0, n_train - 1, max(1, min(n_cal, n_train))).round()
                    .astype(np.int64))
    held = np.arange(n_train, n_snap)
    probe = LESSolver(cfg_les, "none")
    sh = Shells(L, probe.diag.nshell)
    kk = np.arange(1, int(kc) + 1)

    def solver(m: dict, gain: Any = None) -> LESSolver:
        return LESSolver(cfg_les, m["model"], net=net if m["model"] == "net" else None, kc=kc,
                         g_scale=1.0 / g_rms, tau_scale=tau_rms * m["gain"],
                         backscatter=m["bs"], mix_cs=m["mix_cs"], shell_gain=gain)

    def spectra(idx: np.ndarray, solvers: dict[str, LESSolver]) -> dict[str, np.ndarray]:
        tot = {k: np.zeros(sh.nshell) for k in ("target", *solvers)}
        for j, i in enumerate(idx):
            if j and j % 200 == 0:
                log(f"  a-priori transfer: {j}/{len(idx)} snapshots")
            u = U[int(i)].cpu().numpy()
            tot["target"] += sh.transfer(u, _half(target_source(probe, T[int(i)], tau_rms), L))
            for key, s in solvers.items():
                s.set_state_device(U[int(i)])
                s.update_source()
                tot[key] += sh.transfer(u, _half(s.src, L))
        return {k: v / len(idx) for k, v in tot.items()}

    active = [m for m in models if m["model"] != "none"]
    solvers = {m["key"]: solver(m) for m in active}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    spec = {"calibration": spectra(cal, solvers), "held_out": spectra(held, solvers)}
    for s in solvers.values():
        s.close()
    gains, kernel_check = {}, {}
    for key in gain_models:
        m = next(m for m in active if m["key"] == key)
        g = fit_shell_gain(spec["calibration"]["target"], spec["calibration"][key], int(kc))
        gains[key] = g
        gk = f"{key}_shellgain"
        spec["calibration"][gk] = g * spec["calibration"][key]
        spec["held_out"][gk] = g * spec["held_out"][key]
        # the same factor applied by the source kernel, on the held-out window
        sg = solver(m, g)
        via = spectra(held, {gk: sg})[gk]
        sg.close()
        ref = spec["held_out"][gk]
        kernel_check[gk] = float(np.abs(via - ref).max() / np.abs(ref).max())
    probe.close()
    wall = time.perf_counter() - t0

    def errors(w: dict[str, np.ndarray]) -> dict[str, Any]:
        tt = w["target"][kk]
        scale = np.abs(tt).max()
        out = {}
        for key, tm in w.items():
            if key == "target":
                continue
            rel = (tm[kk] - tt) / np.abs(tt)
            out[key] = {"rel_err": rel.tolist(), "rel_err_max": float(np.abs(rel).max()),
                        "rel_err_mean": float(np.abs(rel).mean()),
                        "err_over_peak_max": float(np.abs(tm[kk] - tt).max() / scale),
                        "total_ratio": float(tm[1:].sum() / w["target"][1:].sum())}
        return out

    res = {"calibration_snapshots": cal.tolist(), "held_out_snapshots": [int(held[0]),
                                                                          int(held[-1])],
           "wall_s": wall,
           "T": {w: {k: v.tolist() for k, v in d.items()} for w, d in spec.items()},
           "errors": {w: errors(d) for w, d in spec.items()},
           "shell_gain": {k: g.tolist() for k, g in gains.items()},
           "kernel_check": kernel_check}
    held_err = res["errors"]["held_out"]
    log("a-priori SGS transfer, held-out window (shells 1..%d): " % int(kc) +
        "; ".join(f"{k} max {v['rel_err_max']:.3f} mean {v['rel_err_mean']:.3f}"
                  for k, v in held_err.items()))
    return res


def _les_spectrum(make: Any, t_samples: np.ndarray, run_kw: dict) -> tuple[np.ndarray,
                                                                          list[float]]:
    """Time-averaged E(k) and the enstrophy samples of one LES run from make() that stops at
    each sample time, landing on it as simulate.run lands on its callback times."""
    les = make()
    ek = np.zeros(les.diag.nshell)
    zs: list[float] = []
    for ts in t_samples:
        if ts > les.t + 1e-9:
            les.run_device(float(ts), **run_kw)
        sn = les.state_native()
        ek += _spectrum(les, sn)
        zs.append(les.diag.enstrophy(sn))
    les.close()
    return ek / len(t_samples), zs


def fit_shell_gain_aposteriori(make: Any, ek_f: np.ndarray, kc: float, t_samples: np.ndarray,
                               run_kw: dict, iters: int, kmin: int, alpha: float,
                               log: Any) -> dict[str, Any]:
    """Per-shell factor on a model's source fitted a posteriori on one averaging window:
    starting from 1, g(k) <- g(k) (E_LES(k) / E_ref(k))^alpha on shells kmin..kc (each update
    clipped to [0.8, 1.25]; shells past kc follow shell kc, shells below kmin keep 1), for a
    model whose transfer drains every shell (a larger factor, more drain, less E(k)).  Returns
    the iterate with the smallest maximum |E(k) error| over shells 1..kc, and the history.
    ``make(g)`` builds the LES with factor g."""
    kk = np.arange(1, int(kc) + 1)
    ks = np.arange(int(kmin), int(kc) + 1)
    g = None
    best: dict[str, Any] = {}
    hist = []
    for it in range(iters + 1):
        ek, _ = _les_spectrum(lambda: make(g), t_samples, run_kw)
        err = (ek[kk] - ek_f[kk]) / ek_f[kk]
        emax = float(np.abs(err).max())
        hist.append({"iter": it, "err_max": emax, "err_mean": float(np.abs(err).mean())})
        log(f"  a-posteriori fit, iteration {it}: E(k) error max {emax:.3f} "
            f"mean {hist[-1]['err_mean']:.3f}")
        if not best or emax < best["err_max"]:
            best = {"gain": (np.ones(len(ek)) if g is None else g).tolist(), "err_max": emax,
                    "err": err.tolist(), "iter": it}
        if it == iters:
            break
        g = np.ones(len(ek)) if g is None else g.copy()
        g[ks] *= np.clip((ek[ks] / ek_f[ks]) ** alpha, 0.8, 1.25)
        g[int(kc) + 1:] = g[int(kc)]
    return {**best, "history": hist}


def _train(X: Any, T: Any, n_les: int, n_val: int, train_steps: int, batch: int, lr: float,
           seed: int, log: Any) -> tuple[SgsNet, dict[str, Any]]:
    """Adam at batch `batch` over random snapshots (cosine learning-rate decay); the last
    n_val snapshots are held out."""
    import torch

    L3 = n_les ** 3
    net = SgsNet(n=n_les, seed=seed + 1, lr=lr)
    rng = np.random.default_rng(seed + 2)
    n_train = X.shape[0] - n_val
    Xb = torch.zeros((batch * L3, C), dtype=torch.bfloat16, device="cuda")
    Tb = torch.zeros((batch * L3, TAU_C), dtype=torch.float32, device="cuda")

    def gather(idx: np.ndarray) -> None:
        ii = torch.from_numpy(idx).cuda()
        Xb.view(batch, L3, C).copy_(X.index_select(0, ii))
        Tb.view(batch, L3, TAU_C).copy_(T.index_select(0, ii))

    def val_loss() -> float:
        tot = 0.0
        for b in range(0, n_val // batch * batch, batch):
            gather(np.arange(n_train + b, n_train + b + batch))
            tot += net.eval_loss(Xb, Tb, batch)
        return tot / max(1, n_val // batch)

    hist = []
    v0 = val_loss()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for step in range(train_steps):
        net.lr = lr * 0.5 * (1.0 + math.cos(math.pi * step / max(1, train_steps)))
        gather(rng.choice(n_train, batch, replace=False))
        loss = torch.zeros(1, dtype=torch.float64, device="cuda")
        net.train_step(Xb, Tb, batch, loss)
        if step % 200 == 0 or step == train_steps - 1:
            hist.append((step, loss.item()))
            log(f"  train step {step}: loss {hist[-1][1]:.4f}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    v1 = val_loss()
    info = {"steps": train_steps, "batch": batch, "wall_s": wall,
            "ms_per_step": 1e3 * wall / max(1, train_steps), "loss_history": hist,
            "val_loss_before": v0, "val_loss_after": v1,
            "flop_per_step": 3 * 6 * 2 * batch * L3 * 27 * C * C}
    info["tflops"] = info["flop_per_step"] / (info["ms_per_step"] * 1e-3) / 1e12
    log(f"training: {train_steps} steps in {wall:.1f} s; validation loss {v0:.4f} -> {v1:.4f} "
        f"(normalised; 1 = predicting zero)")
    return net, info


def _save_model(path: str, net: SgsNet, dns: dict[str, Any], filt_state: np.ndarray,
                training: dict[str, Any], config: dict[str, Any],
                apriori: dict[str, Any] | None) -> None:
    st = net.state()
    np.savez(path, w=st["w"], b=st["b"], filt_state=filt_state, dns=json.dumps(dns),
             training=json.dumps(training), config=json.dumps(config),
             apriori=json.dumps(apriori or {}))


def _load_model(path: str, n_les: int) -> tuple[SgsNet, dict, np.ndarray, dict, dict, dict]:
    import torch

    z = np.load(path, allow_pickle=False)
    net = SgsNet(n=n_les)
    net.w.copy_(torch.from_numpy(z["w"]))
    net.b.copy_(torch.from_numpy(z["b"]))
    net._adam(lr=0.0)  # bf16 copies of the loaded weights
    config = json.loads(str(z["config"])) if "config" in z.files else {}
    apriori = json.loads(str(z["apriori"])) if "apriori" in z.files else {}
    return (net, json.loads(str(z["dns"])), z["filt_state"], json.loads(str(z["training"])),
            config, apriori)


def les_models(backscatter: tuple[float, ...], mixed: tuple[tuple[float, ...], ...]
               ) -> list[dict[str, Any]]:
    """The LES runs: no model, Smagorinsky, the network at each backscatter factor b, and mixed
    models b:c[:g] (the network at backscatter b, its whole stress times g, plus a Smagorinsky
    stress of constant c)."""
    def m(key: str, model: str, bs: float = 1.0, mix_cs: float = 0.0, gain: float = 1.0
          ) -> dict[str, Any]:
        return {"key": key, "model": model, "bs": bs, "mix_cs": mix_cs, "gain": gain}

    out = [m("none", "none"), m("smagorinsky", "smagorinsky")]
    out += [m(f"net_bs{b:g}", "net", b) for b in backscatter]
    for t in mixed:
        b, c, g = (*t, 1.0)[:3]
        out.append(m(f"net_bs{b:g}_cs{c:g}" + (f"_x{g:g}" if g != 1.0 else ""), "net", b, c, g))
    return out


def closure_study(*, n_dns: int = 256, n_les: int = 64, kc: float = 21.0, nu: float = 1.6e-3,
                  power: float = 0.1, band: tuple[float, float] = (2.0, 3.0),
                  t_spin: float = 10.0, n_snap: int = 192, snap_every: float = 0.1,
                  n_val: int = 16, train_steps: int = 2000, batch: int = 8, lr: float = 1e-3,
                  t_les: float = 30.0, t_avg: float = 27.0, t_cost: float = 5.0, seed: int = 0,
                  backscatter: tuple[float, ...] = (1.0, 0.5, 0.0),
                  mixed: tuple[tuple[float, ...], ...] = ((1.0, 0.04),),
                  n_cal: int | None = None, gain_models: tuple[str, ...] = ("net_bs1",),
                  fit_models: tuple[str, ...] = (), fit_iters: int = 8,
                  fit_t_les: float = 60.0, fit_t_avg: float = 57.0, fit_kmin: int = 4,
                  fit_alpha: float = 1.0,
                  save_model: str | None = None, load_model: str | None = None,
                  log: Any = print) -> dict[str, Any]:
    import torch

    out: dict[str, Any] = {"config": {k: v for k, v in locals().items()
                                      if k not in ("log", "torch")}}
    cfg_dns = Config(n=n_dns, dim=3, nu=nu, force_klo=band[0], force_khi=band[1],
                     dtype="float32")
    cfg_les = cfg_dns.with_(n=n_les)
    run_kw = dict(cfl=0.5, dt_max=0.05, target_power=power, record_start=False)

    models = les_models(backscatter, mixed)
    if load_model:
        (net, out["dns"], filt_state, out["training"], model_cfg,
         out["apriori"]) = _load_model(load_model, n_les)
        # the DNS, training and a-priori settings are the saved run's, not this call's defaults
        out["config"] = {**{k: out["config"][k] for k in ("t_les", "t_avg", "backscatter", "mixed",
                                                          "load_model", "kc", "nu", "power",
                                                          "band", "n_les", "fit_models",
                                                          "fit_iters", "fit_t_les", "fit_t_avg",
                                                          "fit_kmin", "fit_alpha")},
                         "model_config": model_cfg}
        log(f"loaded the trained network and the filtered-DNS data from {load_model}")
    else:
        d = _dns_and_snapshots(cfg_dns, n_les, kc, nu, t_spin, n_snap, snap_every, t_cost,
                               seed, run_kw, log)
        out["dns"], filt_state = d["info"], d["filt_state"]
        net, out["training"] = _train(d["X"], d["T"], n_les, n_val, train_steps, batch, lr,
                                      seed, log)
        g_rms, tau_rms = out["dns"]["scales"]["g_rms"], out["dns"]["scales"]["tau_rms"]
        out["apriori"] = apriori_transfer(net, d["U"], d["T"], n_val,
                                          n_val if n_cal is None else n_cal, cfg_les, kc, g_rms,
                                          tau_rms, models, gain_models, log)
        del d
        torch.cuda.empty_cache()
        if save_model:
            _save_model(save_model, net, out["dns"], filt_state, out["training"],
                        out["config"], out["apriori"])
    g_rms, tau_rms = out["dns"]["scales"]["g_rms"], out["dns"]["scales"]["tau_rms"]
    ek_f = np.array(out["dns"]["filtered_Ek"])
    dns_per_time = out["dns"]["gpu_s_per_unit_time"]
    # the network with its a-priori per-shell factor, fitted on the calibration window
    for key, g in out.get("apriori", {}).get("shell_gain", {}).items():
        if key not in gain_models:
            continue
        base = next((m for m in models if m["key"] == key), None)
        if base is not None:
            models.append({**base, "key": f"{key}_shellgain", "shell_gain": np.asarray(g)})

    def builder(mdl: dict[str, Any]) -> Any:
        def make(shell_gain: Any = None) -> LESSolver:
            s = LESSolver(cfg_les, mdl["model"], net=net if mdl["model"] == "net" else None,
                          kc=kc, g_scale=1.0 / g_rms, tau_scale=tau_rms * mdl["gain"],
                          backscatter=mdl["bs"], mix_cs=mdl["mix_cs"],
                          shell_gain=mdl.get("shell_gain") if shell_gain is None else shell_gain)
            s.set_state(filt_state)
            return s
        return make

    # a-posteriori per-shell factors, fitted on their own window (which ends before the
    # reported window starts, so the report is on simulated time the fit never saw)
    fits = {}
    fit_samples = np.arange(fit_t_les - fit_t_avg, fit_t_les + 1e-9, 0.25)
    for key in fit_models:
        base = next((m for m in models if m["key"] == key), None)
        if base is None:
            raise ValueError(f"fit model {key!r} is not one of the LES runs")
        log(f"a-posteriori per-shell factor for {key}: window [{fit_samples[0]:g}, "
            f"{fit_samples[-1]:g}]")
        f = fit_shell_gain_aposteriori(builder(base), ek_f, kc, fit_samples, run_kw, fit_iters,
                                       fit_kmin, fit_alpha, log)
        f["window"] = [float(fit_samples[0]), float(fit_samples[-1])]
        f["disjoint_from_report"] = bool(fit_t_les <= t_les - t_avg)
        fits[key] = f
        models.append({**base, "key": f"{key}_fit", "shell_gain": np.asarray(f["gain"])})
    if fits:
        out["aposteriori_fit"] = fits

    # ---------------------------------------------------------------- LES runs
    # Each model: one continuous device-controlled run of t_les for its GPU time per unit
    # simulated time (as the DNS is timed), and one run from the same state that stops at
    # the spectrum sample times of the averaging window (every 0.25).
    res = {}
    t_samples = np.arange(t_les - t_avg, t_les + 1e-9, 0.25)
    for mdl in models:
        key, model, bs, mix_cs, gain = (mdl["key"], mdl["model"], mdl["bs"], mdl["mix_cs"],
                                        mdl["gain"])
        make = builder(mdl)
        les = make()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        steps = les.run_device(t_les, **run_kw)["steps"]
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        les.close()
        ek, zs = _les_spectrum(make, t_samples, run_kw)
        k = np.arange(1, int(kc) + 1)
        err = (ek[k] - ek_f[k]) / ek_f[k]
        r = {"model": model, "backscatter": bs, "mix_cs": mix_cs, "net_gain": gain,
             "shell_gain": "shell_gain" in mdl,
             "steps": steps, "wall_s": wall,
             "gpu_s_per_unit_time": wall / t_les,
             "speedup_vs_dns": dns_per_time / (wall / t_les), "Ek": ek.tolist(),
             "Ek_rel_err": err.tolist(), "Ek_rel_err_max": float(np.abs(err).max()),
             "Ek_rel_err_mean": float(np.abs(err).mean()),
             "2nuZ": 2 * nu * float(np.mean(zs)), "finite": bool(np.all(np.isfinite(ek))),
             "large_eddy_times": t_les / out["dns"]["t_eddy"]}
        r["2nuZ_rel_err"] = abs(r["2nuZ"] - out["dns"]["filtered_2nuZ"]) / \
            out["dns"]["filtered_2nuZ"]
        res[key] = r
        log(f"LES {key:12s}: {steps} steps {wall:.2f} s, x{r['speedup_vs_dns']:.1f} vs DNS; "
            f"E(k) error max {r['Ek_rel_err_max']:.3f} mean {r['Ek_rel_err_mean']:.3f}; "
            f"2nuZ error {r['2nuZ_rel_err']:.3f}")
    out["les"] = res
    return out


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--quick", action="store_true", help="short run (minutes -> seconds)")
    ap.add_argument("--out", default=None)
    ap.add_argument("--train-steps", type=int, default=None)
    ap.add_argument("--snapshots", type=int, default=None)
    ap.add_argument("--snap-every", type=float, default=None)
    ap.add_argument("--validation", type=int, default=None, help="held-out snapshots")
    ap.add_argument("--t-spin", type=float, default=None)
    ap.add_argument("--t-les", type=float, default=None)
    ap.add_argument("--t-avg", type=float, default=None)
    ap.add_argument("--backscatter", default=None,
                    help="comma-separated backscatter factors of the network LES runs")
    ap.add_argument("--mixed", default=None,
                    help="comma-separated b:c[:g] triples: the network (backscatter factor b, "
                         "stress times g, default 1) plus a Smagorinsky stress of constant c")
    ap.add_argument("--calibration", type=int, default=None,
                    help="training snapshots in the a-priori calibration window (default: as "
                         "many as held out)")
    ap.add_argument("--gain-models", default=None,
                    help="comma-separated LES model keys that get an a-priori per-shell factor")
    ap.add_argument("--fit-models", default=None,
                    help="comma-separated LES model keys that get an a-posteriori per-shell "
                         "factor, fitted on the --fit-t-les/--fit-t-avg window")
    ap.add_argument("--fit-iters", type=int, default=None)
    ap.add_argument("--fit-t-les", type=float, default=None)
    ap.add_argument("--fit-t-avg", type=float, default=None)
    ap.add_argument("--fit-kmin", type=int, default=None)
    ap.add_argument("--fit-alpha", type=float, default=None)
    ap.add_argument("--save-model", default=None, help="save the network and the DNS data")
    ap.add_argument("--load-model", default=None, help="skip DNS and training: load them")
    a = ap.parse_args(argv)
    kw: dict[str, Any] = {}
    if a.quick:
        kw.update(t_spin=3.0, n_snap=40, snap_every=0.1, n_val=8, train_steps=200, t_les=6.0,
                  t_avg=3.0)
    for key, val in (("train_steps", a.train_steps), ("n_snap", a.snapshots),
                     ("snap_every", a.snap_every), ("n_val", a.validation),
                     ("t_spin", a.t_spin), ("t_les", a.t_les), ("t_avg", a.t_avg),
                     ("n_cal", a.calibration), ("fit_iters", a.fit_iters),
                     ("fit_t_les", a.fit_t_les), ("fit_t_avg", a.fit_t_avg),
                     ("fit_kmin", a.fit_kmin), ("fit_alpha", a.fit_alpha),
                     ("save_model", a.save_model), ("load_model", a.load_model)):
        if val is not None:
            kw[key] = val
    if a.fit_models:
        kw["fit_models"] = tuple(x for x in a.fit_models.split(",") if x)
    if a.gain_models is not None:
        kw["gain_models"] = tuple(x for x in a.gain_models.split(",") if x)
    if a.backscatter:
        kw["backscatter"] = tuple(float(x) for x in a.backscatter.split(","))
    if a.mixed:
        kw["mixed"] = tuple(tuple(float(v) for v in x.split(":")) for x in a.mixed.split(","))
    res = closure_study(log=lambda m: print(m, flush=True), **kw)
    txt = json.dumps(res)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt)
    print(json.dumps({k: {kk: vv for kk, vv in v.items()
                          if kk not in ("Ek", "filtered_Ek", "loss_history", "T")}
                      if isinstance(v, dict) else v for k, v in res.items()}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
