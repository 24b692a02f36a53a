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
large-eddy time, and GPU time per unit simulated time of the DNS and each LES.  Every kernel
is libpsx's; torch provides device memory and the batch gathers.
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
from pyspectral_amd.les import C, TAU_C, FilterPipeline, LESSolver, SgsNet, rms


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
    return {"info": info, "X": X, "T": T, "filt_state": filt_state}


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
                training: dict[str, Any], config: dict[str, Any]) -> None:
    st = net.state()
    np.savez(path, w=st["w"], b=st["b"], filt_state=filt_state, dns=json.dumps(dns),
             training=json.dumps(training), config=json.dumps(config))


def _load_model(path: str, n_les: int) -> tuple[SgsNet, dict, np.ndarray, dict, dict]:
    import torch

    z = np.load(path, allow_pickle=False)
    net = SgsNet(n=n_les)
    net.w.copy_(torch.from_numpy(z["w"]))
    net.b.copy_(torch.from_numpy(z["b"]))
    net._adam(lr=0.0)  # bf16 copies of the loaded weights
    config = json.loads(str(z["config"])) if "config" in z.files else {}
    return (net, json.loads(str(z["dns"])), z["filt_state"], json.loads(str(z["training"])),
            config)


def closure_study(*, n_dns: int = 256, n_les: int = 64, kc: float = 21.0, nu: float = 1.6e-3,
                  power: float = 0.1, band: tuple[float, float] = (2.0, 3.0),
                  t_spin: float = 10.0, n_snap: int = 192, snap_every: float = 0.1,
                  n_val: int = 16, train_steps: int = 2000, batch: int = 8, lr: float = 1e-3,
                  t_les: float = 30.0, t_avg: float = 27.0, t_cost: float = 5.0, seed: int = 0,
                  backscatter: tuple[float, ...] = (1.0, 0.5, 0.0),
                  mixed: tuple[tuple[float, ...], ...] = ((1.0, 0.04),),
                  save_model: str | None = None, load_model: str | None = None,
                  log: Any = print) -> dict[str, Any]:
    import torch

    out: dict[str, Any] = {"config": {k: v for k, v in locals().items()
                                      if k not in ("log", "torch")}}
    cfg_dns = Config(n=n_dns, dim=3, nu=nu, force_klo=band[0], force_khi=band[1],
                     dtype="float32")
    cfg_les = cfg_dns.with_(n=n_les)
    run_kw = dict(cfl=0.5, dt_max=0.05, target_power=power, record_start=False)

    if load_model:
        net, out["dns"], filt_state, out["training"], model_cfg = _load_model(load_model, n_les)
        # the DNS and training settings are the saved run's, not this call's defaults
        out["config"] = {**{k: out["config"][k] for k in ("t_les", "t_avg", "backscatter", "mixed",
                                                          "load_model", "kc", "nu", "power",
                                                          "band", "n_les")},
                         "model_config": model_cfg}
        log(f"loaded the trained network and the filtered-DNS data from {load_model}")
    else:
        d = _dns_and_snapshots(cfg_dns, n_les, kc, nu, t_spin, n_snap, snap_every, t_cost,
                               seed, run_kw, log)
        out["dns"], filt_state = d["info"], d["filt_state"]
        net, out["training"] = _train(d["X"], d["T"], n_les, n_val, train_steps, batch, lr,
                                      seed, log)
        del d
        torch.cuda.empty_cache()
        if save_model:
            _save_model(save_model, net, out["dns"], filt_state, out["training"],
                        out["config"])
    g_rms, tau_rms = out["dns"]["scales"]["g_rms"], out["dns"]["scales"]["tau_rms"]
    ek_f = np.array(out["dns"]["filtered_Ek"])
    dns_per_time = out["dns"]["gpu_s_per_unit_time"]

    # ---------------------------------------------------------------- LES runs
    # Each model: one continuous device-controlled run of t_les for its GPU time per unit
    # simulated time (as the DNS is timed), and one run from the same state that stops at
    # the spectrum sample times of the averaging window (every 0.25), landing on them as
    # simulate.run lands on its callback times.
    models = [("none", "none", 1.0, 0.0, 1.0), ("smagorinsky", "smagorinsky", 1.0, 0.0, 1.0)]
    models += [(f"net_bs{b:g}", "net", b, 0.0, 1.0) for b in backscatter]
    # mixed models: the network (backscatter factor b, its whole stress times g) plus a
    # Smagorinsky stress of constant c
    for m in mixed:
        b, c, g = (*m, 1.0)[:3]
        models.append((f"net_bs{b:g}_cs{c:g}" + (f"_x{g:g}" if g != 1.0 else ""), "net", b, c,
                       g))
    res = {}
    t_samples = np.arange(t_les - t_avg, t_les + 1e-9, 0.25)
    for key, model, bs, mix_cs, gain in models:
        def make() -> LESSolver:
            s = LESSolver(cfg_les, model, net=net if model == "net" else None, kc=kc,
                          g_scale=1.0 / g_rms, tau_scale=tau_rms * gain, backscatter=bs,
                          mix_cs=mix_cs)
            s.set_state(filt_state)
            return s

        les = make()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        steps = les.run_device(t_les, **run_kw)["steps"]
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        les.close()
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
        ek /= len(t_samples)
        k = np.arange(1, int(kc) + 1)
        err = (ek[k] - ek_f[k]) / ek_f[k]
        r = {"model": model, "backscatter": bs, "mix_cs": mix_cs, "net_gain": gain,
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
                     ("save_model", a.save_model), ("load_model", a.load_model)):
        if val is not None:
            kw[key] = val
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
                          if kk not in ("Ek", "filtered_Ek", "loss_history")}
                      if isinstance(v, dict) else v for k, v in res.items()}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
