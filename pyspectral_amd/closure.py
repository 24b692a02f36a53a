"""Learned-closure LES study: forced 3D DNS -> filtered snapshots -> network training -> coarse
LES with no model, Smagorinsky and the network, judged against the filtered DNS.

    python -m pyspectral_amd.closure [--quick] [--out results.json]

Workload (defaults): DNS 256^3 fp32, constant-power band forcing on 2 <= |k| <= 3 with
P = 0.1, nu = 2.5e-3; snapshots box-filtered (k_c = 21) onto the 64^3 LES grid; six 3x3x3
convolutions of 32 channels trained with Adam at batch 8 on the GPU; 64^3 LES runs from the
filtered DNS field with the same forcing.  Reported: the time-averaged E(k) error of each LES
against the time-averaged filtered DNS spectrum for 1 <= k <= k_c, the resolved dissipation
2 nu Z, and GPU time per unit simulated time of the DNS and each LES.  Every kernel is
libpsx's; torch provides device memory and the batch gathers.
"""

from __future__ import annotations

import argparse
import json
import time
from typing import Any

import numpy as np

from pyspectral_amd import initial
from pyspectral_amd.backends.hip_fused import HipSolver
from pyspectral_amd.config import Config
from pyspectral_amd.les import C, TAU_C, FilterPipeline, LESSolver, SgsNet, rms


def _spectrum(solver: Any, half_state: Any) -> np.ndarray:
    return solver.diag.spectrum(half_state.contiguous())[1]


def closure_study(*, n_dns: int = 256, n_les: int = 64, kc: float = 21.0, nu: float = 2.5e-3,
                  power: float = 0.1, band: tuple[float, float] = (2.0, 3.0),
                  t_spin: float = 10.0, n_snap: int = 192, snap_every: float = 0.1,
                  n_val: int = 16, train_steps: int = 2000, batch: int = 8, lr: float = 1e-3,
                  t_les: float = 30.0, t_avg: float = 10.0, seed: int = 0,
                  log: Any = print) -> dict[str, Any]:
    import torch

    out: dict[str, Any] = {"config": {k: v for k, v in locals().items()
                                      if k not in ("log", "torch")}}
    cfg_dns = Config(n=n_dns, dim=3, nu=nu, force_klo=band[0], force_khi=band[1],
                     dtype="float32")
    cfg_les = cfg_dns.with_(n=n_les)
    run_kw = dict(cfl=0.5, dt_max=0.05, target_power=power, record_start=False)

    # ---------------------------------------------------------------- DNS spin-up
    dns = HipSolver(cfg_dns)
    dns.set_state(initial.to_spectral(initial.random_3d(n_dns, k0=2.5, seed=seed,
                                                        energy=0.5), 3))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    spin = dns.run_device(t_spin, **run_kw)
    torch.cuda.synchronize()
    dns_wall = time.perf_counter() - t0
    dns_time = t_spin
    log(f"DNS {n_dns}^3 spin-up to t={dns.t:.2f}: {spin['steps']} steps, {dns_wall:.1f} s")

    # ---------------------------------------------------------------- snapshots
    L3 = n_les ** 3
    pipe = FilterPipeline(n_dns, n_les, kc)
    X = torch.zeros((n_snap, L3, C), dtype=torch.bfloat16, device="cuda")
    T = torch.zeros((n_snap, L3, TAU_C), dtype=torch.float32, device="cuda")
    pipe(dns.buf.u0_rows, dns.buf.pitch, X[0], T[0])  # unscaled pilot for the scales
    g_rms = rms(torch.view_as_real(pipe.grad[2 * L3: 6 * L3]).reshape(-1), 2, 2)
    tau_rms = rms(T[0].reshape(-1), TAU_C, 6)
    log(f"scales: velocity-gradient rms {g_rms:.4g}, stress rms {tau_rms:.4g}")
    les_probe = LESSolver(cfg_les, "none")
    ek_f = np.zeros(les_probe.diag.nshell)
    zf = []
    for i in range(n_snap):
        t0 = time.perf_counter()
        dns.run_device(dns.t + snap_every, **run_kw)
        torch.cuda.synchronize()
        dns_wall += time.perf_counter() - t0
        dns_time += snap_every
        ubar = pipe(dns.buf.u0_rows, dns.buf.pitch, X[i], T[i], g_scale=1.0 / g_rms,
                    tau_scale=1.0 / tau_rms)
        hs = pipe.half_state(ubar)
        ek_f += _spectrum(les_probe, hs)
        zf.append(les_probe.diag.enstrophy(hs.contiguous()))
    ek_f /= n_snap
    filt_state = pipe.half_state(pipe.ubar).cpu().numpy().astype(np.complex128)
    dns_per_time = dns_wall / dns_time
    out["dns"] = {"wall_s": dns_wall, "simulated_time": dns_time,
                  "gpu_s_per_unit_time": dns_per_time,
                  "filtered_Ek": ek_f.tolist(), "filtered_2nuZ": 2 * nu * float(np.mean(zf)),
                  "scales": {"g_rms": g_rms, "tau_rms": tau_rms}}
    log(f"DNS: {dns_time:.1f} time units in {dns_wall:.1f} s "
        f"({dns_per_time:.3f} GPU s per unit time); {n_snap} filtered snapshots")
    del dns

    # ---------------------------------------------------------------- training
    net = SgsNet(n=n_les, seed=seed + 1, lr=lr)
    rng = np.random.default_rng(seed + 2)
    n_train = n_snap - n_val
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
        gather(rng.choice(n_train, batch, replace=False))
        loss = torch.zeros(1, dtype=torch.float64, device="cuda")
        net.train_step(Xb, Tb, batch, loss)
        if step % 200 == 0 or step == train_steps - 1:
            hist.append((step, loss.item()))
            log(f"  train step {step}: loss {hist[-1][1]:.4f}")
    torch.cuda.synchronize()
    train_wall = time.perf_counter() - t0
    v1 = val_loss()
    out["training"] = {"steps": train_steps, "batch": batch, "wall_s": train_wall,
                       "ms_per_step": 1e3 * train_wall / max(1, train_steps),
                       "loss_history": hist, "val_loss_before": v0, "val_loss_after": v1,
                       "flop_per_step": 3 * 6 * 2 * batch * L3 * 27 * C * C}
    out["training"]["tflops"] = (out["training"]["flop_per_step"] /
                                 (out["training"]["ms_per_step"] * 1e-3) / 1e12)
    log(f"training: {train_steps} steps in {train_wall:.1f} s; validation loss "
        f"{v0:.4f} -> {v1:.4f} (normalised; 1 = predicting zero)")

    # ---------------------------------------------------------------- LES runs
    # Each LES runs on the device controller, like the DNS; it stops at the spectrum sample
    # times of the averaging window (every 0.25), landing on them as simulate.run lands on
    # its callback times.
    res = {}
    t_samples = np.arange(t_les - t_avg, t_les + 1e-9, 0.25)
    for model in ("none", "smagorinsky", "net"):
        les = LESSolver(cfg_les, model, net=net if model == "net" else None, kc=kc,
                        g_scale=1.0 / g_rms, tau_scale=tau_rms)
        les.set_state(filt_state)
        ek = np.zeros(les.diag.nshell)
        zs: list[float] = []
        steps = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for ts in t_samples:
            if ts > les.t + 1e-9:
                steps += les.run_device(float(ts), **run_kw)["steps"]
            sn = les.state_native()
            ek += _spectrum(les, sn)
            zs.append(les.diag.enstrophy(sn))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ek /= len(t_samples)
        k = np.arange(1, int(kc) + 1)
        err = np.abs(ek[k] - ek_f[k]) / ek_f[k]
        res[model] = {"steps": steps, "wall_s": wall, "gpu_s_per_unit_time": wall / t_les,
                      "speedup_vs_dns": dns_per_time / (wall / t_les), "Ek": ek.tolist(),
                      "Ek_rel_err_max": float(err.max()), "Ek_rel_err_mean": float(err.mean()),
                      "2nuZ": 2 * nu * float(np.mean(zs)),
                      "finite": bool(np.all(np.isfinite(ek)))}
        res[model]["2nuZ_rel_err"] = abs(res[model]["2nuZ"] - out["dns"]["filtered_2nuZ"]) / \
            out["dns"]["filtered_2nuZ"]
        log(f"LES {model:12s}: {steps} steps {wall:.1f} s, x{res[model]['speedup_vs_dns']:.1f} "
            f"vs DNS; E(k) error max {err.max():.3f} mean {err.mean():.3f}; 2nuZ error "
            f"{res[model]['2nuZ_rel_err']:.3f}")
        les.close()
    out["les"] = res
    return out


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--quick", action="store_true", help="short run (minutes -> seconds)")
    ap.add_argument("--out", default=None)
    ap.add_argument("--train-steps", type=int, default=None)
    ap.add_argument("--snapshots", type=int, default=None)
    a = ap.parse_args(argv)
    kw: dict[str, Any] = {}
    if a.quick:
        kw.update(t_spin=3.0, n_snap=40, snap_every=0.1, n_val=8, train_steps=200, t_les=6.0,
                  t_avg=3.0)
    if a.train_steps is not None:
        kw["train_steps"] = a.train_steps
    if a.snapshots is not None:
        kw["n_snap"] = a.snapshots
    res = closure_study(log=lambda m: print(m, flush=True), **kw)
    txt = json.dumps(res)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt)
    print(json.dumps({k: {kk: vv for kk, vv in v.items() if kk not in ("Ek", "filtered_Ek")}
                      if isinstance(v, dict) else v for k, v in res.items()}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
