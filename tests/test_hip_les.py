"""LES closure kernels against the CPU oracle (oracle/les_oracle.py) (GPU only).

Complex FFTs against numpy; the filter pipeline (box filter, truncation, subgrid-stress
target, model input) against the oracle's FP64 restatement; the Smagorinsky source and a
whole LES step with it against the oracle solver with the same source; the MFMA convolutions
(forward, input gradient, weight gradient), the loss and Adam against FP64 on the same bf16
operands; and training that fits a teacher network.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from oracle import les_oracle as LO
from oracle import numpy_oracle as O
from pyspectral_amd import _lib, initial
from pyspectral_amd.backends import get
from pyspectral_amd.config import Config
from pyspectral_amd.les import (C, TAU_C, FilterPipeline, LESSolver, SgsNet, conv, fft3,
                                les_op, rms)

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def _bf16(a):
    return torch.from_numpy(np.asarray(a, dtype=np.float32)).to(torch.bfloat16)


@pytest.mark.parametrize("L", [64, 128])
def test_fft3_matches_numpy(L):
    rng = np.random.default_rng(L)
    a = rng.normal(size=(2, L, L, L)) + 1j * rng.normal(size=(2, L, L, L))
    t = torch.from_numpy(a.astype(np.complex64)).cuda()
    fft3(t, L, 2, inverse=False)
    err = _rel(t.cpu().numpy(), np.fft.fftn(a, axes=(1, 2, 3)))
    t2 = torch.from_numpy(a.astype(np.complex64)).cuda()
    fft3(t2, L, 2, inverse=True)
    err2 = _rel(t2.cpu().numpy(), np.fft.ifftn(a, axes=(1, 2, 3)) * L ** 3)
    print(f"  L={L} forward {err:.2e} inverse {err2:.2e}", flush=True)
    assert err < 2e-6 and err2 < 2e-6


def _state3(n, seed, k0=4):
    return initial.to_spectral(initial.random_3d(n, k0=k0, seed=seed), 3)


def test_filter_pipeline_matches_oracle():
    Lf, Lc, kc = 128, 64, 21.0
    s = get("hip/fused").make(Config(n=Lf, dim=3, dtype="float32"))
    st = _state3(Lf, 3)
    s.set_state(st)
    fine = s.get_state()  # masked, fp32-rounded
    pipe = FilterPipeline(Lf, Lc, kc)
    X = torch.zeros((Lc ** 3, C), dtype=torch.bfloat16, device="cuda")
    T = torch.zeros((Lc ** 3, TAU_C), dtype=torch.float32, device="cuda")
    ubar = pipe(s.buf.u0_rows, s.buf.pitch, X, T, g_scale=0.5, tau_scale=2.0)
    ub_ref, tau_ref = LO.sgs_target(LO.physical(fine, Lf), Lc, kc)
    e_u = _rel(pipe.half_state(ubar).cpu().numpy(), ub_ref)
    tau = T.cpu().numpy().reshape(Lc, Lc, Lc, TAU_C)[..., :6].transpose(3, 0, 1, 2) / 2.0
    e_t = _rel(tau, tau_ref)
    g_ref = LO.gradient(ub_ref, Lc).reshape(9, Lc, Lc, Lc)
    x = X.float().cpu().numpy().reshape(Lc, Lc, Lc, C)
    e_x = _rel(x[..., :9].transpose(3, 0, 1, 2) / 0.5, g_ref)
    print(f"  ubar {e_u:.2e} tau {e_t:.2e} features(bf16) {e_x:.2e}", flush=True)
    assert e_u < 2e-6 and e_t < 1e-4 and e_x < 1e-2
    assert np.all(x[..., 9:] == 0)


def _smag_source_ref(state, L, cs, kc):
    g = LO.gradient(state, L)
    tau = LO.smagorinsky(g, (cs * math.pi / kc) ** 2)
    return LO.stress_source(tau, L)


def test_smagorinsky_source_and_step_match_oracle():
    L = 64
    cfg = Config(n=L, dim=3, nu=2e-3, dtype="float32")
    s = LESSolver(cfg, "smagorinsky", cs=0.17, kc=21.0)
    st = _state3(L, 4)
    s.set_state(st)
    u0 = s.get_state()
    s.update_source()
    src = s.src[..., : L // 2 + 1].cpu().numpy()
    ref = _smag_source_ref(u0, L, 0.17, 21.0)
    e_s = _rel(src, ref)
    # one LES step against the oracle holding the same source over the step
    s2 = LESSolver(cfg, "smagorinsky", cs=0.17, kc=21.0)
    s2.set_state(st)
    s2.step(0.01)
    o = O.OracleSolver(cfg.with_(dtype="float64"))
    o.set_state(u0)
    o.source = ref
    o.step(0.01)
    e_step = O.rel_err(s2.get_state(), o.get_state())
    print(f"  source {e_s:.2e} step {e_step:.2e}", flush=True)
    assert e_s < 1e-4 and e_step < 2e-5


def test_les_without_model_is_the_plain_step():
    cfg = Config(n=64, dim=3, nu=2e-3, dtype="float32")
    a = LESSolver(cfg, "none")
    b = get("hip/fused").make(cfg)
    st = _state3(64, 5)
    a.set_state(st)
    b.set_state(st)
    a.advance(3, 0.01)
    b.advance(3, 0.01)
    assert np.array_equal(a.get_state(), b.get_state())


def _rand_bf16(rng, shape, scale=1.0):
    return _bf16(rng.normal(0, scale, shape))


def test_conv_forward_matches_oracle():
    rng = np.random.default_rng(7)
    n, nb = 64, 1
    P = nb * n ** 3
    x = _rand_bf16(rng, (P, C))
    w = _rand_bf16(rng, (27, C, C), 0.2)
    b = torch.from_numpy(rng.normal(0, 0.1, C).astype(np.float32))
    xc, wc, bc = x.cuda(), w.cuda(), b.cuda()
    yf = torch.zeros((P, C), dtype=torch.float32, device="cuda")
    y = torch.zeros((P, C), dtype=torch.bfloat16, device="cuda")
    conv(xc, wc, n=n, nb=nb, bias=bc, yf=yf, y=y, relu=False)
    ref = LO.conv_fwd(x.double().numpy().reshape(nb, n, n, n, C), w.double().numpy(),
                      b.double().numpy()).reshape(P, C)
    e = _rel(yf.cpu().numpy(), ref)
    e_b = _rel(y.float().cpu().numpy(), ref)
    yr = torch.zeros((P, C), dtype=torch.float32, device="cuda")
    conv(xc, wc, n=n, nb=nb, bias=bc, yf=yr, relu=True)
    e_r = _rel(yr.cpu().numpy(), np.maximum(ref, 0))
    print(f"  conv fp32 {e:.2e} bf16-out {e_b:.2e} relu {e_r:.2e}", flush=True)
    assert e < 1e-5 and e_r < 1e-5 and e_b < 1e-2


def test_conv_gradients_match_oracle():
    rng = np.random.default_rng(8)
    n, nb = 64, 1
    P = n ** 3
    net = SgsNet(seed=1)
    W = net.wb[1].float().cpu().numpy().astype(np.float64).reshape(27, C, C)
    dz = _rand_bf16(rng, (P, C), 1e-3)
    mask = _rand_bf16(rng, (P, C))
    out = torch.zeros((P, C), dtype=torch.float32, device="cuda")
    dzc, maskc = dz.cuda(), mask.cuda()
    conv(dzc, net.wt[1], n=n, nb=nb, mask=maskc, yf=out)
    dzn = dz.double().numpy().reshape(nb, n, n, n, C)
    ref = (LO.conv_bwd_data(dzn, W) * (mask.double().numpy().reshape(dzn.shape) > 0)).reshape(P, C)
    e_d = _rel(out.cpu().numpy(), ref)
    x = _rand_bf16(rng, (P, C))
    dw = torch.zeros(27 * C * C, dtype=torch.float32, device="cuda")
    db = torch.zeros(C, dtype=torch.float32, device="cuda")
    lib = _lib.load()
    xg, dzg = x.cuda(), dz.cuda()  # kept alive: a temporary's block could be handed to the next
    _lib.check(lib.psx_cnn_wgrad(xg.data_ptr(), dzg.data_ptr(), dw.data_ptr(),
                                 db.data_ptr(), n, nb, _lib.stream_ptr()))
    dw_ref, db_ref = LO.conv_wgrad(x.double().numpy().reshape(dzn.shape[:-1] + (C,)), dzn)
    e_w = _rel(dw.cpu().numpy().reshape(27, C, C), dw_ref)
    e_b = _rel(db.cpu().numpy(), db_ref)
    print(f"  input grad {e_d:.2e} weight grad {e_w:.2e} bias grad {e_b:.2e}", flush=True)
    assert e_d < 1e-5 and e_w < 1e-4 and e_b < 1e-4


def test_mse_and_adam_match_oracle():
    rng = np.random.default_rng(9)
    P = 4096
    y = torch.from_numpy(rng.normal(size=(P, C)).astype(np.float32)).cuda()
    t = torch.from_numpy(rng.normal(size=(P, TAU_C)).astype(np.float32)).cuda()
    g = torch.zeros((P, C), dtype=torch.bfloat16, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    lib = _lib.load()
    _lib.check(lib.psx_cnn_mse(y.data_ptr(), t.data_ptr(), TAU_C, 6, P, g.data_ptr(),
                               loss.data_ptr(), _lib.stream_ptr()))
    l_ref, g_ref = LO.mse(y.double().cpu().numpy(), t.double().cpu().numpy(), 6)
    assert loss.item() == pytest.approx(l_ref, rel=1e-6)  # y - t formed in fp32
    assert _rel(g.float().cpu().numpy(), g_ref) < 1e-2
    net = SgsNet(seed=2, lr=1e-2)
    w0 = net.w[0].cpu().numpy().astype(np.float64)
    gw = rng.normal(size=w0.shape).astype(np.float32)
    m = v = np.zeros_like(w0)
    w = w0
    for step in (1, 2):
        net.gw[0].copy_(torch.from_numpy(gw))
        net.gb[0].zero_()
        net._adam()
        w, m, v = LO.adam(w, m, v, gw.astype(np.float64), step, 1e-2)
    got = net.w[0].cpu().numpy()
    assert _rel(got - w0, w - w0) < 1e-4
    # bf16 copies: forward layout, and the flipped-transposed layout of the input gradient
    wb = net.wb[0].float().cpu().numpy().reshape(27, C, C)
    wt = net.wt[0].float().cpu().numpy().reshape(27, C, C)
    ref_b = _bf16(got).float().numpy().reshape(27, C, C)
    assert np.array_equal(wb, ref_b)
    assert np.array_equal(wt, ref_b[::-1].transpose(0, 2, 1))


def _bf16r(a):
    return torch.from_numpy(np.asarray(a, dtype=np.float32)).to(torch.bfloat16).double().numpy()


def test_network_gradients_match_oracle():
    """The whole backward chain (loss, 6 weight and bias gradients through 5 ReLU layers)
    against the oracle run on the same bf16 weights, with the same bf16 rounding of the
    stored activations and gradients."""
    rng = np.random.default_rng(13)
    n, nb = 64, 1
    P = n ** 3
    net = SgsNet(seed=14)
    X = torch.zeros((P, C), dtype=torch.bfloat16, device="cuda")
    X[:, :9] = _rand_bf16(rng, (P, 9)).cuda()
    T = torch.from_numpy(rng.normal(size=(P, TAU_C)).astype(np.float32)).cuda()
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    net.gradients(X, T, nb, loss)
    ws = [net.wb[i].double().cpu().numpy().reshape(27, C, C) for i in range(6)]
    bs = [net.b[i].double().cpu().numpy() for i in range(6)]
    shp = (nb, n, n, n, C)
    acts = [X.double().cpu().numpy().reshape(shp)]
    for i in range(6):
        z = LO.conv_fwd(acts[-1], ws[i], bs[i])
        acts.append(_bf16r(np.maximum(z, 0)) if i < 5 else z.astype(np.float32).astype(np.float64))
    tt = T.double().cpu().numpy().reshape(nb, n, n, n, TAU_C)
    l_ref, g = LO.mse(acts[-1], tt, 6)
    dz = _bf16r(g)
    errs = []
    for i in range(5, -1, -1):
        gw, gb = LO.conv_wgrad(acts[i], dz)
        errs.append((i, _rel(net.gw[i].cpu().numpy().reshape(27, C, C), gw),
                     _rel(net.gb[i].cpu().numpy(), gb)))
        if i > 0:
            dz = _bf16r(LO.conv_bwd_data(dz, ws[i]) * (acts[i] > 0))
    print(f"  loss {loss.item():.6e} / {l_ref:.6e}; (layer, dW, db) {errs}", flush=True)
    assert loss.item() == pytest.approx(l_ref, rel=1e-3)
    for i, ew, eb in errs:
        assert ew < 2e-2 and eb < 2e-2, (i, ew, eb)


def test_training_fits_a_teacher_network():
    rng = np.random.default_rng(10)
    nb, n = 2, 64
    P = nb * n ** 3
    X = torch.zeros((P, C), dtype=torch.bfloat16, device="cuda")
    X[:, :9] = _rand_bf16(rng, (P, 9)).cuda()
    teacher = SgsNet(seed=11)
    y = teacher.forward(X, nb)
    T = torch.zeros((P, TAU_C), dtype=torch.float32, device="cuda")
    T[:, :6] = y[:, :6]
    s = rms(T, TAU_C, 6)
    T[:, :6] /= s
    student = SgsNet(seed=12, lr=2e-3)
    losses = []
    for k in range(120):
        loss = torch.zeros(1, dtype=torch.float64, device="cuda")
        student.train_step(X, T, nb, loss)
        if k % 20 == 0 or k == 119:
            losses.append(loss.item())
    print(f"  losses {['%.3e' % x for x in losses]}", flush=True)
    assert all(math.isfinite(x) for x in losses)
    assert losses[-1] < 0.5 * losses[0] and losses[-1] < losses[1]


def test_attached_network_step_equals_explicit_source():
    """A step of the LES with the network attached (psx_set_sgs: the C step evaluates the
    stress first) equals the explicit evaluation followed by a plain step that adds the same
    source, bit for bit; and the device controller's captured chunks equal its eager ones."""
    L = 64
    cfg = Config(n=L, dim=3, nu=2e-3, dtype="float32")
    net = SgsNet(seed=21)
    st = _state3(L, 6)
    a = LESSolver(cfg, "net", net=net, g_scale=0.5, tau_scale=1e-3)
    a.set_state(st)
    a.step(0.01)
    b = LESSolver(cfg, "net", net=net, g_scale=0.5, tau_scale=1e-3)
    b.set_state(st)
    b.update_source()
    plain = get("hip/fused").make(cfg)
    plain.set_state(st)
    _lib.check(plain.lib.psx_set_source(plain.ctx, b.src.data_ptr()))
    plain.step(0.01)
    assert np.array_equal(a.get_state(), plain.get_state())
    assert float(torch.abs(b.src).max()) > 0
    runs = []
    for use_graph in (False, True):
        s = LESSolver(cfg, "net", net=net, g_scale=0.5, tau_scale=1e-3)
        s.set_state(st)
        r = s.run_device(math.inf, cfl=0.5, dt_max=0.05, cfl_every=10, kmax=25, chunk=5,
                         use_graph=use_graph)
        runs.append((r["dt"], s.get_state()))
    (da, sa), (db, sb) = runs
    print(f"  dt {da[0]:.4e} .. {da[-1]:.4e}", flush=True)
    assert da == db and np.array_equal(sa, sb) and np.isfinite(sa).all()


def test_network_stress_backscatter_clipping_matches_oracle():
    """PSX_LES_TAU_FROM_NET with the gradient pairs multiplies the stress by the backscatter
    factor exactly where the oracle's clip_backscatter does (-tau_ij S_ij < 0); without them
    it only scales."""
    L = 64
    L3 = L ** 3
    rng = np.random.default_rng(31)
    g = rng.normal(size=(3, 3, L3)).astype(np.float32)
    vel = rng.normal(size=(3, L3)).astype(np.float32)
    comps = [vel[0], vel[1], vel[2]] + [g[i, j] for i in range(3) for j in range(3)]
    pairs = np.stack([comps[2 * p] + 1j * comps[2 * p + 1] for p in range(6)])
    Y = rng.normal(size=(L3, C)).astype(np.float32)
    yc = torch.from_numpy(Y).cuda()
    pc = torch.from_numpy(pairs.astype(np.complex64).reshape(-1)).cuda()
    t6 = (0.5 * Y[:, :6].T).astype(np.float64)
    ref = LO.clip_backscatter(t6, g.astype(np.float64))
    half = LO.clip_backscatter(t6, g.astype(np.float64), keep=0.5)
    for grad, keep, want in ((pc, 0.0, ref), (pc, 0.5, half), (None, 0.0, t6)):
        tau = torch.zeros(3 * L3, dtype=torch.complex64, device="cuda")
        les_op("tau_from_net", grad, tau, L=L, scale=0.5, scale2=keep, in2=yc)
        tp = tau.cpu().numpy().reshape(3, L3)
        got = np.stack([tp[0].real, tp[0].imag, tp[1].real, tp[1].imag, tp[2].real, tp[2].imag])
        # points whose transfer is within fp32 round-off of zero may fall either way
        S = 0.5 * (g + np.swapaxes(g, 0, 1)).astype(np.float64)
        ts = sum((1.0 if i == j else 2.0) * t6[c] * S[i, j]
                 for c, (i, j) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]))
        sure = np.abs(ts) > 1e-4 * np.abs(ts).max()
        assert np.abs(got - want)[:, sure].max() < 1e-6 * np.abs(want).max()
    frac = float(np.mean(np.all(ref == 0, axis=0)))
    print(f"  clipped fraction {frac:.3f}", flush=True)
    assert 0.3 < frac < 0.7


def test_mixed_model_stress_matches_oracle():
    """PSX_LES_TAU_FROM_NET with scale3 = (Cs D)^2 > 0: the network's stress (after the
    backscatter factor) plus the oracle's Smagorinsky stress of the same gradient."""
    L = 64
    L3 = L ** 3
    rng = np.random.default_rng(37)
    g = rng.normal(size=(3, 3, L3)).astype(np.float32)
    vel = rng.normal(size=(3, L3)).astype(np.float32)
    comps = [vel[0], vel[1], vel[2]] + [g[i, j] for i in range(3) for j in range(3)]
    pairs = np.stack([comps[2 * p] + 1j * comps[2 * p + 1] for p in range(6)])
    Y = rng.normal(size=(L3, C)).astype(np.float32)
    yc = torch.from_numpy(Y).cuda()
    pc = torch.from_numpy(pairs.astype(np.complex64).reshape(-1)).cuda()
    t6 = (0.5 * Y[:, :6].T).astype(np.float64)
    g64 = g.astype(np.float64)
    csd2 = 0.02
    smag = LO.smagorinsky(g64, csd2)
    S = 0.5 * (g64 + np.swapaxes(g64, 0, 1))
    ts = sum((1.0 if i == j else 2.0) * t6[c] * S[i, j]
             for c, (i, j) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]))
    sure = np.abs(ts) > 1e-4 * np.abs(ts).max()
    for keep in (1.0, 0.75):
        want = LO.clip_backscatter(t6, g64, keep=keep) + smag
        tau = torch.zeros(3 * L3, dtype=torch.complex64, device="cuda")
        les_op("tau_from_net", pc, tau, L=L, scale=0.5, scale2=keep, scale3=csd2, in2=yc)
        tp = tau.cpu().numpy().reshape(3, L3)
        got = np.stack([tp[0].real, tp[0].imag, tp[1].real, tp[1].imag, tp[2].real, tp[2].imag])
        err = np.abs(got - want)[:, sure].max() / np.abs(want).max()
        print(f"  keep {keep}: max rel error {err:.2e}", flush=True)
        assert err < 2e-6
    # the Smagorinsky part is a real share of the stress in this test
    assert np.abs(smag).max() > 0.3 * np.abs(t6).max()


def test_mixed_model_les_step():
    """An LES with the network plus a Smagorinsky part (mix_cs > 0) attached: its source
    differs from the network-only source by the divergence of the Smagorinsky stress alone,
    i.e. mixed - net equals the Smagorinsky model's source at the same constant."""
    L = 64
    cfg = Config(n=L, dim=3, nu=2e-3, dtype="float32")
    st = _state3(L, 5)
    net = SgsNet(seed=3)
    srcs = {}
    for key, model, kw in (("net", "net", {}), ("mixed", "net", {"mix_cs": 0.12}),
                           ("smag", "smagorinsky", {"cs": 0.12})):
        s = LESSolver(cfg, model, net=net if model == "net" else None, g_scale=0.5,
                      tau_scale=1e-3, **kw)
        s.set_state(st)
        s.update_source()
        torch.cuda.synchronize()
        srcs[key] = s.src.cpu().numpy().astype(np.complex128)
        s.close()
    diff = srcs["mixed"] - srcs["net"]
    big = max(np.abs(srcs["net"]).max(), np.abs(srcs["smag"]).max())
    err = np.abs(diff - srcs["smag"]).max() / big
    ratio = np.abs(srcs["smag"]).max() / big
    print(f"  mixed - net vs Smagorinsky source: {err:.2e} (Smagorinsky share {ratio:.3f})",
          flush=True)
    assert ratio > 1e-2 and err < 1e-5


def test_sgs_attach_rejects_bad_coefficients():
    """psx_set_sgs refuses a backscatter factor outside [0, 2] and a negative Smagorinsky
    coefficient (the mixed model's constant), with the library's error message."""
    cfg = Config(n=64, dim=3, nu=2e-3, dtype="float32")
    net = SgsNet(seed=5)
    for kw, msg in (({"backscatter": 2.5}, "backscatter factor"),
                    ({"backscatter": -0.1}, "backscatter factor"),
                    ({"mix_cs": float("nan")}, "Smagorinsky")):
        with pytest.raises(RuntimeError, match=msg):
            LESSolver(cfg, "net", net=net, g_scale=0.5, tau_scale=1e-3, **kw)
    ok = LESSolver(cfg, "net", net=net, g_scale=0.5, tau_scale=1e-3, backscatter=2.0,
                   mix_cs=0.04)
    ok.close()
