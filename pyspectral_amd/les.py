"""Large-eddy simulation with a learned subgrid-stress closure (3D, fp32, 64^3 coarse grid).

No counterpart in the reference.  Pieces, all on libpsx kernels:
- ``FilterPipeline``: a fine DNS state (the hip solver's own 256^3 state) -> the box-filtered
  (transfer function prod_i sin(k_i D/2)/(k_i D/2), D = pi/k_c), truncated velocity on the
  coarse grid, the model input (filtered velocity gradient, 9 channels) and the subgrid-stress
  target tau_ij = filt(u_i u_j) - filt(u_i) filt(u_j) (6 channels).
- ``SgsNet``: six periodic 3x3x3 convolutions with 32 channels (ReLU between them), bf16
  weights and activations, fp32 accumulation on MFMA; hand-written forward, backward
  (input and weight gradients) and Adam kernels.
- ``LESSolver``: the fused 64^3 IF-RK4 step plus the Leray-projected divergence of a subgrid
  stress (none, Smagorinsky, or the network), added to the nonlinear term of every stage.
  The C step evaluates the stress from the state at the start of every step and holds it
  over the step, so the device-resident time loop runs the LES as it runs the DNS.
"""

from __future__ import annotations

import ctypes
import math
from typing import Any

import numpy as np

from pyspectral_amd import _lib
from pyspectral_amd.backends.base import BackendUnavailable
from pyspectral_amd.backends.hip_fused import HipSolver
from pyspectral_amd.config import Config

C = 32          # channels of every network tensor
TAU_C = 8       # channels of a stored stress target (6 used)
TAPS = 27
NLAYER = 6
IN_CH, OUT_CH = 9, 6
WSIZE = TAPS * C * C


def _p(t: Any) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr() if t is not None else 0)


def fft3(t: Any, L: int, nbatch: int, inverse: bool) -> None:
    """In-place unnormalised 3D complex FFT of nbatch L^3 cubes (complex64)."""
    lib = _lib.load()
    _lib.check(lib.psx_fft3(_p(t), L, nbatch, int(inverse), _lib.stream_ptr()))


def les_op(name: str, inp: Any, out: Any, *, n: int = 0, L: int = 0, L2: int = 0,
           pitch: int = 0, scale: float = 1.0, scale2: float = 1.0, scale3: float = 0.0,
           in2: Any = None) -> None:
    """One of the LES support kernels (include/psx.h, PSX_LES_*)."""
    lib = _lib.load()
    a = _lib.LesArgs(inp=_p(inp).value, in2=_p(in2).value, out=_p(out).value, n=n, L=L, L2=L2,
                     pitch=pitch, scale=scale, scale2=scale2, scale3=scale3)
    _lib.check(lib.psx_les_op(_lib.LES_OPS[name], ctypes.byref(a), _lib.stream_ptr()))


def rms(t: Any, width: int, count: int) -> float:
    """RMS over the first ``count`` of every ``width`` fp32 values of t (FP64 sum)."""
    import torch

    acc = torch.zeros(1, dtype=torch.float64, device=t.device)
    les_op("sumsq", None, acc, n=t.numel(), L=4, L2=count, pitch=width, in2=t)
    return math.sqrt(float(acc.item()) / (t.numel() // width * count))


def conv(x: Any, w: Any, *, n: int, nb: int, bias: Any = None, mask: Any = None, y: Any = None,
         yf: Any = None, relu: bool = False) -> None:
    lib = _lib.load()
    a = _lib.ConvArgs(x=_p(x).value, w=_p(w).value, bias=_p(bias).value, mask=_p(mask).value,
                      y=_p(y).value, yf=_p(yf).value, n=n, nb=nb, relu=int(relu))
    _lib.check(lib.psx_cnn_conv(ctypes.byref(a), _lib.stream_ptr()))


# --------------------------------------------------------------------------------- network
class SgsNet:
    """The subgrid-stress network and its Adam state (fp32 master weights on the device)."""

    def __init__(self, n: int = 64, seed: int = 0, lr: float = 1e-3,
                 betas: tuple[float, float] = (0.9, 0.999), eps: float = 1e-8) -> None:
        import torch

        self.torch = torch
        self.n = n
        self.lr, self.b1, self.b2, self.eps = lr, betas[0], betas[1], eps
        rng = np.random.default_rng(seed)
        w = np.zeros((NLAYER, TAPS, C, C), dtype=np.float32)
        for layer in range(NLAYER):
            cin = IN_CH if layer == 0 else C
            cout = OUT_CH if layer == NLAYER - 1 else C
            bound = math.sqrt((6.0 if layer < NLAYER - 1 else 3.0) / (TAPS * cin))
            w[layer, :, :cout, :cin] = rng.uniform(-bound, bound, (TAPS, cout, cin))
        dev = "cuda"
        f32, bf16 = torch.float32, torch.bfloat16
        self.w = torch.from_numpy(w.reshape(NLAYER, WSIZE)).to(dev)
        self.b = torch.zeros((NLAYER, C), dtype=f32, device=dev)
        self.mw, self.vw = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.mb, self.vb = torch.zeros_like(self.b), torch.zeros_like(self.b)
        self.gw, self.gb = torch.zeros_like(self.w), torch.zeros_like(self.b)
        self.wb = torch.zeros((NLAYER, WSIZE), dtype=bf16, device=dev)
        self.wt = torch.zeros((NLAYER, WSIZE), dtype=bf16, device=dev)
        self.t = 0
        self._acts: dict[int, list[Any]] = {}
        self._adam(lr=0.0)  # bf16 copies of the initial weights (zero step)

    def _buffers(self, nb: int) -> list[Any]:
        torch = self.torch
        if nb not in self._acts:
            P = nb * self.n ** 3
            acts = [torch.zeros((P, C), dtype=torch.bfloat16, device="cuda")
                    for _ in range(NLAYER - 1)]
            out = torch.zeros((P, C), dtype=torch.float32, device="cuda")
            g = [torch.zeros((P, C), dtype=torch.bfloat16, device="cuda") for _ in range(2)]
            self._acts[nb] = acts + [out] + g
        return self._acts[nb]

    def forward(self, X: Any, nb: int) -> Any:
        """Network output (fp32 [nb n^3][32], channels 0..5 meaningful) for input X."""
        buf = self._buffers(nb)
        x = X
        for layer in range(NLAYER):
            last = layer == NLAYER - 1
            conv(x, self.wb[layer], n=self.n, nb=nb, bias=self.b[layer],
                 y=None if last else buf[layer], yf=buf[NLAYER - 1] if last else None,
                 relu=not last)
            if not last:
                x = buf[layer]
        return buf[NLAYER - 1]

    def eval_loss(self, X: Any, T: Any, nb: int) -> float:
        """Mean-square loss of the current network on (X, T) (no update)."""
        lib = _lib.load()
        loss = self.torch.zeros(1, dtype=self.torch.float64, device="cuda")
        y = self.forward(X, nb)
        g = self._buffers(nb)[NLAYER]
        _lib.check(lib.psx_cnn_mse(_p(y), _p(T), TAU_C, OUT_CH, nb * self.n ** 3, _p(g),
                                   _p(loss), _lib.stream_ptr()))
        return float(loss.item())

    def train_step(self, X: Any, T: Any, nb: int, loss: Any) -> None:
        """Forward, mean-square loss against T (fp32 [nb n^3][8]; accumulated into the device
        scalar ``loss``), backward and one Adam step."""
        self.gradients(X, T, nb, loss)
        self._adam()

    def gradients(self, X: Any, T: Any, nb: int, loss: Any) -> None:
        """Forward, loss and backward: weight and bias gradients into gw, gb."""
        lib = _lib.load()
        st = _lib.stream_ptr()
        buf = self._buffers(nb)
        y = self.forward(X, nb)
        P = nb * self.n ** 3
        g = buf[NLAYER]
        _lib.check(lib.psx_cnn_mse(_p(y), _p(T), TAU_C, OUT_CH, P, _p(g), _p(loss), st))
        self.gw.zero_()
        self.gb.zero_()
        dz, spare = g, buf[NLAYER + 1]
        for layer in range(NLAYER - 1, -1, -1):
            x_in = X if layer == 0 else buf[layer - 1]
            # the trainer runs on one stream, so the library's static partial-sum array is
            # safe here and spares each call a stream-ordered allocation (about 50 us)
            _lib.check(lib.psx_cnn_wgrad_noalloc(_p(x_in), _p(dz), _p(self.gw[layer]),
                                                 _p(self.gb[layer]), self.n, nb, st))
            if layer > 0:
                conv(dz, self.wt[layer], n=self.n, nb=nb, mask=buf[layer - 1], y=spare)
                dz, spare = spare, dz

    def _adam(self, lr: float | None = None) -> None:
        lib = _lib.load()
        st = _lib.stream_ptr()
        lr = self.lr if lr is None else lr
        if lr > 0:
            self.t += 1
        t = max(self.t, 1)
        c1 = 1.0 / (1.0 - self.b1 ** t)
        c2 = 1.0 / (1.0 - self.b2 ** t)
        for layer in range(NLAYER):
            a = _lib.AdamArgs(w=_p(self.w[layer]).value, m=_p(self.mw[layer]).value,
                              v=_p(self.vw[layer]).value, g=_p(self.gw[layer]).value,
                              wb=_p(self.wb[layer]).value, wt=_p(self.wt[layer]).value,
                              n=WSIZE, lr=lr, b1=self.b1, b2=self.b2, eps=self.eps, c1=c1,
                              c2=c2, is_w=1)
            _lib.check(lib.psx_cnn_adam(ctypes.byref(a), st))
            a = _lib.AdamArgs(w=_p(self.b[layer]).value, m=_p(self.mb[layer]).value,
                              v=_p(self.vb[layer]).value, g=_p(self.gb[layer]).value,
                              wb=None, wt=None, n=C, lr=lr, b1=self.b1, b2=self.b2,
                              eps=self.eps, c1=c1, c2=c2, is_w=0)
            _lib.check(lib.psx_cnn_adam(ctypes.byref(a), st))

    def state(self) -> dict[str, np.ndarray]:
        return {"w": self.w.cpu().numpy(), "b": self.b.cpu().numpy()}


# ------------------------------------------------------------------------ filtered DNS
class FilterPipeline:
    """Fine DNS state -> filtered coarse velocity spectra, model input and stress target."""

    def __init__(self, Lf: int, Lc: int = 64, kc: float = 21.0) -> None:
        import torch

        self.torch = torch
        self.Lf, self.Lc, self.kc = Lf, Lc, kc
        c64 = torch.complex64
        Lf3, Lc3 = Lf ** 3, Lc ** 3
        self.full = torch.zeros(3 * Lf3, dtype=c64, device="cuda")
        self.vel = torch.zeros(2 * Lf3, dtype=c64, device="cuda")
        self.prod = torch.zeros(3 * Lf3, dtype=c64, device="cuda")
        self.fprod_s = torch.zeros(6 * Lc3, dtype=c64, device="cuda")
        self.ubar = torch.zeros(3 * Lc3, dtype=c64, device="cuda")
        self.grad = torch.zeros(6 * Lc3, dtype=c64, device="cuda")
        self.fprod = torch.zeros(3 * Lc3, dtype=c64, device="cuda")

    def __call__(self, state_rows: Any, pitch: int, X: Any, T: Any, g_scale: float = 1.0,
                 tau_scale: float = 1.0) -> Any:
        """state_rows: the fine state (3, Lf, Lf, pitch) complex64; writes X (bf16 [Lc^3][32])
        and T (fp32 [Lc^3][8]); returns the filtered coarse velocity, full spectra (3 Lc^3)."""
        Lf, Lc, kc = self.Lf, self.Lc, self.kc
        les_op("expand", state_rows, self.full, n=3, L=Lf, pitch=pitch)
        les_op("gradpack", self.full, self.vel, n=1, L=Lf, scale=1.0 / Lf ** 3)
        fft3(self.vel, Lf, 2, inverse=True)
        les_op("products", self.vel, self.prod, L=Lf)
        fft3(self.prod, Lf, 3, inverse=False)
        les_op("truncate", self.prod, self.fprod_s, n=3, L=Lf, L2=Lc, pitch=1, scale=kc)
        les_op("truncate", self.full, self.ubar, n=3, L=Lf, L2=Lc, pitch=0, scale=kc)
        les_op("gradpack", self.ubar, self.grad, n=6, L=Lc, scale=1.0 / Lc ** 3)
        fft3(self.grad, Lc, 6, inverse=True)
        les_op("pairup", self.fprod_s, self.fprod, n=3, L=Lc, scale=1.0 / Lc ** 3)
        fft3(self.fprod, Lc, 3, inverse=True)
        les_op("target", self.fprod, T, L=Lc, scale=tau_scale, in2=self.grad)
        les_op("features", self.grad, X, L=Lc, scale=g_scale)
        return self.ubar

    def half_state(self, full: Any) -> Any:
        """NumPy-layout (rfftn) view of full coarse spectra: (3, Lc, Lc, Lc/2+1)."""
        Lc = self.Lc
        return full.view(3, Lc, Lc, Lc)[..., : Lc // 2 + 1]


# ------------------------------------------------------------------------------ LES solver
class LESSolver(HipSolver):
    """Coarse 3D fp32 flow with a subgrid stress: model in {"none", "smagorinsky", "net"}.

    The model is attached to the C context (psx_set_sgs): every step -- eager, graph-replayed
    or run by the device controller -- first evaluates the stress from the state at its start
    and holds its divergence over the four stages.  ``run_device`` therefore drives an LES
    exactly as it drives the DNS.  ``backscatter`` multiplies the network's stress wherever it
    would transfer energy to the resolved scales (-tau_ij S_ij < 0): 1 keeps the network's
    output as it is, 0 clips the backscatter, a factor up to 2 amplifies it.  ``mix_cs`` > 0 adds a Smagorinsky stress of that
    constant to the network's (a mixed model); ``cs`` is the constant of the pure Smagorinsky
    model.  ``shell_gain`` (one factor per wavenumber shell (int)(|k| + 0.5), shells past its
    end get 1) multiplies the model's source shell by shell."""

    variant = "les"

    def __init__(self, cfg: Config, model: str = "none", *, net: SgsNet | None = None,
                 cs: float = 0.17, kc: float = 21.0, g_scale: float = 1.0,
                 tau_scale: float = 1.0, backscatter: float = 1.0, mix_cs: float = 0.0,
                 shell_gain: Any = None) -> None:
        if cfg.dim != 3 or cfg.dtype != "float32":
            raise BackendUnavailable("LES: 3D fp32 flows")
        if model not in ("none", "smagorinsky", "net"):
            raise ValueError(f"unknown subgrid model {model!r}")
        if model == "net" and net is None:
            raise ValueError("the network model needs a trained SgsNet")
        super().__init__(cfg)
        torch = self.torch
        L = cfg.n
        self.model, self.net, self.cs, self.kc = model, net, cs, kc
        self.g_scale, self.tau_scale = g_scale, tau_scale
        self.mix_cs = mix_cs
        self.src = torch.zeros_like(self.buf.u0_rows)
        self.full = torch.zeros(3 * L ** 3, dtype=torch.complex64, device="cuda")
        self.grad = torch.zeros(6 * L ** 3, dtype=torch.complex64, device="cuda")
        self.tau = torch.zeros(3 * L ** 3, dtype=torch.complex64, device="cuda")
        self.X = torch.zeros((L ** 3, C), dtype=torch.bfloat16, device="cuda")
        self.gain = None
        if shell_gain is not None:
            g = np.ones(self.diag.nshell, dtype=np.float32)
            sg = np.asarray(shell_gain, dtype=np.float32)[: self.diag.nshell]
            if not np.all(np.isfinite(sg)):
                raise ValueError("shell_gain: non-finite factor")
            g[: len(sg)] = sg
            self.gain = torch.from_numpy(g).cuda()
        if model == "none":
            return
        d = _lib.SgsDesc(model=_lib.SGS_MODELS[model], src=_p(self.src).value,
                         full=_p(self.full).value, grad=_p(self.grad).value,
                         tau=_p(self.tau).value, x=_p(self.X).value,
                         g_scale=g_scale, tau_scale=tau_scale,
                         smag=((cs if model == "smagorinsky" else mix_cs) * math.pi / kc) ** 2,
                         backscatter=backscatter, shell_gain=_p(self.gain).value)
        if model == "net":
            assert net is not None and net.n == L
            acts = net._buffers(1)
            for i in range(NLAYER):
                d.w[i] = _p(net.wb[i]).value
                d.b[i] = _p(net.b[i]).value
            for i in range(NLAYER - 1):
                d.act[i] = _p(acts[i]).value
            d.out = _p(acts[NLAYER - 1]).value
        self._sgs = d  # the C context borrows the pointers; the tensors live as long as self
        _lib.check(self.lib.psx_set_sgs(self.ctx, ctypes.byref(d)))

    def update_source(self) -> None:
        """Subgrid-stress divergence of the current state into the source buffer (the same
        evaluation every step starts with)."""
        if self.model != "none":
            _lib.check(self.lib.psx_sgs_update(self.ctx, _lib.stream_ptr()))
