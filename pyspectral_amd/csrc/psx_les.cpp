// C ABI of the LES support kernels and the subgrid-stress network (include/psx.h, psx_les_*,
// psx_fft3, psx_cnn_*).  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "../../include/psx.h"
#include "psx_host.h"
#include "psx_les.h"

using namespace psx;

namespace {

// Pass twiddles of the c2c plan per length, built once per process on first use.
const cx<float>* c2c_twiddles(int L) {
  static std::mutex mu;
  static std::vector<std::pair<int, void*>> cache;
  std::lock_guard<std::mutex> lock(mu);
  for (auto& e : cache)
    if (e.first == L) return static_cast<const cx<float>*>(e.second);
  FftPlan p{};
  c2c_plan(L, &p);
  std::vector<cx<float>> buf;
  append_pass_twiddles<float>(&p, buf);
  if (buf.empty()) buf.push_back(cx<float>{1.f, 0.f});
  void* d = nullptr;
  if (hipMalloc(&d, buf.size() * sizeof(cx<float>)) != hipSuccess) return nullptr;
  if (hipMemcpy(d, buf.data(), buf.size() * sizeof(cx<float>), hipMemcpyHostToDevice) !=
      hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  cache.emplace_back(L, d);
  return static_cast<const cx<float>*>(d);
}

}  // namespace

namespace psx {

int sgs_prepare(int L) {
  if (!c2c_twiddles(L)) return fail("sgs: twiddle table allocation failed");
  return 0;
}

int enqueue_sgs(const psx_sgs_desc& d, const cx<float>* u0, int pitch, int L, hipStream_t st) {
  const cx<float>* tw = c2c_twiddles(L);
  if (!tw) return fail("sgs: twiddle table allocation failed");
  auto* full = static_cast<cx<float>*>(d.full);
  auto* grad = static_cast<cx<float>*>(d.grad);
  auto* tau = static_cast<cx<float>*>(d.tau);
  auto op = [&](int which, const void* in, const void* in2, void* out, long long n, int pch,
                double scale, double scale2 = 1.0, double scale3 = 0.0) {
    LesArgs a;
    a.cin = static_cast<const cx<float>*>(in);
    a.in2 = in2;
    a.cout = static_cast<cx<float>*>(out);
    a.out = out;
    a.n = n;
    a.L = L;
    a.pitch = pch;
    a.scale = (float)scale;
    a.scale2 = (float)scale2;
    a.scale3 = (float)scale3;
    return launch_les(which, a, st);
  };
  // velocity gradients of the state in physical space, six complex pairs
  PSX_HIP(op(kLesExpand, u0, nullptr, full, 3, pitch, 1.0));
  PSX_HIP(op(kLesGradPack, full, nullptr, grad, 6, 0, 1.0 / ((double)L * L * L)));
  PSX_HIP(launch_c2c3(grad, tw, L, 6, true, st));
  if (d.model == PSX_SGS_SMAGORINSKY) {
    PSX_HIP(op(kLesSmagorinsky, grad, nullptr, tau, 0, 0, d.smag));
  } else {
    PSX_HIP(op(kLesFeatures, grad, nullptr, d.x, 0, 0, d.g_scale));
    for (int layer = 0; layer < 6; ++layer) {
      ConvArgs c;
      c.x = static_cast<const __bf16*>(layer ? d.act[layer - 1] : d.x);
      c.w = static_cast<const __bf16*>(d.w[layer]);
      c.bias = d.b[layer];
      c.n = L;
      c.nb = 1;
      if (layer < 5) {
        c.y = static_cast<__bf16*>(d.act[layer]);
        c.relu = 1;
      } else {
        c.yf = d.out;
      }
      PSX_HIP(launch_conv(c, st));
    }
    // 1 and no Smagorinsky part: the network's stress as it is
    const bool pointwise = d.backscatter != 1.0 || d.smag > 0.0;
    PSX_HIP(op(kLesTauFromNet, pointwise ? grad : nullptr, d.out, tau, 0, 0, d.tau_scale,
               d.backscatter, d.smag > 0.0 ? d.smag : 0.0));
  }
  PSX_HIP(launch_c2c3(tau, tw, L, 3, false, st));
  PSX_HIP(op(kLesSource, tau, d.shell_gain, d.src, 0, pitch, 1.0));
  return 0;
}

}  // namespace psx

extern "C" {

int psx_fft3(void* data, int L, int nbatch, int inverse, void* stream) {
  if (!data || nbatch < 1) return fail("psx_fft3: bad arguments");
  if (!c2c_supported(L)) return fail("psx_fft3: L=%d unsupported (64, 128, 256)", L);
  const cx<float>* tw = c2c_twiddles(L);
  if (!tw) return fail("psx_fft3: twiddle table allocation failed");
  PSX_HIP(launch_c2c3(static_cast<cx<float>*>(data), tw, L, nbatch, inverse != 0,
                      static_cast<hipStream_t>(stream)));
  return 0;
}

int psx_les_op(int which, const psx_les_args* a, void* stream) {
  if (!a) return fail("psx_les_op: null arguments");
  if (a->L < 4 || (a->L & (a->L - 1))) return fail("psx_les_op: L=%d not a power of two", a->L);
  LesArgs x;
  x.cin = static_cast<const cx<float>*>(a->in);
  x.in2 = a->in2;
  x.cout = static_cast<cx<float>*>(a->out);
  x.out = a->out;
  x.n = a->n;
  x.L = a->L;
  x.L2 = a->L2;
  x.pitch = a->pitch;
  x.scale = (float)a->scale;
  x.scale2 = (float)a->scale2;
  x.scale3 = (float)a->scale3;
  PSX_HIP(launch_les(which, x, static_cast<hipStream_t>(stream)));
  return 0;
}

int psx_cnn_conv(const psx_conv_args* a, void* stream) {
  if (!a) return fail("psx_cnn_conv: null arguments");
  ConvArgs c;
  c.x = static_cast<const __bf16*>(a->x);
  c.w = static_cast<const __bf16*>(a->w);
  c.bias = static_cast<const float*>(a->bias);
  c.mask = static_cast<const __bf16*>(a->mask);
  c.y = static_cast<__bf16*>(a->y);
  c.yf = static_cast<float*>(a->yf);
  c.n = a->n;
  c.nb = a->nb;
  c.relu = a->relu;
  if (c.n < 4 || (c.n & (c.n - 1)) || c.nb < 1 || !c.x || !c.w || (!c.y && !c.yf))
    return fail("psx_cnn_conv: bad arguments (n=%d nb=%d)", c.n, c.nb);
  PSX_HIP(launch_conv(c, static_cast<hipStream_t>(stream)));
  return 0;
}

int psx_cnn_wgrad(const void* x, const void* dz, float* dw, float* db, int n, int nb,
                  void* stream) {
  if (!x || !dz || !dw || !db) return fail("psx_cnn_wgrad: null argument");
  if (n != 64) return fail("psx_cnn_wgrad: n must be 64 (got %d)", n);
  if (nb < 1 || nb > kWgradMaxCubes)
    return fail("psx_cnn_wgrad: nb = %d outside [1, %d]", nb, kWgradMaxCubes);
  PSX_HIP(launch_conv_wgrad(static_cast<const __bf16*>(x), static_cast<const __bf16*>(dz), dw,
                            db, n, nb, static_cast<hipStream_t>(stream)));
  return 0;
}

int psx_cnn_wgrad_noalloc(const void* x, const void* dz, float* dw, float* db, int n, int nb,
                          void* stream) {
  if (!x || !dz || !dw || !db) return fail("psx_cnn_wgrad_noalloc: null argument");
  if (n != 64) return fail("psx_cnn_wgrad_noalloc: n must be 64 (got %d)", n);
  if (nb < 1 || nb > kWgradMaxCubes)
    return fail("psx_cnn_wgrad_noalloc: nb = %d outside [1, %d]", nb, kWgradMaxCubes);
  PSX_HIP(launch_conv_wgrad_noalloc(static_cast<const __bf16*>(x),
                                    static_cast<const __bf16*>(dz), dw, db, n, nb,
                                    static_cast<hipStream_t>(stream)));
  return 0;
}

unsigned long psx_cnn_wgrad_scratch_bytes(void) { return wgrad_scratch_bytes(); }

int psx_cnn_mse(const float* y, const float* t, int tstride, int nch, long long npts, void* g,
                double* loss, void* stream) {
  if (!y || !t || !g || !loss || nch < 1 || nch > 32 || tstride < nch || npts < 1)
    return fail("psx_cnn_mse: bad arguments");
  PSX_HIP(launch_mse(y, t, tstride, nch, npts, static_cast<__bf16*>(g), loss,
                     static_cast<hipStream_t>(stream)));
  return 0;
}

int psx_cnn_adam(const psx_adam_args* a, void* stream) {
  if (!a || !a->w || !a->m || !a->v || !a->g || a->n < 1) return fail("psx_cnn_adam: bad arguments");
  if (a->is_w && a->n != 27LL * 32 * 32) return fail("psx_cnn_adam: weight tensors hold 27*32*32");
  AdamArgs x;
  x.w = a->w;
  x.m = a->m;
  x.v = a->v;
  x.g = a->g;
  x.wb = static_cast<__bf16*>(a->wb);
  x.wt = static_cast<__bf16*>(a->wt);
  x.live = nullptr;
  x.n = a->n;
  x.lr = (float)a->lr;
  x.b1 = (float)a->b1;
  x.b2 = (float)a->b2;
  x.eps = (float)a->eps;
  x.c1 = (float)a->c1;
  x.c2 = (float)a->c2;
  x.is_w = a->is_w;
  PSX_HIP(launch_adam(x, static_cast<hipStream_t>(stream)));
  return 0;
}

}  // extern "C"
