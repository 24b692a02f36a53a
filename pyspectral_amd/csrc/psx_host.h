// Host-side helpers shared by the single-flow (psx_api.cpp) and ensemble (psx_ens.cpp)
// contexts: error reporting and the FP64-generated twiddle tables.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/psx.h"
#include "psx_kernels.h"

namespace psx {

// Message of the last failure on this thread (psx_last_error).
inline thread_local std::string g_err;

inline int fail(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return 1;
}

#define PSX_HIP(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return psx::fail("%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                           __FILE__, __LINE__);                         \
  } while (0)

template <typename T>
inline int make_twiddles(int n, void** out) {
  std::vector<cx<T>> h(n);
  for (int i = 0; i < n; ++i) {
    const long double ang = 2.0L * 3.14159265358979323846264338327950288L * i / n;
    h[i].x = (T)(double)std::cos(ang);
    h[i].y = (T)(double)(-std::sin(ang));
  }
  PSX_HIP(hipMalloc(out, n * sizeof(cx<T>)));
  PSX_HIP(hipMemcpy(*out, h.data(), n * sizeof(cx<T>), hipMemcpyHostToDevice));
  return 0;
}

// Append the per-pass twiddle tables of plan p (forward and reversed radix order) to buf and
// record their offsets in p->twoff.
template <typename T>
inline void append_pass_twiddles(FftPlan* p, std::vector<cx<T>>& buf) {
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int order = 0; order < 2; ++order) {
    int Ns = 1;
    for (int q = 0; q < p->np; ++q) {
      const int R = order == 0 ? p->rad[q] : p->rad[p->np - 1 - q];
      p->twoff[order][q] = (int)buf.size();
      if (Ns > 1) {
        for (int m = 1; m < R; ++m)
          for (int k = 0; k < Ns; ++k) {
            const long double ang = two_pi * (long double)(m * k) / (long double)(Ns * R);
            cx<T> w;
            w.x = (T)(double)std::cos(ang);
            w.y = (T)(double)(-std::sin(ang));
            buf.push_back(w);
          }
      }
      Ns *= R;
    }
  }
}

template <typename T>
inline int make_pass_twiddles(Flow<T>* f, void** out) {
  std::vector<cx<T>> buf;
  FftPlan* plans[4] = {&f->prow, &f->pcol, &f->pa, &f->py};
  for (FftPlan* p : plans)
    if (p->np > 0) append_pass_twiddles<T>(p, buf);
  if (buf.empty()) buf.push_back(cx<T>{(T)1, (T)0});
  PSX_HIP(hipMalloc(out, buf.size() * sizeof(cx<T>)));
  PSX_HIP(hipMemcpy(*out, buf.data(), buf.size() * sizeof(cx<T>), hipMemcpyHostToDevice));
  return 0;
}

// Controller state at the start of a run (psx_run_params semantics, include/psx.h).
inline DrvState drv_initial_state(const psx_run_params& rp, int n) {
  DrvState s{};
  s.t = rp.t0;
  s.t_end = rp.t_end;
  s.eps = 1e-12 * std::fmax(1.0, std::isfinite(rp.t_end) ? rp.t_end : 1.0);
  s.dt_fixed = rp.dt_fixed > 0.0 ? rp.dt_fixed : 0.0;
  s.cfl = rp.cfl;
  s.dt_max = rp.dt_max;
  s.dx = 2.0 * M_PI / n;
  s.diag_dt = rp.diag_dt > 0.0 ? rp.diag_dt : 0.0;
  s.next_diag = !(rp.diag_dt > 0.0)        ? INFINITY
                : std::isfinite(rp.next_diag) ? rp.next_diag
                                              : rp.t0 + rp.diag_dt;
  s.power = rp.power;
  s.amp = rp.amp0;
  s.cur_dt = s.dt_fixed;
  s.h = 0.0;
  s.last_rec_t = std::isfinite(rp.last_rec_t) ? rp.last_rec_t : -INFINITY;
  s.k = 0;
  s.kmax = rp.kmax > 0 ? rp.kmax : (long long)9e18;
  s.cfl_every = rp.cfl_every;
  s.nrec = 0;
  s.done = !(s.t < s.t_end - s.eps);
  s.active = !s.done;
  // a run over before its first step still records its end state when the last record of
  // the history is older than t (the end-of-run rule of the step kernels)
  s.record = (rp.record_start || (s.done && s.last_rec_t < s.t - s.eps)) ? 1 : 0;
  return s;
}

inline void drv_export(const DrvState& s, psx_run_state* out) {
  out->t = s.t;
  out->amp = s.amp;
  out->cur_dt = s.cur_dt;
  out->next_diag = s.next_diag;
  out->k = s.k;
  out->nrec = s.nrec;
  out->done = s.done;
}

}  // namespace psx
