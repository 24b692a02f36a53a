// Device-resident time-loop controller (psx_drv.h): one work-group per flow, so one launch
// serves a single flow or a whole ensemble.
//
// Per step the controller runs twice around the four stages:
//   k_drv_pre   after the first stage's line kernel (which left sum_i max|u_i| of the current
//               state in stats): CFL dt and constant-power A every cfl_every steps, the step
//               h that lands on t_end and the next record time, the parameter block for h;
//   k_drv_post  after the last stage: t += h, step history, record / end decisions;
//   k_drv_reduce + k_drv_record   FP64 energy and enstrophy of the new state, only on steps
//               that record.
// Reference semantics: src/pyspectral/simulate.py:31-37 (cfl_dt), :66-96 (the loop) and
// src/pyspectral/diagnostics.py:51-88 (densities, band energy).
#include "psx_drv.h"
#include "psx_kernels.h"

namespace psx {

// Sum over modes of |u_hat_c|^2 at state row `row`, half-axis column j.
__device__ __forceinline__ double mode_sq(const DrvMember& m, long long row, int j) {
  const long long sstride = (long long)m.n * m.n * m.sp;
  const int ncomp = m.dim == 2 ? 1 : 3;
  double a = 0.0;
  for (int c = 0; c < ncomp; ++c) {
    const long long off = c * sstride + row * m.sp + j;
    if (m.fp64) {
      const double2 z = static_cast<const double2*>(m.u0)[off];
      a += z.x * z.x + z.y * z.y;
    } else {
      const float2 z = static_cast<const float2*>(m.u0)[off];
      a += (double)z.x * (double)z.x + (double)z.y * (double)z.y;
    }
  }
  return a;
}

// Deterministic block sum (fixed tree over kDrvThreads values); result valid in thread 0.
__device__ __forceinline__ double block_sum(double x, double* red) {
  const int tid = threadIdx.x;
  red[tid] = x;
  __syncthreads();
  for (int s = kDrvThreads / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// Kinetic energy in klo^2 <= k^2 <= khi^2 of the (masked) state, FP64
// (diagnostics.py:83-87): 0.5 * sum |u_hat|^2 w_h / n^(2d), divided by k^2 in 2D.
__device__ double band_energy(const DrvMember& m, double* red) {
  const int n = m.n, nh = n / 2 + 1;
  int kb = m.khi2 >= 0.0 ? (int)floor(sqrt(m.khi2)) : -1;
  if (kb > n / 2) kb = n / 2;
  double acc = 0.0;
  if (kb >= 0 && m.khi2 >= m.klo2) {
    const int w = 2 * kb + 1;
    const long long count = m.dim == 2 ? (long long)w * (kb + 1) : (long long)w * w * (kb + 1);
    const double inv_norm = 1.0 / pow((double)n, 2.0 * m.dim);
    for (long long i = threadIdx.x; i < count; i += kDrvThreads) {
      const int kz = (int)(i % (kb + 1));
      const long long r = i / (kb + 1);
      int kx, ky;
      if (m.dim == 2) {
        kx = (int)r - kb;
        ky = kz;
      } else {
        kx = (int)(r / w) - kb;
        ky = (int)(r % w) - kb;
      }
      const int az = kz;
      const int ay = ky < 0 ? -ky : ky, ax = kx < 0 ? -kx : kx;
      // retained by the 2/3 rule (the state outside it is zero)
      if (3 * ax >= n || 3 * ay >= n || 3 * az >= n) continue;
      const double k2 = m.dim == 2 ? (double)kx * kx + (double)kz * kz
                                   : (double)kx * kx + (double)ky * ky + (double)kz * kz;
      if (k2 < m.klo2 || k2 > m.khi2) continue;
      const int j = kz;
      long long row;
      if (m.dim == 2) {
        row = kx >= 0 ? kx : kx + n;
      } else {
        const int ix = kx >= 0 ? kx : kx + n, iy = ky >= 0 ? ky : ky + n;
        row = (long long)ix * n + iy;
      }
      const double wh = (j == 0 || j == nh - 1) ? 1.0 : 2.0;
      double d = mode_sq(m, row, j) * wh * inv_norm;
      if (m.dim == 2) d = k2 > 0.0 ? d / k2 : 0.0;
      acc += 0.5 * d;
    }
  }
  return block_sum(acc, red);
}

__global__ void __launch_bounds__(kDrvThreads) k_drv_pre(const DrvMember* __restrict__ mem) {
  __shared__ double red[kDrvThreads];
  __shared__ double sh[2];
  const DrvMember m = mem[blockIdx.x];
  DrvState* s = m.s;
  if (!s->active) return;
  const int tid = threadIdx.x;
  const bool control = s->k % s->cfl_every == 0;
  double eb = 0.0;
  if (control && s->power >= 0.0) eb = band_energy(m, red);
  if (tid == 0) {
    if (control) {
      if (s->dt_fixed <= 0.0) {
        double umax = 0.0;
        for (int c = 0; c < m.dim; ++c) umax += __longlong_as_double((long long)m.stats[c]);
        s->cur_dt = (umax <= 0.0 || !isfinite(umax)) ? s->dt_max
                                                     : fmin(s->dt_max, s->cfl * s->dx / umax);
      }
      if (s->power >= 0.0) s->amp = eb > 0.0 ? s->power / (2.0 * eb) : 0.0;
    }
    const double t = s->t;
    const double h = fmin(fmin(s->cur_dt, s->t_end - t), s->next_diag - t);
    s->h = h;
    sh[0] = h;
    sh[1] = s->amp;
    // maxima of the next step start from zero
    for (int c = 0; c < 3; ++c) m.stats[c] = 0ull;
  }
  __syncthreads();
  write_params(m.prm, m.n / 2 + 1, m.nu, sh[0], m.alpha, sh[1], m.klo2, m.khi2, m.ka2, tid,
               kDrvThreads);
}

__global__ void k_drv_post(const DrvMember* __restrict__ mem) {
  const DrvMember m = mem[blockIdx.x];
  DrvState* s = m.s;
  if (threadIdx.x != 0 || !s->active) return;
  const double h = s->h;
  const double t = s->t + h;
  s->t = t;
  const long long slot = s->k % m.cap_step;
  m.steps[2 * slot] = h;
  m.steps[2 * slot + 1] = s->amp;
  s->k += 1;
  int rec = 0;
  if (t >= s->next_diag - s->eps) {
    rec = 1;
    s->next_diag += s->diag_dt;
  }
  const bool finished = t >= s->t_end - s->eps;
  if (finished || s->k >= s->kmax) {
    s->done = 1;
    s->active = 0;
    m.prm[2 * (m.n / 2 + 1) + PrmLayout::kActive] = 0.0;
    if (finished && s->last_rec_t < t - s->eps) rec = 1;
  }
  s->record = rec;
}

// Per-block FP64 partial energy and enstrophy of every recording flow: grid (blocks, flows).
__global__ void __launch_bounds__(kDrvThreads) k_drv_reduce(const DrvMember* __restrict__ mem) {
  __shared__ double red[kDrvThreads];
  const DrvMember m = mem[blockIdx.y];
  if (!m.s->record) return;
  const int n = m.n, nh = n / 2 + 1;
  const long long rows = m.dim == 2 ? (long long)n : (long long)n * n;
  const long long nm = rows * nh;
  const double inv_norm = 1.0 / pow((double)n, 2.0 * m.dim);
  double e = 0.0, z = 0.0;
  for (long long idx = (long long)blockIdx.x * kDrvThreads + threadIdx.x; idx < nm;
       idx += (long long)gridDim.x * kDrvThreads) {
    const long long row = idx / nh;
    const int j = (int)(idx % nh);
    int kx, ky;
    if (m.dim == 2) {
      kx = (int)row < n / 2 ? (int)row : (int)row - n;
      ky = 0;
    } else {
      const int ix = (int)(row / n), iy = (int)(row % n);
      kx = ix < n / 2 ? ix : ix - n;
      ky = iy < n / 2 ? iy : iy - n;
    }
    const double k2 = (double)kx * kx + (double)ky * ky + (double)j * j;
    const double wh = (j == 0 || j == nh - 1) ? 1.0 : 2.0;
    const double d = mode_sq(m, row, j) * wh * inv_norm;
    if (m.dim == 2) {
      e += k2 > 0.0 ? 0.5 * (d / k2) : 0.0;
      z += 0.5 * d;
    } else {
      e += 0.5 * d;
      z += 0.5 * d * k2;
    }
  }
  const double se = block_sum(e, red);
  const double sz = block_sum(z, red);
  if (threadIdx.x == 0) {
    m.partial[2 * blockIdx.x] = se;
    m.partial[2 * blockIdx.x + 1] = sz;
  }
}

// Fold the partials of every recording flow into one record (t, E, Z).
__global__ void __launch_bounds__(kDrvThreads) k_drv_record(const DrvMember* __restrict__ mem,
                                                            int nblocks) {
  __shared__ double red[kDrvThreads];
  const DrvMember m = mem[blockIdx.x];
  DrvState* s = m.s;
  if (!s->record) return;
  double e = 0.0, z = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kDrvThreads) {
    e += m.partial[2 * b];
    z += m.partial[2 * b + 1];
  }
  const double se = block_sum(e, red);
  const double sz = block_sum(z, red);
  if (threadIdx.x == 0) {
    const long long slot = s->nrec % m.cap_rec;
    m.rec[3 * slot] = s->t;
    m.rec[3 * slot + 1] = se;
    m.rec[3 * slot + 2] = sz;
    s->nrec += 1;
    s->last_rec_t = s->t;
    s->record = 0;
  }
}

// Mark every flow's parameter block active (after a run, before eager stepping) or inactive
// (a run that is over before its first step).
__global__ void k_drv_set_active(const DrvMember* __restrict__ mem, double v) {
  const DrvMember m = mem[blockIdx.x];
  if (threadIdx.x == 0) m.prm[2 * (m.n / 2 + 1) + PrmLayout::kActive] = v;
}

hipError_t launch_drv(int which, const DrvMember* mem, int count, hipStream_t st) {
  if (count < 1) return hipErrorInvalidValue;
  switch (which) {
    case 0:
      hipLaunchKernelGGL(k_drv_pre, dim3(count), dim3(kDrvThreads), 0, st, mem);
      break;
    case 1:
      hipLaunchKernelGGL(k_drv_post, dim3(count), dim3(64), 0, st, mem);
      break;
    case 2:
      hipLaunchKernelGGL(k_drv_reduce, dim3(kDrvRedBlocks, count), dim3(kDrvThreads), 0, st,
                         mem);
      hipLaunchKernelGGL(k_drv_record, dim3(count), dim3(kDrvThreads), 0, st, mem,
                         kDrvRedBlocks);
      break;
    case 3:
    case 4:
      hipLaunchKernelGGL(k_drv_set_active, dim3(count), dim3(64), 0, st, mem,
                         which == 3 ? 1.0 : 0.0);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace psx
