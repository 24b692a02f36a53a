// Device-resident time-loop controller shared by single flows and ensembles.
//
// Restates the host loop of src/pyspectral/simulate.py:66-96 on the device so that steps run
// back to back in captured graphs with no host synchronisation:
//   every cfl_every steps: dt = min(dt_max, cfl * dx / sum_i max|u_i|) (dt_max when the sum
//   is not positive and finite), and with a target power A = P / (2 E_band) (0 if E_band = 0);
//   each step: h = min(dt, t_end - t, next_diag - t), t += h;
//   a record (t, E, Z) after every step that reaches next_diag - eps, and at the end.
// The maxima come from the first stage's line kernel of the same step (it transforms exactly
// the velocity of the current state), the band energy from the state itself.
#pragma once

namespace psx {

// FP64 controller state of one flow (device memory, written by the controller kernels).
struct DrvState {
  double t, t_end, eps;
  double dt_fixed;   // > 0: fixed step; 0: CFL-adaptive
  double cfl, dt_max, dx;
  double next_diag, diag_dt;  // diag_dt = 0: no records between the ends
  double power;      // target power P; negative: no constant-power forcing
  double amp;        // current band growth rate A
  double cur_dt;     // current CFL (or fixed) step
  double h;          // step being taken
  double last_rec_t; // time of the last record
  long long k;       // steps taken in this run
  long long kmax;    // step budget (done when k reaches it)
  long long cfl_every;
  long long nrec;    // records written
  int done, record, active, pad;
};

// One flow as seen by the controller kernels.
struct DrvMember {
  const void* u0;              // state (working precision), row pitch sp
  double* prm;                 // FP64 parameter block of the flow
  unsigned long long* stats;   // CFL maxima (FP64 bit patterns)
  DrvState* s;
  double* rec;                 // ring of cap_rec records (t, E, Z)
  double* steps;               // ring of cap_step entries (h, A)
  double* partial;             // per-block partial sums (E, Z) of the record reduction
  double nu, alpha, klo2, khi2, ka2;
  int n, dim, sp, fp64;
  int cap_rec, cap_step;
  int pad0, pad1;
};

constexpr int kDrvThreads = 256;
constexpr int kDrvRedBlocks = 128;  // record-reduction blocks per flow

}  // namespace psx
