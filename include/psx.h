/* C ABI of the MI355X (gfx950) IF-RK4 pseudo-spectral step library, libpsx.so.
 *
 * The library advances the state of one periodic-box Navier-Stokes flow with the Lawson
 * integrating-factor RK4 scheme on one GPU: hand-written batched r2c/c2r Stockham FFTs with
 * the Basdevant nonlinear term, the 2/3 dealias mask, the Leray projection (3D), the
 * integrating factor exp(L dt) (viscosity, linear drag, constant-power band forcing) and the
 * RK stage combination fused into the transform kernels.
 *
 * Ownership: the caller allocates the state, the accumulator and the work buffer with the
 * sizes psx_query_layout reports and keeps them alive until psx_destroy; the library borrows
 * the pointers and owns only its small tables (twiddles, FP64 integrating-factor table,
 * reduction scratch) and its captured graph.  Every enqueueing call takes the caller's HIP
 * stream (hipStream_t passed as void*); work is ordered on it.  Calls on one context are not
 * re-entrant.  Every function returns 0 on success and a non-zero code on failure, with a
 * message in psx_last_error() (thread-local).
 *
 * Each entry point names the reference interface it replaces
 * (paths in the pyspectral repository).
 */
#ifndef PSX_H_
#define PSX_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psx_ctx psx_ctx;

/* Physical and numerical parameters of one flow.
 * Mirrors src/pyspectral/config.py:28-38 (Config) with the derived quantities of
 * config.py:59-62 (drag_k2) and base.py:114-123 (the band of the linear operator).
 * form: 0 = basdevant (2D and 3D), 1 = rotational (3D only).  fp64: working precision. */
typedef struct psx_flow_params {
  int n;
  int dim;
  int form;
  int fp64;
  double nu;
  double alpha;
  double ka2;  /* drag band upper k^2 (1e30: uniform drag) */
  double klo2; /* forcing band lower k^2 */
  double khi2; /* forcing band upper k^2 */
  int dealias; /* 0: 2/3 rule (cubic, 3 |k_i| < n); 1: phase shift with the spherical mask
                  |k|^2 < 2 (n/3)^2 (config.py:38, grid.py:36-50, numpy_ref.py:68-72,125-153) */
} psx_flow_params;

/* Buffer sizes in complex elements of the working precision. */
typedef struct psx_layout {
  long long state_elems; /* u0 and acc each: rows * state_pitch, rows = n^(d-1) (x3 in 3D) */
  long long work_elems;  /* work buffer: nfields * n^(d-1) * ldw */
  int nfields;           /* work fields: 2 (2D), 5 (3D basdevant), 6 (3D rotational) */
  int ldw;               /* padded retained half-axis width of a work row */
  int jm;                /* retained half-axis modes, 3 j < n */
  int four_step;         /* 1 if the full axes use the two-kernel split */
  int e_line, e_spec, e_outer, e_mid; /* elements per thread of each kernel (0: unused) */
  int t_line, t_spec, t_outer, t_mid; /* threads per work-group of each kernel */
  int state_pitch;       /* row pitch of u0/acc in complex elements (>= n/2+1, 128-byte rows);
                            columns 0..n/2 of each row are the NumPy rfftn coefficients */
} psx_layout;

/* Plan the kernels for p and report the buffer sizes.  Replaces the buffer and plan set-up
 * of src/pyspectral/backends/cupy_cuda.py:40-82,100-112 (CupyCudaSolver.__init__,
 * _make_plans).  Returns non-zero for unsupported sizes (n must be even, >= 8, of the
 * form 2^a, 3*2^a or 5*2^a with a >= 2). */
int psx_query_layout(const psx_flow_params* p, psx_layout* out);

/* Create a context on the current device borrowing u0, acc and work.
 * Replaces cupy_cuda.py:40-82 (buffers, NVRTC module, cuFFT plans). */
int psx_create(const psx_flow_params* p, void* u0, void* acc, void* work, psx_ctx** out);

/* Release the context (not the borrowed buffers). */
int psx_destroy(psx_ctx* c);

/* After the caller wrote a masked spectral state into u0: compute the first stage's
 * transform inputs.  Replaces cupy_cuda.py:134-145 (set_state/_pre, kernel pre2d/pre3d of
 * src/pyspectral/kernels/spectral.cu:93-103,164-173). */
int psx_prepare(psx_ctx* c, void* stream);

/* Set the step size and the band-forcing growth rate for the following steps: the FP64
 * integrating-factor tables are computed on the device into a parameter block the step
 * kernels read, so a captured step graph stays valid.  Replaces cupy_cuda.py:151-161
 * (_update_table) and src/pyspectral/backends/_fused_common.py:10-19 (if_table, if_scalars). */
int psx_set_step(psx_ctx* c, double dt, double amp, void* stream);

/* Phase-shift dealiasing (dealias = 1): the grid shifts of the next step.  shift[d] is the
 * random shift the reference's oracle draws for that step (rng.uniform(0, 2 pi / n), one per
 * axis, numpy_ref.py:125-128); stages 1 and 2 use it, stages 3 and 4 shift + pi / n
 * (numpy_ref.py:130-132, added here in FP64 as there); next[d] is the shift of the step after
 * it, applied to the inputs the last stage prepares.  The prepare call uses shift.  Every
 * stage multiplies its input state by exp(+i k.shift) and its nonlinear term by
 * exp(-i k.shift) (FP64 phases, rounded once).  The values live in the device parameter
 * block, so a captured step graph stays valid.  Replaces numpy_ref.py:68-72,125-153 (the
 * reference's CUDA variant has no phase shift). */
int psx_set_phase_shift(psx_ctx* c, const double* shift, const double* next, void* stream);

/* One IF-RK4 step (4 stages) with the current parameters.  Replaces cupy_cuda.py:163-195
 * (_stage, _step without graph). */
int psx_step(psx_ctx* c, void* stream);

/* Same step replayed from a hipGraph captured on first use (dt and amp live in device
 * memory, so the graph is captured once).  Replaces cupy_cuda.py:181-195 (graph variant). */
int psx_step_graph(psx_ctx* c, void* stream);

/* Kernel kinds reported by psx_timing_end. */
#define PSX_KIND_LINE 0      /* real-axis line kernel: c2r, products, r2c */
#define PSX_KIND_OUTER_FWD 1 /* 2D four-step outer forward pass */
#define PSX_KIND_SPEC 2      /* spectral kernel: forward, physics, inverse along the outer axis */
#define PSX_KIND_OUTER_INV 3 /* 2D four-step outer inverse pass */
#define PSX_KIND_MID_FWD 4   /* 3D middle-axis forward pass */
#define PSX_KIND_MID_INV 5   /* 3D middle-axis inverse pass */
#define PSX_KIND_CTRL 6      /* device time-loop controller and its record reduction */
#define PSX_KIND_SGS 7       /* LES subgrid-stress evaluation at the start of a step */
#define PSX_NKINDS 8

/* Per-launch GPU timing of eager steps with HIP events on the caller's stream: between
 * psx_timing_begin and psx_timing_end every psx_step records an event before each launch;
 * psx_timing_end synchronises and returns, per kernel kind, the summed milliseconds and the
 * launch count.  No reference counterpart (the reference times whole steps with wall clocks,
 * src/pyspectral/benchmark.py:120-142); used by the benchmark's roofline figures. */
int psx_timing_begin(psx_ctx* c);
int psx_timing_end(psx_ctx* c, void* stream, double* ms, int* count);

/* max |u_i| over the physical grid of the current state, per component (2 in 2D, 3 in 3D);
 * synchronises the stream.  Replaces src/pyspectral/diagnostics.py:99-109
 * (max_velocity_sum, which sums these). */
int psx_velocity_max(psx_ctx* c, double* out3, void* stream);

/* FP64 energy, enstrophy, palinstrophy and band energy (klo^2 <= k^2 <= khi^2) of a state
 * with this context's size and precision, rows of n/2+1 coefficients `pitch` elements apart
 * (components n^(d-1) rows apart); synchronises the stream.
 * Replaces src/pyspectral/diagnostics.py:65-88. */
int psx_reduce(psx_ctx* c, const void* state, int pitch, double klo2, double khi2,
               double* out4, void* stream);

/* FP64 shell spectrum E(k), k = 0..nshell-1 (shell = int(|k| + 0.5)), same state layout as
 * psx_reduce; synchronises.  Replaces src/pyspectral/diagnostics.py:90-97. */
int psx_spectrum(psx_ctx* c, const void* state, int pitch, double* out, int nshell,
                 void* stream);

/* ---- device-resident time loop ------------------------------------------------------
 * The loop of src/pyspectral/simulate.py:66-96 (run) with its controller on the device:
 * every cfl_every steps dt = min(dt_max, cfl * (2 pi / n) / sum_i max|u_i|) (:31-37, the
 * maxima taken by the first stage of the same step) and, when power >= 0, the band growth
 * rate A = power / (2 E_band) (:80-83); each step h = min(dt, t_end - t, next_diag - t);
 * a record (t, E, Z) after every step that reaches a diagnostics time and at t_end (:86-96).
 * Steps run in chunks (one captured graph of nsteps controller steps); chunks enqueued past
 * the end are no-ops, so the host can keep several in flight and poll rarely. */
typedef struct psx_run_params {
  double t0;        /* current time */
  double t_end;     /* end time (may be +inf with kmax > 0) */
  double dt_fixed;  /* > 0: fixed dt; otherwise CFL-adaptive with cfl and dt_max */
  double cfl, dt_max;
  double diag_dt;   /* > 0: a record every diag_dt (steps land on it); 0: records at the ends */
  double power;     /* >= 0: constant-power band forcing target P; < 0: A stays amp0 */
  double amp0;      /* band growth rate in force at the start */
  long long kmax;   /* > 0: stop after kmax steps */
  int cfl_every;    /* >= 1 */
  int cap_rec;      /* capacity of the record ring (t, E, Z) */
  int cap_step;     /* capacity of the step ring (h, A) */
  int record_start; /* 1: record the initial state */
  double next_diag; /* first record time when diag_dt > 0 (NaN: t0 + diag_dt) */
  double last_rec_t; /* time of the last record of an earlier part of the same history (NaN:
                        none); a run that begins at t_end records its end state only when that
                        record is older than t0, as simulate.py:93-95 does */
} psx_run_params;

typedef struct psx_run_state {
  double t, amp, cur_dt, next_diag;
  long long k;      /* steps taken; step i is in ring slot i % cap_step */
  long long nrec;   /* records taken; record i is in ring slot i % cap_rec */
  int done;
} psx_run_state;

/* Start a run from the prepared state (after psx_prepare); synchronises the stream. */
int psx_run_begin(psx_ctx* c, const psx_run_params* rp, void* stream);
/* Enqueue nsteps controller steps (as one graph if use_graph, captured on first use),
 * followed by a copy of the controller state to a pinned host slot (8 slots). */
int psx_run_chunk(psx_ctx* c, int nsteps, int use_graph, void* stream);
/* Wait for the chunk enqueued `lag` chunks before the latest (0 <= lag < 8) and return the
 * controller state after it, without synchronising the stream. */
int psx_run_peek(psx_ctx* c, int lag, psx_run_state* out);
/* Synchronise; read the controller state and, when non-null, the rings (3 cap_rec doubles
 * into rec, 2 cap_step doubles into steps). */
int psx_run_poll(psx_ctx* c, psx_run_state* out, double* rec, double* steps, void* stream);
/* End a run: the context steps eagerly again (psx_set_step must follow before psx_step). */
int psx_run_end(psx_ctx* c, void* stream);

/* ---- ensembles of 2D flows ------------------------------------------------------------
 * Independent 2D fp32 Basdevant flows of several sizes (powers of two, 64..2048) advanced
 * together: each phase of a stage is one persistent launch over the tiles of every flow.
 * Each flow follows the single-flow semantics above (the reference has no ensembles; per
 * flow this replaces src/pyspectral/backends/cupy_cuda.py:175-195 and, for the time loop,
 * simulate.py:40-96).  Every member has its own u0/acc/work buffers (sizes from
 * psx_ens_query_layout), parameters, step, forcing amplitude and controller state. */
typedef struct psx_ens psx_ens;

/* Compiled spectral-pass variants for member size n (first = default) into out[0..cap);
 * returns their number (0: n unsupported). */
int psx_ens_variants(int n, int* out, int cap);
/* Buffer sizes of one member planned with spectral variant `variant` (-1: default). */
int psx_ens_query_layout(const psx_flow_params* p, int variant, psx_layout* out);
/* Create an ensemble of `count` members borrowing their buffers. */
int psx_ens_create(int count, const psx_flow_params* params, const int* variants,
                   void* const* u0, void* const* acc, void* const* work, psx_ens** out);
int psx_ens_destroy(psx_ens* e);
/* Persistent work-groups per compute unit of every phase launch (1..16, default 2). */
int psx_ens_set_grid(psx_ens* e, int wg_per_cu);
/* First-stage inputs of every member from its state (after loading the states). */
int psx_ens_prepare(psx_ens* e, void* stream);
/* Per-member step and band growth rate (arrays of count). */
int psx_ens_set_step(psx_ens* e, const double* dt, const double* amp, void* stream);
/* One IF-RK4 step of every member, eagerly / replayed from a graph captured once. */
int psx_ens_step(psx_ens* e, void* stream);
int psx_ens_step_graph(psx_ens* e, void* stream);
/* Per-launch event timing of eager steps, as psx_timing_begin / psx_timing_end. */
int psx_ens_timing_begin(psx_ens* e);
int psx_ens_timing_end(psx_ens* e, void* stream, double* ms, int* count);
/* Device time loop of every member (arrays of count psx_run_params / psx_run_state; the ring
 * capacities and the CFL-or-fixed choice are shared); rings are member-major. */
int psx_ens_run_begin(psx_ens* e, const psx_run_params* rp, void* stream);
int psx_ens_run_chunk(psx_ens* e, int nsteps, int use_graph, void* stream);
int psx_ens_run_peek(psx_ens* e, int lag, psx_run_state* out);
int psx_ens_run_poll(psx_ens* e, psx_run_state* out, double* rec, double* steps, void* stream);
int psx_ens_run_end(psx_ens* e, void* stream);

/* ---- large-eddy simulation support (no counterpart in the reference) ------------------
 * psx_set_source: a 3D context adds src (state layout, or null) to the nonlinear term of every
 * stage; the LES adds the Leray-projected divergence of its subgrid stress this way. */
int psx_set_source(psx_ctx* c, void* src);

/* psx_set_sgs: a 3D fp32 context (L = n = 64, 128 or 256) evaluates a subgrid-stress model
 * from its state at the start of every step -- eager, graph-replayed and controller-driven
 * steps alike -- and adds the result through the source above: velocity gradients by
 * spectral differentiation and inverse 3D FFTs, the stress (Smagorinsky, or the network's six
 * convolutions on MFMA), its forward FFT and the Leray-projected divergence -i k_j tau_ij.
 * Every buffer belongs to the caller and must stay valid while attached; d = NULL detaches
 * (and clears the source).  psx_sgs_update enqueues the same evaluation once, outside a step.
 * Sizes are complex64 counts over the L^3 grid, and [L^3][32] point-major channel rows. */
#define PSX_SGS_SMAGORINSKY 1
#define PSX_SGS_NET 2
typedef struct psx_sgs_desc {
  int model;          /* PSX_SGS_SMAGORINSKY or PSX_SGS_NET */
  void* src;          /* the source, state layout (3 n^2 rows of the state pitch) */
  void* full;         /* scratch: 3 L^3 */
  void* grad;         /* scratch: 6 L^3 */
  void* tau;          /* scratch: 3 L^3 */
  void* x;            /* network input, bf16 [L^3][32] */
  const void* w[6];   /* network weights, bf16 [27][32 out][32 in] per layer */
  const float* b[6];  /* network biases, fp32 [32] per layer */
  void* act[5];       /* hidden activations, bf16 [L^3][32] */
  float* out;         /* network output, fp32 [L^3][32] (channels 0..5 used) */
  double g_scale;     /* input scale: 1 / velocity-gradient rms */
  double tau_scale;   /* output scale: stress rms */
  double smag;        /* Smagorinsky (Cs Delta)^2; with PSX_SGS_NET, > 0 adds the Smagorinsky
                         stress to the network's (a mixed model), 0 leaves it out */
  double backscatter; /* network: factor in [0, 2] on the stress where -tau_ij S_ij < 0, i.e.
                         where it would move energy to the resolved scales (1 keeps it, 0 clips
                         it, > 1 amplifies it) */
  const float* shell_gain; /* optional device array: factor on the source per wavenumber shell
                              (int)(|k| + 0.5), floor(sqrt(3) L / 2 + 0.5) + 1 entries; null:
                              none */
} psx_sgs_desc;
int psx_set_sgs(psx_ctx* c, const psx_sgs_desc* d);
int psx_sgs_update(psx_ctx* c, void* stream);

/* In-place unnormalised 3D complex FFT (complex64) of nbatch contiguous L^3 cubes
 * (L = 64, 128, 256; z fastest); inverse = 1 for exp(+i k.x). */
int psx_fft3(void* data, int L, int nbatch, int inverse, void* stream);

/* Spectral and pointwise LES kernels (fp32 complex unless stated); `which` selects the
 * operation (PSX_LES_*), the fields of psx_les_args are interpreted as documented there. */
typedef struct psx_les_args {
  const void* in;   /* main input */
  const void* in2;  /* second input */
  void* out;        /* output */
  long long n;      /* count (components, pairs or values) */
  int L, L2, pitch; /* grid size, second grid size, row pitch or flag */
  double scale;
  double scale2;    /* second factor (PSX_LES_TAU_FROM_NET: backscatter factor) */
  double scale3;    /* third factor (PSX_LES_TAU_FROM_NET: Smagorinsky (Cs D)^2 added, 0: none) */
} psx_les_args;
#define PSX_LES_EXPAND 0     /* state (pitch) with n components -> full L^3 spectra */
#define PSX_LES_GRADPACK 1   /* 3 full spectra -> n = 6 pairs (u,v)(w,ux)(uy,uz)(vx,vy)(vz,wx)(wy,wz), or n = 1: (u,v)(w,0); times scale */
#define PSX_LES_PRODUCTS 2   /* physical velocity pairs -> pairs (uu,vv)(ww,uv)(uw,vw) */
#define PSX_LES_TRUNCATE 3   /* fine (L) spectra, n arrays, pitch = 1: split pairs -> coarse (L2) spectra, box filter k_c = scale (0: none) */
#define PSX_LES_PAIRUP 4     /* 2n full spectra -> n pairs, times scale */
#define PSX_LES_FEATURES 5   /* physical gradient pairs -> bf16 [L^3][32] model input, times scale */
#define PSX_LES_SMAGORINSKY 6 /* physical gradient pairs -> stress pairs, scale = (Cs D)^2 */
#define PSX_LES_TAU_FROM_NET 7 /* in2: fp32 [L^3][32] model output -> stress pairs, times scale; with in = physical gradient pairs, the stress is multiplied by scale2 where -tau_ij S_ij < 0 (backscatter; 0 clips it), and for scale3 > 0 the Smagorinsky stress with (Cs D)^2 = scale3 is added */
#define PSX_LES_TARGET 8     /* filtered product pairs, in2: velocity pairs -> fp32 [L^3][8] stress, times scale */
#define PSX_LES_SOURCE 9     /* stress pair spectra -> projected -i k_j tau_ij, state layout (pitch); in2: null or per-shell factors (as psx_sgs_desc.shell_gain) */
#define PSX_LES_SUMSQ 10     /* in2: n fp32 values in rows of `pitch`, first L2 of each -> += sum of squares (double) */
int psx_les_op(int which, const psx_les_args* a, void* stream);

/* Convolutional subgrid-stress network: periodic 3x3x3 convolution of nb cubes of n^3 points,
 * 32 bf16 channels in and out (point-major), weights [27][32 out][32 in] bf16, fp32
 * accumulation on MFMA; optional bias, ReLU, and (input-gradient use) a multiply by
 * (mask > 0).  Output bf16 y and/or fp32 yf. */
typedef struct psx_conv_args {
  const void* x;
  const void* w;
  const float* bias;
  const void* mask;
  void* y;
  float* yf;
  int n, nb, relu;
} psx_conv_args;
int psx_cnn_conv(const psx_conv_args* a, void* stream);
/* dw[27][32][32] += sum_p dz[p][o] x[p + tap][i]; db[32] += sum_p dz[p][o] (n = 64,
 * 1 <= nb <= 4194303).  The per-work-group partial sums (56.7 MB,
 * psx_cnn_wgrad_scratch_bytes()) live in a buffer of this call's own, allocated and freed in
 * stream order (hipMallocAsync / hipFreeAsync on `stream`), so calls on different streams do
 * not share it. */
int psx_cnn_wgrad(const void* x, const void* dz, float* dw, float* db, int n, int nb,
                  void* stream);
/* The same weight gradient with no allocation, bit for bit the result of psx_cnn_wgrad.  No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight training or serving loops, harnesses that forbid
 * hipMalloc).  The partial sums go to a static array of psx_cnn_wgrad_scratch_bytes() bytes
 * (56.7 MB) per GPU, held until the process ends; its size is fixed by the kernel's 512
 * work-groups, not by nb, and nb is limited as for psx_cnn_wgrad (nb > 4194303 is rejected
 * with an error).  Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other streams, other threads) race and must be serialised by the caller. */
int psx_cnn_wgrad_noalloc(const void* x, const void* dz, float* dw, float* db, int n, int nb,
                          void* stream);
unsigned long psx_cnn_wgrad_scratch_bytes(void);
/* g = 2 (y - t) / (npts nch) on the first nch channels (bf16 [npts][32]); loss += mean square. */
int psx_cnn_mse(const float* y, const float* t, int tstride, int nch, long long npts, void* g,
                double* loss, void* stream);
/* Adam step on n fp32 values (bias corrections c1 = 1/(1-b1^t), c2 = 1/(1-b2^t)); writes the
 * bf16 copy wb and, for a weight tensor (is_w), the flipped-transposed copy wt used by the
 * input gradient: wt[26 - tap][i][o] = w[tap][o][i]. */
typedef struct psx_adam_args {
  float* w;
  float* m;
  float* v;
  const float* g;
  void* wb;
  void* wt;
  long long n;
  double lr, b1, b2, eps, c1, c2;
  int is_w;
} psx_adam_args;
int psx_cnn_adam(const psx_adam_args* a, void* stream);

/* ---- FFT engine test hooks (no counterpart in the reference, whose FFTs are cuFFT) -------
 * psx_fft_lines: in-place complex fp32 FFT of nrows contiguous rows of length L with the plan,
 * work-group shape and twiddle tables of the line kernel for that length (L = 64..8192 powers
 * of two: the 2D kernels'; 96, 192, 384: the 3D kernels'); inverse = 1 for exp(+2 pi i k n / L),
 * unnormalised.  psx_fft_twiddles: that plan's radices (forward or reversed order) and its
 * per-pass twiddle table in that order, as float pairs (pass q of radix R after Ns points holds
 * exp(-2 pi i m k / (Ns R)) at (m - 1) Ns + k, m < R, k < Ns; passes in order, pass 0 none). */
int psx_fft_lines(void* data, int L, int nrows, int inverse, void* stream);
int psx_fft_twiddles(int L, int inverse, float* out, int cap, int* radices, int* npasses);

/* Message of the last failure on this thread. */
const char* psx_last_error(void);

/* Device name and architecture of the current device into buf; 0 if a gfx950 device is
 * present.  Replaces the availability probe of cupy_cuda.py:84-93. */
int psx_device_info(char* buf, int len);

#ifdef __cplusplus
}
#endif

#endif /* PSX_H_ */
