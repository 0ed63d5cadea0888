/* C-ABI of libfvhoe_hip.so: the MI355X (gfx950) per-RK-stage update of
 * FVHOE's finite-volume Euler solver.
 *
 * Conventions
 *  - Arrays are device pointers to contiguous float64 data in the reference's
 *    layout (nvars, n0, n1, n2), C order; variable order rho, vx|mx, vy|my,
 *    vz|mz, P|E, passives (fvhoe/hydro.py:36-63).
 *  - Axes are numbered 1..3 as the reference's array axes (fvhoe/fv.py:106).
 *  - `stream` is a hipStream_t (NULL = the default stream).  Calls enqueue
 *    work and return; nothing here synchronises with the host.
 *  - The library allocates no device memory: outputs and scratch are the
 *    caller's.
 *  - Return: FVHOE_OK, or an error code; the Python layer maps
 *    FVHOE_ERR_ARG to ValueError, FVHOE_ERR_UNSUPPORTED to
 *    NotImplementedError and FVHOE_ERR_LAUNCH to RuntimeError, the exception
 *    types the reference raises for the same conditions.
 */
#ifndef FVHOE_HIP_H
#define FVHOE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define FVHOE_ABI_VERSION 1

enum {
  FVHOE_OK = 0,
  FVHOE_ERR_ARG = 1,
  FVHOE_ERR_UNSUPPORTED = 2,
  FVHOE_ERR_LAUNCH = 3,
  FVHOE_ERR_TOO_LARGE = 4  /* the problem needs more than the fixed device scratch of fvhoe_stage_rk_static */
};

/* stencil kinds, Riemann solvers, limiters, boundary kinds */
enum { FVHOE_STENCIL_L = 0, FVHOE_STENCIL_C = 1, FVHOE_STENCIL_R = 2, FVHOE_STENCIL_T = 3 };
enum { FVHOE_RS_UPWIND = 0, FVHOE_RS_LLF = 1, FVHOE_RS_HLLC = 2 };
enum { FVHOE_LIM_NONE = 0, FVHOE_LIM_MINMOD = 1, FVHOE_LIM_MONCEN = 2 };
enum { FVHOE_BC_PERIODIC = 0, FVHOE_BC_OUTFLOW = 1, FVHOE_BC_REFLECTIVE = 2, FVHOE_BC_FREE = 3,
       FVHOE_BC_DIRICHLET = 4 /* "dirichlet" and "ic": ghost values from a slab */ };

/* Runge-Kutta combinations (fvhoe/ode.py:245-298), dt read from device */
enum {
  FVHOE_RK_AXPY = 0,           /* u + dt*k0                                   (ode.py:251,263,277,293) */
  FVHOE_RK_SSPRK3_STAGE3 = 1,  /* u + 0.25*dt*k0 + 0.25*dt*k1                 (ode.py:278) */
  FVHOE_RK_SSPRK3_FINAL = 2,   /* u + (1/6)*dt*(k0 + k1 + 4*k2)               (ode.py:279) */
  FVHOE_RK_SSPRK2_FINAL = 3,   /* 0.5*u + 0.5*(k0 + dt*k1), k0 holding u1     (ode.py:265) */
  FVHOE_RK_HALF_AXPY = 4,      /* u + 0.5*dt*k0                               (ode.py:291-292) */
  FVHOE_RK_RK4_FINAL = 5,      /* u + (1/6)*dt*(k0 + 2*k1 + 2*k2 + k3)        (ode.py:294) */
  /* fused stage only: SSP-RK3 with k0 + k1 formed once (the sum the final
   * combine starts with, the same rounding) so the last stage reads one
   * array less */
  FVHOE_RK_SSPRK3_STAGE3_SUM = 6,  /* as SSPRK3_STAGE3; dudt receives k0 + this stage's k */
  FVHOE_RK_SSPRK3_FINAL_SUM = 7    /* u + (1/6)*dt*(k0 + 4*k2), k0 holding k0 + k1 */
};

/* ---------------------------------------------------------------- host-only */

int fvhoe_version(void);

/* Weights of get_conservative_interpolation_stencil_weights (kind l/c/r) and
 * get_transverse_reconstruction_stencil_weights (kind t), fvhoe/stencils.py:25-255.
 * out holds >= 9 doubles; *n receives the tap count 2*ceil(p/2)+1. */
int fvhoe_stencil_weights(int p, int kind, double* out, int* n);

/* ------------------------------------------------------- elementary kernels */

/* stencil_sweep / conservative_interpolation / transverse_reconstruction
 * (fvhoe/fv.py:106-172): out is u's shape shortened by 2*ceil(p/2) on axis. */
int fvhoe_stencil_sweep(const double* u, double* out, int nv, int n0, int n1, int n2, int kind, int p, int axis,
                        void* stream);

/* compute_primitives / compute_conservatives (fvhoe/hydro.py:146-196) */
int fvhoe_compute_primitives(const double* u, double* w, int nv, long long ncell, double gamma, void* stream);
int fvhoe_compute_conservatives(const double* w, double* u, int nv, long long ncell, double gamma, void* stream);

/* compute_sound_speed (fvhoe/hydro.py:199-210); w has >= 5 variables of ncell values */
int fvhoe_compute_sound_speed(const double* w, double* c, long long ncell, double gamma, double csq_floor,
                              void* stream);

/* hydro_dt (fvhoe/hydro.py:270-294): *dt_out = CFL*h / max(|vx|+|vy|+|vz| + ndim*c);
 * dt_out and scratch (1 x u64) are device memory. */
int fvhoe_hydro_dt(const double* w, long long ncell, double h, int ndim, double CFL, double gamma, double csq_floor,
                   double* dt_out, unsigned long long* scratch, void* stream);

/* advection_upwind / llf / hllc (fvhoe/riemann_solvers.py:13,48,91), pointwise
 * on contiguous (nv, ncell) primitive states; dim 0/1/2 = x/y/z. */
int fvhoe_riemann(const double* wl, const double* wr, double* out, int nv, long long ncell, int solver, int dim,
                  double gamma, double csq_floor, void* stream);

/* The Riemann call of hydrofluxes (fvhoe/solver.py:575-606): out[f] =
 * RS(wr_face[f], wl_face[f+1]) along axis, nv = 5. */
int fvhoe_riemann_faces(const double* wr_face, const double* wl_face, double* out, int n0, int n1, int n2, int axis,
                        int solver, double gamma, double csq_floor, void* stream);

/* Two sweeps of one kind (c or t) and degree p in one pass, along axis (1 = x
 * or 2 = y) and along z (3), z first if zfirst: the same result as two
 * fvhoe_stencil_sweep calls (fvhoe/fv.py:106-224), bit for bit.  Needs
 * n2 >= 64 (else FVHOE_ERR_UNSUPPORTED: sweep separately). */
int fvhoe_stencil_sweep_zpair(const double* u, double* out, int nv, int n0, int n1, int n2, int kind, int p, int axis,
                              int zfirst, void* stream);

/* Two sweeps of one kind (c or t) and degree p in one pass along the two
 * strided axes, first_axis (1 = x or 2 = y) first, then the other: the same
 * result as two fvhoe_stencil_sweep calls (fvhoe/fv.py:106-224), bit for
 * bit, without storing the intermediate array.  For arrays of 2^16 to 2^31
 * elements (else FVHOE_ERR_UNSUPPORTED: sweep separately). */
int fvhoe_stencil_sweep_xypair(const double* u, double* out, int nv, int n0, int n1, int n2, int kind, int p,
                               int first_axis, void* stream);
int fvhoe_stencil_sweep_xypair_f32(const float* u, float* out, int nv, int n0, int n1, int n2, int kind, int p,
                                   int first_axis, void* stream);

/* The l / r conservative interpolations along axis and the Riemann call of
 * hydrofluxes in one pass (fvhoe/fv.py:143-159, fvhoe/solver.py:538-606):
 * out[f] = RS(r-sweep(w)[f], l-sweep(w)[f+1]) for the 5-variable primitive
 * array w (5, n0, n1, n2); out has axis length n_axis - (2 ceil(p/2) + 1).
 * Bit-identical to fvhoe_stencil_sweep (kinds r, l) + fvhoe_riemann_faces. */
int fvhoe_face_fluxes(const double* w, double* out, int n0, int n1, int n2, int axis, int p, int solver, double gamma,
                      double csq_floor, void* stream);

/* minmod (limiter FVHOE_LIM_MINMOD) and moncen (FVHOE_LIM_MONCEN) of two
 * difference arrays of n values (fvhoe/slope_limiting.py:56-84). */
int fvhoe_slope_limit(int limiter, const double* du_left, const double* du_right, double* out, long long n,
                      void* stream);

/* second_order_central_difference (fvhoe/fv.py:280-296) of u viewed as
 * (outer, len, inner) along its middle axis: out (outer, len - 2, inner) =
 * 0.5 * (u[l + 2] - u[l]), then / h when has_h. */
int fvhoe_central_difference(const double* u, double* out, long long outer, long long len, long long inner,
                             double h, int has_h, void* stream);

/* compute_fluxes (fvhoe/hydro.py:213-246): the physical flux along dim (0/1/2
 * = x/y/z) of (nv, ncell) primitives w and conservatives u, with or without
 * the pressure terms; passives (nv > 5) are advected. */
int fvhoe_compute_fluxes(const double* w, const double* u, double* out, int nv, long long ncell, int dim,
                         int include_pressure, void* stream);

/* MUSCL_interpolations with minmod / moncen / no limiter (fvhoe/slope_limiting.py:24-84) */
int fvhoe_muscl(const double* u, double* left, double* right, int nv, int n0, int n1, int n2, int axis, int limiter,
                void* stream);

/* BoundaryCondition.apply for periodic / outflow / reflective / free
 * (fvhoe/boundary_conditions.py:306-402): out has shape (nv, n0+2g0, n1+2g1, n2+2g2);
 * bc6 = kinds of (x lo, x hi, y lo, y hi, z lo, z hi). */
int fvhoe_bc_apply(const double* u, double* out, int nv, int n0, int n1, int n2, int g0, int g1, int g2,
                   const int* bc6, void* stream);

/* BoundaryCondition.apply (periodic / outflow / reflective / free) followed
 * by the first stencil sweep of interpolate_cell_centers or
 * interpolate_fv_averages (fvhoe/solver.py:774-789, fvhoe/fv.py:106-224)
 * along axis (1 = x or 2 = y), in one pass: out has the padded shape
 * (nv, n0+2g0, n1+2g1, n2+2g2) shortened by 2*ceil(p/2) on axis, and equals
 * fvhoe_bc_apply + fvhoe_stencil_sweep bit for bit (_f32: the float32 fill +
 * fvhoe_stencil_sweep_f32).  For padded arrays of 2^16 to 2^31 elements and
 * p >= 1; FVHOE_ERR_UNSUPPORTED otherwise (fill and sweep separately). */
int fvhoe_bc_stencil_sweep(const double* u, double* out, int nv, int n0, int n1, int n2, int g0, int g1, int g2,
                           const int* bc6, int kind, int p, int axis, void* stream);
int fvhoe_bc_stencil_sweep_f32(const double* u, float* out, int nv, int n0, int n1, int n2, int g0, int g1, int g2,
                               const int* bc6, int kind, int p, int axis, void* stream);

/* The same with FVHOE_BC_DIRICHLET sides ("dirichlet" / "ic",
 * fvhoe/boundary_conditions.py:18-57,247-304,334-348).  slabs6[s] (device,
 * NULL for the other sides) holds side s's ghost values, shaped as that
 * side's ghost region of out: for an axis-d side nv x g_d along d x
 * (n_e + 2*g_e) along each other axis e, C order (the reference's slab view).
 * Later axes decide corners, as the reference's x, y, z slab order does. */
int fvhoe_bc_apply_dirichlet(const double* u, double* out, int nv, int n0, int n1, int n2, int g0, int g1, int g2,
                             const int* bc6, const double* const* slabs6, void* stream);

/* EulerSolver.euler_equation (fvhoe/solver.py:734-751); c* = -(1/h*); active bit d = axis d+1 */
int fvhoe_euler_equation(const double* F, const double* G, const double* H, double* dudt, int nv, int n0, int n1,
                         int n2, int active, double cx, double cy, double cz, void* stream);

/* FP32 bulk form of the composed stage: the same operations with float32
 * intermediates (no reference counterpart; the reference computes in float64
 * only).  Sums, face states and Riemann solves run in float with the weights
 * rounded to float once; the pressure recovery (fvhoe/hydro.py:146-170), the
 * CFL maximum (fvhoe/hydro.py:270-294) and the divergence
 * (fvhoe/solver.py:734-751) run in double on the float values.  Shapes,
 * argument meaning and errors as the float64 entries above.
 * fvhoe_bc_apply_f32 reads the float64 state and writes its float32 rounding
 * (non-slab kinds); fvhoe_euler_equation_f32 writes a float64 dudt. */
int fvhoe_stencil_sweep_f32(const float* u, float* out, int nv, int n0, int n1, int n2, int kind, int p, int axis,
                            void* stream);
int fvhoe_stencil_sweep_zpair_f32(const float* u, float* out, int nv, int n0, int n1, int n2, int kind, int p,
                                  int axis, int zfirst, void* stream);
int fvhoe_compute_primitives_f32(const float* u, float* w, int nv, long long ncell, double gamma, void* stream);
int fvhoe_hydro_dt_f32(const float* w, long long ncell, double h, int ndim, double CFL, double gamma,
                       double csq_floor, double* dt_out, unsigned long long* scratch, void* stream);
int fvhoe_face_fluxes_f32(const float* w, float* out, int n0, int n1, int n2, int axis, int p, int solver,
                          double gamma, double csq_floor, void* stream);
int fvhoe_bc_apply_f32(const double* u, float* out, int nv, int n0, int n1, int n2, int g0, int g1, int g2,
                       const int* bc6, void* stream);
int fvhoe_euler_equation_f32(const float* F, const float* G, const float* H, double* dudt, int nv, int n0, int n1,
                             int n2, int active, double cx, double cy, double cz, void* stream);

/* EulerSolver.euler_equation (fvhoe/solver.py:734-751) on face-flux arrays
 * given by element strides: fst / gst / hst = (variable, axis-0, axis-1)
 * strides of F / G / H, whose last axis is contiguous (NULL: the contiguous
 * shapes of fvhoe_euler_equation).  The fluxes may be views into larger
 * arrays, as the trimmed fluxes of hydrofluxes (fvhoe/solver.py:609-649), so
 * no trimmed copies are made.  Bit-identical to fvhoe_euler_equation. */
int fvhoe_euler_equation_strided(const double* F, const long long* fst, const double* G, const long long* gst,
                                 const double* H, const long long* hst, double* dudt, int nv, int n0, int n1, int n2,
                                 int active, double cx, double cy, double cz, void* stream);
int fvhoe_euler_equation_strided_f32(const float* F, const long long* fst, const float* G, const long long* gst,
                                     const float* H, const long long* hst, double* dudt, int nv, int n0, int n1,
                                     int n2, int active, double cx, double cy, double cz, void* stream);

/* per-variable (min, max) over ncell values: out[2v] = min, out[2v+1] = max
 * (the global range of fvhoe/slope_limiting.py:168-170,184-185); scratch 2*nv x u64 */
int fvhoe_minmax(const double* u, int nv, long long ncell, double* out, unsigned long long* scratch, void* stream);

/* detect_troubled_cells (fvhoe/slope_limiting.py:87-244) on padded primitive
 * averages w, ws of shape (5, N0, N1, N2) with 3 ghosts on each active axis
 * (dims bit d = axis d+1).  range_minmax = fvhoe_minmax(w) output (mode global).
 * mode 0 global / 1 local; range_type 0 relative / 1 absolute; nad_vars bit
 * mask; pad_var = variable of the last PAD key (-1: none).  Outputs are the
 * interior cells. */
int fvhoe_detect_troubled_cells(const double* w, const double* ws, int N0, int N1, int N2, int dims,
                                const double* range_minmax, int mode, int range_type, double eps, int nad_vars,
                                int pad_var, double pad_lo, double pad_hi, int sed, double sed_eps,
                                signed char* trouble, double* nad_mag, double* pad_mag, void* stream);

/* compute_{1,2,3}d_smooth_extrema_detector (fvhoe/smooth_extrema_detection.py:8-95) */
int fvhoe_smooth_extrema_detector(const double* u, double* alpha, int nv, int N0, int N1, int N2, int dims,
                                  double eps, void* stream);

/* broadcast_to_troubled_interfaces, simple form (fvhoe/slope_limiting.py:364-380) */
int fvhoe_broadcast_simple(const signed char* trouble, double* tx, double* ty, double* tz, int n0, int n1, int n2,
                           int dims, void* stream);

/* out[r] = max(coef * g(s1[r]), s2[r]) over a strided 3-D region, g(x) = x or
 * (x > 0); the update step of the convex broadcast (fvhoe/slope_limiting.py:387-442) */
int fvhoe_shift_max(double* out, const long long* out_strides, long long out_offset, const double* s1,
                    const long long* s1_strides, long long s1_offset, double coef, int gt, const double* s2,
                    const long long* s2_strides, long long s2_offset, const int* extent, void* stream);

/* F = (1 - theta)*F + theta*Fl, theta broadcast over variables (fvhoe/solver.py:729-731) */
int fvhoe_blend(double* F, const double* Fl, const double* theta, int nv, long long nface, void* stream);

/* Runge-Kutta combinations, FVHOE_RK_* kinds; dt is a device scalar.  dt = 0
 * (a step already at its target time) writes out = u exactly, here and in the
 * fused stage's combine. */
int fvhoe_rk_combine(int kind, double* out, const double* u, const double* k0, const double* k1, const double* k2,
                     const double* k3, const double* dt, long long n, void* stream);

/* EulerSolver.log_troubles (fvhoe/solver.py:794-833) on device: mode 0
 * mean = trouble, 1 mean += trouble, 2 mean /= divisor (trouble unused). */
int fvhoe_log_troubles(double* mean, const signed char* trouble, long long n, int mode, double divisor,
                       void* stream);
/* The same, doing nothing when *gate (device: the step's dt) is 0. */
int fvhoe_log_troubles_gated(double* mean, const signed char* trouble, long long n, int mode, double divisor,
                             const double* gate, void* stream);
/* The trouble log kept as counts between resets: mode 0 count = trouble,
 * 1 count += trouble; fvhoe_log_mean then writes mean = count / divisor,
 * which equals the reference's running double sum divided at the reset (sums
 * of 0/1 values are exact).  gate (device, may be NULL) as above. */
int fvhoe_log_count(int* count, const signed char* trouble, long long n, int mode, const double* gate, void* stream);
int fvhoe_log_mean(double* mean, const int* count, long long n, double divisor, const double* gate, void* stream);
/* The "mean NAD mag" / "mean PAD mag" arrays of the trouble log
 * (fvhoe/solver.py:794-833) from the fused stage's magnitude slots
 * (FvhoeStageConfig.mag_*; cap entries per slot): nad_out / pad_out (n
 * cells) = base + slot 0 + slot 1 + ... + slot nslots-1, added in slot
 * order cell by cell, then divided by *divisor when divisor is not NULL.
 * base: nad_base / pad_base, or zeros when they are NULL.  A cell missing
 * from a slot adds 0, so the sums and the quotient are the reference's
 * running double sums bit for bit (magnitudes are >= 0). */
int fvhoe_log_magnitudes(double* nad_out, double* pad_out, const double* nad_base, const double* pad_base,
                         long long n, const int* count, const int* idx, const double* nad, const double* pad,
                         int nslots, long long cap, const double* divisor, void* stream);
/* *p = v unless *gate (device, may be NULL) is 0. */
int fvhoe_set_scalar_gated(double* p, double v, const double* gate, void* stream);

/* device scalars: *p = v; *dt_step = _hit_target(t, dt_stage, target) (fvhoe/ode.py:11-22); *t += *dt */
int fvhoe_set_scalar(double* p, double v, void* stream);
int fvhoe_hit_target(double* dt_step, const double* dt_stage, const double* t, double target, int has_target,
                     void* stream);
int fvhoe_advance_time(double* t, const double* dt, void* stream);
/* For a step that moved (*dt != 0): *t += *dt, times[*steps] = *t while
 * *steps < cap, ++*steps (device scalars; steps after the target time move
 * nothing and count nothing). */
int fvhoe_advance_time_counted(double* t, const double* dt, long long* steps, double* times, long long cap,
                               void* stream);

/* ------------------------------------------------------------ fused stage */

/* Everything one stage needs: the EulerSolver arguments that reach
 * hydrodynamics (fvhoe/solver.py:42-84). */
typedef struct FvhoeStageConfig {
  int n[3];             /* interior cells per axis */
  int p[3];             /* polynomial degree per axis (inactive axis: n = 1, p = 0) */
  double h[3];          /* mesh spacing */
  double gamma, cfl, csq_floor;
  int bc[6];            /* FVHOE_BC_* for x lo, x hi, y lo, y hi, z lo, z hi */
  int riemann;          /* FVHOE_RS_* */
  int a_posteriori;     /* a_posteriori_slope_limiting */
  int limiter;          /* fallback slope limiter, FVHOE_LIM_* */
  int force_trouble;
  double nad_eps;
  int nad_mode;         /* 0 global, 1 local */
  int nad_range;        /* 0 relative, 1 absolute */
  int nad_vars;         /* bit mask over rho, vx, vy, vz, P */
  int pad_var;          /* variable of the last PAD key, -1 for none */
  double pad_lo, pad_hi;
  int sed;
  double sed_eps;
  int fixed_dt_on;
  double fixed_dt;
  int bulk_fp32;        /* 1: FP32 sweeps, Riemann solves and W/F/G storage; FP64 pressure
                           recovery, trouble tests, divergence and RK accumulation */
  /* ghost slabs of the FVHOE_BC_DIRICHLET sides, laid out as for
   * fvhoe_bc_apply_dirichlet (nv = 5, slab_pad[2] = 0); slab_g and the pads
   * of the active axes must reach the stage's deepest ghost
   * (fvhoe_stage_ghost_depth).  Contents are read at launch time, so a
   * time-dependent boundary refreshes them between steps. */
  const double* slab[6];
  int slab_g;
  int slab_pad[3];
  /* Violation magnitudes for the trouble log (fvhoe/slope_limiting.py:225-239,
   * fvhoe/solver.py:794-833).  With mag_idx set, an a-posteriori stage
   * appends (cell, NAD magnitude, PAD magnitude) for every cell whose NAD
   * or PAD test flags to slot mag_slot: mag_count[mag_slot] entries at
   * mag_idx / mag_nad / mag_pad + mag_slot * n0 * n1 (each slot holds
   * n0 * n1 entries, every cell at most once).  The stage clears its slot
   * first.  Under a FvhoeRkFuse a step whose *dt_step is 0 (a step replayed
   * at its target time) leaves the slot as it was.  fvhoe_log_magnitudes
   * turns slots into the log's dense arrays.  mag_idx NULL: nothing logged. */
  int* mag_count;
  int* mag_idx;
  double* mag_nad;
  double* mag_pad;
  int mag_slot;
} FvhoeStageConfig;

/* 1 if fvhoe_stage handles cfg (1-D along x, or 2-D in x-y, nvars = 5,
 * simple face broadcast), else 0. */
int fvhoe_stage_supported(const FvhoeStageConfig* cfg);

/* Deepest ghost cell (cells beyond the interior) the fused stage reads from
 * u for cfg: the slab depth its Dirichlet sides need. */
int fvhoe_stage_ghost_depth(const FvhoeStageConfig* cfg, int* depth);

/* Bytes of device workspace fvhoe_stage needs for cfg. */
int fvhoe_stage_workspace_bytes(const FvhoeStageConfig* cfg, unsigned long long* bytes);

/* One EulerSolver.hydrodynamics call (fvhoe/solver.py:429-454) fused into a
 * few kernels: reads u (5, n0, n1, n2), writes dudt (same shape) and the
 * stage's CFL dt to *dt_stage (device).  trouble (n0*n1*n2 int8, may be NULL)
 * receives the a-posteriori flags.  workspace: fvhoe_stage_workspace_bytes. */
int fvhoe_stage(const FvhoeStageConfig* cfg, const double* u, double* dudt, double* dt_stage, signed char* trouble,
                void* workspace, void* stream);

/* The Runge-Kutta step fused into a stage: the stage's last kernel also
 * writes out = combine(kind, u0, k0, k1, k2, this stage's dudt, *dt_step),
 * this stage's dudt taking the place of the kind's last k (k0 for AXPY and
 * HALF_AXPY, k1 for SSPRK3_STAGE3 and SSPRK2_FINAL, k2 for SSPRK3_FINAL, k3
 * for RK4_FINAL).  With set_dt_step, *dt_step = _hit_target(*t, this stage's
 * dt, target) first (fvhoe/ode.py:11-22, 274-279).  out may alias u0 or the
 * stage input; dudt may be NULL when a combine is given. */
typedef struct FvhoeRkFuse {
  int kind;
  double* out;
  const double* u0;
  const double* k0;
  const double* k1;
  const double* k2;
  double* dt_step;
  int set_dt_step;
  const double* t;
  double target;
  int has_target;
  const double* target_ptr;  /* device target time read at run time instead of target (may be NULL) */
} FvhoeRkFuse;

/* The composed stage's divergence (fvhoe_euler_equation_strided) and the
 * step's RK combination (fvhoe_rk_combine, fvhoe/ode.py:245-298) in one
 * pass: rk->out = combination of kind rk->kind with the stage's dudt as its
 * last k and rk->dt_step as dt; dudt (may be NULL) receives dudt, or
 * rk->k0 + dudt for FVHOE_RK_SSPRK3_STAGE3_SUM.  rk->set_dt_step and the
 * target fields are not used (set dt_step first, e.g. with
 * fvhoe_hit_target).  Bit-identical to the two separate calls. */
int fvhoe_euler_equation_rk(const double* F, const long long* fst, const double* G, const long long* gst,
                            const double* H, const long long* hst, double* dudt, int nv, int n0, int n1, int n2,
                            int active, double cx, double cy, double cz, const FvhoeRkFuse* rk, void* stream);
int fvhoe_euler_equation_rk_f32(const float* F, const long long* fst, const float* G, const long long* gst,
                                const float* H, const long long* hst, double* dudt, int nv, int n0, int n1, int n2,
                                int active, double cx, double cy, double cz, const FvhoeRkFuse* rk, void* stream);

int fvhoe_stage_rk(const FvhoeStageConfig* cfg, const double* u, double* dudt, double* dt_stage,
                   signed char* trouble, void* workspace, const FvhoeRkFuse* rk, void* stream);

/* fvhoe_stage_rk without a workspace argument: the stage's scratch (the
 * arrays fvhoe_stage_workspace_bytes counts) is a fixed device array of
 * FVHOE_STATIC_SCRATCH_BYTES inside the library, one per GPU, held until the
 * process ends, so the call allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Returns FVHOE_ERR_TOO_LARGE when cfg needs more scratch than that (with
 * a-posteriori limiting it fits 2-D grids up to 612 x 612 cells in FP64 and
 * 828 x 828 in FP32 at p=3, 1-D grids up to about 535,000 cells at p=5).  Every call on
 * one GPU shares the array: calls that overlap on one GPU (concurrent streams
 * or threads) race.  Results equal fvhoe_stage_rk's bit for bit. */
#define FVHOE_STATIC_SCRATCH_BYTES (64ull << 20)
int fvhoe_stage_rk_static(const FvhoeStageConfig* cfg, const double* u, double* dudt, double* dt_stage,
                          signed char* trouble, const FvhoeRkFuse* rk, void* stream);

/* Bench instrumentation: with timing on, every fused-stage kernel launch is
 * bracketed by hipEvents on its stream.  fvhoe_stage_timing_read waits for
 * the recorded events and returns, per kernel (prims, flux x, flux y,
 * candidate prims, detect, fallback, final, predictor features + inference),
 * accumulated milliseconds and launch counts (8 entries each).  fvhoe_stage_timing(on) also resets the sums.
 * No counterpart in the reference (its Timer takes host wall-clock,
 * fvhoe/timer.py:6-109). */
int fvhoe_stage_timing(int on);
int fvhoe_stage_timing_read(double* ms, long long* count);

/* ------------------------------------------------------------------------
 * Learned troubled-cell predictor.  No counterpart in the reference: it is the
 * learned part of the stage.  A three-layer valid 5x5 convolutional network,
 * 4 -> 16 -> 16 -> 1 channels, ReLU after the first two layers, on features
 * X (4, n0+12, n1+12) float: the primitive fv averages (rho, vx, vy, P) of the
 * stage input, (w - min) * (1 / (max - min)) in FP64 with the stage's global NAD range,
 * extended by 6 ghost cells per side with the solver's boundary conditions.
 * Parameters theta (FVHOE_PRED_NPARAMS float): w1 (16,4,5,5), b1 (16),
 * w2 (16,16,5,5), b2 (16), w3 (1,16,5,5), b3 (1).  The network sees its
 * weights rounded to bfloat16; inference runs on BF16 MFMA with FP32
 * accumulation, training in FP32 (gradients w.r.t. the rounded weights are
 * applied to the FP32 master copy).
 * ---------------------------------------------------------------------- */
#define FVHOE_PRED_NPARAMS 8433
#define FVHOE_PRED_PACKED_BYTES 24576
#define FVHOE_PRED_HALO 6

/* theta -> packed inference weights (FVHOE_PRED_PACKED_BYTES, device). */
int fvhoe_pred_pack(const float* theta, void* packed, void* stream);

/* theta -> packed weights of the planar network (FVHOE_PRED_PLANAR_BYTES,
 * device): the network on 1-D runs seen as 2-D states constant along y,
 * where each 5x5 kernel acts as its sum over the y taps (the sum in float32
 * of the bfloat16 weights, rounded to bfloat16 again for layers 1 and 2).
 * No counterpart in the reference (the predictor is new). */
#define FVHOE_PRED_PLANAR_BYTES 4608
int fvhoe_pred_pack_planar(const float* theta, void* packed, void* stream);

/* Network on X (4, n0+12, n1+12): flags (n0*n1 int8, may be NULL) = logit >
 * logit_threshold, logits (n0*n1 float, may be NULL). */
int fvhoe_pred_infer(const float* X, int n0, int n1, const void* packed, float logit_threshold,
                     signed char* flags, float* logits, void* stream);

/* Features X of the input of the last fvhoe_stage / fvhoe_stage_rk /
 * fvhoe_stage_predicted call made with the same cfg and workspace (2-D,
 * a-posteriori with the global NAD range). */
int fvhoe_pred_features(const FvhoeStageConfig* cfg, const void* workspace, float* X, void* stream);

/* Training kernels (FP32).  out (B, Cout, H-4, W-4) = conv(in (B, Cin, H, W),
 * bf16(w) (Cout, Cin, 5, 5)) + bias, ReLU when relu != 0. */
int fvhoe_pred_conv_forward(const float* in, const float* w, const float* bias, float* out, int B, int Cin,
                            int Cout, int H, int W, int relu, void* stream);
/* dw, db = gradients of the convolution given dout; din (may be NULL) =
 * dL/d(in) masked by in > 0 (the ReLU of the layer that produced in). */
int fvhoe_pred_conv_backward(const float* in, const float* dout, const float* w, float* din, float* dw, float* db,
                             int B, int Cin, int Cout, int H, int W, void* stream);
/* Class-weighted binary cross-entropy on logits, mean over n: dlogits, and
 * *loss += the loss (device, caller zeroes it). */
int fvhoe_pred_loss(const float* logits, const signed char* labels, float pos_weight, long long n,
                    float* dlogits, float* loss, void* stream);
/* One Adam update of theta (n values), step counted from 1. */
int fvhoe_pred_adam(float* theta, const float* grad, float* m, float* v, int n, float lr, float beta1,
                    float beta2, float eps, int step, void* stream);

/* The stage with the predictor (2-D fused stage only).  With infer = 1 the
 * network runs on this stage's input and its flags are kept in the
 * workspace; with infer = 0 the flags of the last infer = 1 call are reused.
 * Pre-flagged cells get MUSCL fluxes at their faces before the candidate is
 * formed; the NAD/PAD/SED check then runs on every cell of that candidate
 * and the faces of the cells it flags get MUSCL fluxes as well.  trouble
 * receives pre-flags | check flags.  stats (device, may be NULL):
 * [0] += pre-flagged cells, [1] += cells flagged by the check only,
 * [2] += pre-flagged cells the check did not flag.
 * overlap = 1: a call with infer = 1 runs the network on side_stream,
 * forked after the stage's reductions (ev_fork) and concurrent with the
 * flux, candidate, check, fallback and final kernels of that call, which
 * therefore apply no pre-flags; stream waits for the network (ev_join)
 * before the call returns, so later calls (infer = 0) apply its flags.
 * Calls with infer = 0 ignore overlap.  Events come from fvhoe_event_create;
 * the side stream must have the default priority. */
typedef struct FvhoePredict {
  const void* packed;
  float logit_threshold;
  int infer;
  long long* stats;
  int overlap;
  void* side_stream;
  void* ev_fork;
  void* ev_join;
} FvhoePredict;

/* hipEvent_t without timing, for the overlap fork / join (host-only). */
int fvhoe_event_create(void** event);
int fvhoe_event_destroy(void* event);

int fvhoe_stage_predicted_workspace_bytes(const FvhoeStageConfig* cfg, unsigned long long* bytes);
int fvhoe_stage_predicted(const FvhoeStageConfig* cfg, const double* u, double* dudt, double* dt_stage,
                          signed char* trouble, void* workspace, const FvhoeRkFuse* rk, const FvhoePredict* pred,
                          void* stream);

/* ------------------------------------------------------------------------
 * Batched ensemble of independent 1-D runs of uneven size, advanced together
 * by SSP-RK3 (fvhoe/ode.py:271-283), each instance with its own CFL dt, time
 * and _hit_target (fvhoe/ode.py:11-22, 183-193).  No counterpart in the
 * reference, which runs one EulerSolver (fvhoe/solver.py:41-1001) at a time;
 * every instance follows that solver's stage (fvhoe/solver.py:429-454) bit
 * for bit in the FP64 form.  scheme: the shared settings (p[0], gamma, cfl,
 * bc[0..1], riemann, a-posteriori options, bulk_fp32; n[1] = n[2] = 1,
 * p[1] = p[2] = 0; n[0] and h are per instance).  tile: 64, 128 or 256 cells
 * per workgroup.  u: the instances' conservative states, instance-major, each
 * (5, n[i]) contiguous; t: per-instance times (device).  A "dirichlet" / "ic"
 * side (bc[k] = FVHOE_BC_DIRICHLET, fvhoe/boundary_conditions.py:18-57,334-348)
 * reads slab[k]: device values (count, 5, slab_g), layer j of the low side
 * being cell j - slab_g and of the high side cell n[i] + j of instance i;
 * slab_g at least fvhoe_stage_ghost_depth's depth for p[0].
 * ---------------------------------------------------------------------- */
int fvhoe_ensemble_workspace_bytes(const FvhoeStageConfig* scheme, int count, const int* n, int tile,
                                   unsigned long long* bytes);
/* Writes the instance tables into the workspace (synchronous copy; n and h
 * are host arrays). */
int fvhoe_ensemble_setup(const FvhoeStageConfig* scheme, int count, const int* n, const double* h, int tile,
                         void* workspace, void* stream);
/* One SSP-RK3 step of every instance whose time is below target; the others
 * are left untouched.  steps (device, may be NULL): per-instance step
 * counters, incremented for the instances that stepped.  Enqueues only
 * kernels, so it can be captured in a graph. */
int fvhoe_ensemble_step(const FvhoeStageConfig* scheme, int count, const int* n, int tile, double* u, double* t,
                        int* steps, double target, void* workspace, void* stream);
/* The same step with each RK stage as one persistent kernel (three launches
 * per step instead of 22): the grid is as many blocks as are resident at
 * once, the phases of the stage are separated by grid-wide barriers whose
 * waits are bounded; a wait that times out sets the error word that
 * fvhoe_ensemble_status reads (0: no error) instead of hanging.  Results
 * equal fvhoe_ensemble_step's bit for bit.  Captured in a graph like it. */
int fvhoe_ensemble_step_persistent(const FvhoeStageConfig* scheme, int count, const int* n, int tile, double* u,
                                   double* t, int* steps, double target, void* workspace, void* stream);
int fvhoe_ensemble_status(const FvhoeStageConfig* scheme, int count, const int* n, int tile, const void* workspace,
                          int* error_word);

/* fvhoe_ensemble_step with the troubled-cell predictor (no counterpart in
 * the reference): once per step, on the step's input, the planar network
 * (pred->packed from fvhoe_pred_pack_planar) pre-flags the cells whose logit
 * exceeds pred->logit_threshold, from features defined as the 2-D ones
 * ((rho, vx, vy, P) of the stage input's primitive averages over each run
 * extended by 6 cells, ghosts mapped into the interior, scaled by the run's
 * NAD range); in each of the step's three stages those cells get MUSCL
 * fluxes before the candidate is formed, and the plain check then runs on
 * every cell and sends the cells it flags (the pre-flagged ones excepted) to
 * the fallback as well.  pred->infer 0: no network, the plain step (bit for
 * bit) with the counters.  pred->stats (3 long long, may be NULL) += per
 * stage: pre-flagged cells, cells the check listed, pre-flagged cells the
 * check did not flag.  The chain of launches only (not the persistent
 * stage).  Captured in a graph like fvhoe_ensemble_step. */
int fvhoe_ensemble_step_predicted(const FvhoeStageConfig* scheme, int count, const int* n, int tile, double* u,
                                  double* t, int* steps, double target, void* workspace, const FvhoePredict* pred,
                                  void* stream);

/* The planar network's logits on the instances' states u (5 x n_i each,
 * instance-major): logits (sum n_i floats, instance-major).  t: the
 * instances' times (device, count doubles; read only). */
int fvhoe_ensemble_pred_logits(const FvhoeStageConfig* scheme, int count, const int* n, int tile, double* u,
                               double* t, void* workspace, const void* packed, float* logits, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FVHOE_HIP_H */
