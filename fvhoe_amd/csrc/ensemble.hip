// Batched ensemble of independent 1-D runs of uneven size (EulerSolver along
// x, fvhoe/solver.py:41-1001, one per instance), advanced together: every
// SSP-RK3 step of every active instance is one sequence of kernels per stage,
// each launched once for the whole ensemble over a table of (instance, tile)
// pairs, so that the work of one launch is balanced in tiles of TILE cells
// whatever the instance sizes.  Each instance keeps its own CFL dt, time and
// _hit_target (fvhoe/ode.py:11-22, 183-193); an instance whose time has
// reached the target is skipped by every kernel.
//
// Kernels of one stage of the chain (the same arithmetic, in the same order,
// as the fused 1-D stage, so the FP64 form reproduces the reference bit for
// bit):
//   prims+flux : u (BC-mapped on load) -> centres -> FP64 primitives -> fv
//                averages W on the ghost-extended range, with per-instance
//                CFL max and NAD range folded in by 64-bit atomics, and the
//                Riemann fluxes of the tile's faces from its W.
//   scalars    : per-instance dt (fvhoe/hydro.py:270-294), range, in the first
//                stage the step's dt and the active flag; clears the
//                reduction words for the next stage and this stage's counters.
//   cand+detect: the candidate u + dt L(F) -> WS in LDS and its NAD/PAD/SED
//                flags (fvhoe/slope_limiting.py:87-244), ballot-compacted per
//                instance.
//   fallback   : MUSCL fluxes at the faces of flagged cells (fvhoe/solver.py:710-731).
//   final      : divergence and the SSP-RK3 combination (fvhoe/ode.py:271-283);
//                in the last stage also the instance's clock.
// With the troubled-cell predictor the first stage also runs the network
// (pre-flags) and every stage a fallback over the pre-flagged cells before
// the candidate.  The persistent form runs a stage's phases in one kernel.
#include <vector>

#include "fvhoe_common.h"
#include "fvhoe_hip.h"
#include "predictor_infer.h"

using namespace fvhoe;

namespace {

__host__ __device__ constexpr int e_reach(int p) { return p == 1 ? 0 : (p % 2 ? reach(p) - 1 : reach(p)); }
__host__ __device__ constexpr int e_ghosts(int p) { return reach(p) + 1 > 2 * reach(p) ? reach(p) + 1 : 2 * reach(p); }
__host__ __device__ constexpr int e_gext(int p) { return e_ghosts(p) > 3 ? e_ghosts(p) : 3; }

constexpr int NACC = 11;  // [0] max wave speed (raw bits), [1 + 2v] / [2 + 2v] ordered min / max of variable v

// FP32 bulk form: fused multiply-adds in the sweeps, the Riemann solvers' approximate reciprocal / square
// root and one reciprocal of rho in the FP64 pressure recovery (FVHOE_ENS_F32_FAST = 0: correctly rounded
// operations throughout); FP64 keeps the reference's operations
// the chain's fused passes: W with its tile's faces, and the candidate with its cells' tests (1), or the
// separate W / F / WS / test passes of the persistent stage (0)
#ifndef FVHOE_ENS_FUSED
#define FVHOE_ENS_FUSED 1
#endif
#ifndef FVHOE_ENS_F32_FAST
#define FVHOE_ENS_F32_FAST 1
#endif
template <typename T>
constexpr bool ens_fast() { return FVHOE_ENS_F32_FAST && std::is_same<T, float>::value; }

template <int KIND, int P, typename T, typename Get>
__device__ __forceinline__ T esweep(Get&& x) {
  T acc = T(0);
  static_for<ntaps(P)>([&](auto I) {
    constexpr double w = weight(KIND, P, decltype(I)::value);
    if constexpr (w != 0.0) acc = q_fma<ens_fast<T>()>(T(w), x(decltype(I)::value), acc);
  });
  return acc;
}

__device__ __forceinline__ unsigned long long obits(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double from_obits(unsigned long long b) {
  const unsigned long long r = (b & 0x8000000000000000ull) ? (b & 0x7fffffffffffffffull) : ~b;
  return __longlong_as_double((long long)r);
}

// per-instance tables and scalars, all device memory
struct Ens {
  int count, bc_lo, bc_hi, riemann, limiter, apost;
  double gamma, cfl, csq_floor;
  const int* n;                 // cells
  const double* h;              // cell width
  const long long* off_u;       // offset of instance i in u-shaped arrays (x 5 for the variable planes)
  const long long* off_w;       // in W (n + 2 GE cells)
  const long long* off_f;       // in F (n + 1 faces)
  const long long* off_s;       // in WS (n + 6 cells)
  double* dt_stage;
  double* dt_step;
  int* active;
  unsigned long long* acc;      // NACC per instance
  double* range;                // 10 per instance
  int* cnt;                     // flagged cells per instance
  // detection
  int nad_mode, nad_range, nad_vars, pad_var, sed, force;
  double nad_eps, pad_lo, pad_hi, sed_eps;
  // ghost values of "dirichlet" / "ic" sides: (count, 5, slab_g) per side, layer k of the lo side is cell
  // k - slab_g, of the hi side cell n + k
  const double* slab_lo;
  const double* slab_hi;
  int slab_g;
  // learned stage: the step's pre-flags (per cell), and per instance the pre-flagged cells the check does
  // not flag (null without a predictor)
  const signed char* pre;
  int* fpc;
  int* pcnt;          // pre-flagged cells per instance (cleared in the step's first stage)
  long long* stats;   // counters of the learned stage (may be null)
};

// ghost value of a slab side of instance i, or nothing (returns false) for a cell that no slab covers
__device__ __forceinline__ bool slab_at(const Ens& e, int i, int n, int v, int c, double* x) {
  if (c < 0 && e.bc_lo == BC_DIRICHLET) {
    *x = e.slab_lo[((long long)i * 5 + v) * e.slab_g + (c + e.slab_g)];
    return true;
  }
  if (c >= n && e.bc_hi == BC_DIRICHLET) {
    *x = e.slab_hi[((long long)i * 5 + v) * e.slab_g + (c - n)];
    return true;
  }
  return false;
}

// u of instance i at (possibly ghost) cell c, BC applied (fvhoe/boundary_conditions.py:306-402); slab sides
// give their ghost values directly (fvhoe/boundary_conditions.py:334-348)
__device__ __forceinline__ double u_at(const Ens& e, int i, const double* __restrict__ u, int n, int v, int c,
                                       int lo, int hi) {
  double g;
  if (slab_at(e, i, n, v, c, &g)) return g;
  bool f;
  const int s = bc_map(c, n, n, lo, hi, &f);
  const double x = u[v * n + s];
  return (f && v == 1) ? -x : x;
}

// candidate u + dt L(F) at a (possibly ghost) cell (fvhoe/solver.py:678); reflective ghosts negate it whole;
// the candidate's ghosts on slab sides are the slab values (the boundary fill of the candidate state)
template <typename T>
__device__ __forceinline__ double ustar_at(const Ens& e, int i, const double* __restrict__ u, const T* __restrict__ F,
                                           int n, double cx, double dt, int v, int c, int lo, int hi) {
  double gv;
  if (slab_at(e, i, n, v, c, &gv)) return gv;
  bool f;
  const int s = bc_map(c, n, n, lo, hi, &f);
  const double L = 0.0 + cx * ((double)F[v * (n + 1) + s + 1] - (double)F[v * (n + 1) + s]);
  const double x = u[v * n + s] + dt * L;
  return (f && v == 1) ? -x : x;
}

__device__ __forceinline__ void d_ens_reset(const Ens& e, int i) {
  if (i >= e.count) return;
  e.acc[i * NACC] = 0ull;
  for (int k = 1; k < NACC; ++k) e.acc[i * NACC + k] = (k & 1) ? ~0ull : 0ull;
  if (e.fpc) e.fpc[i] = 0;
  e.cnt[i] = 0;
}
__global__ void k_ens_reset(Ens e) { d_ens_reset(e, blockIdx.x * blockDim.x + threadIdx.x); }

// Primitive fv averages of one tile: W (CAND = false, from u; ghost width GE) or WS (CAND = true, from the
// candidate; ghost width 3), outputs a in [tl.y, tl.y + TILE) of the instance's ghost-extended grid.
// XL / XR: outputs computed before / after the tile as well and kept in LDS (SO) for the caller.
template <typename T, int P, int TILE, bool CAND, int XL, int XR>
__device__ __forceinline__ void ens_prims_tile(const Ens& e, int i, int a0, const double* __restrict__ u,
                                               const T* __restrict__ F, T (&SO)[5][TILE + XL + XR]) {
  constexpr int RC = e_reach(P), RT = e_reach(P), GE = e_gext(P);
  constexpr int G0 = CAND ? 3 : GE;
  constexpr int NOUT = TILE + XL + XR;
  constexpr int HU = NOUT + 2 * (RC + RT), HC = NOUT + 2 * RT;
  __shared__ T SU[5][HU];
  __shared__ T SC[5][HC];
  const int n = e.n[i];
  const double cx = -(1.0 / e.h[i]);
  const double dt = CAND ? e.dt_stage[i] : 0.0;
  const double* ui = u + 5 * e.off_u[i];
  const T* Fi = CAND ? F + 5 * e.off_f[i] : nullptr;
  const int c0 = a0 - XL - G0;  // cell of the first output
  const int tid = threadIdx.x;
  const int span = G0 + RC + RT;  // cells beyond [-span, n + span) feed discarded outputs only
  for (int k = tid; k < HU; k += TILE) {
    const int c = max(min(c0 - RC - RT + k, n + span - 1), -span);
#pragma unroll
    for (int v = 0; v < 5; ++v)
      SU[v][k] = T(CAND ? ustar_at<T>(e, i, ui, Fi, n, cx, dt, v, c, e.bc_lo, e.bc_hi)
                        : u_at(e, i, ui, n, v, c, e.bc_lo, e.bc_hi));
  }
  __syncthreads();
  for (int k = tid; k < HC; k += TILE) {
    double uu[5], ww[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      // x sweep, then the degree-0 sweeps of the inactive y and z axes (fvhoe/fv.py:175-196)
      const T sx = esweep<KIND_C, P, T>([&](int t) { return SU[v][k + RC + t - reach(P)]; });
      const T sy = esweep<KIND_C, 0, T>([&](int) { return sx; });
      uu[v] = (double)(T(0) + sy);
    }
    prims_from_cons<ens_fast<T>()>(uu, ww, e.gamma);
#pragma unroll
    for (int v = 0; v < 5; ++v) SC[v][k] = T(ww[v]);
  }
  __syncthreads();
  for (int k = tid; k < NOUT; k += TILE) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const T tx = esweep<KIND_T, P, T>([&](int t) { return SC[v][k + RT + t - reach(P)]; });
      const T ty = esweep<KIND_T, 0, T>([&](int) { return tx; });
      SO[v][k] = T(0) + ty;
    }
  }
  __syncthreads();
}

// W (CAND = false) or WS (CAND = true) of one tile to global memory; W also feeds the per-instance CFL
// and NAD-range reductions, and with FLUX the tile's faces fa = a - GE (a in the tile, 0 <= fa <= n) are
// solved from the tile's W with EL + 1 / EL more outputs on each side (F written, as k_ens_flux does)
template <typename T, int P, int TILE, bool CAND, bool FLUX = false>
__device__ __forceinline__ void d_ens_prims(const Ens& e, int2 tl, const double* __restrict__ u,
                                            const T* __restrict__ F, T* __restrict__ out,
                                            T* __restrict__ Fout = nullptr) {
  constexpr int GE = e_gext(P), G = e_ghosts(P), EL = reach(P);
  constexpr int G0 = CAND ? 3 : GE;
  constexpr int XL = FLUX ? EL + 1 : 0, XR = FLUX ? EL : 0;
  __shared__ T SO[5][TILE + XL + XR];
  const int i = tl.x;
  if (!e.active[i]) return;
  ens_prims_tile<T, P, TILE, CAND, XL, XR>(e, i, tl.y, u, F, SO);
  const int n = e.n[i], NO = n + 2 * G0;
  const int tid = threadIdx.x;
  const int a = tl.y + tid;  // output index (cell a - G0)
  const bool valid = a < NO;
  T w[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    w[v] = SO[v][tid + XL];
    if (valid) out[5 * (CAND ? e.off_s[i] : e.off_w[i]) + v * NO + a] = w[v];
  }
  if (CAND) return;
  if constexpr (FLUX) {
    const int fa = a - GE;
    if (fa >= 0 && fa <= n) {
      T wl[5], wr[5], res[5];
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        // face fa: cells fa - 1 - EL .. fa + EL = tile outputs tid + XL - 1 - EL .. (fvhoe/solver.py:522-568)
        auto xc = [&](int k) {
          const T x = esweep<KIND_C, 0, T>([&](int) { return SO[q][k]; });
          return T(0) + x;
        };
        wl[q] = esweep<KIND_R, P, T>([&](int t) { return xc(tid + XL - 1 - EL + t); });
        wr[q] = esweep<KIND_L, P, T>([&](int t) { return xc(tid + XL - EL + t); });
      }
      riemann<0, ens_fast<T>()>(e.riemann, wl, wr, res, T(e.gamma), T(e.csq_floor));
      T* Fi = Fout + 5 * e.off_f[i];
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        const T x = esweep<KIND_T, 0, T>([&](int) { return res[q]; });
        Fi[q * (n + 1) + fa] = T(0) + x;
      }
    }
  }
  // per-instance CFL max over cells [-G, n+G) and NAD range over [-3, n+3)
  const int c = a - GE;
  unsigned long long val[NACC];
  double m = 0.0;
  val[0] = 0ull;
#pragma unroll
  for (int k = 1; k < NACC; ++k) val[k] = (k & 1) ? ~0ull : 0ull;
  if (valid && c >= -G && c < n + G) {
    const double cs = sound_speed((double)w[0], (double)w[4], e.gamma, e.csq_floor);
    const double s = (fabs((double)w[1]) + fabs((double)w[2])) + fabs((double)w[3]);
    m = np_max(m, s + 1.0 * cs);
  }
  val[0] = (unsigned long long)__double_as_longlong(m);
  if (valid && c >= -3 && c < n + 3) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      val[1 + 2 * v] = obits((double)w[v]);
      val[2 + 2 * v] = val[1 + 2 * v];
    }
  }
  __shared__ unsigned long long part[TILE / 64][NACC];
#pragma unroll
  for (int k = 0; k < NACC; ++k) {
    unsigned long long x = val[k];
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long y = __shfl_down(x, off, 64);
      x = (k == 0 || (k & 1) == 0) ? (x > y ? x : y) : (x < y ? x : y);
    }
    if ((tid & 63) == 0) part[tid >> 6][k] = x;
  }
  __syncthreads();
  if (tid < NACC) {
    unsigned long long x = part[0][tid];
    for (int wv = 1; wv < TILE / 64; ++wv) {
      const unsigned long long y = part[wv][tid];
      x = (tid == 0 || (tid & 1) == 0) ? (x > y ? x : y) : (x < y ? x : y);
    }
    if (tid == 0 || (tid & 1) == 0) atomicMax(&e.acc[i * NACC + tid], x);
    else atomicMin(&e.acc[i * NACC + tid], x);
  }
}
template <typename T, int P, int TILE, bool CAND, bool FLUX = false>
__global__ void __launch_bounds__(TILE) k_ens_prims(Ens e, const int2* __restrict__ tiles,
                                                   const double* __restrict__ u, const T* __restrict__ F,
                                                   T* __restrict__ out, T* __restrict__ Fout) {
  d_ens_prims<T, P, TILE, CAND, FLUX>(e, tiles[blockIdx.x], u, F, out, Fout);
}

// per-instance dt, range; first stage: active flag and the step's dt.  With reset (the chain of launches)
// the reduction words are cleared once read, for the next stage's W pass, and this stage's counters before
// its check (so no separate reset launch is needed; the set-up clears the words for the first stage)
__device__ __forceinline__ void d_ens_scalars(const Ens& e, int i, int first, const double* __restrict__ t,
                                              double target, int reset = 0) {
  if (i >= e.count) return;
  if (first) e.active[i] = t[i] < target ? 1 : 0;
  if (!e.active[i]) return;
  unsigned long long* a = e.acc + i * NACC;
  const double d = (e.cfl * e.h[i]) / __longlong_as_double((long long)a[0]);
  e.dt_stage[i] = d;
  if (first) e.dt_step[i] = (target - t[i] < d) ? target - t[i] : d;
  for (int k = 0; k < 10; ++k) e.range[i * 10 + k] = from_obits(a[1 + k]);
  if (reset) {
    a[0] = 0ull;
    for (int k = 1; k < NACC; ++k) a[k] = (k & 1) ? ~0ull : 0ull;
    e.cnt[i] = 0;
    if (e.fpc) e.fpc[i] = 0;
    if (first && e.pcnt) e.pcnt[i] = 0;
  }
}
__global__ void k_ens_scalars(Ens e, int first, const double* __restrict__ t, double target, int reset) {
  d_ens_scalars(e, blockIdx.x * blockDim.x + threadIdx.x, first, t, target, reset);
}

// faces fa in [0, n] of each instance: l/r values of W, Riemann solver
template <typename T, int P, int TILE>
__device__ __forceinline__ void d_ens_flux(const Ens& e, int2 tl, const T* __restrict__ W, T* __restrict__ F) {
  constexpr int GE = e_gext(P), EL = reach(P);
  const int i = tl.x;
  if (!e.active[i]) return;
  const int n = e.n[i], fa = tl.y + threadIdx.x;
  if (fa > n) return;
  const T* Wi = W + 5 * e.off_w[i];
  const int NW = n + 2 * GE;
  T wl[5], wr[5], res[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    // degree-0 sweeps of the inactive axes around the face values (fvhoe/solver.py:522-568)
    auto xc = [&](int c) {
      const T x = esweep<KIND_C, 0, T>([&](int) { return Wi[q * NW + c + GE]; });
      return T(0) + x;
    };
    wl[q] = esweep<KIND_R, P, T>([&](int t) { return xc(fa - 1 - EL + t); });
    wr[q] = esweep<KIND_L, P, T>([&](int t) { return xc(fa - EL + t); });
  }
  riemann<0, ens_fast<T>()>(e.riemann, wl, wr, res, T(e.gamma), T(e.csq_floor));
  T* Fi = F + 5 * e.off_f[i];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const T x = esweep<KIND_T, 0, T>([&](int) { return res[q]; });
    Fi[q * (n + 1) + fa] = T(0) + x;
  }
}
template <typename T, int P, int TILE>
__global__ void __launch_bounds__(TILE) k_ens_flux(Ens e, const int2* __restrict__ tiles, const T* __restrict__ W,
                                                  T* __restrict__ F) {
  d_ens_flux<T, P, TILE>(e, tiles[blockIdx.x], W, F);
}

// NAD / PAD / SED tests of interior cell c of instance i (fvhoe/slope_limiting.py:87-244); ws(v, o) is the
// candidate's primitive average of variable v at cell c + o (-3 <= o <= 3)
template <typename T, int P, typename WsAt>
__device__ __forceinline__ int ens_detect_cell(const Ens& e, int i, int n, int c, const T* __restrict__ W,
                                               WsAt&& ws) {
  constexpr int GE = e_gext(P);
  const T* Wi = W + 5 * e.off_w[i];
  const int NW = n + 2 * GE;
  const double* rg = e.range + 10 * i;
  bool tvv[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const T* wv = Wi + v * NW + c + GE;
    double M = (double)wv[0], m = M;
#pragma unroll
    for (int d = -1; d <= 1; ++d) {
      const double x = (double)wv[d];
      M = dmax(M, x);
      m = dmin(m, x);
    }
    const double uc = (double)ws(v, 0);
    bool tv;
    if (e.nad_range == 0) {
      const double rng = e.nad_mode == 0 ? rg[2 * v + 1] - rg[2 * v] : M - m;
      tv = dmin(uc - m, M - uc) < -e.nad_eps * rng;
    } else {
      const double lo = e.nad_mode == 0 ? (1.0 - e.nad_eps) * rg[2 * v] : (1.0 - e.nad_eps) * m;
      const double hi = e.nad_mode == 0 ? (1.0 + e.nad_eps) * rg[2 * v + 1] : (1.0 + e.nad_eps) * M;
      tv = dmin(uc - lo, hi - uc) < 0.0;
    }
    tvv[v] = tv && ((e.nad_vars >> v) & 1);
  }
  bool nad = false;
  for (int v = 0; v < 5; ++v) {
    bool tv = tvv[v];
    if (tv && e.sed && !nad)  // smooth-extrema detector only where NAD flags
      tv = sed_alpha_1d([&](int o) { return (double)ws(v, o); }, e.sed_eps) < 1.0;
    nad = nad || tv;
  }
  bool pad = false;
  if (e.pad_var >= 0) {
    const double x = (double)ws(e.pad_var, 0);
    pad = dmin(x - e.pad_lo, e.pad_hi - x) < 0.0;
  }
  return (e.force || pad || nad) ? 1 : 0;
}

// the flagged cells of the tile appended to the instance's list: wave ballot, block prefix, one atomic
template <int TILE>
__device__ __forceinline__ void ens_list_append(const Ens& e, int i, int t, int c, int* __restrict__ list) {
  __shared__ int wbase[TILE / 64];
  __shared__ int bbase;
  const unsigned long long ball = __ballot(t);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wbase[wave] = __popcll(ball);
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < TILE / 64; ++w) {
      const int x = wbase[w];
      wbase[w] = tot;
      tot += x;
    }
    bbase = tot ? atomicAdd(&e.cnt[i], tot) : 0;
  }
  __syncthreads();
  if (t) list[e.off_u[i] + bbase + wbase[wave] + __popcll(ball & ((1ull << lane) - 1ull))] = c;
}

// With a predictor, a pre-flagged cell already has MUSCL fluxes at its faces: the check lists only the
// others, and counts the pre-flagged cells it does not flag (one atomic per wave)
__device__ __forceinline__ int ens_pre_filter(const Ens& e, int i, int n, int c, int t) {
  if (!e.pre) return t;
  const int pf = c < n ? e.pre[e.off_u[i] + c] : 0;
  const unsigned long long b = __ballot(pf && !t);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(&e.fpc[i], (int)__popcll(b));
  return t && !pf;
}

template <typename T, int P, int TILE>
__device__ __forceinline__ void d_ens_detect(const Ens& e, int2 tl, const T* __restrict__ W, const T* __restrict__ WS,
                                             int* __restrict__ list) {
  const int i = tl.x;
  if (!e.active[i]) return;
  const int n = e.n[i], c = tl.y + threadIdx.x;
  int t = 0;
  if (c < n) {
    const T* Si = WS + 5 * e.off_s[i];
    const int NS = n + 6;
    t = ens_detect_cell<T, P>(e, i, n, c, W, [&](int v, int o) { return Si[v * NS + c + 3 + o]; });
  }
  t = ens_pre_filter(e, i, n, c, t);
  ens_list_append<TILE>(e, i, t, c, list);
}

// The candidate tile and the tests of its cells in one pass: WS of the tile's cells and 3 more on each
// side (the SED's reach) stays in LDS, so WS is never stored (the chain's form; the persistent stage keeps
// d_ens_prims<CAND> + d_ens_detect, with the same arithmetic)
template <typename T, int P, int TILE>
__device__ __forceinline__ void d_ens_cand_detect(const Ens& e, int2 tl, const double* __restrict__ u,
                                                  const T* __restrict__ F, const T* __restrict__ W,
                                                  int* __restrict__ list) {
  __shared__ T SO[5][TILE + 6];
  const int i = tl.x;
  if (!e.active[i]) return;
  // outputs: candidate grid (3 ghosts) indices tl.y .. tl.y + TILE + 5 = cells tl.y - 3 .. tl.y + TILE + 2
  ens_prims_tile<T, P, TILE, true, 0, 6>(e, i, tl.y, u, F, SO);
  const int n = e.n[i], c = tl.y + threadIdx.x;
  int t = 0;
  if (c < n) t = ens_detect_cell<T, P>(e, i, n, c, W, [&](int v, int o) { return SO[v][threadIdx.x + 3 + o]; });
  t = ens_pre_filter(e, i, n, c, t);
  ens_list_append<TILE>(e, i, t, c, list);
}
template <typename T, int P, int TILE>
__global__ void __launch_bounds__(TILE) k_ens_cand_detect(Ens e, const int2* __restrict__ tiles,
                                                         const double* __restrict__ u, const T* __restrict__ F,
                                                         const T* __restrict__ W, int* __restrict__ list) {
  d_ens_cand_detect<T, P, TILE>(e, tiles[blockIdx.x], u, F, W, list);
}

template <typename T, int P, int TILE>
__global__ void __launch_bounds__(TILE) k_ens_detect(Ens e, const int2* __restrict__ tiles, const T* __restrict__ W,
                                                    const T* __restrict__ WS, int* __restrict__ list) {
  d_ens_detect<T, P, TILE>(e, tiles[blockIdx.x], W, WS, list);
}

// MUSCL p=1 flux at face fa of instance i from the stage input (fvhoe/solver.py:712-717)
template <typename T>
__device__ void ens_fallback_flux(const Ens& e, int i, const double* __restrict__ ui, int n, int fa, T F[5]) {
  T w[4][5];
  auto cells = [&](auto in_c) {  // in_c: the four cells are interior, read directly (no boundary map)
    constexpr bool IN = decltype(in_c)::value;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double uu[5], wd[5];
#pragma unroll
      for (int q = 0; q < 5; ++q)
        uu[q] = 0.0 + (IN ? ui[q * n + fa - 2 + k] : u_at(e, i, ui, n, q, fa - 2 + k, e.bc_lo, e.bc_hi));
      prims_from_cons(uu, wd, e.gamma);
#pragma unroll
      for (int q = 0; q < 5; ++q) w[k][q] = T(0.0 + wd[q]);
    }
  };
  if (fa >= 2 && fa + 1 < n) cells(std::true_type{});
  else cells(std::false_type{});
  T wl[5], wr[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    if (e.limiter == 3) {  // slope_limiter None: p=1 conservative interpolation
      wl[q] = ((T(0) + T(-0.25) * w[0][q]) + T(1) * w[1][q]) + T(0.25) * w[2][q];
      wr[q] = ((T(0) + T(0.25) * w[1][q]) + T(1) * w[2][q]) + T(-0.25) * w[3][q];
    } else {
      const T s1 = limited_slope(e.limiter, w[1][q] - w[0][q], w[2][q] - w[1][q]);
      const T s2 = limited_slope(e.limiter, w[2][q] - w[1][q], w[3][q] - w[2][q]);
      wl[q] = w[1][q] + T(0.5) * s1;
      wr[q] = w[2][q] - T(0.5) * s2;
    }
  }
  riemann<0>(e.riemann, wl, wr, F, T(e.gamma), T(e.csq_floor));
}

template <typename T, int TILE>
__device__ __forceinline__ void d_ens_fallback(const Ens& e, int2 tl, const double* __restrict__ u,
                                               const int* __restrict__ list, T* __restrict__ F,
                                               const int* __restrict__ counts = nullptr) {
  const int i = tl.x;
  if (!e.active[i]) return;
  const int n = e.n[i], cnt = counts ? counts[i] : e.cnt[i];
  const double* ui = u + 5 * e.off_u[i];
  T* Fi = F + 5 * e.off_f[i];
  for (int k = threadIdx.x; k < 2 * TILE; k += TILE) {
    const int j = tl.y + k / 2;
    if (j >= cnt) continue;
    const int fa = list[e.off_u[i] + j] + (k & 1);
    T Fv[5];
    ens_fallback_flux<T>(e, i, ui, n, fa, Fv);
#pragma unroll
    for (int q = 0; q < 5; ++q) Fi[q * (n + 1) + fa] = Fv[q];
  }
}
template <typename T, int TILE>
__global__ void __launch_bounds__(TILE) k_ens_fallback(Ens e, const int2* __restrict__ tiles,
                                                      const double* __restrict__ u, const int* __restrict__ list,
                                                      T* __restrict__ F, const int* __restrict__ counts) {
  d_ens_fallback<T, TILE>(e, tiles[blockIdx.x], u, list, F, counts);
}

// The chain's fallback over instances instead of tiles: blocks (i, y) walk instance i's list with a stride
// of FVHOE_ENS_FB_BPI blocks, so a launch is one round of count x FVHOE_ENS_FB_BPI blocks whatever the
// number of tiles (the lists hold a few cells per run after the first steps, and a grid of one block per
// tile spent its time on blocks that only find their tile past the list).  The same faces and values as
// d_ens_fallback.
#ifndef FVHOE_ENS_FB_BPI
#define FVHOE_ENS_FB_BPI 4
#endif
template <typename T, int TILE>
__global__ void __launch_bounds__(TILE) k_ens_fallback_inst(Ens e, const double* __restrict__ u,
                                                           const int* __restrict__ list, T* __restrict__ F,
                                                           const int* __restrict__ counts) {
  const int i = blockIdx.x;
  if (!e.active[i]) return;
  const int n = e.n[i], cnt = counts ? counts[i] : e.cnt[i];
  const double* ui = u + 5 * e.off_u[i];
  T* Fi = F + 5 * e.off_f[i];
  const int* li = list + e.off_u[i];
  for (int k = (int)blockIdx.y * TILE + threadIdx.x; k < 2 * cnt; k += FVHOE_ENS_FB_BPI * TILE) {
    const int fa = li[k >> 1] + (k & 1);
    T Fv[5];
    ens_fallback_flux<T>(e, i, ui, n, fa, Fv);
#pragma unroll
    for (int q = 0; q < 5; ++q) Fi[q * (n + 1) + fa] = Fv[q];
  }
}

// divergence (fvhoe/solver.py:744-750) and the SSP-RK3 combination of this stage
// (fvhoe/ode.py:275-279): stage 0 -> k0, s = u + dt k0; stage 1 -> k1,
// s = u + 0.25 dt k0 + 0.25 dt k1; stage 2 -> u + (1/6) dt (k0 + k1 + 4 k2).
// With `clock` (the chain of launches) the instance's first tile also adds this stage's counters to
// e.stats and, in the last stage, moves the instance's clock and step count (d_ens_advance's operations).
template <typename T, int TILE>
__device__ __forceinline__ void d_ens_final(const Ens& e, int2 tl, const T* __restrict__ F, int stage,
                                            double* __restrict__ u, double* __restrict__ s, double* __restrict__ k0,
                                            double* __restrict__ k1, double* __restrict__ t = nullptr,
                                            int* __restrict__ nsteps = nullptr, bool clock = false) {
  const int i = tl.x;
  if (!e.active[i]) return;
  if (clock && tl.y == 0 && threadIdx.x == 0) {
    if (e.stats && e.dt_step[i] != 0.0) {
      if (e.pcnt) atomicAdd((unsigned long long*)&e.stats[0], (unsigned long long)e.pcnt[i]);
      atomicAdd((unsigned long long*)&e.stats[1], (unsigned long long)e.cnt[i]);
      if (e.fpc) atomicAdd((unsigned long long*)&e.stats[2], (unsigned long long)e.fpc[i]);
    }
    if (stage == 2) {
      t[i] = t[i] + e.dt_step[i];
      if (nsteps) nsteps[i] += 1;
    }
  }
  const int n = e.n[i], c = tl.y + threadIdx.x;
  if (c >= n) return;
  const double cx = -(1.0 / e.h[i]);
  const double dt = e.dt_step[i];
  const T* Fi = F + 5 * e.off_f[i];
  const long long b = 5 * e.off_u[i];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const double L = 0.0 + cx * ((double)Fi[q * (n + 1) + c + 1] - (double)Fi[q * (n + 1) + c]);
    const long long o = b + q * n + c;
    if (stage == 0) {
      k0[o] = L;
      s[o] = u[o] + dt * L;
    } else if (stage == 1) {
      k1[o] = L;
      s[o] = u[o] + 0.25 * dt * k0[o] + 0.25 * dt * L;
    } else {
      u[o] = u[o] + (1.0 / 6.0) * dt * (k0[o] + k1[o] + 4.0 * L);
    }
  }
}
template <typename T, int TILE>
__global__ void __launch_bounds__(TILE) k_ens_final(Ens e, const int2* __restrict__ tiles, const T* __restrict__ F,
                                                   int stage, double* __restrict__ u, double* __restrict__ s,
                                                   double* __restrict__ k0, double* __restrict__ k1,
                                                   double* __restrict__ t, int* __restrict__ nsteps) {
  d_ens_final<T, TILE>(e, tiles[blockIdx.x], F, stage, u, s, k0, k1, t, nsteps, true);
}

__device__ __forceinline__ void d_ens_advance(const Ens& e, int i, double* __restrict__ t, int* __restrict__ steps) {
  if (i < e.count && e.active[i]) {
    t[i] = t[i] + e.dt_step[i];
    if (steps) steps[i] += 1;
  }
}

// ---------------------------------------------------------------------------
// Troubled-cell predictor on the 1-D runs: the 2-D network applied to each run seen as a 2-D state that
// does not vary along y.  On such an input every valid 5x5 cross-correlation acts as a 1-D one along x
// with the kernel summed over its y taps, so the network reduces to three 1-D convolutions of 5 taps
// (4 -> 16 -> 16 -> 1 channels) with the summed kernels (fvhoe_pred_pack_planar).  One block per tile of
// TILE cells of an instance: the features of the TILE + 12 cells around it (the 2-D features' definition:
// (rho, vx, vy, P) of W, ghosts mapped into the interior, scaled by the instance's NAD range) are staged
// in LDS as bfloat16; layers 1 and 2 run on v_mfma_f32_16x16x32_bf16 with FP32 accumulation (A = the 16
// output channels' weights x 32 of K, B = 16 consecutive pixels x 32 of K, as the 2-D kernel), layer 3 as
// 80 FP32 products per cell.  Flags (logit > threshold) go to the per-cell pre-flag array and to the
// instance's pre-flag list.
// ---------------------------------------------------------------------------
template <typename T, int TILE>
__device__ __forceinline__ void d_ens_pred(const Ens& e, int2 tl, const T* __restrict__ W, int GE,
                                           const unsigned char* __restrict__ pw, float thr,
                                           signed char* __restrict__ pre, int* __restrict__ plist,
                                           int* __restrict__ pcnt, float* __restrict__ logits) {
  using namespace fvhoe::pred;
  constexpr int NG2 = (TILE + 4 + 15) / 16, NG1 = NG2 + 1;  // 16-pixel groups of layers 2 and 1
  constexpr int NX = 16 * NG1 + 4;                         // feature pixels staged (cells c0 - 6 + p)
  constexpr int NWV = TILE / 64;
  __shared__ __attribute__((aligned(16))) unsigned short xin[NX * 4];
  __shared__ __attribute__((aligned(16))) unsigned short a1[16 * NG1 * 16];
  __shared__ __attribute__((aligned(16))) unsigned short a2[16 * NG2 * 16];
  __shared__ float w3[C2 * KS + 1];
  __shared__ double sc[8];
  const int i = tl.x;
  if (!e.active[i]) return;
  const int n = e.n[i], c0 = tl.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, mrow = lane & 15, kb = lane >> 4;
  if (tid < 4) {  // per-variable (min, 1 / (max - min)) of the instance's NAD range, 0 where max = min
    const int v = tid == 3 ? 4 : tid;
    const double lo = e.range[10 * i + 2 * v], rng = e.range[10 * i + 2 * v + 1] - lo;
    sc[2 * tid] = lo;
    sc[2 * tid + 1] = rng > 0.0 ? 1.0 / rng : 0.0;
  }
  for (int k = tid; k < C2 * KS + 1; k += TILE) w3[k] = k < C2 * KS ? ((const float*)(pw + QW3))[k] : *(const float*)(pw + QB3);
  __syncthreads();
  const T* Wi = W + 5 * e.off_w[i];
  const int NW = n + 2 * GE;
  for (int p = tid; p < NX; p += TILE) {
    const int c = c0 - HALO + p;
    unsigned short f[4] = {0, 0, 0, 0};
    if (c < n + HALO) {  // beyond the extended run (the last tile's overhang): zero
      bool flip;
      const int src = bc_map(c, n, n, e.bc_lo == BC_DIRICHLET ? BC_OUTFLOW : e.bc_lo,
                             e.bc_hi == BC_DIRICHLET ? BC_OUTFLOW : e.bc_hi, &flip);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int v = k == 3 ? 4 : k;
        double x = (double)Wi[v * NW + src + GE];
        if (flip && v == 1) x = -x;
        const double inv = sc[2 * k + 1];
        f[k] = f2bf(inv != 0.0 ? (float)((x - sc[2 * k]) * inv) : 0.0f);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) xin[p * 4 + k] = f[k];
  }
  __syncthreads();
  // layer 1: group g = a1 pixels 16g .. 16g + 15 (cells c0 - 4 + p); lane (mrow, kb): taps 2kb, 2kb + 1
  const bf16x8 wa1 = *(const bf16x8*)(pw + QW1 + 2 * (mrow * 32 + kb * 8));
  float bias1[4], bias2[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    bias1[r] = ((const float*)(pw + QB1))[4 * kb + r];
    bias2[r] = ((const float*)(pw + QB2))[4 * kb + r];
  }
  for (int g = wave; g < NG1; g += NWV) {
    const unsigned short* px = &xin[(16 * g + mrow + 2 * kb) * 4];
    u16x4 lo = {0, 0, 0, 0}, hi = {0, 0, 0, 0};
    if (kb < 3) lo = *(const u16x4*)px;        // tap 2kb (tap 4 for kb = 2)
    if (kb < 2) hi = *(const u16x4*)(px + 4);  // tap 2kb + 1
    const u16x8 b = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa1, __builtin_bit_cast(bf16x8, b), acc, 0, 0, 0);
    const bf16x4v v = {(__bf16)fmaxf(acc[0] + bias1[0], 0.f), (__bf16)fmaxf(acc[1] + bias1[1], 0.f),
                       (__bf16)fmaxf(acc[2] + bias1[2], 0.f), (__bf16)fmaxf(acc[3] + bias1[3], 0.f)};
    *(bf16x4v*)&a1[(16 * g + mrow) * 16 + 4 * kb] = v;
  }
  __syncthreads();
  // layer 2: K step k = taps 2k + kb/2, channels 8 (kb % 2) .. + 7 (tap 5 zero)
  bf16x8 wa2[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) wa2[k] = *(const bf16x8*)(pw + QW2 + 2 * ((k * 16 + mrow) * 32 + kb * 8));
  const u16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int g = wave; g < NG2; g += NWV) {
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int t = 2 * k + (kb >> 1);
      u16x8 b = zero8;
      if (t < KS) b = *(const u16x8*)&a1[(16 * g + mrow + t) * 16 + 8 * (kb & 1)];
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa2[k], __builtin_bit_cast(bf16x8, b), acc, 0, 0, 0);
    }
    const bf16x4v v = {(__bf16)fmaxf(acc[0] + bias2[0], 0.f), (__bf16)fmaxf(acc[1] + bias2[1], 0.f),
                       (__bf16)fmaxf(acc[2] + bias2[2], 0.f), (__bf16)fmaxf(acc[3] + bias2[3], 0.f)};
    *(bf16x4v*)&a2[(16 * g + mrow) * 16 + 4 * kb] = v;
  }
  __syncthreads();
  // layer 3: cell c0 + tid from a2 pixels tid .. tid + 4
  float z = w3[C2 * KS];
#pragma unroll
  for (int t = 0; t < KS; ++t) {
    const u16x8 lo8 = *(const u16x8*)&a2[(tid + t) * 16];
    const u16x8 hi8 = *(const u16x8*)&a2[(tid + t) * 16 + 8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      z = z + w3[c * KS + t] * bf2f(lo8[c]);
      z = z + w3[(c + 8) * KS + t] * bf2f(hi8[c]);
    }
  }
  const int c = c0 + tid;
  const int hit = (c < n && z > thr) ? 1 : 0;
  if (c < n) {
    pre[e.off_u[i] + c] = (signed char)hit;
    if (logits) logits[e.off_u[i] + c] = z;
  }
  // the flagged cells appended to the instance's pre-flag list: wave ballot, block prefix, one atomic
  __shared__ int wbase[NWV];
  __shared__ int bbase;
  const unsigned long long ball = __ballot(hit);
  if (lane == 0) wbase[wave] = __popcll(ball);
  __syncthreads();
  if (tid == 0) {
    int tot = 0;
    for (int w = 0; w < NWV; ++w) {
      const int x = wbase[w];
      wbase[w] = tot;
      tot += x;
    }
    bbase = tot ? atomicAdd(&pcnt[i], tot) : 0;
  }
  __syncthreads();
  if (hit) plist[e.off_u[i] + bbase + wbase[wave] + __popcll(ball & ((1ull << lane) - 1ull))] = c;
}
template <typename T, int TILE>
__global__ void __launch_bounds__(TILE) k_ens_pred(Ens e, const int2* __restrict__ tiles, const T* __restrict__ W,
                                                  int GE, const unsigned char* __restrict__ pw, float thr,
                                                  signed char* __restrict__ pre, int* __restrict__ plist,
                                                  int* __restrict__ pcnt, float* __restrict__ logits) {
  d_ens_pred<T, TILE>(e, tiles[blockIdx.x], W, GE, pw, thr, pre, plist, pcnt, logits);
}

// ---------------------------------------------------------------------------
// The whole RK stage in one persistent kernel: the blocks (all resident at once: the host sizes the grid
// from the occupancy of this kernel) walk each phase's tiles in turn, with a grid-wide barrier between
// phases where one phase needs another's results (per-instance reductions before dt, the flag list
// before the fallback, ...).  The barrier waits a bounded time: a block that waits too long sets the
// error word and goes on, so the kernel always ends (the host reports the error).
// ---------------------------------------------------------------------------
struct GridBar {
  unsigned* count;  // blocks arrived at the current barrier
  unsigned* gen;    // barrier generation
  int* err;         // set when a wait timed out
  unsigned nblocks;
};

constexpr long long BAR_SPIN_LIMIT = 1ll << 24;  // ~1 s of polling before the error word is set

__device__ __forceinline__ void grid_sync(const GridBar& b) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();  // this block's writes are visible to every XCD before it arrives
    const unsigned g = __hip_atomic_load(b.gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned arrived = __hip_atomic_fetch_add(b.count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == b.nblocks - 1) {
      __hip_atomic_store(b.count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(b.gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      long long spins = 0;
      while (__hip_atomic_load(b.gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > BAR_SPIN_LIMIT) {
          __hip_atomic_fetch_or(b.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
    __threadfence();  // and sees every other block's writes after it
  }
  __syncthreads();
}

struct StageTiles {
  const int2 *tw, *tf, *ts, *tc;
  int nw, nf, ns, nc;
};

template <typename T, int P, int TILE>
__global__ void __launch_bounds__(TILE) k_ens_stage(Ens e, StageTiles tl, GridBar bar, int st, const double* in,
                                                   double* u, double* sv, double* k0, double* k1, T* W, T* F, T* WS,
                                                   int* list, double* t, int* nsteps, double target) {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  for (int i = b * TILE + (int)threadIdx.x; i < e.count; i += nb * TILE) d_ens_reset(e, i);
  grid_sync(bar);
  for (int k = b; k < tl.nw; k += nb) d_ens_prims<T, P, TILE, false>(e, tl.tw[k], in, nullptr, W);
  grid_sync(bar);
  for (int i = b * TILE + (int)threadIdx.x; i < e.count; i += nb * TILE) d_ens_scalars(e, i, st == 0, t, target);
  grid_sync(bar);
  for (int k = b; k < tl.nf; k += nb) d_ens_flux<T, P, TILE>(e, tl.tf[k], W, F);
  if (e.apost) {
    grid_sync(bar);
    for (int k = b; k < tl.ns; k += nb) d_ens_prims<T, P, TILE, true>(e, tl.ts[k], in, F, WS);
    grid_sync(bar);
    for (int k = b; k < tl.nc; k += nb) {
      __syncthreads();  // the block's compaction words of the previous tile have been read
      d_ens_detect<T, P, TILE>(e, tl.tc[k], W, WS, list);
    }
    grid_sync(bar);
    for (int k = b; k < tl.nc; k += nb) d_ens_fallback<T, TILE>(e, tl.tc[k], in, list, F);
  }
  grid_sync(bar);
  for (int k = b; k < tl.nc; k += nb) d_ens_final<T, TILE>(e, tl.tc[k], F, st, u, sv, k0, k1);
  if (st == 2)  // the clock moves once per step (no instance's final reads t)
    for (int i = b * TILE + (int)threadIdx.x; i < e.count; i += nb * TILE) d_ens_advance(e, i, t, nsteps);
}

// ------------------------------------------------------------------ host side

size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Plan {
  int count, p, tile, fp32, GE;
  long long cells, wcells, fcells, scells;
  int tw, tf, ts, tc;  // tiles of each kind
  size_t o_n, o_h, o_ou, o_ow, o_of, o_os, o_dts, o_dtp, o_act, o_acc, o_rng, o_cnt, o_tw, o_tf, o_ts, o_tc;
  size_t o_bar, o_W, o_F, o_WS, o_list, o_k0, o_k1, o_s, o_pre, o_plist, o_pcnt, o_fpc, total;
};

int make_plan(const FvhoeStageConfig* c, int count, const int* n, int tile, Plan* P) {
  if (!c || !n || count < 1) return FVHOE_ERR_ARG;
  if (tile != 64 && tile != 128 && tile != 256) return FVHOE_ERR_ARG;
  if (c->p[0] < 1 || c->p[0] > 8 || c->n[1] != 1 || c->p[1] != 0 || c->n[2] != 1 || c->p[2] != 0)
    return FVHOE_ERR_UNSUPPORTED;
  if (c->fixed_dt_on || c->bulk_fp32 < 0 || c->bulk_fp32 > 1) return FVHOE_ERR_UNSUPPORTED;
  if (c->riemann < 0 || c->riemann > 2 || c->limiter < 0 || c->limiter > 3) return FVHOE_ERR_ARG;
  for (int k = 0; k < 2; ++k) {
    if (c->bc[k] < 0 || c->bc[k] > BC_DIRICHLET) return FVHOE_ERR_ARG;
    // slab sides: (count, 5, slab_g) ghost values, deep enough for every ghost the stage reads
    if (c->bc[k] == BC_DIRICHLET &&
        (!c->slab[k] || c->slab_g < e_gext(c->p[0]) + 2 * e_reach(c->p[0])))
      return FVHOE_ERR_ARG;
  }
  P->count = count;
  P->p = c->p[0];
  P->tile = tile;
  P->fp32 = c->bulk_fp32;
  P->GE = e_gext(P->p);
  const int pad = e_ghosts(P->p) + 2 * reach(P->p);
  P->cells = P->wcells = P->fcells = P->scells = 0;
  P->tw = P->tf = P->ts = P->tc = 0;
  for (int i = 0; i < count; ++i) {
    if (n[i] < pad || n[i] < 7) return FVHOE_ERR_UNSUPPORTED;
    P->cells += n[i];
    P->wcells += n[i] + 2 * P->GE;
    P->fcells += n[i] + 1;
    P->scells += n[i] + 6;
    P->tw += (n[i] + 2 * P->GE + tile - 1) / tile;
    P->tf += (n[i] + 1 + tile - 1) / tile;
    P->ts += (n[i] + 6 + tile - 1) / tile;
    P->tc += (n[i] + tile - 1) / tile;
  }
  if (5 * (P->wcells + P->scells) > 2000000000LL) return FVHOE_ERR_UNSUPPORTED;
  const size_t esz = P->fp32 ? 4 : 8;
  size_t o = 0;
  P->o_n = o; o = a256(o + 4 * (size_t)count);
  P->o_h = o; o = a256(o + 8 * (size_t)count);
  P->o_ou = o; o = a256(o + 8 * (size_t)count);
  P->o_ow = o; o = a256(o + 8 * (size_t)count);
  P->o_of = o; o = a256(o + 8 * (size_t)count);
  P->o_os = o; o = a256(o + 8 * (size_t)count);
  P->o_dts = o; o = a256(o + 8 * (size_t)count);
  P->o_dtp = o; o = a256(o + 8 * (size_t)count);
  P->o_act = o; o = a256(o + 4 * (size_t)count);
  P->o_acc = o; o = a256(o + 8 * NACC * (size_t)count);
  P->o_rng = o; o = a256(o + 80 * (size_t)count);
  P->o_cnt = o; o = a256(o + 4 * (size_t)count);
  P->o_tw = o; o = a256(o + 8 * (size_t)P->tw);
  P->o_tf = o; o = a256(o + 8 * (size_t)P->tf);
  P->o_ts = o; o = a256(o + 8 * (size_t)P->ts);
  P->o_tc = o; o = a256(o + 8 * (size_t)P->tc);
  P->o_bar = o; o = a256(o + 16);  // grid-barrier words of the persistent stage (zeroed by the setup copy)
  P->o_W = o; o = a256(o + esz * 5 * (size_t)P->wcells);
  P->o_F = o; o = a256(o + esz * 5 * (size_t)P->fcells);
  P->o_WS = o; o = a256(o + esz * 5 * (size_t)P->scells);
  P->o_list = o; o = a256(o + 4 * (size_t)P->cells);
  P->o_k0 = o; o = a256(o + 8 * 5 * (size_t)P->cells);
  P->o_k1 = o; o = a256(o + 8 * 5 * (size_t)P->cells);
  P->o_s = o; o = a256(o + 8 * 5 * (size_t)P->cells);
  // the learned stage: pre-flags per cell, the pre-flag list, per-instance counts
  P->o_pre = o; o = a256(o + (size_t)P->cells);
  P->o_plist = o; o = a256(o + 4 * (size_t)P->cells);
  P->o_pcnt = o; o = a256(o + 4 * (size_t)count);
  P->o_fpc = o; o = a256(o + 4 * (size_t)count);
  P->total = o;
  return FVHOE_OK;
}

Ens make_ens(const FvhoeStageConfig* c, const Plan& P, char* ws) {
  Ens e{};
  e.count = P.count;
  e.bc_lo = c->bc[0];
  e.bc_hi = c->bc[1];
  e.riemann = c->riemann;
  e.limiter = c->limiter;
  e.apost = c->a_posteriori;
  e.gamma = c->gamma;
  e.cfl = c->cfl;
  e.csq_floor = c->csq_floor;
  e.n = (const int*)(ws + P.o_n);
  e.h = (const double*)(ws + P.o_h);
  e.off_u = (const long long*)(ws + P.o_ou);
  e.off_w = (const long long*)(ws + P.o_ow);
  e.off_f = (const long long*)(ws + P.o_of);
  e.off_s = (const long long*)(ws + P.o_os);
  e.dt_stage = (double*)(ws + P.o_dts);
  e.dt_step = (double*)(ws + P.o_dtp);
  e.active = (int*)(ws + P.o_act);
  e.acc = (unsigned long long*)(ws + P.o_acc);
  e.range = (double*)(ws + P.o_rng);
  e.cnt = (int*)(ws + P.o_cnt);
  e.nad_mode = c->nad_mode;
  e.nad_range = c->nad_range;
  e.nad_vars = c->nad_vars;
  e.pad_var = c->pad_var;
  e.sed = c->sed;
  e.force = c->force_trouble;
  e.nad_eps = c->nad_eps;
  e.pad_lo = c->pad_lo;
  e.pad_hi = c->pad_hi;
  e.sed_eps = c->sed_eps;
  e.slab_lo = c->bc[0] == BC_DIRICHLET ? c->slab[0] : nullptr;
  e.slab_hi = c->bc[1] == BC_DIRICHLET ? c->slab[1] : nullptr;
  e.slab_g = c->slab_g;
  return e;
}

// blocks of the persistent stage that are resident at once on this device (all of them must be)
template <typename T, int P, int TILE>
int persistent_blocks() {
  static int cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  if (cached[dev] == 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ens_stage<T, P, TILE>, TILE, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 0;
    cached[dev] = per_cu * cus;
  }
  return cached[dev];
}

template <typename T, int P, int TILE>
int run_step(const FvhoeStageConfig* c, const Plan& PL, double* u, double* t, int* nsteps, double target, char* ws,
             hipStream_t s, bool persistent, const FvhoePredict* pred, float* logits_only) {
  Ens e = make_ens(c, PL, ws);
  signed char* pre = (signed char*)(ws + PL.o_pre);
  int* plist = (int*)(ws + PL.o_plist);
  int* pcnt = (int*)(ws + PL.o_pcnt);
  const bool learned = pred && pred->infer;
  if (pred) {
    e.fpc = (int*)(ws + PL.o_fpc);
    e.stats = pred->stats;
  }
  if (learned) {
    e.pre = pre;
    e.pcnt = pcnt;
  }
  if (logits_only) {  // the network on the stage input u alone (its W, range and features): logits per cell
    const int2* tw = (const int2*)(ws + PL.o_tw);
    const int2* tc = (const int2*)(ws + PL.o_tc);
    T* W = (T*)(ws + PL.o_W);
    const int gi = (PL.count + 255) / 256;
    k_ens_reset<<<gi, 256, 0, s>>>(e);
    k_ens_prims<T, P, TILE, false><<<PL.tw, TILE, 0, s>>>(e, tw, u, nullptr, W, nullptr);
    k_ens_scalars<<<gi, 256, 0, s>>>(e, 1, t, target, 1);
    k_ens_pred<T, TILE><<<PL.tc, TILE, 0, s>>>(e, tc, W, PL.GE, (const unsigned char*)pred->packed,
                                               pred->logit_threshold, pre, plist, pcnt, logits_only);
    return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
  }
  const int2* tw = (const int2*)(ws + PL.o_tw);
  const int2* tf = (const int2*)(ws + PL.o_tf);
  const int2* ts = (const int2*)(ws + PL.o_ts);
  const int2* tc = (const int2*)(ws + PL.o_tc);
  T* W = (T*)(ws + PL.o_W);
  T* F = (T*)(ws + PL.o_F);
  T* WS = (T*)(ws + PL.o_WS);
  int* list = (int*)(ws + PL.o_list);
  double* k0 = (double*)(ws + PL.o_k0);
  double* k1 = (double*)(ws + PL.o_k1);
  double* sv = (double*)(ws + PL.o_s);
  const int gi = (PL.count + 255) / 256;
  if (persistent) {
    if (pred) return FVHOE_ERR_UNSUPPORTED;  // the learned stage runs as the chain of launches
    int nb = persistent_blocks<T, P, TILE>();
    if (nb < 1) return FVHOE_ERR_LAUNCH;
    const int most = PL.tw > PL.ts ? PL.tw : PL.ts;
    if (nb > most) nb = most;
    unsigned* bw = (unsigned*)(ws + PL.o_bar);
    GridBar bar{bw, bw + 1, (int*)(bw + 2), (unsigned)nb};
    StageTiles tls{tw, tf, ts, tc, PL.tw, PL.tf, PL.ts, PL.tc};
    for (int st = 0; st < 3; ++st) {
      const double* in = st == 0 ? u : sv;
      k_ens_stage<T, P, TILE><<<nb, TILE, 0, s>>>(e, tls, bar, st, in, u, sv, k0, k1, W, F, WS, list, t, nsteps,
                                                  target);
    }
    return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
  }
  for (int st = 0; st < 3; ++st) {
    const double* in = st == 0 ? u : sv;
    // (the reduction words were cleared by the set-up or by the previous stage's scalars)
    // W with the faces of each W tile solved in the same pass (FVHOE_ENS_FUSED), or W then F
    if (FVHOE_ENS_FUSED) {
      k_ens_prims<T, P, TILE, false, true><<<PL.tw, TILE, 0, s>>>(e, tw, in, nullptr, W, F);
    } else {
      k_ens_prims<T, P, TILE, false><<<PL.tw, TILE, 0, s>>>(e, tw, in, nullptr, W, nullptr);
    }
    k_ens_scalars<<<gi, 256, 0, s>>>(e, st == 0, t, target, 1);
    if (!FVHOE_ENS_FUSED) k_ens_flux<T, P, TILE><<<PL.tf, TILE, 0, s>>>(e, tf, W, F);
    if (learned && c->a_posteriori) {
      if (st == 0)  // one inference per step, on the step's input; its flags serve the three stages
        k_ens_pred<T, TILE><<<PL.tc, TILE, 0, s>>>(e, tc, W, PL.GE, (const unsigned char*)pred->packed,
                                                   pred->logit_threshold, pre, plist, pcnt, nullptr);
      // MUSCL fluxes at the faces of the pre-flagged cells, before the candidate is formed
      k_ens_fallback_inst<T, TILE><<<dim3(PL.count, FVHOE_ENS_FB_BPI), TILE, 0, s>>>(e, in, plist, F, pcnt);
    }
    if (c->a_posteriori) {
      if (FVHOE_ENS_FUSED) {  // the candidate and the tests of each interior tile in one pass
        k_ens_cand_detect<T, P, TILE><<<PL.tc, TILE, 0, s>>>(e, tc, in, F, W, list);
      } else {
        k_ens_prims<T, P, TILE, true><<<PL.ts, TILE, 0, s>>>(e, ts, in, F, WS, nullptr);
        k_ens_detect<T, P, TILE><<<PL.tc, TILE, 0, s>>>(e, tc, W, WS, list);
      }
      k_ens_fallback_inst<T, TILE><<<dim3(PL.count, FVHOE_ENS_FB_BPI), TILE, 0, s>>>(e, in, list, F, nullptr);
    }
    // the divergence and the RK combination; the counters and, in the last stage, the clock as well
    k_ens_final<T, TILE><<<PL.tc, TILE, 0, s>>>(e, tc, F, st, u, sv, k0, k1, t, nsteps);
  }
  return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
}

template <typename T, int TILE>
int by_p(int p, const FvhoeStageConfig* c, const Plan& PL, double* u, double* t, int* nsteps, double target, char* ws,
         hipStream_t s, bool pers, const FvhoePredict* pr, float* lo) {
  switch (p) {
    case 1: return run_step<T, 1, TILE>(c, PL, u, t, nsteps, target, ws, s, pers, pr, lo);
    case 2: return run_step<T, 2, TILE>(c, PL, u, t, nsteps, target, ws, s, pers, pr, lo);
    case 3: return run_step<T, 3, TILE>(c, PL, u, t, nsteps, target, ws, s, pers, pr, lo);
    case 4: return run_step<T, 4, TILE>(c, PL, u, t, nsteps, target, ws, s, pers, pr, lo);
    case 5: return run_step<T, 5, TILE>(c, PL, u, t, nsteps, target, ws, s, pers, pr, lo);
    case 6: return run_step<T, 6, TILE>(c, PL, u, t, nsteps, target, ws, s, pers, pr, lo);
    case 7: return run_step<T, 7, TILE>(c, PL, u, t, nsteps, target, ws, s, pers, pr, lo);
    case 8: return run_step<T, 8, TILE>(c, PL, u, t, nsteps, target, ws, s, pers, pr, lo);
  }
  return FVHOE_ERR_UNSUPPORTED;
}

template <typename T>
int by_tile(const FvhoeStageConfig* c, const Plan& PL, double* u, double* t, int* nsteps, double target, char* ws,
            hipStream_t s, bool pers, const FvhoePredict* pr, float* lo) {
  switch (PL.tile) {
    case 64: return by_p<T, 64>(PL.p, c, PL, u, t, nsteps, target, ws, s, pers, pr, lo);
    case 128: return by_p<T, 128>(PL.p, c, PL, u, t, nsteps, target, ws, s, pers, pr, lo);
    case 256: return by_p<T, 256>(PL.p, c, PL, u, t, nsteps, target, ws, s, pers, pr, lo);
  }
  return FVHOE_ERR_ARG;
}

}  // namespace

namespace {
int ensemble_step(const FvhoeStageConfig* cfg, int count, const int* n, int tile, double* u, double* t, int* steps,
                  double target, void* workspace, void* stream, bool persistent, const FvhoePredict* pred = nullptr,
                  float* logits_only = nullptr) {
  Plan P;
  const int st = make_plan(cfg, count, n, tile, &P);
  if (st != FVHOE_OK) return st;
  if (!u || !t || !workspace || ((uintptr_t)workspace & 255) != 0) return FVHOE_ERR_ARG;
  if (pred && pred->infer && (!pred->packed || !cfg->a_posteriori)) return FVHOE_ERR_ARG;
  char* ws = (char*)workspace;
  hipStream_t s = (hipStream_t)stream;
  return P.fp32 ? by_tile<float>(cfg, P, u, t, steps, target, ws, s, persistent, pred, logits_only)
                 : by_tile<double>(cfg, P, u, t, steps, target, ws, s, persistent, pred, logits_only);
}
}  // namespace

extern "C" {

int fvhoe_ensemble_workspace_bytes(const FvhoeStageConfig* cfg, int count, const int* n, int tile,
                                   unsigned long long* bytes) {
  Plan P;
  const int st = make_plan(cfg, count, n, tile, &P);
  if (st != FVHOE_OK) return st;
  if (!bytes) return FVHOE_ERR_ARG;
  *bytes = P.total;
  return FVHOE_OK;
}

int fvhoe_ensemble_setup(const FvhoeStageConfig* cfg, int count, const int* n, const double* h, int tile,
                         void* workspace, void* stream) {
  Plan P;
  const int st = make_plan(cfg, count, n, tile, &P);
  if (st != FVHOE_OK) return st;
  if (!h || !workspace || ((uintptr_t)workspace & 255) != 0) return FVHOE_ERR_ARG;
  for (int i = 0; i < count; ++i)
    if (!(h[i] > 0.0)) return FVHOE_ERR_ARG;
  // host tables, copied once (synchronously: the host arrays are released on return)
  std::vector<char> host(P.o_W, 0);
  int* hn = (int*)(host.data() + P.o_n);
  double* hh = (double*)(host.data() + P.o_h);
  long long* ou = (long long*)(host.data() + P.o_ou);
  long long* ow = (long long*)(host.data() + P.o_ow);
  long long* of = (long long*)(host.data() + P.o_of);
  long long* os = (long long*)(host.data() + P.o_os);
  int* act = (int*)(host.data() + P.o_act);
  unsigned long long* acc0 = (unsigned long long*)(host.data() + P.o_acc);  // the first W pass's reduction words
  int2* tw = (int2*)(host.data() + P.o_tw);
  int2* tf = (int2*)(host.data() + P.o_tf);
  int2* ts = (int2*)(host.data() + P.o_ts);
  int2* tc = (int2*)(host.data() + P.o_tc);
  long long a = 0, b = 0, f = 0, sn = 0;
  int kw = 0, kf = 0, ks = 0, kc = 0;
  for (int i = 0; i < count; ++i) {
    hn[i] = n[i];
    hh[i] = h[i];
    ou[i] = a;
    ow[i] = b;
    of[i] = f;
    os[i] = sn;
    act[i] = 1;
    for (int k = 0; k < NACC; ++k) acc0[(size_t)i * NACC + k] = (k > 0 && (k & 1)) ? ~0ull : 0ull;
    a += n[i];
    b += n[i] + 2 * P.GE;
    f += n[i] + 1;
    sn += n[i] + 6;
    for (int x = 0; x < n[i] + 2 * P.GE; x += tile) tw[kw++] = make_int2(i, x);
    for (int x = 0; x < n[i] + 1; x += tile) tf[kf++] = make_int2(i, x);
    for (int x = 0; x < n[i] + 6; x += tile) ts[ks++] = make_int2(i, x);
    for (int x = 0; x < n[i]; x += tile) tc[kc++] = make_int2(i, x);
  }
  (void)stream;
  if (hipMemcpy(workspace, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) return FVHOE_ERR_LAUNCH;
  return FVHOE_OK;
}


int fvhoe_ensemble_step(const FvhoeStageConfig* cfg, int count, const int* n, int tile, double* u, double* t,
                        int* steps, double target, void* workspace, void* stream) {
  return ensemble_step(cfg, count, n, tile, u, t, steps, target, workspace, stream, false);
}

int fvhoe_ensemble_step_persistent(const FvhoeStageConfig* cfg, int count, const int* n, int tile, double* u,
                                   double* t, int* steps, double target, void* workspace, void* stream) {
  return ensemble_step(cfg, count, n, tile, u, t, steps, target, workspace, stream, true);
}

int fvhoe_ensemble_step_predicted(const FvhoeStageConfig* cfg, int count, const int* n, int tile, double* u,
                                  double* t, int* steps, double target, void* workspace, const FvhoePredict* pred,
                                  void* stream) {
  if (!pred) return FVHOE_ERR_ARG;
  return ensemble_step(cfg, count, n, tile, u, t, steps, target, workspace, stream, false, pred);
}

int fvhoe_ensemble_pred_logits(const FvhoeStageConfig* cfg, int count, const int* n, int tile, double* u,
                               double* t, void* workspace, const void* packed, float* logits, void* stream) {
  if (!packed || !logits) return FVHOE_ERR_ARG;
  FvhoePredict pr{};
  pr.packed = packed;
  pr.logit_threshold = 0.0f;
  pr.infer = 1;
  return ensemble_step(cfg, count, n, tile, u, t, nullptr, __builtin_huge_val(), workspace, stream, false, &pr,
                       logits);
}

int fvhoe_ensemble_status(const FvhoeStageConfig* cfg, int count, const int* n, int tile, const void* workspace,
                          int* error_word) {
  Plan P;
  const int st = make_plan(cfg, count, n, tile, &P);
  if (st != FVHOE_OK) return st;
  if (!workspace || !error_word) return FVHOE_ERR_ARG;
  if (hipMemcpy(error_word, (const char*)workspace + P.o_bar + 8, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return FVHOE_ERR_LAUNCH;
  return FVHOE_OK;
}

}  // extern "C"
