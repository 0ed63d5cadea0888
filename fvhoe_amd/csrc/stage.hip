// Fused per-RK-stage update (EulerSolver.hydrodynamics, fvhoe/solver.py:429-454)
// for 1-D (x) and 2-D (x-y) grids, nvars = 5, same degree p on the active axes.
//
// Kernels of one stage (all FP64, reference operation order):
//   prims  : u (BC-mapped on load) -> centres (c-sweeps x, y) -> primitives ->
//            fv averages (t-sweeps x, y) = W on the ghost-extended region, plus
//            the CFL max (hydro_dt over w_bc incl. ghosts) and the per-variable
//            min/max of the global NAD range (fvhoe/solver.py:753-792,
//            fvhoe/hydro.py:270-294, fvhoe/slope_limiting.py:168-170).
//   flux<D>: W -> c-sweep across the face, l/r sweeps along it, Riemann,
//            t-sweep across it = F (D=0) or G (D=1) (fvhoe/solver.py:519-649).
//   prims again on u* = u + dt*L(F,G), formed on load, -> WS (candidate
//            primitive averages, fvhoe/solver.py:678-687).
//   detect : NAD/PAD/SED flags from W and WS (fvhoe/slope_limiting.py:87-244).
//   final  : per cell, faces touching a flagged cell take the MUSCL p=1 flux
//            recomputed locally from the stage input (fvhoe/solver.py:710-731);
//            dudt = divergence (fvhoe/solver.py:734-751).
// All intermediate state is LDS-tiled; blocks are 256 threads.
#include "fvhoe_common.h"
#include "fvhoe_hip.h"

using namespace fvhoe;

namespace {

constexpr int TPB = 256;

// effective reach of the c and t stencils (outermost non-zero tap)
__host__ __device__ constexpr int ereach_c(int p) { return p == 1 ? 0 : (p % 2 ? reach(p) - 1 : reach(p)); }
__host__ __device__ constexpr int ereach_t(int p) { return ereach_c(p); }
__host__ __device__ constexpr int flux_ghosts(int p) {
  return reach(p) + 1 > 2 * reach(p) ? reach(p) + 1 : 2 * reach(p);  // gw of hydrofluxes, fvhoe/solver.py:493
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Give each XCD a contiguous run of tiles so neighbouring tiles (which share
// halo data) are served by the same L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int per = nb / 8, rem = nb % 8, x = b % 8, k = b / 8;
  return x * per + (x < rem ? x : rem) + k;
}

struct Geo {
  int n0, n1;        // interior cells (n1 = 1 in 1-D)
  int ge0, ge1;      // ghost width of W / WS per axis (0 on an inactive axis)
  int bc[4];         // x lo, x hi, y lo, y hi
  double cx, cy;     // -(1/hx), -(1/hy)
};

__device__ __forceinline__ int cell_map(int c, int n, int lo, int hi, bool* flip) {
  *flip = false;
  if (c < 0) {
    if (lo == BC_PERIODIC) return n + c;
    if (lo == BC_FREE) return 0;
    *flip = (lo == BC_REFLECTIVE);
    return -1 - c;
  }
  if (c >= n) {
    if (hi == BC_PERIODIC) return c - n;
    if (hi == BC_FREE) return n - 1;
    *flip = (hi == BC_REFLECTIVE);
    return 2 * n - 1 - c;
  }
  return c;
}

// conservative state at a (possibly ghost) cell: the BC fill of
// fvhoe/boundary_conditions.py:306-402 evaluated on demand
struct SrcU {
  const double* __restrict__ u;
  __device__ __forceinline__ double get(const Geo& g, int v, int cx, int cy) const {
    bool fx, fy;
    const int sx = cell_map(cx, g.n0, g.bc[0], g.bc[1], &fx);
    const int sy = cell_map(cy, g.n1, g.bc[2], g.bc[3], &fy);
    const double x = u[((long long)v * g.n0 + sx) * g.n1 + sy];
    return ((fx && v == 1) || (fy && v == 2)) ? -x : x;
  }
};

// candidate u* = u + dt * euler_equation(F, G) at a (possibly ghost) cell
// (fvhoe/solver.py:678); reflective ghosts negate the whole candidate value,
// which equals negating u and L first (IEEE rounding is sign-symmetric).
template <bool TWO_D>
struct SrcUstar {
  const double* __restrict__ u;
  const double* __restrict__ F;
  const double* __restrict__ G;
  const double* __restrict__ dt;
  __device__ __forceinline__ double get(const Geo& g, int v, int cx, int cy) const {
    bool fx, fy;
    const int sx = cell_map(cx, g.n0, g.bc[0], g.bc[1], &fx);
    const int sy = cell_map(cy, g.n1, g.bc[2], g.bc[3], &fy);
    const long long fi = ((long long)v * (g.n0 + 1) + sx) * g.n1 + sy;
    double L = 0.0 + g.cx * (F[fi + g.n1] - F[fi]);
    if (TWO_D) {
      const long long gi = ((long long)v * g.n0 + sx) * (g.n1 + 1) + sy;
      L = L + g.cy * (G[gi + 1] - G[gi]);
    }
    const double x = u[((long long)v * g.n0 + sx) * g.n1 + sy] + dt[0] * L;
    return ((fx && v == 1) || (fy && v == 2)) ? -x : x;
  }
};

// Per-block partial reductions, 11 words per block: [0] max of
// |vx|+|vy|+|vz|+ndim*c (raw bits of a non-negative double), [1 + 2v] / [2 + 2v]
// ordered bits of the min / max of variable v.  Reduced by k_stage_reduce.
constexpr int NPART = 11;
struct Reduce {
  unsigned long long* partials;   // null: no reduction in this launch
  int do_dt, do_range;
  int dt_lo0, dt_hi0, dt_lo1, dt_hi1;  // W index box of the dt reduction
  int rg_lo0, rg_hi0, rg_lo1, rg_hi1;  // W index box of the range reduction
  int ndim;
  double gamma, csq_floor;
};

// ---------------------------------------------------------------------------
// prims tile: W[a][b] for W indices a in [oa, oa+TX), b in [ob, ob+TY).
// ---------------------------------------------------------------------------
template <int PX, int PY, int TX, int TY, typename Src, int NT = TPB>
__global__ void __launch_bounds__(NT) k_prims(Src src, Geo g, double* __restrict__ W, int NW0, int NW1,
                                              int tiles1, Reduce red, double gamma) {
  constexpr int RCX = ereach_c(PX), RTX = ereach_t(PX), RCY = ereach_c(PY), RTY = ereach_t(PY);
  constexpr int HX0 = TX + 2 * (RCX + RTX), HY0 = TY + 2 * (RCY + RTY);
  constexpr int HX1 = TX + 2 * RTX, HY1 = TY + 2 * RTY;
  constexpr int NPT = (TX * TY + NT - 1) / NT;
  __shared__ double S0[HX0 * HY0];
  __shared__ double S1[HX1 * HY0];
  __shared__ double C[5][HX1 * HY1];

  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int oa = (tile / tiles1) * TX, ob = (tile % tiles1) * TY;
  const int cx0 = oa - g.ge0 - (RCX + RTX), cy0 = ob - g.ge1 - (RCY + RTY);
  const int tid = threadIdx.x;

  for (int v = 0; v < 5; ++v) {
    for (int e = tid; e < HX0 * HY0; e += NT) S0[e] = src.get(g, v, cx0 + e / HY0, cy0 + e % HY0);
    __syncthreads();
    for (int e = tid; e < HX1 * HY0; e += NT) {
      const int i = e / HY0, j = e % HY0;
      S1[e] = sweep<KIND_C, PX>([&](int t) { return S0[(i + RCX + t - reach(PX)) * HY0 + j]; });
    }
    __syncthreads();
    for (int e = tid; e < HX1 * HY1; e += NT) {
      const int i = e / HY1, j = e % HY1;
      const double x = sweep<KIND_C, PY>([&](int t) { return S1[i * HY0 + j + RCY + t - reach(PY)]; });
      C[v][e] = 0.0 + x;  // degree-0 z sweep (fvhoe/fv.py:186-196)
    }
    __syncthreads();
  }
  for (int e = tid; e < HX1 * HY1; e += NT) {
    double uu[5], ww[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) uu[q] = C[q][e];
    prims_from_cons(uu, ww, gamma);
#pragma unroll
    for (int q = 0; q < 5; ++q) C[q][e] = ww[q];
  }
  __syncthreads();
  double wr[5][NPT];
  for (int v = 0; v < 5; ++v) {
    for (int e = tid; e < TX * HY1; e += NT) {
      const int i = e / HY1, j = e % HY1;
      S1[e] = sweep<KIND_T, PX>([&](int t) { return C[v][(i + RTX + t - reach(PX)) * HY1 + j]; });
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
      const int e = tid + q * NT;
      if (e < TX * TY) {
        const int i = e / TY, j = e % TY;
        const double x = sweep<KIND_T, PY>([&](int t) { return S1[i * HY1 + j + RTY + t - reach(PY)]; });
        wr[v][q] = 0.0 + x;  // z-axis sweep of degree 0 (fvhoe/fv.py:221-223)
        const int a = oa + i, b = ob + j;
        if (a < NW0 && b < NW1) W[((long long)v * NW0 + a) * NW1 + b] = wr[v][q];
      }
    }
    __syncthreads();
  }
  // reductions over this tile's cells of the dt box and range box
  double m = 0.0;
  unsigned long long rmin[5], rmax[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) { rmin[q] = ~0ull; rmax[q] = 0ull; }
#pragma unroll
  for (int q = 0; q < NPT; ++q) {
    const int e = tid + q * NT;
    if (e < TX * TY) {
      const int a = oa + e / TY, b = ob + e % TY;
      if (red.do_dt && a >= red.dt_lo0 && a < red.dt_hi0 && b >= red.dt_lo1 && b < red.dt_hi1) {
        const double c = sound_speed(wr[0][q], wr[4][q], red.gamma, red.csq_floor);
        const double s = (fabs(wr[1][q]) + fabs(wr[2][q])) + fabs(wr[3][q]);
        m = dmax(m, s + (double)red.ndim * c);
      }
      if (red.do_range && a >= red.rg_lo0 && a < red.rg_hi0 && b >= red.rg_lo1 && b < red.rg_hi1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          const unsigned long long o = ordered_bits(wr[v][q]);
          rmin[v] = o < rmin[v] ? o : rmin[v];
          rmax[v] = o > rmax[v] ? o : rmax[v];
        }
      }
    }
  }
  if (red.partials) {
    __shared__ unsigned long long part[NT / 64][NPART];
    unsigned long long val[NPART];
    val[0] = (unsigned long long)__double_as_longlong(m);
#pragma unroll
    for (int v = 0; v < 5; ++v) { val[1 + 2 * v] = rmin[v]; val[2 + 2 * v] = rmax[v]; }
#pragma unroll
    for (int k = 0; k < NPART; ++k) {
      unsigned long long x = val[k];
      for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long y = __shfl_down(x, off, 64);
        x = (k == 0 || (k & 1) == 0) ? (y > x ? y : x) : (y < x ? y : x);
      }
      if ((tid & 63) == 0) part[tid >> 6][k] = x;
    }
    __syncthreads();
    if (tid < NPART) {
      unsigned long long x = part[0][tid];
      for (int w = 1; w < NT / 64; ++w) {
        const unsigned long long y = part[w][tid];
        x = (tid == 0 || (tid & 1) == 0) ? (y > x ? y : x) : (y < x ? y : x);
      }
      red.partials[(long long)blockIdx.x * NPART + tid] = x;
    }
  }
}

// ---------------------------------------------------------------------------
// flux tile.  D = 0: x faces fa in [f0, f0+TA) (face fa = left face of cell
// fa), rows b in [b0, b0+TB).  D = 1: y faces, rows along x.  "normal" is the
// face-normal axis with degree PN, "trans" the other with degree PT.
// ---------------------------------------------------------------------------
template <int D, int PN, int PT, int TA, int TB, int NT = TPB>
__global__ void __launch_bounds__(NT) k_flux(const double* __restrict__ W, int NW0, int NW1, Geo g,
                                             double* __restrict__ Fout, int tilesB, int solver, double gamma,
                                             double csq_floor) {
  constexpr int EL = reach(PN);
  constexpr int RC = ereach_c(PT), RT = ereach_t(PT);
  constexpr int HN = TA + 1 + 2 * EL;
  constexpr int HT0 = TB + 2 * (RC + RT), HT1 = TB + 2 * RT;
  constexpr int SZ_S0 = HN * HT0, SZ_FC = 5 * TA * HT1;
  __shared__ double X[5][HN * HT1];
  __shared__ double SF[SZ_S0 > SZ_FC ? SZ_S0 : SZ_FC];  // S0 during loads, FC after the Riemann solves

  const int nN = D == 0 ? g.n0 : g.n1, nT = D == 0 ? g.n1 : g.n0;
  const int geN = D == 0 ? g.ge0 : g.ge1, geT = D == 0 ? g.ge1 : g.ge0;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int f0 = (tile / tilesB) * TA, b0 = (tile % tilesB) * TB;
  const int tid = threadIdx.x;
  // W index of normal cell (f0 - 1 - EL) and transverse cell (b0 - RC - RT)
  const int wn0 = f0 - 1 - EL + geN, wt0 = b0 - RC - RT + geT;

  for (int v = 0; v < 5; ++v) {
    for (int e = tid; e < SZ_S0; e += NT) {
      int r, c;
      if (D == 0) { r = e / HT0; c = e % HT0; } else { c = e / HN; r = e % HN; }  // contiguous W axis fastest
      const int a = D == 0 ? wn0 + r : wt0 + c, b = D == 0 ? wt0 + c : wn0 + r;
      SF[r * HT0 + c] = (a < NW0 && b < NW1) ? W[((long long)v * NW0 + a) * NW1 + b] : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < HN * HT1; e += NT) {
      const int r = e / HT1, c = e % HT1;
      const double x = sweep<KIND_C, PT>([&](int t) { return SF[r * HT0 + c + RC + t - reach(PT)]; });
      X[v][e] = 0.0 + x;  // degree-0 z sweep precedes/follows (fvhoe/solver.py:524,541)
    }
    __syncthreads();
  }
  // Riemann problems at faces f0+fr, transverse rows c in [0, HT1)
  double* FC = SF;
  for (int e = tid; e < TA * HT1; e += NT) {
    const int fr = e / HT1, c = e % HT1;
    double wl[5], wr[5], F[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      // right face of cell f0+fr-1: centre row fr+EL; left face of cell f0+fr: centre row fr+EL+1
      wl[q] = sweep<KIND_R, PN>([&](int t) { return X[q][(fr + t) * HT1 + c]; });
      wr[q] = sweep<KIND_L, PN>([&](int t) { return X[q][(fr + 1 + t) * HT1 + c]; });
    }
    riemann<D>(solver, wl, wr, F, gamma, csq_floor);
#pragma unroll
    for (int q = 0; q < 5; ++q) FC[(q * TA + fr) * HT1 + c] = F[q];
  }
  __syncthreads();
  const int nF0 = D == 0 ? g.n0 + 1 : g.n0, nF1 = D == 0 ? g.n1 : g.n1 + 1;
  for (int e = tid; e < 5 * TA * TB; e += NT) {
    int q, fr, c;
    if (D == 0) { q = e / (TA * TB); fr = (e / TB) % TA; c = e % TB; }
    else { q = e / (TA * TB); c = (e / TA) % TB; fr = e % TA; }
    const int fa = f0 + fr, b = b0 + c;
    if (fa > nN || b >= nT) continue;
    const double x = sweep<KIND_T, PT>([&](int t) { return FC[(q * TA + fr) * HT1 + c + RT + t - reach(PT)]; });
    const int i0 = D == 0 ? fa : b, i1 = D == 0 ? b : fa;
    Fout[((long long)q * nF0 + i0) * nF1 + i1] = 0.0 + x;  // degree-0 z sweep (fvhoe/solver.py:633,639)
  }
}

// ---------------------------------------------------------------------------
// troubled-cell flags from W (3^d max/min) and WS (candidate, SED)
// ---------------------------------------------------------------------------
struct DetectP {
  int nad_mode, nad_range, nad_vars, pad_var, sed, force;
  double eps, pad_lo, pad_hi, sed_eps;
};

__global__ void __launch_bounds__(TPB) k_detect(const double* __restrict__ W, int NW0, int NW1,
                                               const double* __restrict__ WS, int NS0, int NS1, Geo g,
                                               const double* __restrict__ range_mm, DetectP a,
                                               signed char* __restrict__ trouble, int* __restrict__ count,
                                               int* __restrict__ list) {
  const long long nc = (long long)g.n0 * g.n1;
  const bool two_d = g.ge1 > 0;
  const long long c = blockIdx.x * (long long)TPB + threadIdx.x;
  int t = 0;
  if (c < nc) {
    const int i = (int)(c / g.n1), j = (int)(c % g.n1);
    const long long wb = (long long)(i + g.ge0) * NW1 + (j + g.ge1);
    const long long sb = (long long)(i + 3) * NS1 + (j + (two_d ? 3 : 0));
    const long long NWC = (long long)NW0 * NW1, NSC = (long long)NS0 * NS1;
    bool nad = false;
    for (int v = 0; v < 5; ++v) {
      const double* wv = W + v * NWC;
      double M = wv[wb], m = wv[wb];
#pragma unroll
      for (int di = -1; di <= 1; ++di) {
        if (two_d) {
#pragma unroll
          for (int dj = -1; dj <= 1; ++dj) {
            const double x = wv[wb + di * NW1 + dj];
            M = dmax(M, x);
            m = dmin(m, x);
          }
        } else {
          const double x = wv[wb + di * NW1];
          M = dmax(M, x);
          m = dmin(m, x);
        }
      }
      const double uc = WS[v * NSC + sb];
      double ind;
      bool tv;
      if (a.nad_range == 0) {
        const double rng = a.nad_mode == 0 ? range_mm[2 * v + 1] - range_mm[2 * v] : M - m;
        ind = dmin(uc - m, M - uc);
        tv = ind < -a.eps * rng;
      } else {
        const double lo = a.nad_mode == 0 ? (1.0 - a.eps) * range_mm[2 * v] : (1.0 - a.eps) * m;
        const double hi = a.nad_mode == 0 ? (1.0 + a.eps) * range_mm[2 * v + 1] : (1.0 + a.eps) * M;
        ind = dmin(uc - lo, hi - uc);
        tv = ind < 0.0;
      }
      if (!((a.nad_vars >> v) & 1)) tv = false;
      if (tv && a.sed) {
        const double* sv = WS + v * NSC;
        double alpha = sed_alpha_1d([&](int o) { return sv[sb + (long long)o * NS1]; }, a.sed_eps);
        if (two_d) alpha = dmin(alpha, sed_alpha_1d([&](int o) { return sv[sb + o]; }, a.sed_eps));
        tv = alpha < 1.0;
      }
      nad = nad || tv;
    }
    bool pad = false;
    if (a.pad_var >= 0) {
      const double x = WS[a.pad_var * NSC + sb];
      pad = dmin(x - a.pad_lo, a.pad_hi - x) < 0.0;
    }
    t = (a.force || pad || nad) ? 1 : 0;
    trouble[c] = (signed char)t;
  }
  // compaction of the flagged cells: wave ballot, block prefix, one atomic per block
  __shared__ int wbase[TPB / 64];
  __shared__ int bbase;
  const unsigned long long ball = __ballot(t);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wbase[wave] = __popcll(ball);
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < TPB / 64; ++w) {
      const int x = wbase[w];
      wbase[w] = tot;
      tot += x;
    }
    bbase = tot ? atomicAdd(count, tot) : 0;
  }
  __syncthreads();
  if (t) list[bbase + wbase[wave] + __popcll(ball & ((1ull << lane) - 1ull))] = (int)c;
}

// ---------------------------------------------------------------------------
// final: revised fluxes at faces touching flagged cells, dudt = divergence
// ---------------------------------------------------------------------------
struct FinalP {
  int solver, limiter;  // limiter: FVHOE_LIM_* or 3 = unlimited p=1 interpolation (slope_limiter=None)
  double gamma, csq_floor;
};

// primitive fv average of the p=1 fallback at a (ghost) cell: the p=1 c/t
// sweeps and the degree-0 sweeps are identities apart from 0.0 + x
__device__ __forceinline__ void p1_prims(const SrcU& src, const Geo& g, int cx, int cy, double w[5], double gamma) {
  double u[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) u[q] = 0.0 + src.get(g, q, cx, cy);
  prims_from_cons(u, w, gamma);
#pragma unroll
  for (int q = 0; q < 5; ++q) w[q] = 0.0 + w[q];
}

template <int D>
__device__ void fallback_flux(const SrcU& src, const Geo& g, int i, int j, const FinalP& p, double F[5]) {
  // face between cells (i-1, j) and (i, j) for D=0, (i, j-1) and (i, j) for D=1
  const int di = D == 0 ? 1 : 0, dj = D == 0 ? 0 : 1;
  double w0[5], w1[5], w2[5], w3[5], wl[5], wr[5];
  p1_prims(src, g, i - 2 * di, j - 2 * dj, w0, p.gamma);
  p1_prims(src, g, i - di, j - dj, w1, p.gamma);
  p1_prims(src, g, i, j, w2, p.gamma);
  p1_prims(src, g, i + di, j + dj, w3, p.gamma);
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    if (p.limiter == 3) {  // slope_limiter None: p=1 conservative interpolation (fvhoe/solver.py:531-536)
      // r weights are the reversed l weights: (-1/4, 1, 1/4)
      wl[q] = ((0.0 + (-0.25) * w0[q]) + 1.0 * w1[q]) + 0.25 * w2[q];
      wr[q] = ((0.0 + 0.25 * w1[q]) + 1.0 * w2[q]) + (-0.25) * w3[q];
    } else {
      const double s1 = limited_slope(p.limiter, w1[q] - w0[q], w2[q] - w1[q]);
      const double s2 = limited_slope(p.limiter, w2[q] - w1[q], w3[q] - w2[q]);
      wl[q] = w1[q] + 0.5 * s1;  // right face of the left cell
      wr[q] = w2[q] - 0.5 * s2;  // left face of the right cell
    }
  }
  riemann<D>(p.solver, wl, wr, F, p.gamma, p.csq_floor);
}

// MUSCL fluxes at the faces of every flagged cell, written over the
// high-order fluxes (theta = 1 faces of the simple broadcast); a face shared by
// two flagged cells is written twice with the same value.
template <bool TWO_D>
__global__ void __launch_bounds__(TPB) k_fallback(const int* __restrict__ count, const int* __restrict__ list,
                                                 SrcU src, Geo g, FinalP p, double* __restrict__ F,
                                                 double* __restrict__ G) {
  const long long n = count[0];
  constexpr int NF = TWO_D ? 4 : 2;
  const long long nF = (long long)(g.n0 + 1) * g.n1, nG = (long long)g.n0 * (g.n1 + 1);
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n * NF;
       e += (long long)gridDim.x * blockDim.x) {
    const int f = (int)(e / n);
    const int c = list[e % n];
    const int i = c / g.n1, j = c % g.n1;
    double Fv[5];
    if (f < 2) {
      fallback_flux<0>(src, g, i + f, j, p, Fv);
      const long long fi = (long long)(i + f) * g.n1 + j;
#pragma unroll
      for (int q = 0; q < 5; ++q) F[q * nF + fi] = Fv[q];
    } else if (TWO_D) {
      fallback_flux<1>(src, g, i, j + (f - 2), p, Fv);
      const long long gi = (long long)i * (g.n1 + 1) + j + (f - 2);
#pragma unroll
      for (int q = 0; q < 5; ++q) G[q * nG + gi] = Fv[q];
    }
  }
}

// Runge-Kutta combination written by the stage's last kernel; the stage's own
// dudt is the last k operand of the kind (fvhoe/ode.py:245-298).
struct RkArgs {
  int kind;                     // FVHOE_RK_*, -1 = none
  double* out;
  const double *u0, *k0, *k1, *k2;
  const double* dt;             // step dt (device)
};

__device__ __forceinline__ double rk_combine(const RkArgs& r, long long i, double k_this) {
  const double dt = r.dt[0];
  switch (r.kind) {
    case FVHOE_RK_AXPY: return r.u0[i] + dt * k_this;
    case FVHOE_RK_SSPRK3_STAGE3: return r.u0[i] + 0.25 * dt * r.k0[i] + 0.25 * dt * k_this;
    case FVHOE_RK_SSPRK3_FINAL: return r.u0[i] + (1.0 / 6.0) * dt * (r.k0[i] + r.k1[i] + 4.0 * k_this);
    case FVHOE_RK_SSPRK2_FINAL: return 0.5 * r.u0[i] + 0.5 * (r.k0[i] + dt * k_this);
    case FVHOE_RK_HALF_AXPY: return r.u0[i] + 0.5 * dt * k_this;
    case FVHOE_RK_RK4_FINAL: return r.u0[i] + (1.0 / 6.0) * dt * (r.k0[i] + 2.0 * r.k1[i] + 2.0 * r.k2[i] + k_this);
  }
  return 0.0;
}

// dudt = 0.0 + cx (F[i+1]-F[i]) + cy (G[j+1]-G[j])  (fvhoe/solver.py:744-750), then the
// optional RK combination.  rk.out may alias the stage input or rk.u0 (pointwise).
template <bool TWO_D>
__global__ void __launch_bounds__(TPB) k_final(const double* __restrict__ F, const double* __restrict__ G, Geo g,
                                              double* dudt, RkArgs rk) {
  const long long nc = (long long)g.n0 * g.n1;
  const long long nF = (long long)(g.n0 + 1) * g.n1, nG = (long long)g.n0 * (g.n1 + 1);
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < nc; c += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(c / g.n1), j = (int)(c % g.n1);
    const long long fi = (long long)i * g.n1 + j;
    const long long gi = (long long)i * (g.n1 + 1) + j;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      double acc = 0.0 + g.cx * (F[q * nF + fi + g.n1] - F[q * nF + fi]);
      if (TWO_D) acc = acc + g.cy * (G[q * nG + gi + 1] - G[q * nG + gi]);
      if (dudt) dudt[q * nc + c] = acc;
      if (rk.kind >= 0) rk.out[q * nc + c] = rk_combine(rk, q * nc + c, acc);
    }
  }
}

__global__ void k_stage_init(int* count) {
  if (threadIdx.x == 0) count[0] = 0;
}

// final reduction of the prims partials: dt (fvhoe/hydro.py:290-293) and the
// per-variable global (min, max) (fvhoe/slope_limiting.py:168-170)
__global__ void __launch_bounds__(1024) k_stage_reduce(const unsigned long long* __restrict__ partials, int nblocks,
                                                       double* range, double cfl_h, int fixed_on, double fixed_dt,
                                                       double* dt, double* dt_step, const double* t, double target,
                                                       int has_target) {
  __shared__ unsigned long long part[16][NPART];
  const int tid = threadIdx.x;
  unsigned long long val[NPART];
  val[0] = 0ull;
#pragma unroll
  for (int k = 1; k < NPART; ++k) val[k] = (k & 1) ? ~0ull : 0ull;
  for (int b = tid; b < nblocks; b += 1024) {
#pragma unroll
    for (int k = 0; k < NPART; ++k) {
      const unsigned long long y = partials[(long long)b * NPART + k];
      val[k] = (k == 0 || (k & 1) == 0) ? (y > val[k] ? y : val[k]) : (y < val[k] ? y : val[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < NPART; ++k) {
    unsigned long long x = val[k];
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long y = __shfl_down(x, off, 64);
      x = (k == 0 || (k & 1) == 0) ? (y > x ? y : x) : (y < x ? y : x);
    }
    if ((tid & 63) == 0) part[tid >> 6][k] = x;
  }
  __syncthreads();
  if (tid < NPART) {
    unsigned long long x = part[0][tid];
    for (int w = 1; w < 16; ++w) {
      const unsigned long long y = part[w][tid];
      x = (tid == 0 || (tid & 1) == 0) ? (y > x ? y : x) : (y < x ? y : x);
    }
    if (tid == 0) {
      const double d = fixed_on ? fixed_dt : cfl_h / __longlong_as_double((long long)x);
      dt[0] = d;
      // _hit_target (fvhoe/ode.py:11-22) for the first stage of a step
      if (dt_step) dt_step[0] = (has_target && target - t[0] < d) ? target - t[0] : d;
    }
    else range[tid - 1] = from_ordered_bits(x);
  }
}

// ------------------------------------------------------------------ host side

// Optional per-kernel event timing (bench instrumentation): each launch of a
// stage kernel is bracketed by a pair of events from a fixed pool.
enum { TK_PRIMS = 0, TK_FLUX_X = 1, TK_FLUX_Y = 2, TK_PRIMS_CAND = 3, TK_DETECT = 4, TK_FALLBACK = 5, TK_FINAL = 6,
       TK_COUNT = 7 };
constexpr int TIMING_POOL = 8192;
struct Timing {
  bool on = false;
  int used = 0;
  int kind[TIMING_POOL];
  hipEvent_t a[TIMING_POOL], b[TIMING_POOL];
  bool created = false;
  double ms[TK_COUNT] = {};
  long long count[TK_COUNT] = {};
};
Timing g_timing;

struct TimedLaunch {
  int slot = -1;
  hipStream_t s;
  TimedLaunch(int k, hipStream_t st) : s(st) {
    if (g_timing.on && g_timing.used < TIMING_POOL) {
      slot = g_timing.used++;
      g_timing.kind[slot] = k;
      (void)hipEventRecord(g_timing.a[slot], s);
    }
  }
  ~TimedLaunch() {
    if (slot >= 0) (void)hipEventRecord(g_timing.b[slot], s);
  }
};

struct Layout {
  int two_d, p, G, GE, n0, n1, NW0, NW1, NS0, NS1;
  int prims_blocks;
  size_t offW, offF, offG, offWS, offFlags, offList, offPart, offScal, total;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int make_layout(const FvhoeStageConfig* c, Layout* L) {
  if (!c) return FVHOE_ERR_ARG;
  const bool ax = !(c->n[0] == 1 && c->p[0] == 0), ay = !(c->n[1] == 1 && c->p[1] == 0),
             az = !(c->n[2] == 1 && c->p[2] == 0);
  if (!ax || az) return FVHOE_ERR_UNSUPPORTED;
  if (ay && c->p[1] != c->p[0]) return FVHOE_ERR_UNSUPPORTED;
  if (c->p[0] < 1 || c->p[0] > 8) return FVHOE_ERR_UNSUPPORTED;
  if (c->riemann < 0 || c->riemann > 2 || c->limiter < 0 || c->limiter > 3) return FVHOE_ERR_ARG;
  L->two_d = ay;
  L->p = c->p[0];
  L->G = flux_ghosts(L->p);
  L->GE = L->G > 3 ? L->G : 3;
  L->n0 = c->n[0];
  L->n1 = c->n[1];
  // the reference pads u by gw + 2 ceil(p/2) with slabs copied from the interior
  const int pad = L->G + 2 * reach(L->p);
  if (L->n0 < pad || (ay && L->n1 < pad)) return FVHOE_ERR_UNSUPPORTED;
  if (c->a_posteriori && (L->n0 < 7 || (ay && L->n1 < 7))) return FVHOE_ERR_UNSUPPORTED;
  L->NW0 = L->n0 + 2 * L->GE;
  L->NW1 = ay ? L->n1 + 2 * L->GE : 1;
  L->NS0 = L->n0 + 6;
  L->NS1 = ay ? L->n1 + 6 : 1;
  size_t o = 0;
  L->offW = o; o = align256(o + sizeof(double) * 5 * (size_t)L->NW0 * L->NW1);
  L->offF = o; o = align256(o + sizeof(double) * 5 * (size_t)(L->n0 + 1) * L->n1);
  L->offG = o; o = align256(o + (ay ? sizeof(double) * 5 * (size_t)L->n0 * (L->n1 + 1) : 0));
  L->offWS = o; o = align256(o + (c->a_posteriori ? sizeof(double) * 5 * (size_t)L->NS0 * L->NS1 : 0));
  L->offFlags = o; o = align256(o + (c->a_posteriori ? (size_t)L->n0 * L->n1 : 0));
  L->offList = o; o = align256(o + (c->a_posteriori ? sizeof(int) * (size_t)L->n0 * L->n1 : 0));
  {
    const int tx = L->two_d ? 16 : 256, ty = L->two_d ? 32 : 1;
    L->prims_blocks = ((L->NW0 + tx - 1) / tx) * ((L->NW1 + ty - 1) / ty);
  }
  L->offPart = o; o = align256(o + sizeof(unsigned long long) * NPART * (size_t)L->prims_blocks);
  L->offScal = o; o = align256(o + 256);
  L->total = o;
  return FVHOE_OK;
}

template <int P, bool TWO_D>
int run_stage(const FvhoeStageConfig* c, const FvhoeRkFuse* rkf, const Layout& L, const double* u, double* dudt,
              double* dt_stage,
              signed char* trouble_out, char* ws, hipStream_t s) {
  constexpr int PY = TWO_D ? P : 0;
  constexpr int TX = TWO_D ? 16 : 256, TY = TWO_D ? 32 : 1;
  constexpr int NTILE = TWO_D ? 512 : 256;  // threads of the LDS-tiled kernels
  double* W = (double*)(ws + L.offW);
  double* F = (double*)(ws + L.offF);
  double* G = TWO_D ? (double*)(ws + L.offG) : nullptr;
  double* WS = (double*)(ws + L.offWS);
  signed char* flags = (signed char*)(ws + L.offFlags);
  unsigned long long* partials = (unsigned long long*)(ws + L.offPart);
  double* range = (double*)(ws + L.offScal);
  int* count = (int*)(range + 10);
  int* list = (int*)(ws + L.offList);

  Geo g{};
  g.n0 = L.n0;
  g.n1 = L.n1;
  g.ge0 = L.GE;
  g.ge1 = TWO_D ? L.GE : 0;
  g.bc[0] = c->bc[0]; g.bc[1] = c->bc[1]; g.bc[2] = c->bc[2]; g.bc[3] = c->bc[3];
  g.cx = -(1.0 / c->h[0]);
  g.cy = -(1.0 / c->h[1]);

  const bool apost = c->a_posteriori != 0;
  const bool need_range = apost && c->nad_mode == 0;
  k_stage_init<<<1, 64, 0, s>>>(count);

  Reduce red{};
  red.partials = partials;
  red.do_dt = !c->fixed_dt_on;
  red.do_range = need_range;
  red.dt_lo0 = L.GE - L.G; red.dt_hi0 = L.GE + L.n0 + L.G;
  red.dt_lo1 = TWO_D ? L.GE - L.G : 0; red.dt_hi1 = TWO_D ? L.GE + L.n1 + L.G : 1;
  red.rg_lo0 = L.GE - 3; red.rg_hi0 = L.GE + L.n0 + 3;
  red.rg_lo1 = TWO_D ? L.GE - 3 : 0; red.rg_hi1 = TWO_D ? L.GE + L.n1 + 3 : 1;
  int nd = 0;
  for (int d = 0; d < 3; ++d) nd += !(c->n[d] == 1 && c->p[d] == 0);
  red.ndim = nd;
  red.gamma = c->gamma;
  red.csq_floor = c->csq_floor;

  {
    const int t0 = (L.NW0 + TX - 1) / TX, t1 = (L.NW1 + TY - 1) / TY;
    TimedLaunch tl(TK_PRIMS, s);
    k_prims<P, PY, TX, TY, SrcU, NTILE><<<t0 * t1, NTILE, 0, s>>>(SrcU{u}, g, W, L.NW0, L.NW1, t1, red, c->gamma);
  }
  const double hmin = c->h[0] < c->h[1] ? (c->h[0] < c->h[2] ? c->h[0] : c->h[2]) : (c->h[1] < c->h[2] ? c->h[1] : c->h[2]);
  RkArgs rk{-1, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  const bool set_step = rkf && rkf->set_dt_step;
  if (rkf && rkf->kind >= 0) rk = RkArgs{rkf->kind, rkf->out, rkf->u0, rkf->k0, rkf->k1, rkf->k2, rkf->dt_step};
  k_stage_reduce<<<1, 1024, 0, s>>>(partials, L.prims_blocks, range, c->cfl * hmin, c->fixed_dt_on, c->fixed_dt,
                                    dt_stage, set_step ? rkf->dt_step : nullptr, set_step ? rkf->t : nullptr,
                                    set_step ? rkf->target : 0.0, set_step ? rkf->has_target : 0);

  if (TWO_D) {
    constexpr int TA0 = 16, TB0 = 32, TA1 = 32, TB1 = 16;
    const int tx = (L.n0 + 1 + TA0 - 1) / TA0, ty = (L.n1 + TB0 - 1) / TB0;
    { TimedLaunch tl(TK_FLUX_X, s);
    k_flux<0, P, PY, TA0, TB0, NTILE><<<tx * ty, NTILE, 0, s>>>(W, L.NW0, L.NW1, g, F, ty, c->riemann, c->gamma, c->csq_floor); }
    const int fy = (L.n1 + 1 + TA1 - 1) / TA1, rx = (L.n0 + TB1 - 1) / TB1;
    { TimedLaunch tl(TK_FLUX_Y, s);
    k_flux<1, PY, P, TA1, TB1, NTILE><<<fy * rx, NTILE, 0, s>>>(W, L.NW0, L.NW1, g, G, rx, c->riemann, c->gamma, c->csq_floor); }
  } else {
    const int tx = (L.n0 + 1 + 255) / 256;
    TimedLaunch tl(TK_FLUX_X, s);
    k_flux<0, P, 0, 256, 1><<<tx, TPB, 0, s>>>(W, L.NW0, L.NW1, g, F, 1, c->riemann, c->gamma, c->csq_floor);
  }

  FinalP fp{c->riemann, c->limiter, c->gamma, c->csq_floor};
  const long long nc = (long long)L.n0 * L.n1;
  const int gridc = (int)((nc + TPB - 1) / TPB < 65536 ? (nc + TPB - 1) / TPB : 65536);
  if (apost) {
    Geo gs = g;
    gs.ge0 = 3;
    gs.ge1 = TWO_D ? 3 : 0;
    const int t0 = (L.NS0 + TX - 1) / TX, t1 = (L.NS1 + TY - 1) / TY;
    Reduce none{};
    { TimedLaunch tl(TK_PRIMS_CAND, s);
    k_prims<P, PY, TX, TY, SrcUstar<TWO_D>, NTILE><<<t0 * t1, NTILE, 0, s>>>(SrcUstar<TWO_D>{u, F, G, dt_stage}, gs, WS,
                                                                      L.NS0, L.NS1, t1, none, c->gamma); }
    DetectP dp{c->nad_mode, c->nad_range, c->nad_vars, c->pad_var, c->sed, c->force_trouble,
               c->nad_eps, c->pad_lo, c->pad_hi, c->sed_eps};
    const int gridd = (int)((nc + TPB - 1) / TPB);
    { TimedLaunch tl(TK_DETECT, s);
    k_detect<<<gridd, TPB, 0, s>>>(W, L.NW0, L.NW1, WS, L.NS0, L.NS1, g, range, dp, flags, count, list); }
    { TimedLaunch tl(TK_FALLBACK, s);
    k_fallback<TWO_D><<<1024, TPB, 0, s>>>(count, list, SrcU{u}, g, fp, F, G); }
    if (trouble_out && hipMemcpyAsync(trouble_out, flags, (size_t)nc, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return FVHOE_ERR_LAUNCH;
  }
  {
    TimedLaunch tl(TK_FINAL, s);
    k_final<TWO_D><<<gridc, TPB, 0, s>>>(F, G, g, dudt, rk);
  }
  return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
}

template <bool TWO_D>
int dispatch(int p, const FvhoeStageConfig* c, const FvhoeRkFuse* rk, const Layout& L, const double* u, double* dudt,
             double* dt,
             signed char* t, char* ws, hipStream_t s) {
  switch (p) {
    case 1: return run_stage<1, TWO_D>(c, rk, L, u, dudt, dt, t, ws, s);
    case 2: return run_stage<2, TWO_D>(c, rk, L, u, dudt, dt, t, ws, s);
    case 3: return run_stage<3, TWO_D>(c, rk, L, u, dudt, dt, t, ws, s);
    case 4: return run_stage<4, TWO_D>(c, rk, L, u, dudt, dt, t, ws, s);
    case 5: return run_stage<5, TWO_D>(c, rk, L, u, dudt, dt, t, ws, s);
    case 6: return run_stage<6, TWO_D>(c, rk, L, u, dudt, dt, t, ws, s);
    case 7: return run_stage<7, TWO_D>(c, rk, L, u, dudt, dt, t, ws, s);
    case 8: return run_stage<8, TWO_D>(c, rk, L, u, dudt, dt, t, ws, s);
  }
  return FVHOE_ERR_UNSUPPORTED;
}

}  // namespace

extern "C" {

int fvhoe_stage_supported(const FvhoeStageConfig* cfg) {
  Layout L;
  return make_layout(cfg, &L) == FVHOE_OK ? 1 : 0;
}

int fvhoe_stage_workspace_bytes(const FvhoeStageConfig* cfg, unsigned long long* bytes) {
  Layout L;
  const int st = make_layout(cfg, &L);
  if (st != FVHOE_OK) return st;
  if (!bytes) return FVHOE_ERR_ARG;
  *bytes = L.total;
  return FVHOE_OK;
}

int fvhoe_stage_timing(int on) {
  if (on && !g_timing.created) {
    for (int i = 0; i < TIMING_POOL; ++i) {
      if (hipEventCreate(&g_timing.a[i]) != hipSuccess || hipEventCreate(&g_timing.b[i]) != hipSuccess)
        return FVHOE_ERR_LAUNCH;
    }
    g_timing.created = true;
  }
  g_timing.on = on != 0;
  g_timing.used = 0;
  for (int k = 0; k < TK_COUNT; ++k) { g_timing.ms[k] = 0.0; g_timing.count[k] = 0; }
  return FVHOE_OK;
}

int fvhoe_stage_timing_read(double* ms, long long* count) {
  if (!ms || !count) return FVHOE_ERR_ARG;
  for (int i = 0; i < g_timing.used; ++i) {
    float t = 0.0f;
    if (hipEventSynchronize(g_timing.b[i]) != hipSuccess) return FVHOE_ERR_LAUNCH;
    if (hipEventElapsedTime(&t, g_timing.a[i], g_timing.b[i]) != hipSuccess) return FVHOE_ERR_LAUNCH;
    g_timing.ms[g_timing.kind[i]] += t;
    g_timing.count[g_timing.kind[i]] += 1;
  }
  g_timing.used = 0;
  for (int k = 0; k < TK_COUNT; ++k) { ms[k] = g_timing.ms[k]; count[k] = g_timing.count[k]; }
  return FVHOE_OK;
}

int fvhoe_stage_rk(const FvhoeStageConfig* cfg, const double* u, double* dudt, double* dt_stage,
                   signed char* trouble, void* workspace, const FvhoeRkFuse* rk, void* stream) {
  Layout L;
  const int st = make_layout(cfg, &L);
  if (st != FVHOE_OK) return st;
  if (!u || !dt_stage || !workspace) return FVHOE_ERR_ARG;
  if (!dudt && !(rk && rk->kind >= 0)) return FVHOE_ERR_ARG;
  if (rk && rk->kind >= 0) {
    if (rk->kind > FVHOE_RK_RK4_FINAL || !rk->out || !rk->u0 || !rk->dt_step) return FVHOE_ERR_ARG;
    if ((rk->kind == FVHOE_RK_SSPRK3_STAGE3 || rk->kind == FVHOE_RK_SSPRK2_FINAL) && !rk->k0) return FVHOE_ERR_ARG;
    if (rk->kind == FVHOE_RK_SSPRK3_FINAL && (!rk->k0 || !rk->k1)) return FVHOE_ERR_ARG;
    if (rk->kind == FVHOE_RK_RK4_FINAL && (!rk->k0 || !rk->k1 || !rk->k2)) return FVHOE_ERR_ARG;
    if (rk->out == dudt) return FVHOE_ERR_ARG;
  }
  if (rk && rk->set_dt_step && (!rk->dt_step || !rk->t)) return FVHOE_ERR_ARG;
  if (((uintptr_t)workspace & 255) != 0) return FVHOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  return L.two_d ? dispatch<true>(L.p, cfg, rk, L, u, dudt, dt_stage, trouble, ws, s)
                 : dispatch<false>(L.p, cfg, rk, L, u, dudt, dt_stage, trouble, ws, s);
}

int fvhoe_stage(const FvhoeStageConfig* cfg, const double* u, double* dudt, double* dt_stage, signed char* trouble,
                void* workspace, void* stream) {
  if (!dudt) return FVHOE_ERR_ARG;
  return fvhoe_stage_rk(cfg, u, dudt, dt_stage, trouble, workspace, nullptr, stream);
}

}  // extern "C"
