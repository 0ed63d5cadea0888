// Fused per-RK-stage update (EulerSolver.hydrodynamics, fvhoe/solver.py:429-454)
// for 1-D (x) and 2-D (x-y) grids, nvars = 5, same degree p on the active axes.
//
// Kernels of one stage (reference operation order; T = double reproduces the
// reference bit for bit, T = float is the FP32 bulk form):
//   prims  : u (BC-mapped on load) -> centres (c-sweeps x, y) -> primitives ->
//            fv averages (t-sweeps x, y) = W on the ghost-extended region, plus
//            per-block partials of the CFL max (hydro_dt over w_bc incl. ghosts)
//            and of the global NAD range (fvhoe/solver.py:753-792,
//            fvhoe/hydro.py:270-294, fvhoe/slope_limiting.py:168-170).
//   flux<D>: W -> c-sweep across the face, l/r sweeps along it, Riemann,
//            t-sweep across it = F (D=0) or G (D=1) (fvhoe/solver.py:519-649).
//   prims again on u* = u + dt*L(F,G), formed on load -> WS (candidate
//            primitive averages, fvhoe/solver.py:678-687); in 2-D the same
//            tile runs the NAD/PAD tests of its cells from W and WS.
//   detect : NAD/PAD/SED flags from W and WS (fvhoe/slope_limiting.py:87-244)
//            and a ballot compaction of the flagged cells; in 2-D only for the
//            cells whose NAD test flagged (the SED reads WS beyond the tile).
//   fallback: MUSCL p=1 fluxes at the faces of flagged cells, recomputed
//            locally from the stage input and written over F/G
//            (fvhoe/solver.py:710-731).
//   final  : dudt = divergence (fvhoe/solver.py:734-751) and, optionally, the
//            stepper's RK combination (fvhoe/ode.py:245-298).
// In the FP32 form the sweeps, Riemann solves and W/F/G/WS storage are float;
// the conservative-to-primitive pressure recovery, the NAD/PAD/SED tests, the
// divergence, the candidate and the RK accumulation stay FP64.
#include <mutex>
#include <type_traits>

#include "fvhoe_common.h"
#include "fvhoe_hip.h"
#include "predictor.h"
#include "predictor_infer.h"

using namespace fvhoe;

namespace {

constexpr int TPB = 256;

// Launch shapes, chosen by timing KH 2048^2 p=3 on MI355X (each can be overridden at build
// time for a sweep): W rows per lane of the column-strip prims kernel, LDS tile of the tiled
// prims kernel (candidate state and p > 3) per arithmetic type, face tile of the flux kernels
// (faces x rows), rows per thread of the 2-D check.
#ifndef FVHOE_STRIP_ROWS_F64
#define FVHOE_STRIP_ROWS_F64 4
#endif
#ifndef FVHOE_STRIP_ROWS_F32
#define FVHOE_STRIP_ROWS_F32 4
#endif
#ifndef FVHOE_TILE_X
#define FVHOE_TILE_X 16
#endif
#ifndef FVHOE_TILE_Y
#define FVHOE_TILE_Y 32
#endif
#ifndef FVHOE_TILE32_X
#define FVHOE_TILE32_X 16
#endif
#ifndef FVHOE_TILE32_Y
#define FVHOE_TILE32_Y 64
#endif
// candidate-state tiles (rows x columns per arithmetic type) and their threads.  FP64 at degrees up to
// FVHOE_CAND_MAX_P: 32 x 32 on 512 threads (21 KB of LDS at p = 3 with the centre values held in
// registers) measured 191 us per launch at KH 2048^2 p = 3 against 208 for 16 x 32; higher degrees
// (whose halo pre-loads need 112-222 VGPRs at 32 x 32) keep the prims tile.  FP32 keeps the prims
// tile's 16 x 64 (141 us; 165 at 32 x 64)
#ifndef FVHOE_CAND_MAX_P
#define FVHOE_CAND_MAX_P 3
#endif
// LDS-tiled prims / candidate: 1 = two buffers per sweep stage and one barrier per variable, 0 = one
// buffer and three (two in the t-sweeps) barriers per variable
#ifndef FVHOE_TILE_PIPE
#define FVHOE_TILE_PIPE 1
#endif
// flux tiles: 1 = two staging buffers for the transverse sweep (one barrier per variable), 0 = one.
// Two buffers measured 107 / 108 us per FP64 launch at KH 2048^2 against 104 / 105 for one, and no
// change in FP32 (60 / 62 us)
#ifndef FVHOE_FLUX_PIPE_F64
#define FVHOE_FLUX_PIPE_F64 0
#endif
#ifndef FVHOE_FLUX_PIPE_F32
#define FVHOE_FLUX_PIPE_F32 0
#endif
#ifndef FVHOE_CAND_X
#define FVHOE_CAND_X 32
#endif
#ifndef FVHOE_CAND_Y
#define FVHOE_CAND_Y 32
#endif
#ifndef FVHOE_CAND32_X
#define FVHOE_CAND32_X FVHOE_TILE32_X
#endif
#ifndef FVHOE_CAND32_Y
#define FVHOE_CAND32_Y FVHOE_TILE32_Y
#endif
#ifndef FVHOE_CAND_NT_F64
#define FVHOE_CAND_NT_F64 512
#endif
#ifndef FVHOE_CAND_NT_F32
#define FVHOE_CAND_NT_F32 512
#endif
#ifndef FVHOE_FLUX_A
#define FVHOE_FLUX_A 32
#endif
#ifndef FVHOE_FLUX_B
#define FVHOE_FLUX_B 32
#endif
// FP64 flux tiles (faces x rows) and threads of the FP64 / FP32 flux tiles.  FP64: 24 x 16 faces on
// 256 threads (27 KB of LDS, 5 tiles per CU) measured 105 / 108 us per launch at KH 2048^2 against
// 114 / 113 for 32 x 32 on 512 threads (61 KB, 2 tiles per CU)
#ifndef FVHOE_FLUX64_A
#define FVHOE_FLUX64_A 24
#endif
#ifndef FVHOE_FLUX64_B
#define FVHOE_FLUX64_B 16
#endif
// FP64 y-flux tiles (faces along y x rows along x): 32 x 8 writes G in 256-byte row runs (16-face runs
// straddle cache lines), 104 us per launch at KH 2048^2 against 110 for the transposed x-flux tile
#ifndef FVHOE_FLUX64_YA
#define FVHOE_FLUX64_YA 32
#endif
#ifndef FVHOE_FLUX64_YB
#define FVHOE_FLUX64_YB 8
#endif
// FP32 y-flux tiles (faces along y x rows along x); default: the x-flux tile transposed
#ifndef FVHOE_FLUX32_YA
#define FVHOE_FLUX32_YA FVHOE_FLUX_B
#endif
#ifndef FVHOE_FLUX32_YB
#define FVHOE_FLUX32_YB FVHOE_FLUX_A
#endif
#ifndef FVHOE_FLUX_NT_F64
#define FVHOE_FLUX_NT_F64 256
#endif
#ifndef FVHOE_FLUX_NT_F32
#define FVHOE_FLUX_NT_F32 512
#endif
// 2-D fluxes: one kernel for both face axes from one W box (FVHOE_FLUX_XY = 1), tile of A (x) x B (y)
// face positions on NT threads per arithmetic type, or one kernel per axis (0)
// occupancy floor of the FP32 two-axis flux kernel (1 = the compiler's choice)
#ifndef FVHOE_FLUXXY_WAVES_F32
#define FVHOE_FLUXXY_WAVES_F32 1
#endif
#ifndef FVHOE_FLUX_XY
#define FVHOE_FLUX_XY 1
#endif
// (KH 2048^2 p=3, both axes per launch: FP64 32 x 16 on 256 threads holding the box in registers 202 us
// against 104 + 103 for the two per-axis kernels, 212 fetching it again for the y faces, 214 at 16 x 16;
// FP32 16 x 32 on 256 threads fetching it again 114 us against 58 + 61, 119 for 32 x 32 on 512 holding it)
#ifndef FVHOE_FXY_RELOAD_F64
#define FVHOE_FXY_RELOAD_F64 0
#endif
#ifndef FVHOE_FXY_RELOAD_F32
#define FVHOE_FXY_RELOAD_F32 1
#endif
#ifndef FVHOE_FXY64_A
#define FVHOE_FXY64_A 32
#endif
#ifndef FVHOE_FXY64_B
#define FVHOE_FXY64_B 16
#endif
#ifndef FVHOE_FXY64_NT
#define FVHOE_FXY64_NT 256
#endif
#ifndef FVHOE_FXY32_A
#define FVHOE_FXY32_A 16
#endif
#ifndef FVHOE_FXY32_B
#define FVHOE_FXY32_B 32
#endif
#ifndef FVHOE_FXY32_NT
#define FVHOE_FXY32_NT 256
#endif
// 2-D: the troubled-cell test inside the candidate tile (1) or its own pass over W and WS (0)
#ifndef FVHOE_CAND_CHECK
#define FVHOE_CAND_CHECK 1
#endif
#ifndef FVHOE_DETECT_ROWS
#define FVHOE_DETECT_ROWS 2
#endif

// FP32 bulk form: fused multiply-adds in the stencil sweeps (FVHOE_F32_FMA) and the Riemann solvers'
// approximate reciprocal / square root (FVHOE_F32_FAST); 0 gives the correctly rounded operations
#ifndef FVHOE_F32_FMA
#define FVHOE_F32_FMA 1
#endif
#ifndef FVHOE_F32_FAST
#define FVHOE_F32_FAST 1
#endif

// effective reach of the c and t stencils (outermost non-zero tap)
__host__ __device__ constexpr int ereach_c(int p) { return p == 1 ? 0 : (p % 2 ? reach(p) - 1 : reach(p)); }
__host__ __device__ constexpr int ereach_t(int p) { return ereach_c(p); }
__host__ __device__ constexpr int flux_ghosts(int p) {
  return reach(p) + 1 > 2 * reach(p) ? reach(p) + 1 : 2 * reach(p);  // gw of hydrofluxes, fvhoe/solver.py:493
}

// stencil sweep in arithmetic type T (weights rounded to T once, at compile time)
// (FP64: a product and a sum per tap in the reference's order; FP32 bulk form: one fused multiply-add)
template <int KIND, int P, typename T, typename Get>
__device__ __forceinline__ T sweepT(Get&& x) {
  T acc = T(0);
  static_for<ntaps(P)>([&](auto I) {
    constexpr double w = weight(KIND, P, decltype(I)::value);
    if constexpr (w != 0.0) acc = q_fma<FVHOE_F32_FMA>(T(w), x(decltype(I)::value), acc);
  });
  return acc;
}

// The same visit unrolled, with the visit's ordinal k (0 <= k < ceil(H*Wd/NT))
// passed as well, so per-thread register arrays can be indexed by it.
template <int H, int Wd, int NT, typename F>
__device__ __forceinline__ void tile_for_k(F&& f) {
  constexpr int DR = NT / Wd, DC = NT % Wd, NK = (H * Wd + NT - 1) / NT;
  int r = (int)threadIdx.x / Wd, c = (int)threadIdx.x % Wd;
#pragma unroll
  for (int k = 0; k < NK; ++k) {
    const int e = (int)threadIdx.x + k * NT;
    if (e < H * Wd) f(k, e, r, c);
    r += DR;
    c += DC;
    if (c >= Wd) {
      c -= Wd;
      ++r;
    }
  }
}

// Visit the H x Wd elements of a row-major LDS tile, NT threads, element
// e = r*Wd + c, advancing (r, c) incrementally instead of dividing.
template <int H, int Wd, int NT, typename F>
__device__ __forceinline__ void tile_for(F&& f) {
  constexpr int DR = NT / Wd, DC = NT % Wd;
  int r = (int)threadIdx.x / Wd, c = (int)threadIdx.x % Wd;
  for (int e = threadIdx.x; e < H * Wd; e += NT) {
    f(e, r, c);
    r += DR;
    c += DC;
    if (c >= Wd) {
      c -= Wd;
      ++r;
    }
  }
}


struct Geo {
  int n0, n1;        // interior cells (n1 = 1 in 1-D)
  int ge0, ge1;      // ghost width of W / WS per axis (0 on an inactive axis)
  int bc[4];         // x lo, x hi, y lo, y hi
  double cx, cy;     // -(1/hx), -(1/hy)
};

// ghost value through the Dirichlet-aware composition (2-D grid, z of extent 1)
template <typename At>
__device__ __forceinline__ double geo_fetch(const Geo& g, const Slabs& sl, int v, int cx, int cy, At&& at) {
  const int n[3] = {g.n0, g.n1, 1};
  const int bc[6] = {g.bc[0], g.bc[1], g.bc[2], g.bc[3], BC_PERIODIC, BC_PERIODIC};
  return bc_fetch3(n, bc, sl, v, cx, cy, 0, [&](int vv, int a, int b, int) { return at(vv, a, b); });
}

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
// fvhoe/boundary_conditions.py:306-402 evaluated on demand.  The slab-aware
// variants (SrcUD, SrcUstarD) are separate types, so kernels of grids without
// Dirichlet sides carry none of the slab code.
struct SrcU {
  const double* __restrict__ u;
  __device__ __forceinline__ double at(const Geo& g, int v, int sx, int sy) const {
    return u[(v * g.n0 + sx) * g.n1 + sy];
  }
  __device__ __forceinline__ double get(const Geo& g, int v, int cx, int cy) const {
    bool fx, fy;
    const int sx = cell_map(cx, g.n0, g.bc[0], g.bc[1], &fx);
    const int sy = cell_map(cy, g.n1, g.bc[2], g.bc[3], &fy);
    const double x = at(g, v, sx, sy);
    return ((fx && v == 1) || (fy && v == 2)) ? -x : x;
  }
};

// candidate u* = u + dt * euler_equation(F, G) at a (possibly ghost) cell
// (fvhoe/solver.py:678); reflective ghosts negate the whole candidate value,
// which equals negating u and L first (IEEE rounding is sign-symmetric).
template <bool TWO_D, typename T>
struct SrcUstar {
  const double* __restrict__ u;
  const T* __restrict__ F;
  const T* __restrict__ G;
  const double* __restrict__ dt;
  __device__ __forceinline__ double at(const Geo& g, int v, int sx, int sy) const {
    const int fi = (v * (g.n0 + 1) + sx) * g.n1 + sy;
    double L = 0.0 + g.cx * ((double)F[fi + g.n1] - (double)F[fi]);
    if (TWO_D) {
      const int gi = (v * g.n0 + sx) * (g.n1 + 1) + sy;
      L = L + g.cy * ((double)G[gi + 1] - (double)G[gi]);
    }
    return u[(v * g.n0 + sx) * g.n1 + sy] + dt[0] * L;
  }
  __device__ __forceinline__ double get(const Geo& g, int v, int cx, int cy) const {
    bool fx, fy;
    const int sx = cell_map(cx, g.n0, g.bc[0], g.bc[1], &fx);
    const int sy = cell_map(cy, g.n1, g.bc[2], g.bc[3], &fy);
    const double x = at(g, v, sx, sy);
    return ((fx && v == 1) || (fy && v == 2)) ? -x : x;
  }
};

// the same sources with Dirichlet / "ic" sides read from the slabs (which
// travel with these types only, so the other kernels' arguments stay small)
struct SrcUD : SrcU {
  Slabs sl;
  __device__ __forceinline__ double get(const Geo& g, int v, int cx, int cy) const {
    return geo_fetch(g, sl, v, cx, cy, [&](int vv, int a, int b) { return at(g, vv, a, b); });
  }
};
template <bool TWO_D, typename T>
struct SrcUstarD : SrcUstar<TWO_D, T> {
  Slabs sl;
  __device__ __forceinline__ double get(const Geo& g, int v, int cx, int cy) const {
    return geo_fetch(g, sl, v, cx, cy, [&](int vv, int a, int b) { return this->at(g, vv, a, b); });
  }
};

template <typename S>
S make_src(const double* u, const Slabs& sl) {
  S s{};
  s.u = u;
  if constexpr (std::is_same_v<S, SrcUD>) s.sl = sl;
  return s;
}
template <typename S, typename T>
S make_srcstar(const double* u, const T* F, const T* G, const double* dt, const Slabs& sl) {
  S s{};
  s.u = u;
  s.F = F;
  s.G = G;
  s.dt = dt;
  if constexpr (!std::is_same_v<S, SrcUstar<true, T>> && !std::is_same_v<S, SrcUstar<false, T>>) s.sl = sl;
  return s;
}

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

__device__ __forceinline__ unsigned long long red_op(int k, unsigned long long a, unsigned long long b) {
  return (k == 0 || (k & 1) == 0) ? (a > b ? a : b) : (a < b ? a : b);
}

// ---------------------------------------------------------------------------
// troubled-cell tests (fvhoe/slope_limiting.py:87-244), in FP64
// ---------------------------------------------------------------------------
struct DetectP {
  int nad_mode, nad_range, nad_vars, pad_var, sed, force;
  double eps, pad_lo, pad_hi, sed_eps;
};

// Violation magnitudes for the trouble log (fvhoe/slope_limiting.py:225-239): every cell whose NAD or PAD
// test flags is appended to this stage's slot as (cell, NAD magnitude, PAD magnitude).  NAD magnitude:
// -min over all five variables of the NAD indicator where NAD flags (after SED and the NAD_vars mask),
// else 0; PAD magnitude: -(PAD indicator of the last PAD key) where PAD flags, else 0.
struct MagLog {
  int* count;          // entries in this stage's slot (cleared by k_stage_reduce)
  int* idx;
  double* nad;
  double* pad;
  const double* gate;  // the step's dt: a step at its target time (dt = 0) appends nothing
};
__device__ __forceinline__ bool mag_on(const MagLog& m) { return m.idx && !(m.gate && m.gate[0] == 0.0); }

// append the lanes' flagged cells to the slot: one atomic per wave
__device__ __forceinline__ void mag_append(const MagLog& m, bool f, int cell, double nad, double pad) {
  const unsigned long long b = __ballot(f);
  if (b == 0ull) return;
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (lane == 0) base = atomicAdd(m.count, (int)__popcll(b));
  base = __shfl(base, 0, 64);
  if (f) {
    const int e = base + (int)__popcll(b & ((1ull << lane) - 1ull));
    m.idx[e] = cell;
    m.nad[e] = nad;
    m.pad[e] = pad;
  }
}

// The troubled-cell test fused into the 2-D candidate tile: the tile's candidate values (WS) are still in
// registers and the stage's W over the tile's cells plus one ring comes through LDS.  A cell whose NAD
// test flags while the smooth-extrema detector is on goes to a pending list for k_detect_list (its SED
// reads candidate values up to 3 cells beyond the tile); every other cell is decided in the tile: flag,
// fallback list, pre-flag count and the trouble log's magnitudes, with k_detect_strip's operations.
template <typename T>
struct CheckArgs {
  const T* __restrict__ W;
  int NW0, NW1, wg0, wg1;  // W's extents and ghost widths
  const double* __restrict__ range;
  DetectP dp;
  signed char* __restrict__ trouble;
  int* __restrict__ count;       // cells in the fallback list
  int* __restrict__ list;
  int* __restrict__ pend_count;  // cells left to k_detect_list
  int* __restrict__ pend;        // with pack: cell | (NAD bits << PEND_SHIFT), the NAD indicator in pend_imin
  double* __restrict__ pend_imin;
  int pack;
  const signed char* __restrict__ pre;
  int* __restrict__ fp_count;
  MagLog ml;
};

// Append up to K cells per thread to two lists (bit k of sel0 / sel1 selects cell[k]): wave ballots, a
// block prefix over the waves and one atomic per list and block.  Every thread of the block calls it.
template <int K, int NT>
__device__ __forceinline__ void block_append2(const int (&cell)[K], unsigned sel0, unsigned sel1, int* count0,
                                              int* list0, int* count1, int* list1, const int* entry1 = nullptr,
                                              const double* aux1 = nullptr, double* auxlist1 = nullptr) {
  __shared__ int wsum[NT / 64][2];
  __shared__ int bbase[2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long lt = (1ull << lane) - 1ull;
  int o0[K], o1[K], n0 = 0, n1 = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const unsigned long long b0 = __ballot((sel0 >> k) & 1u), b1 = __ballot((sel1 >> k) & 1u);
    o0[k] = n0 + (int)__popcll(b0 & lt);
    o1[k] = n1 + (int)__popcll(b1 & lt);
    n0 += (int)__popcll(b0);
    n1 += (int)__popcll(b1);
  }
  if (lane == 0) {
    wsum[wave][0] = n0;
    wsum[wave][1] = n1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int a0 = 0, a1 = 0;
    for (int w = 0; w < NT / 64; ++w) {
      const int x0 = wsum[w][0], x1 = wsum[w][1];
      wsum[w][0] = a0;
      wsum[w][1] = a1;
      a0 += x0;
      a1 += x1;
    }
    bbase[0] = a0 ? atomicAdd(count0, a0) : 0;
    bbase[1] = a1 ? atomicAdd(count1, a1) : 0;
  }
  __syncthreads();
  const int p0 = bbase[0] + wsum[wave][0], p1 = bbase[1] + wsum[wave][1];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if ((sel0 >> k) & 1u) list0[p0 + o0[k]] = cell[k];
    if ((sel1 >> k) & 1u) {  // list1 entries may carry another value and an aux double (entry1 / aux1)
      list1[p1 + o1[k]] = entry1 ? entry1[k] : cell[k];
      if (auxlist1) auxlist1[p1 + o1[k]] = aux1[k];
    }
  }
}
// a pending-list entry carries the cell's NAD bits above this bit (grids of fewer cells than 1 << PEND_SHIFT)
constexpr int PEND_SHIFT = 27;

// ---------------------------------------------------------------------------
// prims tile: W[a][b] for W indices a in [oa, oa+TX), b in [ob, ob+TY).
// ---------------------------------------------------------------------------
// The fused check's FP64 tiles ask for 6 waves per SIMD (three 512-thread tiles per CU, <= 80 VGPRs).
#ifndef FVHOE_CHECK_WAVES_F64
#define FVHOE_CHECK_WAVES_F64 6
#endif
// the candidate check's pending cells carry their NAD bits and indicator to the SED pass (1) or the SED
// pass repeats the NAD test (0)
#ifndef FVHOE_PEND_PACK
#define FVHOE_PEND_PACK 1
#endif
// FP32 check tiles: the compiler's occupancy unless a floor is asked for (1 = none)
#ifndef FVHOE_CHECK_WAVES_F32
#define FVHOE_CHECK_WAVES_F32 1
#endif
template <typename T, int PX, int PY, int TX, int TY, typename Src, int NT, bool CHECK = false>
__global__ void __launch_bounds__(NT, !CHECK ? 1 : sizeof(T) == 8 ? FVHOE_CHECK_WAVES_F64 : FVHOE_CHECK_WAVES_F32) k_prims(Src src, Geo g, T* __restrict__ W, int NW0, int NW1, int tiles1,
                                             Reduce red, double gamma, CheckArgs<T> ck) {
  constexpr int RCX = ereach_c(PX), RTX = ereach_t(PX), RCY = ereach_c(PY), RTY = ereach_t(PY);
  constexpr int HX0 = TX + 2 * (RCX + RTX), HY0 = TY + 2 * (RCY + RTY);
  constexpr int HX1 = TX + 2 * RTX, HY1 = TY + 2 * RTY;
  constexpr int NPT = (TX * TY + NT - 1) / NT;
  constexpr int NC = (HX1 * HY1 + NT - 1) / NT;  // centre values per thread
  static_assert(HX1 * HY1 <= HX0 * HY0 && TX * HY1 <= HX1 * HY0, "S0 / S1 hold the t-sweep stages");
  constexpr int NB = FVHOE_TILE_PIPE ? 2 : 1;  // buffers per sweep stage
  __shared__ T S0[NB][HX0 * HY0];
  __shared__ T S1[NB][HX1 * HY0];

  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int oa = (tile / tiles1) * TX, ob = (tile % tiles1) * TY;
  const int cx0 = oa - g.ge0 - (RCX + RTX), cy0 = ob - g.ge1 - (RCY + RTY);
  const int tid = threadIdx.x;
  const bool inside = cx0 >= 0 && cx0 + HX0 <= g.n0 && cy0 >= 0 && cy0 + HY0 <= g.n1;
  // the last tiles overhang W; their extra rows / columns are never stored, and their fetches are
  // clamped to the ghost box W needs so every source index stays in its array (the box's low
  // corner is never crossed: cx0 and cy0 start there for the first tiles)
  const int xhi = g.n0 + g.ge0 + RCX + RTX - 1, yhi = g.n1 + g.ge1 + RCY + RTY - 1;

  // all five variables' tile elements are fetched up front (one memory latency per block)
  constexpr int NE = (HX0 * HY0 + NT - 1) / NT;
  T pre[5][NE];
  {
    constexpr int DR = NT / HY0, DC = NT % HY0;
    int r = tid / HY0, c = tid % HY0;
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      if (tid + k * NT < HX0 * HY0) {
#pragma unroll
        for (int v = 0; v < 5; ++v)
          pre[v][k] = T(inside ? src.at(g, v, cx0 + r, cy0 + c)
                               : src.get(g, v, min(cx0 + r, xhi), min(cy0 + c, yhi)));
      }
      r += DR;
      c += DC;
      if (c >= HY0) {
        c -= HY0;
        ++r;
      }
    }
  }
  // centre values of the HX1 x HY1 box stay in registers: thread tid holds elements tid + k * NT
  T cen[5][NC];
  T wr[5][NPT];
  auto put_pre = [&](int v, int b) {
#pragma unroll
    for (int k = 0; k < NE; ++k)
      if (tid + k * NT < HX0 * HY0) S0[b][tid + k * NT] = pre[v][k];
  };
  auto put_cen = [&](int v, int b) {  // row stride HY1 for the t-sweeps
#pragma unroll
    for (int k = 0; k < NC; ++k)
      if (tid + k * NT < HX1 * HY1) S0[b][tid + k * NT] = cen[v][k];
  };
  auto csweep_x = [&](int b) {
    tile_for<HX1, HY0, NT>([&](int e, int i, int j) {
      S1[b][e] = sweepT<KIND_C, PX, T>([&](int t) { return S0[b][(i + RCX + t - reach(PX)) * HY0 + j]; });
    });
  };
  auto csweep_y = [&](int v, int b) {
    tile_for_k<HX1, HY1, NT>([&](int k, int e, int i, int j) {
      const T x = sweepT<KIND_C, PY, T>([&](int t) { return S1[b][i * HY0 + j + RCY + t - reach(PY)]; });
      cen[v][k] = T(0) + x;  // degree-0 z sweep (fvhoe/fv.py:186-196)
    });
  };
  auto tsweep_x = [&](int b) {
    tile_for<TX, HY1, NT>([&](int e, int i, int j) {
      S1[b][e] = sweepT<KIND_T, PX, T>([&](int t) { return S0[b][(i + RTX + t - reach(PX)) * HY1 + j]; });
    });
  };
  auto tsweep_y = [&](int v, int b) {
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
      const int e = tid + q * NT;
      if (e < TX * TY) {
        const int i = e / TY, j = e % TY;
        const T x = sweepT<KIND_T, PY, T>([&](int t) { return S1[b][i * HY1 + j + RTY + t - reach(PY)]; });
        wr[v][q] = T(0) + x;  // z-axis sweep of degree 0 (fvhoe/fv.py:221-223)
        const int a = oa + i, bb = ob + j;
        if (a < NW0 && bb < NW1) W[(v * NW0 + a) * NW1 + bb] = wr[v][q];
      }
    }
  };
  auto recover = [&]() {  // pressure recovery in FP64 (fvhoe/hydro.py:146-170)
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      if (tid + k * NT < HX1 * HY1) {
        double uu[5], ww[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) uu[q] = (double)cen[q][k];
        prims_from_cons<FVHOE_F32_FAST && std::is_same<T, float>::value>(uu, ww, gamma);
#pragma unroll
        for (int q = 0; q < 5; ++q) cen[q][k] = T(ww[q]);
      }
    }
  };
  // the check's W box: tile cell (ti, tj) is interior cell (oa + ti - ge0, ob + tj - ge1) of the candidate
  // grid; box element (r, c) is W's (oa - 1 - ge0 + wg0 + r, ob - 1 - ge1 + wg1 + c), clamped into W (the
  // clamped elements belong to no interior cell).  Fetched after the c-sweeps, into the registers the
  // tile's own fetches held, so the loads overlap the t-sweeps.
  // FP32 prefetches all five variables' boxes; FP64 only the first (all five would hold 30 more VGPRs
  // through the t-sweeps and cost occupancy), the others are fetched one variable ahead in the check
  constexpr int BX = TX + 2, BY = TY + 2, NBE = (BX * BY + NT - 1) / NT;
  constexpr int NPF = sizeof(T) == 8 ? 1 : 5;
  T wbx[CHECK ? NPF : 1][CHECK ? NBE : 1];
  auto fetch_box = [&](int v, T (&dst)[CHECK ? NBE : 1]) {
    if constexpr (CHECK) {
      const int r0 = oa - 1 - g.ge0 + ck.wg0, c0 = ob - 1 - g.ge1 + ck.wg1;
      tile_for_k<BX, BY, NT>([&](int k, int e, int r, int c) {
        const int wa = min(max(r0 + r, 0), ck.NW0 - 1), wb = min(max(c0 + c, 0), ck.NW1 - 1);
        dst[k] = ck.W[(v * ck.NW0 + wa) * ck.NW1 + wb];
      });
    }
  };
  auto prefetch_box = [&]() {
    if constexpr (CHECK) {
#pragma unroll
      for (int v = 0; v < NPF; ++v) fetch_box(v, wbx[v]);
    }
  };
#if FVHOE_TILE_PIPE
  // Two buffers per sweep stage: variable v's second sweep and variable v+1's first sweep share a
  // barrier interval, one barrier per variable.  Each step writes a buffer whose last reader ran
  // before the previous barrier (buffer v & 1 holds variable v).
  put_pre(0, 0);
  __syncthreads();
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    csweep_x(v & 1);
    if (v + 1 < 5) put_pre(v + 1, (v + 1) & 1);
    __syncthreads();
    csweep_y(v, v & 1);
  }
  prefetch_box();
  recover();
  put_cen(0, 0);
  __syncthreads();
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    tsweep_x(v & 1);
    if (v + 1 < 5) put_cen(v + 1, (v + 1) & 1);
    __syncthreads();
    tsweep_y(v, v & 1);
  }
#else
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    put_pre(v, 0);
    __syncthreads();
    csweep_x(0);
    __syncthreads();
    csweep_y(v, 0);
    __syncthreads();
  }
  prefetch_box();
  recover();
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    put_cen(v, 0);
    __syncthreads();
    tsweep_x(0);
    __syncthreads();
    tsweep_y(v, 0);
    // no barrier here: the next variable's S0 stores touch no S1 element, and its barrier comes
    // before S1 is written again
  }
#endif
  if constexpr (CHECK) {
    // two W box buffers: S0's two buffers where they are large enough (free after the last t-sweep
    // barrier), else LDS of their own
    constexpr bool REUSE = NB == 2 && BX * BY <= HX0 * HY0;
    __shared__ T SBx[REUSE ? 1 : 2][REUSE ? 1 : BX * BY];
    T* const SB0 = REUSE ? &S0[0][0] : &SBx[0][0];
    T* const SB1 = REUSE ? &S0[NB - 1][0] : &SBx[REUSE ? 0 : 1][0];
    auto put_box = [&](const T (&src)[NBE], T* b) {
      tile_for_k<BX, BY, NT>([&](int k, int e, int r, int c) { b[e] = src[k]; });
    };
    const DetectP& a = ck.dp;
    unsigned nadb[NPT];
    double imin[NPT];
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
      nadb[q] = 0u;
      imin[q] = 0.0;
    }
    put_box(wbx[0], SB0);
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      T nxt[NBE];  // the next variable's box, fetched before this variable's tests
      if (v + 1 < 5 && v + 1 >= NPF) fetch_box(v + 1, nxt);
      const T* b = (v & 1) ? SB1 : SB0;
      const double rlo = ck.range[2 * v], rhi = ck.range[2 * v + 1];
#pragma unroll
      for (int q = 0; q < NPT; ++q) {
        const int e = tid + q * NT;
        if (e < TX * TY) {
          const int ti = e / TY, tj = e % TY;
          // max / min of the 3 x 3 box in the storage type, then widened: widening is exact and monotonic,
          // so this equals widening first (FP32: two conversions per cell instead of nine)
          T hM[3], hm[3];
#pragma unroll
          for (int r = 0; r < 3; ++r) {  // rows -1, 0, 1 of the cell, each over columns -1, 0, 1
            const T* row = b + (ti + r) * BY + tj;
            const T x0 = row[0], x1 = row[1], x2 = row[2];
            hM[r] = dmax(dmax(x0, x1), x2);
            hm[r] = dmin(dmin(x0, x1), x2);
          }
          const double M = (double)dmax(dmax(hM[0], hM[1]), hM[2]);
          const double m = (double)dmin(dmin(hm[0], hm[1]), hm[2]);
          const double uc = (double)wr[v][q];
          bool tv;
          double ind;
          if (a.nad_range == 0) {
            const double rng = a.nad_mode == 0 ? rhi - rlo : M - m;
            ind = dmin(uc - m, M - uc);
            tv = ind < -a.eps * rng;
          } else {
            const double lo = a.nad_mode == 0 ? (1.0 - a.eps) * rlo : (1.0 - a.eps) * m;
            const double hi = a.nad_mode == 0 ? (1.0 + a.eps) * rhi : (1.0 + a.eps) * M;
            ind = dmin(uc - lo, hi - uc);
            tv = ind < 0.0;
          }
          if (tv && ((a.nad_vars >> v) & 1)) nadb[q] |= 1u << v;
          imin[q] = v == 0 ? ind : dmin(imin[q], ind);
        }
      }
      if (v + 1 < 5) put_box(v + 1 < NPF ? wbx[v + 1 < NPF ? v + 1 : 0] : nxt, ((v + 1) & 1) ? SB1 : SB0);
      __syncthreads();
    }
    int cell[NPT];
    unsigned lsel = 0u, psel = 0u, msel = 0u, fpbits = 0u;
    double mnad[NPT], mpad[NPT];
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
      const int e = tid + q * NT;
      const int ti = e / TY, tj = e % TY;
      const int i = oa + ti - g.ge0, j = ob + tj - g.ge1;
      cell[q] = i * g.n1 + j;
      mnad[q] = 0.0;
      mpad[q] = 0.0;
      if (e >= TX * TY || i < 0 || i >= g.n0 || j < 0 || j >= g.n1) continue;
      if (a.sed && nadb[q]) {  // the SED decides: k_detect_list
        psel |= 1u << q;
        continue;
      }
      const bool nad = nadb[q] != 0u;
      bool pad = false;
      double pind = 0.0;
      if (a.pad_var >= 0) {
        double x = 0.0;
#pragma unroll
        for (int v = 0; v < 5; ++v)
          if (v == a.pad_var) x = (double)wr[v][q];
        pind = dmin(x - a.pad_lo, a.pad_hi - x);
        pad = pind < 0.0;
      }
      if (nad || pad) msel |= 1u << q;
      mnad[q] = nad ? -imin[q] : 0.0;
      mpad[q] = pad ? -pind : 0.0;
      int t = (a.force || pad || nad) ? 1 : 0;
      if (ck.pre) {  // pre-flagged cells already carry MUSCL fluxes: list only the others
        const int pf = ck.pre[cell[q]];
        ck.trouble[cell[q]] = (signed char)(t | pf);
        if (pf && !t) fpbits |= 1u << q;
        t = t && !pf;
      } else {
        ck.trouble[cell[q]] = (signed char)t;
      }
      if (t) lsel |= 1u << q;
    }
    if (ck.pre) {
      int f = __popc(fpbits);
      for (int off = 32; off > 0; off >>= 1) f += __shfl_down(f, off, 64);
      if ((tid & 63) == 0 && f) atomicAdd(ck.fp_count, f);
    }
    if (ck.pack) {  // the pending cells carry their NAD bits and indicator to k_detect_list
      int ent[NPT];
#pragma unroll
      for (int q = 0; q < NPT; ++q) ent[q] = (int)((unsigned)cell[q] | (nadb[q] << PEND_SHIFT));
      block_append2<NPT, NT>(cell, lsel, psel, ck.count, ck.list, ck.pend_count, ck.pend, ent, imin, ck.pend_imin);
    } else {
      block_append2<NPT, NT>(cell, lsel, psel, ck.count, ck.list, ck.pend_count, ck.pend);
    }
    if (mag_on(ck.ml)) {
#pragma unroll
      for (int q = 0; q < NPT; ++q) mag_append(ck.ml, (msel >> q) & 1u, cell[q], mnad[q], mpad[q]);
    }
    return;
  }
  if (!red.partials) return;
  // partial reductions over this tile's cells of the dt box and the range box
  double m = 0.0;
  unsigned long long val[NPART];
#pragma unroll
  for (int k = 1; k < NPART; ++k) val[k] = (k & 1) ? ~0ull : 0ull;
#pragma unroll
  for (int q = 0; q < NPT; ++q) {
    const int e = tid + q * NT;
    if (e < TX * TY) {
      const int a = oa + e / TY, b = ob + e % TY;
      if (red.do_dt && a >= red.dt_lo0 && a < red.dt_hi0 && b >= red.dt_lo1 && b < red.dt_hi1) {
        const double c = sound_speed((double)wr[0][q], (double)wr[4][q], red.gamma, red.csq_floor);
        const double s = (fabs((double)wr[1][q]) + fabs((double)wr[2][q])) + fabs((double)wr[3][q]);
        m = np_max(m, s + (double)red.ndim * c);  // a NaN wave speed gives a NaN dt, as np.max does
      }
      if (red.do_range && a >= red.rg_lo0 && a < red.rg_hi0 && b >= red.rg_lo1 && b < red.rg_hi1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          const unsigned long long o = ordered_bits((double)wr[v][q]);
          val[1 + 2 * v] = red_op(1, val[1 + 2 * v], o);
          val[2 + 2 * v] = red_op(2, val[2 + 2 * v], o);
        }
      }
    }
  }
  val[0] = (unsigned long long)__double_as_longlong(m);
  __shared__ unsigned long long part[NT / 64][NPART];
#pragma unroll
  for (int k = 0; k < NPART; ++k) {
    unsigned long long x = val[k];
    for (int off = 32; off > 0; off >>= 1) x = red_op(k, x, __shfl_down(x, off, 64));
    if ((tid & 63) == 0) part[tid >> 6][k] = x;
  }
  __syncthreads();
  if (tid < NPART) {
    unsigned long long x = part[0][tid];
    for (int w = 1; w < NT / 64; ++w) x = red_op(tid, x, part[w][tid]);
    red.partials[(long long)tid * gridDim.x + blockIdx.x] = x;  // [quantity][block]: coalesced for the reduction
  }
}

// ---------------------------------------------------------------------------
// 2-D prims over column strips: lane l of a wave owns W column b and R
// consecutive W rows.  The x sweeps run down the strip in registers; the y
// sweeps take their neighbours from adjacent lanes with DPP wavefront shifts,
// so the kernel needs no LDS and no barrier outside the final reduction.  Each
// wave yields 64 - 2H valid columns (H = c and t reaches).  Same operations in
// the same order as k_prims, so W is identical.
// ---------------------------------------------------------------------------
// stencil sweep whose taps are compile-time offsets (for lane_at)
template <int KIND, int P, typename T, typename Get>
__device__ __forceinline__ T sweepL(Get&& x) {
  T acc = T(0);
  static_for<ntaps(P)>([&](auto I) {
    constexpr double w = weight(KIND, P, decltype(I)::value);
    if constexpr (w != 0.0)
      acc = q_fma<FVHOE_F32_FMA>(T(w), x(std::integral_constant<int, decltype(I)::value - reach(P)>{}), acc);
  });
  return acc;
}

__host__ __device__ constexpr int strip_halo(int p) { return ereach_c(p) + ereach_t(p); }
// W rows per lane in k_prims_strip (register budget: 5 variables x (rows + 2) centres per lane)
template <typename T>
constexpr int strip_rows() { return sizeof(T) == 8 ? FVHOE_STRIP_ROWS_F64 : FVHOE_STRIP_ROWS_F32; }
// degrees that use k_prims_strip per arithmetic type (others: the LDS-tiled k_prims).  FP64 uses the
// tiles throughout: 32 x 32 tiles measured 133 us per launch at KH 2048^2 p = 3 against 152 for the
// strips; FP32 strips measured 98 against 118 for 16 x 64 tiles
#ifndef FVHOE_STRIP_MAX_P_F64
#define FVHOE_STRIP_MAX_P_F64 0
#endif
#ifndef FVHOE_STRIP_MAX_P_F32
#define FVHOE_STRIP_MAX_P_F32 3
#endif
template <typename T>
constexpr bool use_strip(int p) {
  return p <= (sizeof(T) == 8 ? FVHOE_STRIP_MAX_P_F64 : FVHOE_STRIP_MAX_P_F32);
}

// minimum waves per SIMD of k_prims_strip (1: the compiler's choice; more caps its registers).  FP32: 4
// (107 VGPRs, no spills) runs the KH 2048^2 strip in 98 us against 111 at 3 waves (135 VGPRs); FP64 at 3
// waves spills (168 VGPRs, 172 B of scratch) and runs 192 us against 151 at 2 waves (211 VGPRs)
#ifndef FVHOE_STRIP_WAVES_F64
#define FVHOE_STRIP_WAVES_F64 1
#endif
#ifndef FVHOE_STRIP_WAVES_F32
#define FVHOE_STRIP_WAVES_F32 4
#endif
template <typename T, int P, typename Src>
__global__ void __launch_bounds__(256, sizeof(T) == 8 ? FVHOE_STRIP_WAVES_F64 : FVHOE_STRIP_WAVES_F32)
    k_prims_strip(Src src, Geo g, T* __restrict__ W, int NW0, int NW1,
                                                       int cwins, Reduce red, double gamma) {
  constexpr int RC = ereach_c(P), RT = ereach_t(P), H = RC + RT, VC = 64 - 2 * H;
  constexpr int R = strip_rows<T>(), NU = R + 2 * H, NC = R + 2 * RT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int win = blockIdx.x % cwins, rb = blockIdx.x / cwins;
  const int b = win * VC - H + lane;   // W column of this lane
  const int a0 = (rb * 4 + wave) * R;  // first W row of this wave
  const int cy = b - g.ge1, cx0 = a0 - g.ge0;
  const int cyl = min(max(cy, -(g.ge1 + H)), g.n1 + g.ge1 + H - 1);
  const bool inside = cx0 - H >= 0 && cx0 + R - 1 + H < g.n0 && win * VC - H - g.ge1 >= 0 &&
                      win * VC - H + 63 - g.ge1 < g.n1;
  T C[5][NC];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    T uc[NU];
#pragma unroll
    for (int r = 0; r < NU; ++r) {
      const int cx = min(max(cx0 - H + r, -(g.ge0 + H)), g.n0 + g.ge0 + H - 1);
      uc[r] = T(inside ? src.at(g, v, cx, cyl) : src.get(g, v, cx, cyl));
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const T s1 = sweepT<KIND_C, P, T>([&](int t) { return uc[k + RC + t - reach(P)]; });
      const T x = sweepL<KIND_C, P, T>([&](auto o) { return lane_at<decltype(o)::value>(s1); });
      C[v][k] = T(0) + x;  // degree-0 z sweep (fvhoe/fv.py:186-196)
    }
  }
  // pressure recovery in FP64 (fvhoe/hydro.py:146-170)
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    double uu[5], ww[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) uu[q] = (double)C[q][k];
    prims_from_cons<FVHOE_F32_FAST && std::is_same<T, float>::value>(uu, ww, gamma);
#pragma unroll
    for (int q = 0; q < 5; ++q) C[q][k] = T(ww[q]);
  }
  const bool lane_ok = lane >= H && lane < 64 - H && b < NW1;
  double m = 0.0;
  unsigned long long val[NPART];
  val[0] = 0ull;
#pragma unroll
  for (int k = 1; k < NPART; ++k) val[k] = (k & 1) ? ~0ull : 0ull;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    T w[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const T tx = sweepT<KIND_T, P, T>([&](int t) { return C[v][i + RT + t - reach(P)]; });
      const T y = sweepL<KIND_T, P, T>([&](auto o) { return lane_at<decltype(o)::value>(tx); });
      w[v] = T(0) + y;  // z-axis sweep of degree 0 (fvhoe/fv.py:221-223)
    }
    const int a = a0 + i;
    if (lane_ok && a < NW0) {
#pragma unroll
      for (int v = 0; v < 5; ++v) W[(v * NW0 + a) * NW1 + b] = w[v];
      if (red.partials) {
        if (red.do_dt && a >= red.dt_lo0 && a < red.dt_hi0 && b >= red.dt_lo1 && b < red.dt_hi1) {
          const double c = sound_speed((double)w[0], (double)w[4], red.gamma, red.csq_floor);
          const double sp = (fabs((double)w[1]) + fabs((double)w[2])) + fabs((double)w[3]);
          m = np_max(m, sp + (double)red.ndim * c);  // a NaN wave speed gives a NaN dt, as np.max does
        }
        if (red.do_range && a >= red.rg_lo0 && a < red.rg_hi0 && b >= red.rg_lo1 && b < red.rg_hi1) {
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            const unsigned long long o = ordered_bits((double)w[v]);
            val[1 + 2 * v] = red_op(1, val[1 + 2 * v], o);
            val[2 + 2 * v] = red_op(2, val[2 + 2 * v], o);
          }
        }
      }
    }
  }
  if (!red.partials) return;
  val[0] = (unsigned long long)__double_as_longlong(m);
  __shared__ unsigned long long part[4][NPART];
#pragma unroll
  for (int k = 0; k < NPART; ++k) {
    unsigned long long x = val[k];
    for (int off = 32; off > 0; off >>= 1) x = red_op(k, x, __shfl_down(x, off, 64));
    if (lane == 0) part[wave][k] = x;
  }
  __syncthreads();
  if (threadIdx.x < NPART) {
    unsigned long long x = part[0][threadIdx.x];
    for (int w = 1; w < 4; ++w) x = red_op(threadIdx.x, x, part[w][threadIdx.x]);
    red.partials[(long long)threadIdx.x * gridDim.x + blockIdx.x] = x;
  }
}

inline int strip_cwins(int p, int NW1) { return (NW1 + (64 - 2 * strip_halo(p)) - 1) / (64 - 2 * strip_halo(p)); }
inline int strip_blocks(int p, int rows, int NW0, int NW1) {
  return strip_cwins(p, NW1) * ((NW0 + 4 * rows - 1) / (4 * rows));
}

// ---------------------------------------------------------------------------
// flux tile.  D = 0: x faces fa in [f0, f0+TA) (face fa = left face of cell
// fa), rows b in [b0, b0+TB).  D = 1: y faces, rows along x.  "normal" is the
// face-normal axis with degree PN, "trans" the other with degree PT.
// ---------------------------------------------------------------------------
template <typename T, int D, int PN, int PT, int TA, int TB, int NT, int RS>
__global__ void __launch_bounds__(NT) k_flux(const T* __restrict__ W, int NW0, int NW1, Geo g, T* __restrict__ Fout,
                                            int tilesB, double gamma_d, double csq_floor_d) {
  constexpr int EL = reach(PN);
  constexpr int RC = ereach_c(PT), RT = ereach_t(PT);
  constexpr int HN = TA + 1 + 2 * EL;
  constexpr int HT0 = TB + 2 * (RC + RT), HT1 = TB + 2 * RT;
  constexpr int HS = HT0 | 1;  // S0 row stride, odd: the transposed stores of D = 1 hit distinct banks
  constexpr int NR = (TA * HT1 + NT - 1) / NT;   // Riemann problems per thread
  static_assert(TA * HT1 <= HN * HT1, "flux results reuse the X buffer");
  constexpr int NB = (sizeof(T) == 8 ? FVHOE_FLUX_PIPE_F64 : FVHOE_FLUX_PIPE_F32) ? 2 : 1;  // S0 buffers
  __shared__ T X[5][HN * HT1];
  __shared__ T S0[NB][HN * HS];
  const T gamma = T(gamma_d), csq_floor = T(csq_floor_d);

  const int nN = D == 0 ? g.n0 : g.n1, nT = D == 0 ? g.n1 : g.n0;
  const int geN = D == 0 ? g.ge0 : g.ge1, geT = D == 0 ? g.ge1 : g.ge0;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int f0 = (tile / tilesB) * TA, b0 = (tile % tilesB) * TB;
  const int tid = threadIdx.x;
  // W index of normal cell (f0 - 1 - EL) and transverse cell (b0 - RC - RT); both >= 0
  const int wn0 = f0 - 1 - EL + geN, wt0 = b0 - RC - RT + geT;

  // all five variables' tile elements are fetched up front (one memory latency per block);
  // for D = 1 the contiguous W axis (y = normal) runs fastest over the threads
  constexpr int NE = (HN * HT0 + NT - 1) / NT;
  constexpr int WFAST = D == 0 ? HT0 : HN;
  T pre[5][NE];
  int dst[NE];
  {
    constexpr int DR = NT / WFAST, DC = NT % WFAST;
    int r = tid / WFAST, c = tid % WFAST;  // (slow, fast) index pair
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      if (tid + k * NT < HN * HT0) {
        // clamped reads feed unused outputs only
        const int a = D == 0 ? min(wn0 + r, NW0 - 1) : min(wt0 + r, NW0 - 1);
        const int b = D == 0 ? min(wt0 + c, NW1 - 1) : min(wn0 + c, NW1 - 1);
        dst[k] = D == 0 ? r * HS + c : c * HS + r;
#pragma unroll
        for (int v = 0; v < 5; ++v) pre[v][k] = W[(v * NW0 + a) * NW1 + b];
      }
      r += DR;
      c += DC;
      if (c >= WFAST) {
        c -= WFAST;
        ++r;
      }
    }
  }
  auto put = [&](int v, int bf) {
#pragma unroll
    for (int k = 0; k < NE; ++k)
      if (tid + k * NT < HN * HT0) S0[bf][dst[k]] = pre[v][k];
  };
  auto csweep = [&](int v, int bf) {
    tile_for<HN, HT1, NT>([&](int e, int r, int c) {
      const T x = sweepT<KIND_C, PT, T>([&](int t) { return S0[bf][r * HS + c + RC + t - reach(PT)]; });
      X[v][e] = T(0) + x;  // degree-0 z sweep precedes/follows (fvhoe/solver.py:524,541)
    });
  };
  if constexpr (NB == 2) {
    // variable v + 1 is staged in the other buffer while v is swept: one barrier per variable (the
    // buffer each store writes was last read before the previous barrier)
    put(0, 0);
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      csweep(v, v & 1);
      if (v + 1 < 5) put(v + 1, (v + 1) & 1);
      __syncthreads();
    }
  } else {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      put(v, 0);
      __syncthreads();
      csweep(v, 0);
      __syncthreads();
    }
  }
  // Riemann problems at faces f0+fr, transverse rows c in [0, HT1)
  T res[NR][5];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int e = tid + k * NT;
    if (e < TA * HT1) {
      const int fr = e / HT1, c = e % HT1;
      T wl[5], wr[5];
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        // right face of cell f0+fr-1: centre row fr+EL; left face of cell f0+fr: centre row fr+EL+1
        wl[q] = sweepT<KIND_R, PN, T>([&](int t) { return X[q][(fr + t) * HT1 + c]; });
        wr[q] = sweepT<KIND_L, PN, T>([&](int t) { return X[q][(fr + 1 + t) * HT1 + c]; });
      }
      riemann_t<RS, D, FVHOE_F32_FAST && std::is_same<T, float>::value>(wl, wr, res[k], gamma, csq_floor);  // the solver is a compile-time choice
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int e = tid + k * NT;
    if (e < TA * HT1) {
#pragma unroll
      for (int q = 0; q < 5; ++q) X[q][e] = res[k][q];  // FC[q][fr][c] in X's space
    }
  }
  __syncthreads();
  const int nF0 = D == 0 ? g.n0 + 1 : g.n0, nF1 = D == 0 ? g.n1 : g.n1 + 1;
  for (int q = 0; q < 5; ++q) {
    T* Fq = Fout + q * nF0 * nF1;
    if (D == 0) {
      tile_for<TA, TB, NT>([&](int e, int fr, int c) {
        const int fa = f0 + fr, b = b0 + c;
        if (fa > nN || b >= nT) return;
        const T x = sweepT<KIND_T, PT, T>([&](int t) { return X[q][fr * HT1 + c + RT + t - reach(PT)]; });
        Fq[fa * nF1 + b] = T(0) + x;  // degree-0 z sweep (fvhoe/solver.py:633)
      });
    } else {
      tile_for<TB, TA, NT>([&](int e, int c, int fr) {
        const int fa = f0 + fr, b = b0 + c;
        if (fa > nN || b >= nT) return;
        const T x = sweepT<KIND_T, PT, T>([&](int t) { return X[q][fr * HT1 + c + RT + t - reach(PT)]; });
        Fq[b * nF1 + fa] = T(0) + x;  // (fvhoe/solver.py:639)
      });
    }
  }
}

// ---------------------------------------------------------------------------
// x and y faces of one tile from one W box (2-D, both axes of degree P): the
// box W[i0 - LO, i0 + A + HI) x [j0 - LO, j0 + B + HI) is fetched once into
// registers and staged twice, once per face axis, through k_flux's steps (the
// same operations per face, so F and G are identical to k_flux's).  x faces
// fa in [i0, i0 + A) for rows [j0, j0 + B); y faces fb in [j0, j0 + B) for
// rows [i0, i0 + A); the tiles cover (n0 + 1) x (n1 + 1) face positions.
// ---------------------------------------------------------------------------
template <typename T, int P, int A, int B, int NT, int RS>
__global__ void __launch_bounds__(NT, sizeof(T) == 8 ? 1 : FVHOE_FLUXXY_WAVES_F32) k_flux_xy(const T* __restrict__ W, int NW0, int NW1, Geo g, T* __restrict__ F,
                                               T* __restrict__ Gout, int tilesB, double gamma_d, double csq_floor_d) {
  constexpr int EL = reach(P), RC = ereach_c(P), RT = ereach_t(P);
  constexpr int LO = 1 + EL > RC + RT ? 1 + EL : RC + RT;  // box cells before the tile on each axis
  constexpr int HI = EL > RC + RT ? EL : RC + RT;          // and after it
  constexpr int BX = A + LO + HI, BY = B + LO + HI;
  constexpr int NE = (BX * BY + NT - 1) / NT;
  // per face axis D: TA faces x TB rows, HN normal cells, HT0 / HT1 transverse cells before / after the c-sweep
  constexpr int HN_0 = A + 1 + 2 * EL, HT1_0 = B + 2 * RT, HS_0 = (B + 2 * (RC + RT)) | 1;
  constexpr int HN_1 = B + 1 + 2 * EL, HT1_1 = A + 2 * RT, HS_1 = (A + 2 * (RC + RT)) | 1;
  constexpr int XSZ = HN_0 * HT1_0 > HN_1 * HT1_1 ? HN_0 * HT1_0 : HN_1 * HT1_1;
  constexpr int SSZ = HN_0 * HS_0 > HN_1 * HS_1 ? HN_0 * HS_0 : HN_1 * HS_1;
  __shared__ T X[5][XSZ];
  __shared__ T S0[SSZ];
  const T gamma = T(gamma_d), csq_floor = T(csq_floor_d);
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int i0 = (tile / tilesB) * A, j0 = (tile % tilesB) * B;
  const int tid = threadIdx.x;
  const int wr0 = i0 - LO + g.ge0, wc0 = j0 - LO + g.ge1;  // W index of box element (0, 0); >= 0 as ge >= LO

  // the box, y (W's contiguous axis) fastest over the threads; clamped reads feed unused outputs only.
  // RELOAD: fetched again for the second axis (the block's own lines, hot in the caches) instead of held
  // in registers through the first axis
  constexpr bool RELOAD = sizeof(T) == 8 ? FVHOE_FXY_RELOAD_F64 : FVHOE_FXY_RELOAD_F32;
  T pre[5][NE];
  auto fetch = [&]() {
    tile_for_k<BX, BY, NT>([&](int k, int e, int r, int c) {
      const int a = min(wr0 + r, NW0 - 1), b = min(wc0 + c, NW1 - 1);
#pragma unroll
      for (int v = 0; v < 5; ++v) pre[v][k] = W[(v * NW0 + a) * NW1 + b];
    });
  };
  fetch();

  auto faces = [&](auto dc) {
    constexpr int D = decltype(dc)::value;
    constexpr int TA = D == 0 ? A : B, TB = D == 0 ? B : A;
    constexpr int HN = TA + 1 + 2 * EL, HT0 = TB + 2 * (RC + RT), HT1 = TB + 2 * RT, HS = HT0 | 1;
    constexpr int NR = (TA * HT1 + NT - 1) / NT;  // Riemann problems per thread
    // S0[nrm * HS + trn]: normal cell f0 - 1 - EL + nrm, transverse cell b0 - RC - RT + trn, i.e. box
    // element (nrm + ON, trn + OT) along (normal, transverse)
    constexpr int ON = LO - 1 - EL, OT = LO - (RC + RT);
    const int f0 = D == 0 ? i0 : j0, b0 = D == 0 ? j0 : i0;
    const int nN = D == 0 ? g.n0 : g.n1, nT = D == 0 ? g.n1 : g.n0;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      tile_for_k<BX, BY, NT>([&](int k, int e, int r, int c) {
        const int nrm = (D == 0 ? r : c) - ON, trn = (D == 0 ? c : r) - OT;
        if (nrm >= 0 && nrm < HN && trn >= 0 && trn < HT0) S0[nrm * HS + trn] = pre[v][k];
      });
      __syncthreads();
      tile_for<HN, HT1, NT>([&](int e, int r, int c) {
        const T x = sweepT<KIND_C, P, T>([&](int t) { return S0[r * HS + c + RC + t - reach(P)]; });
        X[v][e] = T(0) + x;  // degree-0 z sweep precedes/follows (fvhoe/solver.py:524,541)
      });
      __syncthreads();
    }
    T res[NR][5];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int e = tid + k * NT;
      if (e < TA * HT1) {
        const int fr = e / HT1, c = e % HT1;
        T wl[5], wr[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          wl[q] = sweepT<KIND_R, P, T>([&](int t) { return X[q][(fr + t) * HT1 + c]; });
          wr[q] = sweepT<KIND_L, P, T>([&](int t) { return X[q][(fr + 1 + t) * HT1 + c]; });
        }
        riemann_t<RS, D, FVHOE_F32_FAST && std::is_same<T, float>::value>(wl, wr, res[k], gamma, csq_floor);
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int e = tid + k * NT;
      if (e < TA * HT1) {
#pragma unroll
        for (int q = 0; q < 5; ++q) X[q][e] = res[k][q];
      }
    }
    __syncthreads();
    const int nF0 = D == 0 ? g.n0 + 1 : g.n0, nF1 = D == 0 ? g.n1 : g.n1 + 1;
    T* const out = D == 0 ? F : Gout;
    for (int q = 0; q < 5; ++q) {
      T* Fq = out + q * nF0 * nF1;
      if (D == 0) {
        tile_for<TA, TB, NT>([&](int e, int fr, int c) {
          const int fa = f0 + fr, b = b0 + c;
          if (fa > nN || b >= nT) return;
          const T x = sweepT<KIND_T, P, T>([&](int t) { return X[q][fr * HT1 + c + RT + t - reach(P)]; });
          Fq[fa * nF1 + b] = T(0) + x;  // degree-0 z sweep (fvhoe/solver.py:633)
        });
      } else {
        tile_for<TB, TA, NT>([&](int e, int c, int fr) {
          const int fa = f0 + fr, b = b0 + c;
          if (fa > nN || b >= nT) return;
          const T x = sweepT<KIND_T, P, T>([&](int t) { return X[q][fr * HT1 + c + RT + t - reach(P)]; });
          Fq[b * nF1 + fa] = T(0) + x;  // (fvhoe/solver.py:639)
        });
      }
    }
    __syncthreads();  // S0 and X are restaged by the next face axis
  };
  faces(std::integral_constant<int, 0>{});
  if (RELOAD) fetch();
  faces(std::integral_constant<int, 1>{});
}

// ---------------------------------------------------------------------------
// troubled-cell flags from W (3^d max/min) and WS (candidate, SED), in FP64
// ---------------------------------------------------------------------------
// The tests of one interior cell c: t = flag before pre-flags, mflag / mnad / mpad for the trouble log.
// detect_nad: the NAD test of each variable (bit v of the result: variable v flags and is one of the
// tested variables) and the min over the variables of the NAD indicator; detect_sed_pad: the smooth-
// extrema detector over the flagged variables and the PAD test.
template <typename T>
__device__ __forceinline__ unsigned detect_nad(const T* __restrict__ W, int NW1, int NWC, const T* __restrict__ WS,
                                               int NS1, int NSC, const Geo& g, bool two_d,
                                               const double* __restrict__ range_mm, const DetectP& a, int c,
                                               double* imin_out) {
  const int i = c / g.n1, j = c - (c / g.n1) * g.n1;
  const int wb = (i + g.ge0) * NW1 + (j + g.ge1);
  const int sb = (i + 3) * NS1 + (j + (two_d ? 3 : 0));
  unsigned tvv = 0u;
  double imin = 0.0;  // min over the variables of the NAD indicator
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const T* wv = W + v * NWC;
    // max / min in the storage type, widened once (widening is exact and monotonic)
    T Mt = wv[wb], mt = Mt;
#pragma unroll
    for (int di = -1; di <= 1; ++di) {
      if (two_d) {
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj) {
          const T x = wv[wb + di * NW1 + dj];
          Mt = dmax(Mt, x);
          mt = dmin(mt, x);
        }
      } else {
        const T x = wv[wb + di * NW1];
        Mt = dmax(Mt, x);
        mt = dmin(mt, x);
      }
    }
    const double M = (double)Mt, m = (double)mt;
    const double uc = (double)WS[v * NSC + sb];
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
    if (tv && ((a.nad_vars >> v) & 1)) tvv |= 1u << v;
    imin = v == 0 ? ind : dmin(imin, ind);
  }
  *imin_out = imin;
  return tvv;
}

template <typename T>
__device__ __forceinline__ int detect_sed_pad(const T* __restrict__ WS, int NS1, int NSC, const Geo& g, bool two_d,
                                              const DetectP& a, int c, unsigned tvv, double imin, bool* mflag,
                                              double* mnad, double* mpad) {
  const int i = c / g.n1, j = c - (c / g.n1) * g.n1;
  const int sb = (i + 3) * NS1 + (j + (two_d ? 3 : 0));
  bool nad = false;
  for (int v = 0; v < 5; ++v) {
    bool tv = (tvv >> v) & 1u;
    if (tv && a.sed && !nad) {  // smooth-extrema detector only where NAD flags (it can only clear flags)
      const T* sv = WS + v * NSC;
      double alpha = sed_alpha_1d([&](int o) { return (double)sv[sb + o * NS1]; }, a.sed_eps);
      if (two_d) alpha = dmin(alpha, sed_alpha_1d([&](int o) { return (double)sv[sb + o]; }, a.sed_eps));
      tv = alpha < 1.0;
    }
    nad = nad || tv;
  }
  bool pad = false;
  double pind = 0.0;
  if (a.pad_var >= 0) {
    const double x = (double)WS[a.pad_var * NSC + sb];
    pind = dmin(x - a.pad_lo, a.pad_hi - x);
    pad = pind < 0.0;
  }
  *mflag = nad || pad;
  *mnad = nad ? -imin : 0.0;
  *mpad = pad ? -pind : 0.0;
  return (a.force || pad || nad) ? 1 : 0;
}

template <typename T>
__device__ __forceinline__ int detect_cell(const T* __restrict__ W, int NW1, int NWC, const T* __restrict__ WS,
                                           int NS1, int NSC, const Geo& g, bool two_d,
                                           const double* __restrict__ range_mm, const DetectP& a, int c,
                                           bool* mflag, double* mnad, double* mpad) {
  double imin;
  const unsigned tvv = detect_nad(W, NW1, NWC, WS, NS1, NSC, g, two_d, range_mm, a, c, &imin);
  return detect_sed_pad(WS, NS1, NSC, g, two_d, a, c, tvv, imin, mflag, mnad, mpad);
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_detect(const T* __restrict__ W, int NW0, int NW1,
                                               const T* __restrict__ WS, int NS0, int NS1, Geo g,
                                               const double* __restrict__ range_mm, DetectP a,
                                               signed char* __restrict__ trouble, int* __restrict__ count,
                                               int* __restrict__ list, const signed char* __restrict__ pre,
                                               int* __restrict__ fp_count, MagLog ml) {
  const int nc = g.n0 * g.n1;
  const int c = blockIdx.x * TPB + threadIdx.x;
  int t = 0, fp = 0;
  bool mflag = false;
  double mnad = 0.0, mpad = 0.0;
  if (c < nc) {
    t = detect_cell(W, NW1, NW0 * NW1, WS, NS1, NS0 * NS1, g, g.ge1 > 0, range_mm, a, c, &mflag, &mnad, &mpad);
    if (pre) {  // pre-flagged cells already carry MUSCL fluxes: list only the others
      const int pf = pre[c];
      trouble[c] = (signed char)(t | pf);
      fp = pf && !t;
      t = t && !pf;
    } else {
      trouble[c] = (signed char)t;
    }
  }
  if (pre) {
    const unsigned long long bf = __ballot(fp);
    if ((threadIdx.x & 63) == 0 && bf) atomicAdd(fp_count, (int)__popcll(bf));
  }
  const int cell[1] = {c};
  block_append2<1, TPB>(cell, (unsigned)t, 0u, count, list, count, list);
  if (mag_on(ml)) mag_append(ml, mflag, c, mnad, mpad);
}

// The cells the fused candidate check left pending (NAD flagged with the SED on), decided as k_detect
// decides them.  The block loop's trip count is uniform over the block (block_append2 has barriers).
template <typename T>
__global__ void __launch_bounds__(TPB) k_detect_list(const int* __restrict__ pend_count, const int* __restrict__ pend,
                                                    const T* __restrict__ W, int NW0, int NW1,
                                                    const T* __restrict__ WS, int NS0, int NS1, Geo g,
                                                    const double* __restrict__ range_mm, DetectP a,
                                                    signed char* __restrict__ trouble, int* __restrict__ count,
                                                    int* __restrict__ list, const signed char* __restrict__ pre,
                                                    int* __restrict__ fp_count, MagLog ml,
                                                    const double* __restrict__ pend_imin, int pack) {
  const int n = pend_count[0];
  for (int base = blockIdx.x * TPB; base < n; base += gridDim.x * TPB) {
    const int k = base + (int)threadIdx.x;
    int t = 0, fp = 0, c = 0;
    bool mflag = false;
    double mnad = 0.0, mpad = 0.0;
    if (k < n) {
      if (pack) {  // the candidate tile's NAD bits and indicator: only the SED and PAD are left
        const unsigned e = (unsigned)pend[k];
        c = (int)(e & ((1u << PEND_SHIFT) - 1u));
        t = detect_sed_pad(WS, NS1, NS0 * NS1, g, g.ge1 > 0, a, c, e >> PEND_SHIFT, pend_imin[k], &mflag, &mnad,
                           &mpad);
      } else {
        c = pend[k];
        t = detect_cell(W, NW1, NW0 * NW1, WS, NS1, NS0 * NS1, g, g.ge1 > 0, range_mm, a, c, &mflag, &mnad, &mpad);
      }
      if (pre) {
        const int pf = pre[c];
        trouble[c] = (signed char)(t | pf);
        fp = pf && !t;
        t = t && !pf;
      } else {
        trouble[c] = (signed char)t;
      }
    }
    if (pre) {
      const unsigned long long bf = __ballot(fp);
      if ((threadIdx.x & 63) == 0 && bf) atomicAdd(fp_count, (int)__popcll(bf));
    }
    const int cell[1] = {c};
    block_append2<1, TPB>(cell, (unsigned)t, 0u, count, list, count, list);
    if (mag_on(ml)) mag_append(ml, mflag, c, mnad, mpad);
  }
}

// 2-D detection over strips of R rows (x) per thread, one column (y) per
// lane: the 3x3 max/min of W comes from per-row 3-wide maxima slid down the
// strip (3.75 loads per cell and variable instead of 9).  The flags equal
// k_detect's: max/min of the same nine values, in another order.
template <typename T, int R>
__global__ void __launch_bounds__(TPB) k_detect_strip(const T* __restrict__ W, int NW0, int NW1,
                                                     const T* __restrict__ WS, int NS0, int NS1, Geo g,
                                                     const double* __restrict__ range_mm, DetectP a,
                                                     signed char* __restrict__ trouble, int* __restrict__ count,
                                                     int* __restrict__ list, const signed char* __restrict__ pre,
                                                     int* __restrict__ fp_count, int jblocks, MagLog ml) {
  const int tid = threadIdx.x;
  const int j = (blockIdx.x % jblocks) * TPB + tid, i0 = (blockIdx.x / jblocks) * R;
  const bool col = j < g.n1;
  const int NWC = NW0 * NW1, NSC = NS0 * NS1;
  const int jc = col ? j : g.n1 - 1;  // idle lanes read a valid column, their results are dropped
  unsigned nadbits[R];  // bit v: NAD test of variable v (masked by nad_vars)
  double imin[R];       // min over the variables of the NAD indicator
#pragma unroll
  for (int k = 0; k < R; ++k) nadbits[k] = 0u;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    T hM[R + 2], hm[R + 2];  // in the storage type, widened once per cell (exact: widening is monotonic)
#pragma unroll
    for (int r = 0; r < R + 2; ++r) {
      const int i = min(i0 - 1 + r, g.n0);  // W row of cell i (ghosts: g.ge0 >= 3)
      const T* wr = W + v * NWC + (i + g.ge0) * NW1 + jc + g.ge1;
      const T x0 = wr[-1], x1 = wr[0], x2 = wr[1];
      hM[r] = dmax(dmax(x0, x1), x2);
      hm[r] = dmin(dmin(x0, x1), x2);
    }
    const double rlo = range_mm[2 * v], rhi = range_mm[2 * v + 1];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int i = min(i0 + k, g.n0 - 1);
      const double M = (double)dmax(dmax(hM[k], hM[k + 1]), hM[k + 2]);
      const double m = (double)dmin(dmin(hm[k], hm[k + 1]), hm[k + 2]);
      const double uc = (double)WS[v * NSC + (i + 3) * NS1 + jc + 3];
      bool tv;
      double ind;
      if (a.nad_range == 0) {
        const double rng = a.nad_mode == 0 ? rhi - rlo : M - m;
        ind = dmin(uc - m, M - uc);
        tv = ind < -a.eps * rng;
      } else {
        const double lo = a.nad_mode == 0 ? (1.0 - a.eps) * rlo : (1.0 - a.eps) * m;
        const double hi = a.nad_mode == 0 ? (1.0 + a.eps) * rhi : (1.0 + a.eps) * M;
        ind = dmin(uc - lo, hi - uc);
        tv = ind < 0.0;
      }
      if (tv && ((a.nad_vars >> v) & 1)) nadbits[k] |= 1u << v;
      imin[k] = v == 0 ? ind : dmin(imin[k], ind);
    }
  }
  unsigned tbits = 0u, fpbits = 0u, mbits = 0u;
  double mnad[R], mpad[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    mnad[k] = 0.0;
    mpad[k] = 0.0;
    const int i = i0 + k;
    if (!col || i >= g.n0) continue;
    const int sb = (i + 3) * NS1 + j + 3;
    bool nad = false;
    for (int v = 0; v < 5; ++v) {
      bool tv = (nadbits[k] >> v) & 1u;
      if (tv && a.sed && !nad) {  // smooth-extrema detector only where NAD flags (it can only clear flags)
        const T* sv = WS + v * NSC;
        double alpha = sed_alpha_1d([&](int o) { return (double)sv[sb + o * NS1]; }, a.sed_eps);
        alpha = dmin(alpha, sed_alpha_1d([&](int o) { return (double)sv[sb + o]; }, a.sed_eps));
        tv = alpha < 1.0;
      }
      nad = nad || tv;
    }
    bool pad = false;
    double pind = 0.0;
    if (a.pad_var >= 0) {
      const double x = (double)WS[a.pad_var * NSC + sb];
      pind = dmin(x - a.pad_lo, a.pad_hi - x);
      pad = pind < 0.0;
    }
    if (nad || pad) mbits |= 1u << k;
    mnad[k] = nad ? -imin[k] : 0.0;
    mpad[k] = pad ? -pind : 0.0;
    int t = (a.force || pad || nad) ? 1 : 0;
    const int c = i * g.n1 + j;
    if (pre) {  // pre-flagged cells already carry MUSCL fluxes: list only the others
      const int pf = pre[c];
      trouble[c] = (signed char)(t | pf);
      if (pf && !t) fpbits |= 1u << k;
      t = t && !pf;
    } else {
      trouble[c] = (signed char)t;
    }
    if (t) tbits |= 1u << k;
  }
  if (pre) {
    int f = __popc(fpbits);
    for (int off = 32; off > 0; off >>= 1) f += __shfl_down(f, off, 64);
    if ((tid & 63) == 0 && f) atomicAdd(fp_count, f);
  }
  // compaction: block scan of the per-thread counts, one atomic per block
  __shared__ int sc[TPB];
  __shared__ int bbase;
  const int cnt = __popc(tbits);
  sc[tid] = cnt;
  __syncthreads();
  for (int off = 1; off < TPB; off <<= 1) {
    const int x = tid >= off ? sc[tid - off] : 0;
    __syncthreads();
    sc[tid] += x;
    __syncthreads();
  }
  if (tid == TPB - 1) bbase = sc[TPB - 1] ? atomicAdd(count, sc[TPB - 1]) : 0;
  __syncthreads();
  int pos = bbase + sc[tid] - cnt;
#pragma unroll
  for (int k = 0; k < R; ++k)
    if ((tbits >> k) & 1u) list[pos++] = (i0 + k) * g.n1 + j;
  if (mag_on(ml)) {
#pragma unroll
    for (int k = 0; k < R; ++k) mag_append(ml, (mbits >> k) & 1u, (i0 + k) * g.n1 + j, mnad[k], mpad[k]);
  }
}

// ---------------------------------------------------------------------------
// MUSCL fallback fluxes and the final divergence
// ---------------------------------------------------------------------------
struct FinalP {
  int solver, limiter;  // limiter: FVHOE_LIM_* or 3 = unlimited p=1 interpolation (slope_limiter=None)
  double gamma, csq_floor;
};

// primitive fv average of the p=1 fallback at a (ghost) cell: the p=1 c/t
// sweeps and the degree-0 sweeps are identities apart from 0.0 + x
template <typename T, typename Src>
__device__ __forceinline__ void p1_prims(const Src& src, const Geo& g, int cx, int cy, T w[5], double gamma) {
  double u[5], wd[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) u[q] = 0.0 + src.get(g, q, cx, cy);
  prims_from_cons(u, wd, gamma);
#pragma unroll
  for (int q = 0; q < 5; ++q) w[q] = T(0.0 + wd[q]);
}

template <int D, typename T, typename Src>
__device__ void fallback_flux(const Src& src, const Geo& g, int i, int j, const FinalP& p, T F[5]) {
  // face between cells (i-1, j) and (i, j) for D=0, (i, j-1) and (i, j) for D=1
  const int di = D == 0 ? 1 : 0, dj = D == 0 ? 0 : 1;
  T w0[5], w1[5], w2[5], w3[5], wl[5], wr[5];
  p1_prims(src, g, i - 2 * di, j - 2 * dj, w0, p.gamma);
  p1_prims(src, g, i - di, j - dj, w1, p.gamma);
  p1_prims(src, g, i, j, w2, p.gamma);
  p1_prims(src, g, i + di, j + dj, w3, p.gamma);
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    if (p.limiter == 3) {  // slope_limiter None: p=1 conservative interpolation (fvhoe/solver.py:531-536)
      // r weights are the reversed l weights: (-1/4, 1, 1/4)
      wl[q] = ((T(0) + T(-0.25) * w0[q]) + T(1) * w1[q]) + T(0.25) * w2[q];
      wr[q] = ((T(0) + T(0.25) * w1[q]) + T(1) * w2[q]) + T(-0.25) * w3[q];
    } else {
      const T s1 = limited_slope(p.limiter, w1[q] - w0[q], w2[q] - w1[q]);
      const T s2 = limited_slope(p.limiter, w2[q] - w1[q], w3[q] - w2[q]);
      wl[q] = w1[q] + T(0.5) * s1;  // right face of the left cell
      wr[q] = w2[q] - T(0.5) * s2;  // left face of the right cell
    }
  }
  riemann<D>(p.solver, wl, wr, F, T(p.gamma), T(p.csq_floor));
}

// MUSCL fluxes at the faces of every flagged cell, written over the
// high-order fluxes (theta = 1 faces of the simple broadcast); a face shared by
// two flagged cells is written twice with the same value.
template <bool TWO_D, typename T, typename Src>
__global__ void __launch_bounds__(TPB) k_fallback(const int* __restrict__ count, const int* __restrict__ list,
                                                 Src src, Geo g, FinalP p, T* __restrict__ F, T* __restrict__ G) {
  const int n = count[0];
  constexpr int NF = TWO_D ? 4 : 2;
  const int nF = (g.n0 + 1) * g.n1, nG = g.n0 * (g.n1 + 1);
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * NF; e += gridDim.x * blockDim.x) {
    const int f = e / n;
    const int c = list[e - f * n];
    const int i = c / g.n1, j = c - (c / g.n1) * g.n1;
    T Fv[5];
    if (f < 2) {
      fallback_flux<0>(src, g, i + f, j, p, Fv);
      const int fi = (i + f) * g.n1 + j;
#pragma unroll
      for (int q = 0; q < 5; ++q) F[q * nF + fi] = Fv[q];
    } else if (TWO_D) {
      fallback_flux<1>(src, g, i, j + (f - 2), p, Fv);
      const int gi = i * (g.n1 + 1) + j + (f - 2);
#pragma unroll
      for (int q = 0; q < 5; ++q) G[q * nG + gi] = Fv[q];
    }
  }
}

// Runge-Kutta combination written by the stage's last kernel; the stage's own
// dudt is the last k operand of the kind (fvhoe/ode.py:245-298).
// dudt = 0.0 + cx (F[i+1]-F[i]) + cy (G[j+1]-G[j])  (fvhoe/solver.py:744-750) in
// FP64, then the optional RK combination.  rk.out may alias the stage input or
// rk.u0 (pointwise).
// stats (learned stage, may be null): thread 0 of block 0 adds the stage's counters, [0] pre-flagged cells
// (count[1]), [1] cells the check listed (count[0]), [2] pre-flagged cells the check did not flag (count[2]),
// unless the step is at its target time
template <bool TWO_D, typename T>
__global__ void __launch_bounds__(TPB) k_final(const T* __restrict__ F, const T* __restrict__ G, Geo g, double* dudt,
                                              RkArgs rk, const int* __restrict__ count, long long* __restrict__ stats,
                                              int with_pre) {
  if (stats && blockIdx.x == 0 && threadIdx.x == 0 && !(rk.kind >= 0 && rk.dt[0] == 0.0)) {
    if (with_pre) {  // a step at its target time counts nothing
      stats[0] += count[1];
      stats[2] += count[2];
    }
    stats[1] += count[0];
  }
  const int nc = g.n0 * g.n1;
  const int nF = (g.n0 + 1) * g.n1, nG = g.n0 * (g.n1 + 1);
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
    const int i = c / g.n1, j = c - (c / g.n1) * g.n1;
    const int fi = i * g.n1 + j;
    const int gi = i * (g.n1 + 1) + j;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      double acc = 0.0 + g.cx * ((double)F[q * nF + fi + g.n1] - (double)F[q * nF + fi]);
      if (TWO_D) acc = acc + g.cy * ((double)G[q * nG + gi + 1] - (double)G[q * nG + gi]);
      if (dudt) dudt[q * nc + c] = rk.kind == FVHOE_RK_SSPRK3_STAGE3_SUM ? rk.k0[q * nc + c] + acc : acc;
      if (rk.kind >= 0) rk.out[q * nc + c] = rk_combine(rk, q * nc + c, acc);
    }
  }
}

// Predictor features: (rho, vx, vy, P) of W at the cells of the grid extended
// by 6 per side (boundary conditions applied by index mapping), scaled to the
// global range as (w - min) * (1 / (max - min)) in FP64; 0 where max = min.
template <typename T>
__global__ void __launch_bounds__(TPB) k_features(const T* __restrict__ W, int NW0, int NW1, Geo g,
                                                 const double* __restrict__ range, float* __restrict__ X) {
  const int HX = g.n0 + 12, WX = g.n1 + 12, total = HX * WX;
  const int e = blockIdx.x * TPB + threadIdx.x;
  if (e >= total) return;
  const int a = e / WX, b = e - (e / WX) * WX;
  bool fx, fy;
  const int sx = cell_map(a - 6, g.n0, g.bc[0], g.bc[1], &fx);
  const int sy = cell_map(b - 6, g.n1, g.bc[2], g.bc[3], &fy);
  const int wi = (sx + g.ge0) * NW1 + sy + g.ge1, NWC = NW0 * NW1;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int v = k == 3 ? 4 : k;
    double x = (double)W[v * NWC + wi];
    if ((fx && v == 1) || (fy && v == 2)) x = -x;
    const double lo = range[2 * v], rng = range[2 * v + 1] - lo;
    X[k * total + e] = rng > 0.0 ? (float)((x - lo) * (1.0 / rng)) : 0.0f;
  }
}

// The same features read straight from W inside the inference kernel; pixels
// beyond the extended grid (the overhang of the last tiles) are 0.
template <typename T>
struct FeatW {
  const T* __restrict__ W;
  int NW0, NW1;
  Geo g;
  const double* __restrict__ range;
  // per-variable (min, 1 / (max - min)) into sh[0..7], once per block
  __device__ __forceinline__ void init(double* sh) const {
    if (threadIdx.x < 4) {
      const int k = threadIdx.x, v = k == 3 ? 4 : k;
      const double lo = range[2 * v], rng = range[2 * v + 1] - lo;
      sh[2 * k] = lo;
      sh[2 * k + 1] = rng > 0.0 ? 1.0 / rng : 0.0;
    }
  }
  __device__ __forceinline__ void load(int a, int b, float f[4], const double* sh) const {
    if (a >= g.n0 + 12 || b >= g.n1 + 12) {
      f[0] = f[1] = f[2] = f[3] = 0.0f;
      return;
    }
    bool fx, fy;
    const int sx = cell_map(a - 6, g.n0, g.bc[0], g.bc[1], &fx);
    const int sy = cell_map(b - 6, g.n1, g.bc[2], g.bc[3], &fy);
    const int wi = (sx + g.ge0) * NW1 + sy + g.ge1, NWC = NW0 * NW1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int v = k == 3 ? 4 : k;
      double x = (double)W[v * NWC + wi];
      if ((fx && v == 1) || (fy && v == 2)) x = -x;
      const double inv = sh[2 * k + 1];
      f[k] = inv != 0.0 ? (float)((x - sh[2 * k]) * inv) : 0.0f;
    }
  }
};

// The features of the extended grid rounded to bfloat16 once per cell (channel
// last), so the network's tiles load 8 bytes per pixel instead of forming the
// features of their halos again; the values are FeatW's, converted as the
// network converts them.
template <typename T>
__global__ void __launch_bounds__(TPB) k_feat_pack(FeatW<T> feat, pred::bf16x4v* __restrict__ XB, int HX, int WX) {
  __shared__ double sh[8];
  feat.init(sh);
  __syncthreads();
  const int total = HX * WX;
  for (int e = blockIdx.x * TPB + threadIdx.x; e < total; e += gridDim.x * TPB) {
    const int a = e / WX, b = e - (e / WX) * WX;
    XB[e] = pred::load_bf16(feat, a, b, sh);
  }
}

// final reduction of the prims partials: dt (fvhoe/hydro.py:290-293, and
// _hit_target, fvhoe/ode.py:11-22, when asked) and the per-variable global
// (min, max) (fvhoe/slope_limiting.py:168-170)
// One block per quantity k (0: the CFL denominator, 1 + 2v / 2 + 2v: min / max of variable v) over the
// per-block partials, stored [quantity][block] so every block reads its row coalesced.  Max / min of
// the ordered bits are exact in any order.
__global__ void __launch_bounds__(1024) k_stage_reduce(const unsigned long long* __restrict__ partials, int nblocks,
                                                       double* range, double cfl_h, int fixed_on, double fixed_dt,
                                                       double* dt, double* dt_step, const double* t, double target,
                                                       int has_target, const double* target_ptr, int* count,
                                                       int reset_pre, int* mag_count, const double* mag_gate) {
  __shared__ unsigned long long part[16];
  const int tid = threadIdx.x, k = blockIdx.x;
  if (k == 0 && tid == 0) {  // this stage's flag counters (the previous stage's consumers ran before)
    count[0] = 0;
    count[2] = 0;
    count[3] = 0;
    if (reset_pre) count[1] = 0;
  }
  unsigned long long x = k == 0 ? 0ull : ((k & 1) ? ~0ull : 0ull);
  const unsigned long long* row = partials + (long long)k * nblocks;
  for (int b = tid; b < nblocks; b += 1024) x = red_op(k, x, row[b]);
  for (int off = 32; off > 0; off >>= 1) x = red_op(k, x, __shfl_down(x, off, 64));
  if ((tid & 63) == 0) part[tid >> 6] = x;
  __syncthreads();
  if (tid == 0) {
    x = part[0];
    for (int w = 1; w < 16; ++w) x = red_op(k, x, part[w]);
    if (k == 0) {
      const double d = fixed_on ? fixed_dt : cfl_h / __longlong_as_double((long long)x);
      dt[0] = d;
      const double tg = target_ptr ? target_ptr[0] : target;  // a device target serves every replay
      if (dt_step) dt_step[0] = (has_target && tg - t[0] < d) ? tg - t[0] : d;
      // this stage's magnitude slot, after the step's dt is known (a step at its target keeps the slot)
      if (mag_count && !(mag_gate && mag_gate[0] == 0.0)) mag_count[0] = 0;
    } else {
      range[k - 1] = from_ordered_bits(x);
    }
  }
}

// ------------------------------------------------------------------ host side

// Optional per-kernel event timing (bench instrumentation): each launch of a
// stage kernel is bracketed by a pair of events from a fixed pool.
enum { TK_PRIMS = 0, TK_FLUX_X = 1, TK_FLUX_Y = 2, TK_PRIMS_CAND = 3, TK_DETECT = 4, TK_FALLBACK = 5, TK_FINAL = 6,
       TK_PREDICT = 7, TK_COUNT = 8 };
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
  int two_d, p, fp32, G, GE, n0, n1, NW0, NW1, NS0, NS1, prims_blocks, pred, dirichlet;
  size_t offW, offF, offG, offWS, offFlags, offList, offPend, offPart, offScal, offPre, offPreList, offFeat, total;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr int TX1 = 256;
// LDS-tiled prims shape per arithmetic type and degree (FP64 at degrees up to FVHOE_CAND_MAX_P: the
// candidate's 32 x 32)
template <typename T>
constexpr int tile_x(int p) {
  return sizeof(T) == 8 ? (p <= FVHOE_CAND_MAX_P ? FVHOE_CAND_X : FVHOE_TILE_X) : FVHOE_TILE32_X;
}
template <typename T>
constexpr int tile_y(int p) {
  return sizeof(T) == 8 ? (p <= FVHOE_CAND_MAX_P ? FVHOE_CAND_Y : FVHOE_TILE_Y) : FVHOE_TILE32_Y;
}

// deepest ghost read from u: W's ghosts plus the prims stencils' reach
// (the candidate's 3 ghosts and the fallback's MUSCL stencil reach less)
int ghost_depth(int p) { return (flux_ghosts(p) > 3 ? flux_ghosts(p) : 3) + ereach_c(p) + ereach_t(p); }

int make_layout(const FvhoeStageConfig* c, Layout* L, bool pred = false) {
  if (!c) return FVHOE_ERR_ARG;
  const bool ax = !(c->n[0] == 1 && c->p[0] == 0), ay = !(c->n[1] == 1 && c->p[1] == 0),
             az = !(c->n[2] == 1 && c->p[2] == 0);
  if (!ax || az) return FVHOE_ERR_UNSUPPORTED;
  if (ay && c->p[1] != c->p[0]) return FVHOE_ERR_UNSUPPORTED;
  if (c->p[0] < 1 || c->p[0] > 8) return FVHOE_ERR_UNSUPPORTED;
  if (c->riemann < 0 || c->riemann > 2 || c->limiter < 0 || c->limiter > 3) return FVHOE_ERR_ARG;
  if (c->bulk_fp32 < 0 || c->bulk_fp32 > 1) return FVHOE_ERR_ARG;
  bool dirichlet = false;
  for (int k = 0; k < 4; ++k) {
    if (c->bc[k] < 0 || c->bc[k] > BC_DIRICHLET) return FVHOE_ERR_ARG;
    if (c->bc[k] == BC_DIRICHLET) {
      if (k >= 2 && !ay) return FVHOE_ERR_ARG;
      if (!c->slab[k]) return FVHOE_ERR_ARG;
      dirichlet = true;
    }
  }
  {  // the slabs must cover the deepest ghost any kernel reads
    const int need = ghost_depth(c->p[0]);
    if (dirichlet && (c->slab_g < need || (ay && c->slab_pad[1] < need) || (ay && c->slab_pad[0] < need) ||
                      c->slab_pad[2] != 0))
      return FVHOE_ERR_ARG;
  }
  if (dirichlet && pred) return FVHOE_ERR_UNSUPPORTED;  // the predictor's features map ghosts into the interior
  if (c->mag_idx && (!c->mag_count || !c->mag_nad || !c->mag_pad || c->mag_slot < 0)) return FVHOE_ERR_ARG;
  L->dirichlet = dirichlet;
  L->two_d = ay;
  L->p = c->p[0];
  L->fp32 = c->bulk_fp32;
  L->G = flux_ghosts(L->p);
  L->GE = L->G > 3 ? L->G : 3;
  L->n0 = c->n[0];
  L->n1 = c->n[1];
  // the reference pads u by gw + 2 ceil(p/2) with slabs copied from the interior
  const int pad = L->G + 2 * reach(L->p);
  if (L->n0 < pad || (ay && L->n1 < pad)) return FVHOE_ERR_UNSUPPORTED;
  if (c->a_posteriori && (L->n0 < 7 || (ay && L->n1 < 7))) return FVHOE_ERR_UNSUPPORTED;
  // 32-bit element offsets inside the kernels
  if ((double)5 * (L->n0 + 2 * L->GE + 1) * (ay ? L->n1 + 2 * L->GE + 1 : 1) > 2.0e9) return FVHOE_ERR_UNSUPPORTED;
  L->NW0 = L->n0 + 2 * L->GE;
  L->NW1 = ay ? L->n1 + 2 * L->GE : 1;
  L->NS0 = L->n0 + 6;
  L->NS1 = ay ? L->n1 + 6 : 1;
  const size_t esz = L->fp32 ? sizeof(float) : sizeof(double);
  size_t o = 0;
  L->offW = o; o = align256(o + esz * 5 * (size_t)L->NW0 * L->NW1);
  L->offF = o; o = align256(o + esz * 5 * (size_t)(L->n0 + 1) * L->n1);
  L->offG = o; o = align256(o + (ay ? esz * 5 * (size_t)L->n0 * (L->n1 + 1) : 0));
  L->offWS = o; o = align256(o + (c->a_posteriori ? esz * 5 * (size_t)L->NS0 * L->NS1 : 0));
  L->offFlags = o; o = align256(o + (c->a_posteriori ? (size_t)L->n0 * L->n1 : 0));
  L->offList = o; o = align256(o + (c->a_posteriori ? sizeof(int) * (size_t)L->n0 * L->n1 : 0));
  // the cells the fused 2-D candidate check leaves to the SED pass
  // the pending list (cells the candidate check leaves to the SED pass) and, after it, their NAD indicators
  L->offPend = o;
  if (FVHOE_CAND_CHECK && ay && c->a_posteriori && c->sed)
    o = align256(align256(o + sizeof(int) * (size_t)L->n0 * L->n1) + sizeof(double) * (size_t)L->n0 * L->n1);
  if (L->two_d && (L->fp32 ? use_strip<float>(L->p) : use_strip<double>(L->p))) {
    L->prims_blocks = strip_blocks(L->p, L->fp32 ? strip_rows<float>() : strip_rows<double>(), L->NW0, L->NW1);
  } else if (L->two_d) {
    const int tx = L->fp32 ? tile_x<float>(L->p) : tile_x<double>(L->p);
    const int ty = L->fp32 ? tile_y<float>(L->p) : tile_y<double>(L->p);
    L->prims_blocks = ((L->NW0 + tx - 1) / tx) * ((L->NW1 + ty - 1) / ty);
  } else {
    L->prims_blocks = (L->NW0 + TX1 - 1) / TX1;
  }
  L->offPart = o; o = align256(o + sizeof(unsigned long long) * NPART * (size_t)L->prims_blocks);
  L->offScal = o; o = align256(o + 256);
  L->pred = pred;
  L->offPre = L->offPreList = L->offFeat = o;
  if (pred) {  // pre-flags, their list and the network's bfloat16 features of the extended grid
    if (!L->two_d || !c->a_posteriori) return FVHOE_ERR_UNSUPPORTED;
    L->offPre = o; o = align256(o + (size_t)L->n0 * L->n1);
    L->offPreList = o; o = align256(o + sizeof(int) * (size_t)L->n0 * L->n1);
    L->offFeat = o; o = align256(o + 8 * (size_t)(L->n0 + 12) * (L->n1 + 12));
  }
  L->total = o;
  return FVHOE_OK;
}

template <typename T, int P, bool TWO_D, bool DIR>
int run_stage(const FvhoeStageConfig* c, const FvhoeRkFuse* rkf, const FvhoePredict* pred, const Layout& L,
              const double* u, double* dudt, double* dt_stage, signed char* trouble_out, char* ws, hipStream_t s) {
  // sources of the conservative state and of the candidate, slab-aware only where a side needs it
  using SU = std::conditional_t<DIR, SrcUD, SrcU>;
  using SUS = std::conditional_t<DIR, SrcUstarD<TWO_D, T>, SrcUstar<TWO_D, T>>;
  constexpr int PY = TWO_D ? P : 0;
  constexpr int TX = TWO_D ? tile_x<T>(P) : TX1, TY = TWO_D ? tile_y<T>(P) : 1;
  constexpr int NTILE = TWO_D ? 512 : 256;  // threads of the LDS-tiled kernels
  constexpr bool CBIG = sizeof(T) == 8 && P <= FVHOE_CAND_MAX_P;
  constexpr int CX = TWO_D ? (CBIG ? FVHOE_CAND_X : sizeof(T) == 8 ? FVHOE_TILE_X : FVHOE_CAND32_X) : TX1;
  constexpr int CY = TWO_D ? (CBIG ? FVHOE_CAND_Y : sizeof(T) == 8 ? FVHOE_TILE_Y : FVHOE_CAND32_Y) : 1;
  constexpr int NCAND = TWO_D ? (sizeof(T) == 8 ? FVHOE_CAND_NT_F64 : FVHOE_CAND_NT_F32) : 256;
  T* W = (T*)(ws + L.offW);
  T* F = (T*)(ws + L.offF);
  T* G = TWO_D ? (T*)(ws + L.offG) : nullptr;
  T* WS = (T*)(ws + L.offWS);
  // the flags go straight to the caller's array when one is given (no copy node in a captured graph)
  signed char* flags = trouble_out ? trouble_out : (signed char*)(ws + L.offFlags);
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
  Slabs sl{};
  if (DIR) {
    for (int k = 0; k < 6; ++k) sl.s[k] = k < 4 && c->bc[k] == BC_DIRICHLET ? c->slab[k] : nullptr;
    sl.g[0] = sl.g[1] = c->slab_g;
    sl.g[2] = 0;
    sl.pad[0] = c->slab_pad[0];
    sl.pad[1] = c->slab_pad[1];
    sl.pad[2] = 0;
  }
  g.cx = -(1.0 / c->h[0]);
  g.cy = -(1.0 / c->h[1]);

  const bool apost = c->a_posteriori != 0;
  const bool use_pred = TWO_D && apost && pred != nullptr;
  const bool overlap = use_pred && pred->overlap;
  // a stage that runs the network beside its own kernels applies no pre-flags itself
  const bool pre_now = use_pred && !(overlap && pred->infer);
  const bool need_range = apost && (c->nad_mode == 0 || use_pred);
  signed char* preflags = pre_now ? (signed char*)(ws + L.offPre) : nullptr;
  signed char* preflags_out = use_pred ? (signed char*)(ws + L.offPre) : nullptr;
  int* pre_list = (int*)(ws + L.offPreList);

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
    TimedLaunch tl(TK_PRIMS, s);
    if constexpr (TWO_D && use_strip<T>(P)) {
      k_prims_strip<T, P, SU><<<strip_blocks(P, strip_rows<T>(), L.NW0, L.NW1), 256, 0, s>>>(
          make_src<SU>(u, sl), g, W, L.NW0, L.NW1, strip_cwins(P, L.NW1), red, c->gamma);
    } else {
      const int t0 = (L.NW0 + TX - 1) / TX, t1 = (L.NW1 + TY - 1) / TY;
      k_prims<T, P, PY, TX, TY, SU, NTILE><<<t0 * t1, NTILE, 0, s>>>(make_src<SU>(u, sl), g, W, L.NW0, L.NW1, t1, red,
                                                                     c->gamma, CheckArgs<T>{});
    }
  }
  const double hmin = c->h[0] < c->h[1] ? (c->h[0] < c->h[2] ? c->h[0] : c->h[2]) : (c->h[1] < c->h[2] ? c->h[1] : c->h[2]);
  RkArgs rk{-1, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  const bool set_step = rkf && rkf->set_dt_step;
  MagLog ml{};
  if (apost && c->mag_idx) {
    const size_t off = (size_t)c->mag_slot * (size_t)L.n0 * (size_t)L.n1;
    ml = MagLog{c->mag_count + c->mag_slot, c->mag_idx + off, c->mag_nad + off, c->mag_pad + off,
                rkf ? rkf->dt_step : nullptr};
  }
  if (rkf && rkf->kind >= 0) rk = RkArgs{rkf->kind, rkf->out, rkf->u0, rkf->k0, rkf->k1, rkf->k2, rkf->dt_step};
  k_stage_reduce<<<NPART, 1024, 0, s>>>(partials, L.prims_blocks, range, c->cfl * hmin, c->fixed_dt_on, c->fixed_dt,
                                    dt_stage, set_step ? rkf->dt_step : nullptr, set_step ? rkf->t : nullptr,
                                    set_step ? rkf->target : 0.0, set_step ? rkf->has_target : 0,
                                    set_step ? rkf->target_ptr : nullptr, count, use_pred && pred->infer,
                                    ml.idx ? ml.count : nullptr, ml.gate);
  if (use_pred && pred->infer) {  // the network on this stage's W, features formed on load
    hipStream_t ps = s;
    if (overlap) {  // fork: the network runs beside the rest of this stage
      ps = (hipStream_t)pred->side_stream;
      if (hipEventRecord((hipEvent_t)pred->ev_fork, s) != hipSuccess ||
          hipStreamWaitEvent(ps, (hipEvent_t)pred->ev_fork, 0) != hipSuccess)
        return FVHOE_ERR_LAUNCH;
    }
    {
      TimedLaunch tl(TK_PREDICT, ps);
      const int HX = L.n0 + 12, WX = L.n1 + 12;
      pred::bf16x4v* XB = (pred::bf16x4v*)(ws + L.offFeat);
      const int fb = (HX * WX + TPB - 1) / TPB;
      k_feat_pack<T><<<fb < 4096 ? fb : 4096, TPB, 0, ps>>>(FeatW<T>{W, L.NW0, L.NW1, g, range}, XB, HX, WX);
      const int st = pred::launch_infer(pred::FeatBF16{XB, HX, WX}, L.n0, L.n1, pred->packed,
                                        pred->logit_threshold, preflags_out, nullptr, count + 1, pre_list, ps,
                                        overlap);
      if (st != FVHOE_OK) return st;
    }
    if (overlap && hipEventRecord((hipEvent_t)pred->ev_join, ps) != hipSuccess) return FVHOE_ERR_LAUNCH;
  }

  // the flux tiles with the Riemann solver as a template argument (no code or registers of the others)
  auto flux = [&](auto rs) {
    constexpr int RS = decltype(rs)::value;
    if (TWO_D) {
      constexpr bool F64 = sizeof(T) == 8;
      constexpr int FA = F64 ? FVHOE_FLUX64_A : FVHOE_FLUX_A, FB = F64 ? FVHOE_FLUX64_B : FVHOE_FLUX_B;
      constexpr int TA0 = FA, TB0 = FB, TA1 = F64 ? FVHOE_FLUX64_YA : FVHOE_FLUX32_YA,
                    TB1 = F64 ? FVHOE_FLUX64_YB : FVHOE_FLUX32_YB;
      constexpr int NF = F64 ? FVHOE_FLUX_NT_F64 : FVHOE_FLUX_NT_F32;
      if constexpr (FVHOE_FLUX_XY) {
        constexpr int XA = F64 ? FVHOE_FXY64_A : FVHOE_FXY32_A, XB = F64 ? FVHOE_FXY64_B : FVHOE_FXY32_B;
        constexpr int XNT = F64 ? FVHOE_FXY64_NT : FVHOE_FXY32_NT;
        const int ta = (L.n0 + 1 + XA - 1) / XA, tb = (L.n1 + 1 + XB - 1) / XB;
        TimedLaunch tl(TK_FLUX_X, s);  // both axes, timed as one
        k_flux_xy<T, P, XA, XB, XNT, RS><<<ta * tb, XNT, 0, s>>>(W, L.NW0, L.NW1, g, F, G, tb, c->gamma,
                                                                 c->csq_floor);
        return;
      }
      const int tx = (L.n0 + 1 + TA0 - 1) / TA0, ty = (L.n1 + TB0 - 1) / TB0;
      {
        TimedLaunch tl(TK_FLUX_X, s);
        k_flux<T, 0, P, PY, TA0, TB0, NF, RS><<<tx * ty, NF, 0, s>>>(W, L.NW0, L.NW1, g, F, ty, c->gamma,
                                                                      c->csq_floor);
      }
      const int fy = (L.n1 + 1 + TA1 - 1) / TA1, rx = (L.n0 + TB1 - 1) / TB1;
      {
        TimedLaunch tl(TK_FLUX_Y, s);
        k_flux<T, 1, PY, P, TA1, TB1, NF, RS><<<fy * rx, NF, 0, s>>>(W, L.NW0, L.NW1, g, G, rx, c->gamma,
                                                                      c->csq_floor);
      }
    } else {
      const int tx = (L.n0 + 1 + 255) / 256;
      TimedLaunch tl(TK_FLUX_X, s);
      k_flux<T, 0, P, 0, 256, 1, 256, RS><<<tx, 256, 0, s>>>(W, L.NW0, L.NW1, g, F, 1, c->gamma, c->csq_floor);
    }
  };
  if (c->riemann == RS_HLLC) flux(std::integral_constant<int, RS_HLLC>{});
  else if (c->riemann == RS_LLF) flux(std::integral_constant<int, RS_LLF>{});
  else flux(std::integral_constant<int, RS_UPWIND>{});

  FinalP fp{c->riemann, c->limiter, c->gamma, c->csq_floor};
  const int nc = L.n0 * L.n1;
  if (pre_now) {  // MUSCL fluxes at the faces of the pre-flagged cells, before the candidate
    TimedLaunch tl(TK_FALLBACK, s);
    k_fallback<TWO_D, T, SU><<<1024, TPB, 0, s>>>(count + 1, pre_list, make_src<SU>(u, sl), g, fp, F, G);
  }
  const int gridc = (nc + TPB - 1) / TPB < 65536 ? (nc + TPB - 1) / TPB : 65536;
  if (apost) {
    Geo gs = g;
    gs.ge0 = 3;
    gs.ge1 = TWO_D ? 3 : 0;
    const int t0 = (L.NS0 + CX - 1) / CX, t1 = (L.NS1 + CY - 1) / CY;
    Reduce none{};
    DetectP dp{c->nad_mode, c->nad_range, c->nad_vars, c->pad_var, c->sed, c->force_trouble,
               c->nad_eps, c->pad_lo, c->pad_hi, c->sed_eps};
    constexpr bool FUSED_CHECK = TWO_D && FVHOE_CAND_CHECK;
    int* pend = (int*)(ws + L.offPend);
    double* pend_imin = (double*)(ws + align256(L.offPend + sizeof(int) * (size_t)nc));
    const int pack = FVHOE_PEND_PACK && (size_t)nc < ((size_t)1 << PEND_SHIFT) ? 1 : 0;
    {
      TimedLaunch tl(TK_PRIMS_CAND, s);
      // the LDS tiles form each candidate value once per tile (the strips would form it (R + 2H) / R times)
      CheckArgs<T> ck{};
      if (FUSED_CHECK)
        ck = CheckArgs<T>{W, L.NW0, L.NW1, g.ge0, g.ge1, range, dp, flags, count, list, count + 3, pend,
                          pend_imin, pack, preflags, count + 2, ml};
      k_prims<T, P, PY, CX, CY, SUS, NCAND, FUSED_CHECK><<<t0 * t1, NCAND, 0, s>>>(
          make_srcstar<SUS>(u, F, G, dt_stage, sl), gs, WS, L.NS0, L.NS1, t1, none, c->gamma, ck);
    }
    const int gridd = (nc + TPB - 1) / TPB;
    constexpr int DR = FVHOE_DETECT_ROWS;  // rows per thread of the 2-D detection
    const int jblocks = (L.n1 + TPB - 1) / TPB, gridd2 = jblocks * ((L.n0 + DR - 1) / DR);
    {
      TimedLaunch tl(TK_DETECT, s);
      if (FUSED_CHECK) {
        if (c->sed)  // the cells whose NAD test flagged: the SED reads WS beyond the tiles
          k_detect_list<T><<<512, TPB, 0, s>>>(count + 3, pend, W, L.NW0, L.NW1, WS, L.NS0, L.NS1, g, range, dp,
                                               flags, count, list, preflags, count + 2, ml, pend_imin, pack);
      } else if (TWO_D) {
        k_detect_strip<T, DR><<<gridd2, TPB, 0, s>>>(W, L.NW0, L.NW1, WS, L.NS0, L.NS1, g, range, dp, flags, count,
                                                      list, preflags, count + 2, jblocks, ml);
      } else {
        k_detect<T><<<gridd, TPB, 0, s>>>(W, L.NW0, L.NW1, WS, L.NS0, L.NS1, g, range, dp, flags, count, list,
                                        preflags, count + 2, ml);
      }
    }
    {
      TimedLaunch tl(TK_FALLBACK, s);
      k_fallback<TWO_D, T, SU><<<1024, TPB, 0, s>>>(count, list, make_src<SU>(u, sl), g, fp, F, G);
    }
  }
  {
    TimedLaunch tl(TK_FINAL, s);
    // the learned stage's counters ride on the final kernel (after the fallback, which reads the same counts)
    long long* stats = apost && use_pred ? pred->stats : nullptr;
    k_final<TWO_D, T><<<gridc, TPB, 0, s>>>(F, G, g, dudt, rk, count, stats, pre_now ? 1 : 0);
  }
  // join: the caller's stream waits for the network before anything after this call
  if (overlap && pred->infer && hipStreamWaitEvent(s, (hipEvent_t)pred->ev_join, 0) != hipSuccess)
    return FVHOE_ERR_LAUNCH;
  return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
}

template <typename T, bool TWO_D, bool DIR>
int dispatch_p(int p, const FvhoeStageConfig* c, const FvhoeRkFuse* rk, const FvhoePredict* pre, const Layout& L,
               const double* u, double* dudt, double* dt, signed char* t, char* ws, hipStream_t s) {
  switch (p) {
    case 1: return run_stage<T, 1, TWO_D, DIR>(c, rk, pre, L, u, dudt, dt, t, ws, s);
    case 2: return run_stage<T, 2, TWO_D, DIR>(c, rk, pre, L, u, dudt, dt, t, ws, s);
    case 3: return run_stage<T, 3, TWO_D, DIR>(c, rk, pre, L, u, dudt, dt, t, ws, s);
    case 4: return run_stage<T, 4, TWO_D, DIR>(c, rk, pre, L, u, dudt, dt, t, ws, s);
    case 5: return run_stage<T, 5, TWO_D, DIR>(c, rk, pre, L, u, dudt, dt, t, ws, s);
    case 6: return run_stage<T, 6, TWO_D, DIR>(c, rk, pre, L, u, dudt, dt, t, ws, s);
    case 7: return run_stage<T, 7, TWO_D, DIR>(c, rk, pre, L, u, dudt, dt, t, ws, s);
    case 8: return run_stage<T, 8, TWO_D, DIR>(c, rk, pre, L, u, dudt, dt, t, ws, s);
  }
  return FVHOE_ERR_UNSUPPORTED;
}

template <typename T, bool TWO_D>
int dispatch(int p, const FvhoeStageConfig* c, const FvhoeRkFuse* rk, const FvhoePredict* pre, const Layout& L,
             const double* u, double* dudt, double* dt, signed char* t, char* ws, hipStream_t s) {
  if (L.dirichlet) return dispatch_p<T, TWO_D, true>(p, c, rk, pre, L, u, dudt, dt, t, ws, s);
  return dispatch_p<T, TWO_D, false>(p, c, rk, pre, L, u, dudt, dt, t, ws, s);
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

namespace {

int stage_entry(const FvhoeStageConfig* cfg, const double* u, double* dudt, double* dt_stage, signed char* trouble,
                void* workspace, const FvhoeRkFuse* rk, const FvhoePredict* pred, void* stream) {
  Layout L;
  const int st = make_layout(cfg, &L, pred != nullptr);
  if (st != FVHOE_OK) return st;
  if (pred && !pred->packed) return FVHOE_ERR_ARG;
  if (pred && pred->overlap && (!pred->side_stream || !pred->ev_fork || !pred->ev_join)) return FVHOE_ERR_ARG;
  if (!u || !dt_stage || !workspace) return FVHOE_ERR_ARG;
  if (!dudt && !(rk && rk->kind >= 0)) return FVHOE_ERR_ARG;
  if (rk && rk->kind >= 0) {
    if (rk->kind > FVHOE_RK_SSPRK3_FINAL_SUM || !rk->out || !rk->u0 || !rk->dt_step) return FVHOE_ERR_ARG;
    if ((rk->kind == FVHOE_RK_SSPRK3_STAGE3 || rk->kind == FVHOE_RK_SSPRK2_FINAL ||
         rk->kind == FVHOE_RK_SSPRK3_FINAL_SUM) && !rk->k0)
      return FVHOE_ERR_ARG;
    if (rk->kind == FVHOE_RK_SSPRK3_STAGE3_SUM && (!rk->k0 || !dudt)) return FVHOE_ERR_ARG;
    if (rk->kind == FVHOE_RK_SSPRK3_FINAL && (!rk->k0 || !rk->k1)) return FVHOE_ERR_ARG;
    if (rk->kind == FVHOE_RK_RK4_FINAL && (!rk->k0 || !rk->k1 || !rk->k2)) return FVHOE_ERR_ARG;
    if (rk->out == dudt) return FVHOE_ERR_ARG;
  }
  if (rk && rk->set_dt_step && (!rk->dt_step || !rk->t)) return FVHOE_ERR_ARG;
  if (((uintptr_t)workspace & 255) != 0) return FVHOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  if (L.fp32)
    return L.two_d ? dispatch<float, true>(L.p, cfg, rk, pred, L, u, dudt, dt_stage, trouble, ws, s)
                   : dispatch<float, false>(L.p, cfg, rk, pred, L, u, dudt, dt_stage, trouble, ws, s);
  return L.two_d ? dispatch<double, true>(L.p, cfg, rk, pred, L, u, dudt, dt_stage, trouble, ws, s)
                 : dispatch<double, false>(L.p, cfg, rk, pred, L, u, dudt, dt_stage, trouble, ws, s);
}

}  // namespace

int fvhoe_stage_rk(const FvhoeStageConfig* cfg, const double* u, double* dudt, double* dt_stage,
                   signed char* trouble, void* workspace, const FvhoeRkFuse* rk, void* stream) {
  return stage_entry(cfg, u, dudt, dt_stage, trouble, workspace, rk, nullptr, stream);
}

int fvhoe_stage_ghost_depth(const FvhoeStageConfig* cfg, int* depth) {
  if (!cfg || !depth) return FVHOE_ERR_ARG;
  if (cfg->p[0] < 1 || cfg->p[0] > 8) return FVHOE_ERR_UNSUPPORTED;
  *depth = ghost_depth(cfg->p[0]);
  return FVHOE_OK;
}

int fvhoe_event_create(void** event) {
  if (!event) return FVHOE_ERR_ARG;
  hipEvent_t e;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return FVHOE_ERR_LAUNCH;
  *event = (void*)e;
  return FVHOE_OK;
}

int fvhoe_event_destroy(void* event) {
  if (!event) return FVHOE_ERR_ARG;
  return hipEventDestroy((hipEvent_t)event) == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
}

int fvhoe_stage_predicted_workspace_bytes(const FvhoeStageConfig* cfg, unsigned long long* bytes) {
  Layout L;
  const int st = make_layout(cfg, &L, true);
  if (st != FVHOE_OK) return st;
  if (!bytes) return FVHOE_ERR_ARG;
  *bytes = L.total;
  return FVHOE_OK;
}

int fvhoe_stage_predicted(const FvhoeStageConfig* cfg, const double* u, double* dudt, double* dt_stage,
                          signed char* trouble, void* workspace, const FvhoeRkFuse* rk, const FvhoePredict* pred,
                          void* stream) {
  if (!pred) return FVHOE_ERR_ARG;
  return stage_entry(cfg, u, dudt, dt_stage, trouble, workspace, rk, pred, stream);
}

int fvhoe_pred_features(const FvhoeStageConfig* cfg, const void* workspace, float* X, void* stream) {
  Layout L;
  const int st = make_layout(cfg, &L);
  if (st != FVHOE_OK) return st;
  if (!workspace || !X) return FVHOE_ERR_ARG;
  if (!L.two_d || !cfg->a_posteriori || cfg->nad_mode != 0 || L.dirichlet) return FVHOE_ERR_UNSUPPORTED;
  const char* ws = (const char*)workspace;
  Geo g{};
  g.n0 = L.n0;
  g.n1 = L.n1;
  g.ge0 = L.GE;
  g.ge1 = L.GE;
  for (int k = 0; k < 4; ++k) g.bc[k] = cfg->bc[k];
  const double* range = (const double*)(ws + L.offScal);
  const int nx = (L.n0 + 12) * (L.n1 + 12);
  hipStream_t s = (hipStream_t)stream;
  if (L.fp32)
    k_features<float><<<(nx + TPB - 1) / TPB, TPB, 0, s>>>((const float*)(ws + L.offW), L.NW0, L.NW1, g, range, X);
  else
    k_features<double><<<(nx + TPB - 1) / TPB, TPB, 0, s>>>((const double*)(ws + L.offW), L.NW0, L.NW1, g, range, X);
  return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
}

int fvhoe_stage(const FvhoeStageConfig* cfg, const double* u, double* dudt, double* dt_stage, signed char* trouble,
                void* workspace, void* stream) {
  if (!dudt) return FVHOE_ERR_ARG;
  return fvhoe_stage_rk(cfg, u, dudt, dt_stage, trouble, workspace, nullptr, stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The stage without a caller workspace: its W / F / G / WS arrays, flags, cell
// list, reduction partials and counters live in a file-scope device array of
// FVHOE_STATIC_SCRATCH_BYTES.  Each GPU the library runs on holds its own copy
// (allocated when the code object loads, kept until the process ends); host
// code reaches it only through the per-device address hipGetSymbolAddress
// returns.  Nothing in it must start at zero: every stage writes what it reads
// (k_stage_reduce clears the stage's counters).  All calls on one GPU share the
// array, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace {

static __device__ __align__(256) char scratch[FVHOE_STATIC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[MAX_DEVICES] = {};

// this device's address of the scratch array, looked up once per device with that device current
int scratch_for_current_device(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return FVHOE_ERR_LAUNCH;
  if (dev < 0 || dev >= MAX_DEVICES) return FVHOE_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FVHOE_ERR_LAUNCH;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return FVHOE_OK;
}

}  // namespace

extern "C" {

int fvhoe_stage_rk_static(const FvhoeStageConfig* cfg, const double* u, double* dudt, double* dt_stage,
                          signed char* trouble, const FvhoeRkFuse* rk, void* stream) {
  Layout L;
  const int st = make_layout(cfg, &L);
  if (st != FVHOE_OK) return st;
  if (L.total > (size_t)FVHOE_STATIC_SCRATCH_BYTES) return FVHOE_ERR_TOO_LARGE;
  void* ws = nullptr;
  const int sst = scratch_for_current_device(&ws);
  if (sst != FVHOE_OK) return sst;
  return stage_entry(cfg, u, dudt, dt_stage, trouble, ws, rk, nullptr, stream);
}

}  // extern "C"
