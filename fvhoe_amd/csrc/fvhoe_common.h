// Shared device code for the FVHOE per-stage kernels (gfx950).
//
// Every arithmetic expression below keeps the reference's floating-point
// operation order (the NumPy expressions cited next to each function), and
// the library is compiled with -ffp-contract=off, so FP64 results match the
// reference bit for bit.  Division and sqrt are the compiler's IEEE
// correctly rounded FP64 operations.
#pragma once
#include <type_traits>

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <utility>

#include "fvhoe_hip.h"

namespace fvhoe {

// ---------------------------------------------------------------------------
// stencil weights (fvhoe/stencils.py:25-255).  Each entry is a rational
// written as num/den and folded at compile time with round-to-nearest, which
// is the double the reference's `num / den` literals evaluate to.
// kind: 0 = left face "l", 1 = centre "c", 2 = right face "r" (reversed l),
//       3 = transverse (averages from centre values).
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int reach(int p) { return (p + 1) / 2; }
__host__ __device__ constexpr int ntaps(int p) { return 2 * reach(p) + 1; }

__host__ __device__ constexpr double w_left(int p, int i) {
  switch (p) {
    case 0: return 1.0;
    case 1: { const double w[] = {1.0 / 4.0, 1.0, -1.0 / 4.0}; return w[i]; }
    case 2: { const double w[] = {1.0 / 3.0, 5.0 / 6.0, -1.0 / 6.0}; return w[i]; }
    case 3: { const double w[] = {-1.0 / 24.0, 5.0 / 12.0, 5.0 / 6.0, -1.0 / 4.0, 1.0 / 24.0}; return w[i]; }
    case 4: { const double w[] = {-1.0 / 20.0, 9.0 / 20.0, 47.0 / 60.0, -13.0 / 60.0, 1.0 / 30.0}; return w[i]; }
    case 5: { const double w[] = {1.0 / 120.0, -1.0 / 12.0, 59.0 / 120.0, 47.0 / 60.0, -31.0 / 120.0, 1.0 / 15.0,
                                  -1.0 / 120.0}; return w[i]; }
    case 6: { const double w[] = {1.0 / 105.0, -19.0 / 210.0, 107.0 / 210.0, 319.0 / 420.0, -101.0 / 420.0,
                                  5.0 / 84.0, -1.0 / 140.0}; return w[i]; }
    case 7: { const double w[] = {-1.0 / 560.0, 17.0 / 840.0, -97.0 / 840.0, 449.0 / 840.0, 319.0 / 420.0,
                                  -223.0 / 840.0, 71.0 / 840.0, -1.0 / 56.0, 1.0 / 560.0}; return w[i]; }
    case 8: { const double w[] = {-1.0 / 504.0, 11.0 / 504.0, -61.0 / 504.0, 275.0 / 504.0, 1879.0 / 2520.0,
                                  -641.0 / 2520.0, 199.0 / 2520.0, -41.0 / 2520.0, 1.0 / 630.0}; return w[i]; }
  }
  return 0.0;
}

__host__ __device__ constexpr double w_centre(int p, int i) {
  switch (p) {
    case 0: return 1.0;
    case 1: { const double w[] = {0.0, 1.0, 0.0}; return w[i]; }
    case 2: { const double w[] = {-1.0 / 24.0, 13.0 / 12.0, -1.0 / 24.0}; return w[i]; }
    case 3: { const double w[] = {0.0, -1.0 / 24.0, 13.0 / 12.0, -1.0 / 24.0, 0.0}; return w[i]; }
    case 4: { const double w[] = {3.0 / 640.0, -29.0 / 480.0, 1067.0 / 960.0, -29.0 / 480.0, 3.0 / 640.0}; return w[i]; }
    case 5: { const double w[] = {0.0, 3.0 / 640.0, -29.0 / 480.0, 1067.0 / 960.0, -29.0 / 480.0, 3.0 / 640.0, 0.0};
              return w[i]; }
    case 6: { const double w[] = {-5.0 / 7168.0, 159.0 / 17920.0, -7621.0 / 107520.0, 30251.0 / 26880.0,
                                  -7621.0 / 107520.0, 159.0 / 17920.0, -5.0 / 7168.0}; return w[i]; }
    case 7: { const double w[] = {0.0, -5.0 / 7168.0, 159.0 / 17920.0, -7621.0 / 107520.0, 30251.0 / 26880.0,
                                  -7621.0 / 107520.0, 159.0 / 17920.0, -5.0 / 7168.0, 0.0}; return w[i]; }
    case 8: { const double w[] = {35.0 / 294912.0, -425.0 / 258048.0, 31471.0 / 2580480.0, -100027.0 / 1290240.0,
                                  5851067.0 / 5160960.0, -100027.0 / 1290240.0, 31471.0 / 2580480.0,
                                  -425.0 / 258048.0, 35.0 / 294912.0}; return w[i]; }
  }
  return 0.0;
}

__host__ __device__ constexpr double w_transverse(int p, int i) {
  switch (p) {
    case 0: return 1.0;
    case 1: { const double w[] = {0.0, 1.0, 0.0}; return w[i]; }
    case 2: { const double w[] = {1.0 / 24.0, 11.0 / 12.0, 1.0 / 24.0}; return w[i]; }
    case 3: { const double w[] = {0.0, 1.0 / 24.0, 11.0 / 12.0, 1.0 / 24.0, 0.0}; return w[i]; }
    case 4: { const double w[] = {-17.0 / 5760.0, 77.0 / 1440.0, 863.0 / 960.0, 77.0 / 1440.0, -17.0 / 5760.0};
              return w[i]; }
    case 5: { const double w[] = {0.0, -17.0 / 5760.0, 77.0 / 1440.0, 863.0 / 960.0, 77.0 / 1440.0, -17.0 / 5760.0,
                                  0.0}; return w[i]; }
    case 6: { const double w[] = {367.0 / 967680.0, -281.0 / 53760.0, 6361.0 / 107520.0, 215641.0 / 241920.0,
                                  6361.0 / 107520.0, -281.0 / 53760.0, 367.0 / 967680.0}; return w[i]; }
    case 7: { const double w[] = {0.0, 367.0 / 967680.0, -281.0 / 53760.0, 6361.0 / 107520.0, 215641.0 / 241920.0,
                                  6361.0 / 107520.0, -281.0 / 53760.0, 367.0 / 967680.0, 0.0}; return w[i]; }
    case 8: { const double w[] = {-27859.0 / 464486400.0, 49879.0 / 58060800.0, -801973.0 / 116121600.0,
                                  3629953.0 / 58060800.0, 41208059.0 / 46448640.0, 3629953.0 / 58060800.0,
                                  -801973.0 / 116121600.0, 49879.0 / 58060800.0, -27859.0 / 464486400.0};
              return w[i]; }
  }
  return 0.0;
}

enum { KIND_L = 0, KIND_C = 1, KIND_R = 2, KIND_T = 3 };

__host__ __device__ constexpr double weight(int kind, int p, int i) {
  return kind == KIND_L   ? w_left(p, i)
         : kind == KIND_R ? w_left(p, ntaps(p) - 1 - i)
         : kind == KIND_C ? w_centre(p, i)
                          : w_transverse(p, i);
}

// compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// out = sum_i w_i * x(i), accumulated from +0.0 left to right, zero taps
// skipped (bitwise neutral for finite data: the accumulator is never -0.0
// once it starts from +0.0), as fvhoe/fv.py:135-139.
template <int KIND, int P, typename Get>
__device__ __forceinline__ double sweep(Get&& x) {
  double acc = 0.0;
  static_for<ntaps(P)>([&](auto I) {
    constexpr double w = weight(KIND, P, decltype(I)::value);
    if constexpr (w != 0.0) acc = acc + w * x(decltype(I)::value);
  });
  return acc;
}

// ---------------------------------------------------------------------------
// thermodynamics (fvhoe/hydro.py:146-210); w = (rho, vx, vy, vz, P).
// T is the arithmetic type: double for the reference-exact path, float for
// the FP32 bulk form (where pressure recovery still runs in double).
// ---------------------------------------------------------------------------
// ONE_DIV (the FP32 bulk forms, whose recovery still runs in FP64): the velocities as products with one
// reciprocal of rho instead of three quotients
template <bool ONE_DIV = false, typename T>
__device__ __forceinline__ void prims_from_cons(const T u[5], T w[5], T gamma) {
  const T rho = u[0];
  w[0] = rho;
  if constexpr (ONE_DIV) {
    const T inv = T(1) / rho;
    w[1] = u[1] * inv;
    w[2] = u[2] * inv;
    w[3] = u[3] * inv;
  } else {
    w[1] = u[1] / rho;
    w[2] = u[2] / rho;
    w[3] = u[3] / rho;
  }
  const T msum = ((u[1] * w[1]) + (u[2] * w[2])) + (u[3] * w[3]);
  w[4] = (gamma - T(1)) * (u[4] - T(0.5) * msum);
}

// Arithmetic of the Riemann solvers.  FAST (the fused stage's FP32 bulk form only): float quotients as a
// product with the hardware reciprocal (v_rcp_f32, 1 ulp) and square roots as v_sqrt_f32 (1 ulp) instead
// of the correctly rounded multi-instruction sequences, and fused multiply-adds in the products' sums.
// Every FP64 path, and the composed FP32 form that reproduces the oracle's float32 arithmetic, keeps
// FAST = false: IEEE quotients and square roots, no contraction (the library's -ffp-contract=off).
template <bool FAST, typename T>
__device__ __forceinline__ T q_div(T a, T b) {
  if constexpr (FAST && std::is_same<T, float>::value) return a * __builtin_amdgcn_rcpf(b);
  else return a / b;
}
template <bool FAST, typename T>
__device__ __forceinline__ T q_sqrt(T x) {
  if constexpr (FAST && std::is_same<T, float>::value) return __builtin_amdgcn_sqrtf(x);
  else return sqrt(x);
}
template <bool FAST, typename T>
__device__ __forceinline__ T q_fma(T a, T b, T c) {  // a * b + c
  if constexpr (FAST && std::is_same<T, float>::value) return __builtin_fmaf(a, b, c);
  else return a * b + c;
}

// total energy of compute_conservatives: P/(gamma-1) + 0.5*sum(m*v), m = rho*v
template <bool FAST = false, typename T>
__device__ __forceinline__ T energy(const T w[5], T gamma) {
  const T mx = w[0] * w[1], my = w[0] * w[2], mz = w[0] * w[3];
  return q_div<FAST>(w[4], gamma - T(1)) + T(0.5) * q_fma<FAST>(mz, w[3], q_fma<FAST>(my, w[2], mx * w[1]));
}

template <bool FAST = false, typename T>
__device__ __forceinline__ T sound_speed(T rho, T P, T gamma, T csq_floor) {
  const T csq = q_div<FAST>(gamma * P, rho);
  return q_sqrt<FAST>(csq > csq_floor ? csq : csq_floor);
}

// max / min as the hardware v_max / v_min: for non-NaN operands they equal
// np.maximum / np.minimum except possibly in the sign of a zero result, which
// no caller's result depends on (differences, comparisons, alpha < 1 tests).
template <typename T>
__device__ __forceinline__ T dmax(T a, T b) { return fmax(a, b); }
template <typename T>
__device__ __forceinline__ T dmin(T a, T b) { return fmin(a, b); }

// np.minimum / np.maximum exactly (NumPy 2.x on x86-64): a NaN operand
// propagates, and equal operands (+0 and -0) give the second one.  Used where
// the value itself is an output of a drop-in function.
template <typename T>
__device__ __forceinline__ T np_min(T a, T b) { return a != a ? a : (a < b ? a : b); }
template <typename T>
__device__ __forceinline__ T np_max(T a, T b) { return a != a ? a : (a > b ? a : b); }

// ---------------------------------------------------------------------------
// Riemann solvers (fvhoe/riemann_solvers.py).  D = normal axis 0/1/2;
// wl is the state left of the face, wr right of it.
// ---------------------------------------------------------------------------
enum { RS_UPWIND = 0, RS_LLF = 1, RS_HLLC = 2 };

// HLLC "David variation", fvhoe/riemann_solvers.py:110-183
template <int D, bool FAST = false, typename T>
__device__ __forceinline__ void hllc(const T wl[5], const T wr[5], T F[5], T gamma, T csq_floor,
                                     T* v_godunov = nullptr) {
  const T El = energy<FAST>(wl, gamma), Er = energy<FAST>(wr, gamma);
  const T cl = sound_speed<FAST>(wl[0], wl[4], gamma, csq_floor);
  const T cr = sound_speed<FAST>(wr[0], wr[4], gamma, csq_floor);
  const T cmax = dmax(cl, cr);
  const T vl = wl[D + 1], vr = wr[D + 1];
  const T rhol = wl[0], rhor = wr[0], Pl = wl[4], Pr = wr[4];
  const T Sl = dmin(vl, vr) - cmax;
  const T Sr = dmax(vl, vr) + cmax;
  const T rcL = rhol * (vl - Sl);
  const T rcR = rhor * (Sr - vr);
  T vs, Ps;
  if constexpr (FAST && std::is_same<T, float>::value) {  // one reciprocal for both quotients
    const T inv = __builtin_amdgcn_rcpf(rcL + rcR);
    vs = (q_fma<FAST>(rcL, vl, rcR * vr) + Pl - Pr) * inv;
    Ps = q_fma<FAST>(rcL * rcR, vl - vr, q_fma<FAST>(rcL, Pr, rcR * Pl)) * inv;
  } else {
    vs = (rcL * vl + rcR * vr + Pl - Pr) / (rcL + rcR);
    Ps = (rcL * Pr + rcR * Pl + rcL * rcR * (vl - vr)) / (rcL + rcR);
  }
  T r, v, P, E;
  if (Sl > T(0)) {
    r = rhol; v = vl; P = Pl; E = El;
  } else if (vs > T(0)) {
    if constexpr (FAST && std::is_same<T, float>::value) {
      const T inv = __builtin_amdgcn_rcpf(Sl - vs);
      r = rhol * (Sl - vl) * inv;
      E = q_fma<FAST>(Ps, vs, q_fma<FAST>(Sl - vl, El, -(Pl * vl))) * inv;
    } else {
      r = rhol * (Sl - vl) / (Sl - vs);
      E = ((Sl - vl) * El - Pl * vl + Ps * vs) / (Sl - vs);
    }
    v = vs; P = Ps;
  } else if (Sr > T(0)) {
    if constexpr (FAST && std::is_same<T, float>::value) {
      const T inv = __builtin_amdgcn_rcpf(Sr - vs);
      r = rhor * (Sr - vr) * inv;
      E = q_fma<FAST>(Ps, vs, q_fma<FAST>(Sr - vr, Er, -(Pr * vr))) * inv;
    } else {
      r = rhor * (Sr - vr) / (Sr - vs);
      E = ((Sr - vr) * Er - Pr * vr + Ps * vs) / (Sr - vs);
    }
    v = vs; P = Ps;
  } else {
    r = rhor; v = vr; P = Pr; E = Er;
  }
  const T rv = r * v;
  F[0] = rv;
#pragma unroll
  for (int k = 1; k < 4; ++k) {
    if (k == D + 1) {
      F[k] = q_fma<FAST>(rv, v, P);
    } else {
      const T up = v > T(0) ? wl[k] : (v < T(0) ? wr[k] : T(0));
      F[k] = rv * up;
    }
  }
  F[4] = v * (E + P);
  if (v_godunov) *v_godunov = v;
}

// physical flux along D (fvhoe/hydro.py:213-246)
template <int D, typename T>
__device__ __forceinline__ void phys_flux(const T w[5], const T u[5], T F[5], bool pressure) {
  const T v = w[D + 1];
  F[0] = v * w[0];
  F[1] = v * u[1];
  F[2] = v * u[2];
  F[3] = v * u[3];
  F[4] = v * u[4];
  if (pressure) {
    F[D + 1] = F[D + 1] + w[4];
    F[4] = F[4] + v * w[4];
  }
}

template <bool FAST = false, typename T>
__device__ __forceinline__ void cons_from_prims(const T w[5], T u[5], T gamma) {
  u[0] = w[0];
  u[1] = w[0] * w[1];
  u[2] = w[0] * w[2];
  u[3] = w[0] * w[3];
  u[4] = q_div<FAST>(w[4], gamma - T(1)) + T(0.5) * q_fma<FAST>(u[3], w[3], q_fma<FAST>(u[2], w[2], u[1] * w[1]));
}

// local Lax-Friedrichs (Toro p.331; fvhoe/riemann_solvers.py:68-88)
template <int D, bool FAST = false, typename T>
__device__ __forceinline__ void llf(const T wl[5], const T wr[5], T F[5], T gamma, T csq_floor) {
  T ul[5], ur[5], Fl[5], Fr[5];
  cons_from_prims<FAST>(wl, ul, gamma);
  cons_from_prims<FAST>(wr, ur, gamma);
  phys_flux<D>(wl, ul, Fl, true);
  phys_flux<D>(wr, ur, Fr, true);
  const T sl = fabs(wl[D + 1]) + sound_speed<FAST>(wl[0], wl[4], gamma, csq_floor);
  const T sr = fabs(wr[D + 1]) + sound_speed<FAST>(wr[0], wr[4], gamma, csq_floor);
  const T hs = T(0.5) * dmax(sl, sr);
#pragma unroll
  for (int k = 0; k < 5; ++k) F[k] = T(0.5) * (Fl[k] + Fr[k]) - hs * (ur[k] - ul[k]);
}

// upwinding without pressure (fvhoe/riemann_solvers.py:31-45)
template <int D, typename T>
__device__ __forceinline__ void upwind(const T wl[5], const T wr[5], T F[5], T gamma) {
  T ul[5], ur[5], Fl[5], Fr[5];
  cons_from_prims(wl, ul, gamma);
  cons_from_prims(wr, ur, gamma);
  phys_flux<D>(wl, ul, Fl, false);
  phys_flux<D>(wr, ur, Fr, false);
  const T v = wl[D + 1];
#pragma unroll
  for (int k = 0; k < 5; ++k) F[k] = v > T(0) ? Fl[k] : (v < T(0) ? Fr[k] : T(0));
}

template <int D, bool FAST = false, typename T>
__device__ __forceinline__ void riemann(int solver, const T wl[5], const T wr[5], T F[5], T gamma, T csq_floor,
                                        T* v_godunov = nullptr) {
  if (solver == RS_HLLC) hllc<D, FAST>(wl, wr, F, gamma, csq_floor, v_godunov);
  else if (solver == RS_LLF) llf<D, FAST>(wl, wr, F, gamma, csq_floor);
  else upwind<D>(wl, wr, F, gamma);
}

// compile-time solver choice
template <int RS, int D, bool FAST = false, typename T>
__device__ __forceinline__ void riemann_t(const T wl[5], const T wr[5], T F[5], T gamma, T csq_floor) {
  if constexpr (RS == RS_HLLC) hllc<D, FAST>(wl, wr, F, gamma, csq_floor);
  else if constexpr (RS == RS_LLF) llf<D, FAST>(wl, wr, F, gamma, csq_floor);
  else upwind<D>(wl, wr, F, gamma);
}

// ---------------------------------------------------------------------------
// limiters (fvhoe/slope_limiting.py:56-84)
// ---------------------------------------------------------------------------
enum { LIM_NONE = 0, LIM_MINMOD = 1, LIM_MONCEN = 2 };

template <typename T>
__device__ __forceinline__ T minmod(T dl, T dr) {
  const T lo = T(1e-16);
  const T den = dl > T(0) ? (dl > lo ? dl : lo) : (dl < -lo ? dl : -lo);
  T r = dr / den;
  r = r < T(1) ? r : T(1);
  return (r > T(0) ? r : T(0)) * dl;
}

template <typename T>
__device__ __forceinline__ T dsign(T x) {  // np.sign: +0 for either zero, NaN stays NaN
  return x > T(0) ? T(1) : (x < T(0) ? T(-1) : (x == x ? T(0) : x));
}

template <typename T>
__device__ __forceinline__ T moncen(T dl, T dr) {
  const T dc = T(0.5) * (dl + dr);
  T s = dmin(fabs(T(2) * dl), fabs(T(2) * dr));
  s = dsign(dc) * dmin(s, fabs(dc));
  return dl * dr >= T(0) ? s : T(0);
}

template <typename T>
__device__ __forceinline__ T limited_slope(int limiter, T dl, T dr) {
  return limiter == LIM_MINMOD ? minmod(dl, dr) : limiter == LIM_MONCEN ? moncen(dl, dr) : T(0.5) * (dl + dr);
}

// ---------------------------------------------------------------------------
// smooth-extrema detector along one axis (fvhoe/smooth_extrema_detection.py:26-49)
// x(o) returns the value at offset o in [-3, 3] from the cell.  EXACT: the
// minima as np.minimum (the alpha value is an output), else hardware min (only
// alpha < 1 is used).
// ---------------------------------------------------------------------------
template <bool EXACT = false, typename Get>
__device__ __forceinline__ double sed_alpha_1d(Get&& x, double eps) {
  auto mn = [](double a, double b) { return EXACT ? np_min(a, b) : dmin(a, b); };
  double dU[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) dU[j] = 0.5 * (x(j - 1) - x(j - 3));  // central differences at offsets -2..2
  double a[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {  // alpha at offsets -1, 0, 1 -> dU index j+1
    double dv = 0.5 * (0.5 * (dU[j + 2] - dU[j]));
    dv = fabs(dv) <= eps ? (dv >= 0.0 ? eps : -eps) : dv;
    const double vL = dU[j] - dU[j + 1];
    double aL = -(dv < 0.0 ? (vL > 0.0 ? vL : 0.0) : (vL < 0.0 ? vL : 0.0)) / dv;
    aL = aL < 1.0 ? aL : 1.0;
    const double vR = dU[j + 2] - dU[j + 1];
    double aR = (dv > 0.0 ? (vR > 0.0 ? vR : 0.0) : (vR < 0.0 ? vR : 0.0)) / dv;
    aR = fabs(dv) <= eps ? 1.0 : aR;
    aR = aR < aL ? aR : aL;
    a[j] = mn(aL, aR);
  }
  return mn(mn(a[2], a[1]), a[0]);
}

// ---------------------------------------------------------------------------
// order-preserving double <-> uint64 maps for atomic min/max reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long ordered_bits(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ double from_ordered_bits(unsigned long long b) {
  const unsigned long long r = (b & 0x8000000000000000ull) ? (b & 0x7fffffffffffffffull) : ~b;
#ifdef __HIP_DEVICE_COMPILE__
  return __longlong_as_double((long long)r);
#else
  double d;
  __builtin_memcpy(&d, &r, 8);
  return d;
#endif
}

// Values of neighbouring lanes of a 64-lane wave through DPP wavefront shifts
// (0x130: lane l receives lane l + 1; 0x138: lane l - 1; lanes shifted in
// from outside the wave read 0).
__device__ __forceinline__ int dpp_next(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_prev(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ float lane_next(float x) { return __int_as_float(dpp_next(__float_as_int(x))); }
__device__ __forceinline__ float lane_prev(float x) { return __int_as_float(dpp_prev(__float_as_int(x))); }
__device__ __forceinline__ double lane_next(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_next((int)(b & 0xffffffffll)), hi = dpp_next((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_prev(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_prev((int)(b & 0xffffffffll)), hi = dpp_prev((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// x of lane l + O
template <int O, typename T>
__device__ __forceinline__ T lane_at(T x) {
  if constexpr (O == 0) return x;
  else if constexpr (O > 0) return lane_at<O - 1>(lane_next(x));
  else return lane_at<O + 1>(lane_prev(x));
}

// boundary kinds
enum { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_REFLECTIVE = 2, BC_FREE = 3, BC_DIRICHLET = 4 };

// Ghost values of the Dirichlet sides ("dirichlet" and "ic" boundaries,
// fvhoe/boundary_conditions.py:18-57,306-348), one slab per side s (x lo, x hi,
// y lo, y hi, z lo, z hi), null elsewhere.  The slab of an axis-d side has
// g[d] layers along d and n[e] + 2*pad[e] cells along each other axis e, C
// order (v, axis 0, axis 1, axis 2); layer k of a lo side is cell k - g[d], of
// a hi side cell n[d] + k (the slab coordinates of fvhoe/fv.py:53-84).
struct Slabs {
  const double* s[6];
  int g[3];
  int pad[3];
};

// Value of variable v at interior-relative cell c (ghosts allowed), composed
// like the reference's fill (x, y, z slabs in turn, so a later axis decides a
// corner): a Dirichlet side of the last axis on which c is a ghost gives its
// slab value directly; the other kinds map c into the interior first.  at(v,
// c0, c1, c2) reads an interior cell.  Reflective ghosts negate the normal
// momentum of every axis mapped after the slab that gave the value.
template <typename At>
__device__ __forceinline__ double bc_fetch3(const int n[3], const int bc[6], const Slabs& sl, int v, int c0, int c1,
                                            int c2, At&& at) {
  int c[3] = {c0, c1, c2};
  bool flip[3] = {false, false, false};
  for (int d = 2; d >= 0; --d) {
    if (c[d] >= 0 && c[d] < n[d]) continue;
    const int side = 2 * d + (c[d] >= 0 ? 1 : 0);
    if (bc[side] == BC_DIRICHLET) {
      const int layer = c[d] < 0 ? sl.g[d] + c[d] : c[d] - n[d];
      long long idx = v;
      for (int e = 0; e < 3; ++e) {
        const int ext = e == d ? sl.g[d] : n[e] + 2 * sl.pad[e];
        idx = idx * ext + (e == d ? layer : c[e] + sl.pad[e]);
      }
      double x = sl.s[side][idx];
      for (int e = d + 1; e < 3; ++e)
        if (flip[e] && v == e + 1) x = -x;
      return x;
    }
    const int lo = bc[2 * d], hi = bc[2 * d + 1];
    if (c[d] < 0) {
      c[d] = lo == BC_PERIODIC ? n[d] + c[d] : lo == BC_FREE ? 0 : -1 - c[d];
      flip[d] = lo == BC_REFLECTIVE;
    } else {
      c[d] = hi == BC_PERIODIC ? c[d] - n[d] : hi == BC_FREE ? n[d] - 1 : 2 * n[d] - 1 - c[d];
      flip[d] = hi == BC_REFLECTIVE;
    }
  }
  double x = at(v, c[0], c[1], c[2]);
  if ((flip[0] && v == 1) || (flip[1] && v == 2) || (flip[2] && v == 3)) x = -x;
  return x;
}

// Source interior index and momentum sign for padded index a on an axis of
// interior size n, pad g, boundary kinds (lo, hi) (fvhoe/boundary_conditions.py:60-151).
__host__ __device__ __forceinline__ int bc_source(int a, int n, int g, int lo, int hi, bool* flip) {
  *flip = false;
  if (a < g) {
    if (lo == BC_PERIODIC) return n - g + a;
    if (lo == BC_FREE) return 0;
    *flip = (lo == BC_REFLECTIVE);
    return g - 1 - a;
  }
  if (a >= n + g) {
    const int b = a - n - g;
    if (hi == BC_PERIODIC) return b;
    if (hi == BC_FREE) return n - 1;
    *flip = (hi == BC_REFLECTIVE);
    return n - 1 - b;
  }
  return a - g;
}

// Same map for a signed interior-relative index c in [-g, n+g).
__host__ __device__ __forceinline__ int bc_map(int c, int n, int g, int lo, int hi, bool* flip) {
  return bc_source(c + g, n, g, lo, hi, flip);
}

// The RK combination of a stage's derivative k_this (fvhoe/ode.py:245-298), kinds FVHOE_RK_*; the
// *_SUM kinds read k0 + k1 stored once (the same rounding as the reference's (k0 + k1)).
struct RkArgs {
  int kind;                     // FVHOE_RK_*, -1 = none
  double* out;
  const double *u0, *k0, *k1, *k2;
  const double* dt;             // step dt (device)
};

__device__ __forceinline__ double rk_combine(const RkArgs& r, int i, double k_this) {
  const double dt = r.dt[0];
  if (dt == 0.0) return r.u0[i];  // a step at its target time leaves the state exactly as it is
  switch (r.kind) {
    case FVHOE_RK_AXPY: return r.u0[i] + dt * k_this;
    case FVHOE_RK_SSPRK3_STAGE3: return r.u0[i] + 0.25 * dt * r.k0[i] + 0.25 * dt * k_this;
    case FVHOE_RK_SSPRK3_FINAL: return r.u0[i] + (1.0 / 6.0) * dt * (r.k0[i] + r.k1[i] + 4.0 * k_this);
    case FVHOE_RK_SSPRK2_FINAL: return 0.5 * r.u0[i] + 0.5 * (r.k0[i] + dt * k_this);
    case FVHOE_RK_HALF_AXPY: return r.u0[i] + 0.5 * dt * k_this;
    case FVHOE_RK_RK4_FINAL: return r.u0[i] + (1.0 / 6.0) * dt * (r.k0[i] + 2.0 * r.k1[i] + 2.0 * r.k2[i] + k_this);
    case FVHOE_RK_SSPRK3_STAGE3_SUM: return r.u0[i] + 0.25 * dt * r.k0[i] + 0.25 * dt * k_this;
    case FVHOE_RK_SSPRK3_FINAL_SUM: return r.u0[i] + (1.0 / 6.0) * dt * (r.k0[i] + 4.0 * k_this);
  }
  return 0.0;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Give each XCD a contiguous run of tiles so neighbouring tiles (which share
// halo data) are served by the same L2: a bijection of [0, nb).
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int per = nb / 8, rem = nb % 8, x = b % 8, k = b / 8;
  return x * per + (x < rem ? x : rem) + k;
}

}  // namespace fvhoe
