// Elementary kernels: one HIP kernel per reference array function, behind
// the C-ABI declared in include/fvhoe_hip.h.  Arrays are contiguous
// (nvars, n0, n1, n2) float64 in the reference's layout.  The sweep, primitive,
// CFL, face-flux, ghost-fill and divergence kernels are also instantiated for
// float32 intermediates (the FP32 bulk form of the composed stage, entries
// ending in _f32): their sums and Riemann solves run in float, the pressure
// recovery, the CFL maximum and the divergence in double.
#include "fvhoe_common.h"
#include "fvhoe_hip.h"

using namespace fvhoe;

namespace {

constexpr int TPB = 256;

inline int grid_for(long long n) {
  long long b = (n + TPB - 1) / TPB;
  if (b > 65536) b = 65536;
  return (int)(b < 1 ? 1 : b);
}

inline int launch_status() { return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH; }

struct Weights9 {
  double w[9];
  int k;
};

// (a, j, k) with idx = (a * m1 + j) * m2 + k, in 32-bit arithmetic when idx fits
// (64-bit integer division is a long software sequence on the GPU)
__device__ __forceinline__ void split3(long long idx, int m1, int m2, int* a, int* j, int* k) {
  if (idx <= 0x7fffffffLL) {
    const unsigned r = (unsigned)idx, q = r / (unsigned)m2, a32 = q / (unsigned)m1;
    *k = (int)(r - q * (unsigned)m2);
    *j = (int)(q - a32 * (unsigned)m1);
    *a = (int)a32;
  } else {
    const long long q = idx / m2, aa = q / m1;
    *k = (int)(idx - q * m2);
    *j = (int)(q - aa * m1);
    *a = (int)aa;
  }
}
// (v, i, j, k) with idx = ((v * m0 + i) * m1 + j) * m2 + k
__device__ __forceinline__ void split4(long long idx, int m0, int m1, int m2, int* v, int* i, int* j, int* k) {
  int a;
  split3(idx, m1, m2, &a, j, k);
  *v = a / m0;
  *i = a - *v * m0;
}

template <typename T>
__global__ void k_sweep(const T* __restrict__ u, T* __restrict__ out, int nv, int n0, int n1, int n2, int axis,
                        Weights9 w) {
  const int m0 = axis == 1 ? n0 - (w.k - 1) : n0;
  const int m1 = axis == 2 ? n1 - (w.k - 1) : n1;
  const int m2 = axis == 3 ? n2 - (w.k - 1) : n2;
  const long long total = (long long)nv * m0 * m1 * m2;
  const long long s = axis == 1 ? (long long)n1 * n2 : axis == 2 ? n2 : 1;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    int v, i, j, k;
    split4(idx, m0, m1, m2, &v, &i, &j, &k);
    const long long base = (((long long)v * n0 + i) * n1 + j) * n2 + k;
    T acc = T(0);
    for (int t = 0; t < w.k; ++t)
      if (w.w[t] != 0.0) acc = acc + T(w.w[t]) * u[base + t * s];
    out[idx] = acc;
  }
}

// The same sweep over a 2-D launch: blockIdx.y = v * m0 + i, x = j * m2 + k
// (32-bit index math, taps unrolled; zero taps skipped as above).
template <typename T>
__global__ void __launch_bounds__(256) k_sweep2(const T* __restrict__ u, T* __restrict__ out, int n0, int n1, int n2,
                                               int axis, Weights9 w, int m0, int m1, int m2) {
  const int jk = blockIdx.x * blockDim.x + threadIdx.x;
  if (jk >= m1 * m2) return;
  const int vi = blockIdx.y;
  const int v = vi / m0, i = vi - (vi / m0) * m0;
  const int j = m2 == 1 ? jk : jk / m2, k = m2 == 1 ? 0 : jk - (jk / m2) * m2;
  const long long base = (((long long)v * n0 + i) * n1 + j) * n2 + k;
  const long long s = axis == 1 ? (long long)n1 * n2 : axis == 2 ? n2 : 1;
  T acc = T(0);
#pragma unroll
  for (int t = 0; t < 9; ++t)
    if (t < w.k && w.w[t] != 0.0) acc = acc + T(w.w[t]) * u[base + t * s];
  out[(long long)vi * (m1 * m2) + jk] = acc;
}

// The same sweep with compile-time weights, each thread sliding a register
// window over R consecutive outputs along the sweep axis: every input is
// loaded (R + K - 1) / R times instead of K times (R = 16 along x and y, 2
// along z).  Strided axes (x, y):
// threads run along the contiguous z axis, so every load is coalesced; the
// contiguous axis (z): each thread walks its own run of z.  Zero taps are
// skipped and the accumulation order is the reference's (bit-identical).
template <typename T, int KIND, int P, int R, bool ZAX>
__global__ void __launch_bounds__(256) k_sweep_slide(const T* __restrict__ u, T* __restrict__ out, int n0, int n1,
                                                     int n2, int axis, int nblk_fast, int nb_axis) {
  constexpr int K = ntaps(P), NW = R + K - 1;
  const int na = axis == 1 ? n0 : axis == 2 ? n1 : n2, ma = na - (K - 1);
  const int m0 = axis == 1 ? ma : n0, m1 = axis == 2 ? ma : n1, m2 = axis == 3 ? ma : n2;
  int v, i, j, k, a0;
  if (!ZAX) {  // threads over the (other axis, z) plane, z fastest; blocks: window along the axis fastest,
    // then the plane, then v (neighbouring windows, which share K - 1 input rows, on one XCD)
    const unsigned blk = (unsigned)xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int ab = (int)(blk % (unsigned)nb_axis);
    const unsigned rest = blk / (unsigned)nb_axis;
    const int no = axis == 1 ? n1 : n0;
    const int t = (int)(rest % (unsigned)nblk_fast) * 256 + (int)threadIdx.x;
    v = (int)(rest / (unsigned)nblk_fast);
    if (t >= no * n2) return;
    const int o = t / n2;
    k = t - o * n2;
    a0 = ab * R;
    i = axis == 1 ? a0 : o;
    j = axis == 1 ? o : a0;
  } else {  // fast index: block of R along z; slow: (v, i, j)
    const unsigned g = blockIdx.x * 256u + threadIdx.x;
    const int kb = (int)(g % (unsigned)nb_axis);
    unsigned rest = g / (unsigned)nb_axis;
    j = (int)(rest % (unsigned)n1);
    rest /= (unsigned)n1;
    i = (int)(rest % (unsigned)n0);
    v = (int)(rest / (unsigned)n0);
    if (v >= nblk_fast) return;  // nblk_fast carries nv here
    a0 = kb * R;
    k = a0;
  }
  const int s = axis == 1 ? n1 * n2 : axis == 2 ? n2 : 1;
  const int in0 = ((v * n0 + i) * n1 + j) * n2 + k;
  T win[NW];
#pragma unroll
  for (int r = 0; r < NW; ++r) win[r] = a0 + r < na ? u[in0 + r * s] : T(0);
  const int os = axis == 1 ? m1 * m2 : axis == 2 ? m2 : 1;
  const int o0 = ((v * m0 + (axis == 1 ? a0 : i)) * m1 + (axis == 2 ? a0 : j)) * m2 + (axis == 3 ? a0 : k);
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (a0 + q < ma) {
      T acc = T(0);
      static_for<K>([&](auto I) {
        constexpr double wt = weight(KIND, P, decltype(I)::value);
        if constexpr (wt != 0.0) acc = acc + T(wt) * win[q + decltype(I)::value];
      });
      out[o0 + q * os] = acc;
    }
  }
}

// outputs per thread along x / y and along z (measured on the 3-D p=7 wave; overridable for a sweep)
#ifndef FVHOE_SLIDE_R
#define FVHOE_SLIDE_R 16
#endif
#ifndef FVHOE_SLIDE_RZ
#define FVHOE_SLIDE_RZ 2
#endif

template <typename T, int KIND, int P>
int launch_slide(const T* u, T* out, int nv, int n0, int n1, int n2, int axis, hipStream_t st) {
  constexpr int K = ntaps(P);
  const int na = axis == 1 ? n0 : axis == 2 ? n1 : n2, ma = na - (K - 1);
  if (axis != 3) {
    constexpr int R = FVHOE_SLIDE_R;
    const long long plane = (long long)(axis == 1 ? n1 : n0) * n2;
    const int nbp = (int)((plane + 255) / 256), nb = (ma + R - 1) / R;
    const long long blocks = (long long)nbp * nb * nv;
    if (blocks >= (1ll << 31) || plane >= (1ll << 31)) return -1;
    k_sweep_slide<T, KIND, P, R, false><<<(unsigned)blocks, 256, 0, st>>>(u, out, n0, n1, n2, axis, nbp, nb);
  } else {
    constexpr int R = FVHOE_SLIDE_RZ;
    const int nb = (ma + R - 1) / R;
    const long long threads = (long long)nb * n1 * n0 * nv;
    if (threads >= (1ll << 31)) return -1;
    k_sweep_slide<T, KIND, P, R, true><<<(unsigned)((threads + 255) / 256), 256, 0, st>>>(u, out, n0, n1, n2, axis, nv, nb);
  }
  return 0;
}

template <typename T, int KIND>
int launch_slide_p(int p, const T* u, T* out, int nv, int n0, int n1, int n2, int axis, hipStream_t st) {
  switch (p) {
    case 0: return launch_slide<T, KIND, 0>(u, out, nv, n0, n1, n2, axis, st);
    case 1: return launch_slide<T, KIND, 1>(u, out, nv, n0, n1, n2, axis, st);
    case 2: return launch_slide<T, KIND, 2>(u, out, nv, n0, n1, n2, axis, st);
    case 3: return launch_slide<T, KIND, 3>(u, out, nv, n0, n1, n2, axis, st);
    case 4: return launch_slide<T, KIND, 4>(u, out, nv, n0, n1, n2, axis, st);
    case 5: return launch_slide<T, KIND, 5>(u, out, nv, n0, n1, n2, axis, st);
    case 6: return launch_slide<T, KIND, 6>(u, out, nv, n0, n1, n2, axis, st);
    case 7: return launch_slide<T, KIND, 7>(u, out, nv, n0, n1, n2, axis, st);
    case 8: return launch_slide<T, KIND, 8>(u, out, nv, n0, n1, n2, axis, st);
  }
  return -1;
}

template <typename T>
__global__ void k_prims(const T* __restrict__ u, T* __restrict__ w, int nv, long long nc, double gamma) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < nc; c += (long long)gridDim.x * blockDim.x) {
    double uu[5], ww[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) uu[q] = (double)u[q * nc + c];
    prims_from_cons(uu, ww, gamma);
#pragma unroll
    for (int q = 0; q < 5; ++q) w[q * nc + c] = T(ww[q]);
    for (int q = 5; q < nv; ++q) w[q * nc + c] = T((double)u[q * nc + c] / uu[0]);
  }
}

__global__ void k_cons(const double* __restrict__ w, double* __restrict__ u, int nv, long long nc, double gamma) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < nc; c += (long long)gridDim.x * blockDim.x) {
    double uu[5], ww[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) ww[q] = w[q * nc + c];
    cons_from_prims(ww, uu, gamma);
#pragma unroll
    for (int q = 0; q < 5; ++q) u[q * nc + c] = uu[q];
    for (int q = 5; q < nv; ++q) u[q * nc + c] = ww[0] * w[q * nc + c];
  }
}

__global__ void k_sound(const double* __restrict__ w, double* __restrict__ c, long long nc, double gamma,
                        double csq_floor) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nc; i += (long long)gridDim.x * blockDim.x)
    c[i] = sound_speed(w[i], w[4 * nc + i], gamma, csq_floor);
}

__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// block max of a non-negative double, one atomic per block (on the raw bits:
// the IEEE order of non-negative doubles is the unsigned order of their bits)
__device__ __forceinline__ void block_atomic_max_pos(double x, unsigned long long* acc) {
  __shared__ double red[TPB / 64];
  for (int off = 32; off > 0; off >>= 1) x = np_max(x, __shfl_down(x, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = np_max(m, red[i]);
    atomicMax(acc, (unsigned long long)__double_as_longlong(m));
  }
}

template <typename T>
__global__ void k_dt_partial(const T* __restrict__ w, long long nc, int ndim, double gamma, double csq_floor,
                             unsigned long long* acc) {
  double m = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nc; i += (long long)gridDim.x * blockDim.x) {
    const double c = sound_speed((double)w[i], (double)w[4 * nc + i], gamma, csq_floor);
    const double s = (fabs((double)w[nc + i]) + fabs((double)w[2 * nc + i])) + fabs((double)w[3 * nc + i]);
    m = np_max(m, s + (double)ndim * c);
  }
  block_atomic_max_pos(m, acc);
}

__global__ void k_dt_final(const unsigned long long* acc, double cfl_h, double* dt) {
  dt[0] = cfl_h / __longlong_as_double((long long)acc[0]);
}

template <int D>
__global__ void k_riemann_pt(const double* __restrict__ wl, const double* __restrict__ wr, double* __restrict__ out,
                             int nv, long long nc, int solver, double gamma, double csq_floor) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < nc; c += (long long)gridDim.x * blockDim.x) {
    double a[5], b[5], F[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) { a[q] = wl[q * nc + c]; b[q] = wr[q * nc + c]; }
    double vg = 0.0;
    riemann<D>(solver, a, b, F, gamma, csq_floor, &vg);
#pragma unroll
    for (int q = 0; q < 5; ++q) out[q * nc + c] = F[q];
    if (nv > 5) {
      const double rv = F[0];
      if (solver == RS_HLLC) {  // passives upwinded by the Godunov velocity (fvhoe/riemann_solvers.py:186-195)
        const double v = vg;
        for (int q = 5; q < nv; ++q) {
          const double up = v > 0.0 ? wl[q * nc + c] : (v < 0.0 ? wr[q * nc + c] : 0.0);
          out[q * nc + c] = rv * up;
        }
      } else {
        // llf / upwind on passives: conservative rho*s, flux v*rho*s
        const double vl = a[D + 1], vr = b[D + 1];
        for (int q = 5; q < nv; ++q) {
          const double ul = a[0] * wl[q * nc + c], ur = b[0] * wr[q * nc + c];
          const double fl = vl * ul, fr = vr * ur;
          if (solver == RS_LLF) {
            const double sl = fabs(vl) + sound_speed(a[0], a[4], gamma, csq_floor);
            const double sr = fabs(vr) + sound_speed(b[0], b[4], gamma, csq_floor);
            out[q * nc + c] = 0.5 * (fl + fr) - 0.5 * np_max(sl, sr) * (ur - ul);
          } else {
            out[q * nc + c] = vl > 0.0 ? fl : (vl < 0.0 ? fr : 0.0);
          }
        }
      }
    }
  }
}

// out[f] = RS(wr_face[f], wl_face[f + e_axis]); out shape = face arrays' shape with axis shortened by 1
template <int D>
__global__ void k_riemann_faces(const double* __restrict__ wr_face, const double* __restrict__ wl_face,
                                double* __restrict__ out, int n0, int n1, int n2, int axis, int solver, double gamma,
                                double csq_floor) {
  const int m0 = axis == 1 ? n0 - 1 : n0, m1 = axis == 2 ? n1 - 1 : n1, m2 = axis == 3 ? n2 - 1 : n2;
  const long long nin = (long long)n0 * n1 * n2, nout = (long long)m0 * m1 * m2;
  const long long s = axis == 1 ? (long long)n1 * n2 : axis == 2 ? n2 : 1;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < nout;
       c += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    split3(c, m1, m2, &i, &j, &k);
    const long long base = ((long long)i * n1 + j) * n2 + k;
    double a[5], b[5], F[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) { a[q] = wr_face[q * nin + base]; b[q] = wl_face[q * nin + base + s]; }
    riemann<D>(solver, a, b, F, gamma, csq_floor);
#pragma unroll
    for (int q = 0; q < 5; ++q) out[q * nout + c] = F[q];
  }
}

// The face states and the Riemann solve in one pass: for face f between cells
// f and f+1 along the axis, left = r-sweep of w at cell f, right = l-sweep at
// cell f+1 (conservative_interpolation "r" / "l", fvhoe/fv.py:143-159), then
// F = RS(left, right) (fvhoe/solver.py:573-606).  Each thread slides a
// register window of the five variables over R consecutive faces, so w is
// read once and only the fluxes are written; the sums and the solver are the
// separate kernels' operations in the same order (bit-identical).
// FP64 face states formed one variable at a time (1) or from all five variables' windows at once (0)
// (3-D wave 256^3 p=7, per launch FP64: 547-576 -> 493-532 us; FP32 keeps the five windows: its y-face
// launch went 389 -> 444 us one variable at a time)
#ifndef FVHOE_FACE_SPLIT
#define FVHOE_FACE_SPLIT 1
#endif
// waves per SIMD the face-flux kernel is compiled for (1: the compiler's choice)
#ifndef FVHOE_FACE_WAVES
#define FVHOE_FACE_WAVES 1
#endif
template <typename T, int P, int D, int R, bool ZAX, int RS>
__global__ void __launch_bounds__(256, FVHOE_FACE_WAVES) k_face_flux(const T* __restrict__ w, T* __restrict__ out, int n0, int n1,
                                                   int n2, int nblk_fast, int nb_axis, double gamma_d,
                                                   double csq_floor_d) {
  const T gamma = T(gamma_d), csq_floor = T(csq_floor_d);
  constexpr int K = ntaps(P), NW = R + K;
  const int axis = D + 1;
  const int na = axis == 1 ? n0 : axis == 2 ? n1 : n2, mf = na - K;
  const int m0 = axis == 1 ? mf : n0, m1 = axis == 2 ? mf : n1, m2 = axis == 3 ? mf : n2;
  int i, j, k, a0;
  if (!ZAX) {
    // threads over the plane of the two other indices, z fastest (so no wave idles at the end of a z row);
    // the windows along the axis fastest across blocks, neighbours (sharing K - 1 rows) on one XCD
    const unsigned blk = (unsigned)xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int ab = (int)(blk % (unsigned)nb_axis);
    const int t = (int)(blk / (unsigned)nb_axis) * 256 + (int)threadIdx.x;
    if (t >= (axis == 1 ? n1 : n0) * n2) return;
    const int o = t / n2;
    k = t - o * n2;
    a0 = ab * R;
    i = axis == 1 ? a0 : o;
    j = axis == 1 ? o : a0;
  } else {
    const unsigned g = blockIdx.x * 256u + threadIdx.x;
    const int kb = (int)(g % (unsigned)nb_axis);
    const unsigned rest = g / (unsigned)nb_axis;
    j = (int)(rest % (unsigned)n1);
    i = (int)(rest / (unsigned)n1);
    if (i >= n0) return;
    a0 = kb * R;
    k = a0;
  }
  const int s = axis == 1 ? n1 * n2 : axis == 2 ? n2 : 1;
  const int nc = n0 * n1 * n2, no = m0 * m1 * m2;
  const int in0 = (i * n1 + j) * n2 + k;
  const int os = axis == 1 ? m1 * m2 : axis == 2 ? m2 : 1;
  const int o0 = ((axis == 1 ? a0 : i) * m1 + (axis == 2 ? a0 : j)) * m2 + (axis == 3 ? a0 : k);
  // FP64: one variable's window at a time, reduced to its R left and right face values before the next
  // variable's loads are issued (fewer live registers); FP32: all five windows at once
  constexpr bool SPLIT = FVHOE_FACE_SPLIT && sizeof(T) == 8;
  constexpr int RS5 = SPLIT ? R : 1, NW5 = SPLIT ? 1 : NW;
  T lf[RS5][5], rf[RS5][5];
  T win[5][NW5];
  if constexpr (SPLIT) {
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      T wq[NW];
#pragma unroll
      for (int r = 0; r < NW; ++r) wq[r] = a0 + r < na ? w[q * nc + in0 + r * s] : T(0);
#pragma unroll
      for (int f = 0; f < R; ++f) {
        T a = T(0), b = T(0);
        static_for<K>([&](auto I) {
          constexpr int t = decltype(I)::value;
          constexpr double wr = weight(KIND_R, P, t), wl = weight(KIND_L, P, t);
          if constexpr (wr != 0.0) a = a + T(wr) * wq[f + t];
          if constexpr (wl != 0.0) b = b + T(wl) * wq[f + 1 + t];
        });
        lf[f][q] = a;
        rf[f][q] = b;
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
#pragma unroll
    for (int q = 0; q < 5; ++q)
#pragma unroll
      for (int r = 0; r < NW; ++r) win[q][r] = a0 + r < na ? w[q * nc + in0 + r * s] : T(0);
  }
#pragma unroll
  for (int f = 0; f < R; ++f) {
    if (a0 + f < mf) {
      T left[5], right[5], F[5];
      if constexpr (SPLIT) {
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          left[q] = lf[f][q];
          right[q] = rf[f][q];
        }
      } else {
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          T a = T(0), b = T(0);
          static_for<K>([&](auto I) {
            constexpr int t = decltype(I)::value;
            constexpr double wr = weight(KIND_R, P, t), wl = weight(KIND_L, P, t);
            if constexpr (wr != 0.0) a = a + T(wr) * win[q][f + t];
            if constexpr (wl != 0.0) b = b + T(wl) * win[q][f + 1 + t];
          });
          left[q] = a;
          right[q] = b;
        }
      }
      riemann_t<RS, D>(left, right, F, gamma, csq_floor);  // the solver is a compile-time choice
#pragma unroll
      for (int q = 0; q < 5; ++q) out[q * no + o0 + f * os] = F[q];
    }
  }
}

// Two sweeps, along a strided axis S (x or y) and along z, in one pass: the
// lanes of a wave hold 64 consecutive z (the z taps come from neighbouring
// lanes through DPP shifts; 64 - (Kz - 1) lanes produce outputs) and each
// thread slides a register window over R outputs along S.  ZFIRST: the z-sweep
// is applied first (to every window row), else the S-sweep first.  Every
// intermediate value is the double the separate sweeps would store, summed in
// the same order (bit-identical to two fvhoe_stencil_sweep calls).
template <int P, int KIND, typename T>
__device__ __forceinline__ T zsweep(T x) {
  constexpr int K = ntaps(P);
  T acc = T(0), sh = x;
  static_for<K>([&](auto I) {
    constexpr int t = decltype(I)::value;
    if constexpr (t > 0) sh = lane_next(sh);  // sh = x of lane l + t
    constexpr double wt = weight(KIND, P, t);
    if constexpr (wt != 0.0) acc = acc + T(wt) * sh;
  });
  return acc;
}

template <typename T, int KIND, int PS, int PZ, bool ZFIRST, int R>
__global__ void __launch_bounds__(256) k_sweep_zpair(const T* __restrict__ u, T* __restrict__ out, int n0, int n1,
                                                     int n2, int axis, int nzw, int nsb, int nother, int nv) {
  constexpr int KS = ntaps(PS), KZ = ntaps(PZ), NW = R + KS - 1, VC = 64 - (KZ - 1);
  const int lane = (int)threadIdx.x & 63;
  // one wave: one z window, one run along S; neighbouring runs (sharing KS - 1 rows) on one XCD
  const unsigned gw = (unsigned)xcd_tile((int)blockIdx.x, (int)gridDim.x) * 4u + (threadIdx.x >> 6);
  const int zw = (int)(gw % (unsigned)nzw);
  unsigned rest = gw / (unsigned)nzw;
  const int sb = (int)(rest % (unsigned)nsb);
  rest /= (unsigned)nsb;
  const int o = (int)(rest % (unsigned)nother);
  const int v = (int)(rest / (unsigned)nother);
  if (v >= nv) return;  // whole waves only: the DPP shifts below need every lane
  const int ns = axis == 1 ? n0 : n1, ms = ns - (KS - 1), m2 = n2 - (KZ - 1);
  const int z = zw * VC + lane, s0 = sb * R;
  const int i = axis == 1 ? s0 : o, j = axis == 1 ? o : s0;
  const int ss = axis == 1 ? n1 * n2 : n2;
  const int in0 = ((v * n0 + i) * n1 + j) * n2 + z;
  const bool zin = z < n2;
  T win[NW];
#pragma unroll
  for (int r = 0; r < NW; ++r) win[r] = (zin && s0 + r < ns) ? u[in0 + r * ss] : T(0);
  const int m0 = axis == 1 ? ms : n0, m1 = axis == 1 ? n1 : ms;
  const int os = axis == 1 ? m1 * m2 : m2;
  const int o0 = ((v * m0 + (axis == 1 ? s0 : i)) * m1 + (axis == 1 ? j : s0)) * m2 + z;
  const bool zok = lane < VC && z < m2;
  if constexpr (ZFIRST) {
#pragma unroll
    for (int r = 0; r < NW; ++r) win[r] = zsweep<PZ, KIND>(win[r]);
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    T acc = T(0);
    static_for<KS>([&](auto I) {
      constexpr double wt = weight(KIND, PS, decltype(I)::value);
      if constexpr (wt != 0.0) acc = acc + T(wt) * win[q + decltype(I)::value];
    });
    if constexpr (!ZFIRST) acc = zsweep<PZ, KIND>(acc);
    if (zok && s0 + q < ms) out[o0 + q * os] = acc;
  }
}

// outputs per thread along the strided axis of the z-pair kernel
// (3-D wave 256^3 p=7, per launch FP64 / FP32: R = 4: 317-479 / -, 8: 282-424 / 155-232, 16: 268-395 /
// 141-222 us)
#ifndef FVHOE_ZPAIR_R
#define FVHOE_ZPAIR_R 16
#endif

template <typename T, int KIND, int P, bool ZFIRST>
int launch_zpair(const T* u, T* out, int nv, int n0, int n1, int n2, int axis, hipStream_t st) {
  constexpr int K = ntaps(P), R = FVHOE_ZPAIR_R, VC = 64 - (K - 1);
  const int ns = axis == 1 ? n0 : n1, ms = ns - (K - 1), m2 = n2 - (K - 1);
  const int nzw = (m2 + VC - 1) / VC, nsb = (ms + R - 1) / R, nother = axis == 1 ? n1 : n0;
  const long long waves = (long long)nzw * nsb * nother * nv;
  if (waves * 64 >= (1ll << 31)) return -1;
  k_sweep_zpair<T, KIND, P, P, ZFIRST, R><<<(unsigned)((waves + 3) / 4), 256, 0, st>>>(u, out, n0, n1, n2, axis, nzw, nsb,
                                                                                     nother, nv);
  return 0;
}

template <typename T, int KIND, bool ZFIRST>
int launch_zpair_p(int p, const T* u, T* out, int nv, int n0, int n1, int n2, int axis, hipStream_t st) {
  switch (p) {
    case 1: return launch_zpair<T, KIND, 1, ZFIRST>(u, out, nv, n0, n1, n2, axis, st);
    case 2: return launch_zpair<T, KIND, 2, ZFIRST>(u, out, nv, n0, n1, n2, axis, st);
    case 3: return launch_zpair<T, KIND, 3, ZFIRST>(u, out, nv, n0, n1, n2, axis, st);
    case 4: return launch_zpair<T, KIND, 4, ZFIRST>(u, out, nv, n0, n1, n2, axis, st);
    case 5: return launch_zpair<T, KIND, 5, ZFIRST>(u, out, nv, n0, n1, n2, axis, st);
    case 6: return launch_zpair<T, KIND, 6, ZFIRST>(u, out, nv, n0, n1, n2, axis, st);
    case 7: return launch_zpair<T, KIND, 7, ZFIRST>(u, out, nv, n0, n1, n2, axis, st);
    case 8: return launch_zpair<T, KIND, 8, ZFIRST>(u, out, nv, n0, n1, n2, axis, st);
  }
  return -1;
}

// Two sweeps along the strided axes, first along A then along B ({A, B} = {x, y}), in one pass:
// threads along the contiguous z; each thread owns R consecutive outputs along A and a run of RB along
// B.  It slides down the input rows along B, sweeps each row along A (R + K - 1 loads, coalesced over
// z) and keeps, for each of its R outputs along A, a register window of the last K A-swept rows, from
// which the B-sweep is formed.  Every intermediate is the value the separate sweeps would store, summed
// in the same order (bit-identical to two fvhoe_stencil_sweep calls), and the intermediate array is
// never written.
template <typename T, int KIND, int P, int R, int RB>
__global__ void __launch_bounds__(256) k_sweep_xypair(const T* __restrict__ u, T* __restrict__ out, int n0, int n1,
                                                      int n2, int afirst, int nblk_fast, int nbA, int nbB) {
  constexpr int K = ntaps(P), NA = R + K - 1, NB = RB + K - 1;
  // threads over (v, z), z fastest; blocks: run along A fastest (neighbour runs on one XCD), then B, then
  // the (v, z) chunk
  const unsigned blk = (unsigned)xcd_tile((int)blockIdx.x, (int)gridDim.x);
  const int ia = (int)(blk % (unsigned)nbA);
  unsigned rest = blk / (unsigned)nbA;
  const int ib = (int)(rest % (unsigned)nbB);
  const int t = (int)(rest / (unsigned)nbB) * 256 + (int)threadIdx.x;
  if (t >= nblk_fast * n2) return;  // nblk_fast carries nv here
  const int v = t / n2, k = t - (t / n2) * n2;
  const bool a_is_x = afirst == 1;
  const int nA = a_is_x ? n0 : n1, nB = a_is_x ? n1 : n0, mA = nA - (K - 1), mB = nB - (K - 1);
  const int m0 = n0 - (K - 1), m1 = n1 - (K - 1);
  const int sA = a_is_x ? n1 * n2 : n2, sB = a_is_x ? n2 : n1 * n2;      // input strides
  const int osA = a_is_x ? m1 * n2 : n2, osB = a_is_x ? n2 : m1 * n2;   // output strides
  const int a0 = ia * R, b0 = ib * RB;
  const int in0 = v * n0 * n1 * n2 + a0 * sA + b0 * sB + k;
  const int o0 = v * m0 * m1 * n2 + a0 * osA + b0 * osB + k;
  T ring[R][K];
#pragma unroll
  for (int rb = 0; rb < NB; ++rb) {
    if (b0 + rb < nB) {
      T win[NA];
#pragma unroll
      for (int j = 0; j < NA; ++j) win[j] = a0 + j < nA ? u[in0 + j * sA + rb * sB] : T(0);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        T acc = T(0);
        static_for<K>([&](auto I) {
          constexpr double wt = weight(KIND, P, decltype(I)::value);
          if constexpr (wt != 0.0) acc = acc + T(wt) * win[r + decltype(I)::value];
        });
        ring[r][rb % K] = acc;  // the A-swept value of row b0 + rb
      }
    }
    if (rb >= K - 1) {
      const int q = rb - (K - 1);  // output row b0 + q along B: rows q .. q + K - 1 are in the windows
      if (b0 + q < mB) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (a0 + r < mA) {
            T acc = T(0);
            static_for<K>([&](auto I) {
              constexpr int t = decltype(I)::value;
              constexpr double wt = weight(KIND, P, t);
              if constexpr (wt != 0.0) acc = acc + T(wt) * ring[r][(q + t) % K];
            });
            out[o0 + r * osA + q * osB] = acc;
          }
        }
      }
    }
  }
}

// outputs per thread along the first and the second axis of the x-y pair kernel, FP64 and FP32 (3-D wave
// 256^3 p=7, per launch FP64 / FP32: 4 x 16: 656-676 / 363, 4 x 32: 638-639 / 343-361, 8 x 16: 627-657 /
// 316, 8 x 8: 733-769 / 340-345, 2 x 32: 790-809 us)
#ifndef FVHOE_XY_R
#define FVHOE_XY_R 4
#endif
#ifndef FVHOE_XY_RB
#define FVHOE_XY_RB 32
#endif
#ifndef FVHOE_XY32_R
#define FVHOE_XY32_R 8
#endif
#ifndef FVHOE_XY32_RB
#define FVHOE_XY32_RB 16
#endif

template <typename T, int KIND, int P>
int launch_xypair(const T* u, T* out, int nv, int n0, int n1, int n2, int afirst, hipStream_t st) {
  constexpr bool F64 = sizeof(T) == 8;
  constexpr int K = ntaps(P), R = F64 ? FVHOE_XY_R : FVHOE_XY32_R, RB = F64 ? FVHOE_XY_RB : FVHOE_XY32_RB;
  const int nA = afirst == 1 ? n0 : n1, nB = afirst == 1 ? n1 : n0;
  const int nbk = (int)(((long long)nv * n2 + 255) / 256), nbA = (nA - (K - 1) + R - 1) / R,
            nbB = (nB - (K - 1) + RB - 1) / RB;
  const long long blocks = (long long)nbk * nbA * nbB;
  if (blocks >= (1ll << 31)) return -1;
  k_sweep_xypair<T, KIND, P, R, RB><<<(unsigned)blocks, 256, 0, st>>>(u, out, n0, n1, n2, afirst, nv, nbA, nbB);
  return 0;
}

template <typename T, int KIND>
int launch_xypair_p(int p, const T* u, T* out, int nv, int n0, int n1, int n2, int afirst, hipStream_t st) {
  switch (p) {
    case 1: return launch_xypair<T, KIND, 1>(u, out, nv, n0, n1, n2, afirst, st);
    case 2: return launch_xypair<T, KIND, 2>(u, out, nv, n0, n1, n2, afirst, st);
    case 3: return launch_xypair<T, KIND, 3>(u, out, nv, n0, n1, n2, afirst, st);
    case 4: return launch_xypair<T, KIND, 4>(u, out, nv, n0, n1, n2, afirst, st);
    case 5: return launch_xypair<T, KIND, 5>(u, out, nv, n0, n1, n2, afirst, st);
    case 6: return launch_xypair<T, KIND, 6>(u, out, nv, n0, n1, n2, afirst, st);
    case 7: return launch_xypair<T, KIND, 7>(u, out, nv, n0, n1, n2, afirst, st);
    case 8: return launch_xypair<T, KIND, 8>(u, out, nv, n0, n1, n2, afirst, st);
  }
  return -1;
}

// faces per thread of the fused face-flux kernel
#ifndef FVHOE_FACE_R
#define FVHOE_FACE_R 4
#endif

template <typename T, int P, int D, int RS>
int launch_face_flux_rs(const T* w, T* out, int n0, int n1, int n2, double gamma, double csq_floor, hipStream_t st) {
  constexpr int K = ntaps(P), R = FVHOE_FACE_R;
  const int axis = D + 1, na = axis == 1 ? n0 : axis == 2 ? n1 : n2, mf = na - K;
  const int nb = (mf + R - 1) / R;
  if (axis != 3) {
    const long long plane = (long long)(axis == 1 ? n1 : n0) * n2, blocks = (plane + 255) / 256 * nb;
    if (blocks >= (1ll << 31) || plane >= (1ll << 31)) return -1;
    k_face_flux<T, P, D, R, false, RS><<<(unsigned)blocks, 256, 0, st>>>(w, out, n0, n1, n2, 0, nb, gamma,
                                                                       csq_floor);
  } else {
    const long long threads = (long long)nb * n1 * n0;
    if (threads >= (1ll << 31)) return -1;
    k_face_flux<T, P, D, R, true, RS><<<(unsigned)((threads + 255) / 256), 256, 0, st>>>(w, out, n0, n1, n2, 0, nb,
                                                                                      gamma, csq_floor);
  }
  return 0;
}

template <typename T, int P, int D>
int launch_face_flux(const T* w, T* out, int n0, int n1, int n2, int solver, double gamma, double csq_floor,
                     hipStream_t st) {
  if (solver == RS_HLLC) return launch_face_flux_rs<T, P, D, RS_HLLC>(w, out, n0, n1, n2, gamma, csq_floor, st);
  if (solver == RS_LLF) return launch_face_flux_rs<T, P, D, RS_LLF>(w, out, n0, n1, n2, gamma, csq_floor, st);
  return launch_face_flux_rs<T, P, D, RS_UPWIND>(w, out, n0, n1, n2, gamma, csq_floor, st);
}

template <typename T, int D>
int launch_face_flux_p(int p, const T* w, T* out, int n0, int n1, int n2, int solver, double gamma, double csq_floor,
                       hipStream_t st) {
  switch (p) {
    case 0: return launch_face_flux<T, 0, D>(w, out, n0, n1, n2, solver, gamma, csq_floor, st);
    case 1: return launch_face_flux<T, 1, D>(w, out, n0, n1, n2, solver, gamma, csq_floor, st);
    case 2: return launch_face_flux<T, 2, D>(w, out, n0, n1, n2, solver, gamma, csq_floor, st);
    case 3: return launch_face_flux<T, 3, D>(w, out, n0, n1, n2, solver, gamma, csq_floor, st);
    case 4: return launch_face_flux<T, 4, D>(w, out, n0, n1, n2, solver, gamma, csq_floor, st);
    case 5: return launch_face_flux<T, 5, D>(w, out, n0, n1, n2, solver, gamma, csq_floor, st);
    case 6: return launch_face_flux<T, 6, D>(w, out, n0, n1, n2, solver, gamma, csq_floor, st);
    case 7: return launch_face_flux<T, 7, D>(w, out, n0, n1, n2, solver, gamma, csq_floor, st);
    case 8: return launch_face_flux<T, 8, D>(w, out, n0, n1, n2, solver, gamma, csq_floor, st);
  }
  return -1;
}

// minmod / moncen of two difference arrays (fvhoe/slope_limiting.py:56-84), elementwise
__global__ void k_limit(int limiter, const double* __restrict__ dl, const double* __restrict__ dr,
                        double* __restrict__ out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = limiter == LIM_MINMOD ? minmod(dl[i], dr[i]) : moncen(dl[i], dr[i]);
}

// 0.5 (u[l+1] - u[l-1]) along the middle axis of u viewed as (A, L, B), divided by h when asked
// (fvhoe/fv.py:280-296)
__global__ void k_central_diff(const double* __restrict__ u, double* __restrict__ out, long long A, long long L,
                               long long B, double h, int has_h) {
  const long long Lo = L - 2, n = A * Lo * B;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const long long b = i % B, r = i / B, l = r % Lo, a = r / Lo;
    const long long s = (a * L + l) * B + b;
    double d = 0.5 * (u[s + 2 * B] - u[s]);
    if (has_h) d = d / h;
    out[i] = d;
  }
}

// physical Euler flux along dim from primitives w and conservatives u (fvhoe/hydro.py:213-246)
__global__ void k_fluxes(const double* __restrict__ w, const double* __restrict__ u, double* __restrict__ out,
                         int nv, long long nc, int dim, int include_pressure) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nc; i += (long long)gridDim.x * blockDim.x) {
    const double v = w[(1 + dim) * nc + i];
    out[i] = v * w[i];
    for (int q = 1; q < 4; ++q) out[q * nc + i] = v * u[q * nc + i];
    double E = v * u[4 * nc + i];
    if (include_pressure) {
      const double P = w[4 * nc + i];
      out[(1 + dim) * nc + i] = out[(1 + dim) * nc + i] + P;
      E = E + v * P;
    }
    out[4 * nc + i] = E;
    for (int q = 5; q < nv; ++q) out[q * nc + i] = v * u[q * nc + i];
  }
}

__global__ void k_muscl(const double* __restrict__ u, double* __restrict__ left, double* __restrict__ right, int nv,
                        int n0, int n1, int n2, int axis, int limiter) {
  const int m0 = axis == 1 ? n0 - 2 : n0, m1 = axis == 2 ? n1 - 2 : n1, m2 = axis == 3 ? n2 - 2 : n2;
  const long long total = (long long)nv * m0 * m1 * m2;
  const long long s = axis == 1 ? (long long)n1 * n2 : axis == 2 ? n2 : 1;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    int v, i, j, k;
    split4(idx, m0, m1, m2, &v, &i, &j, &k);
    const long long base = (((long long)v * n0 + i) * n1 + j) * n2 + k;
    const double u0 = u[base], u1 = u[base + s], u2 = u[base + 2 * s];
    const double sl = limited_slope(limiter, u1 - u0, u2 - u1);
    left[idx] = u1 - 0.5 * sl;
    right[idx] = u1 + 0.5 * sl;
  }
}

struct BcArgs {
  int n[3], g[3], lo[3], hi[3];
};

template <typename TO>
__global__ void k_bc_apply(const double* __restrict__ u, TO* __restrict__ out, int nv, BcArgs a) {
  const int N0 = a.n[0] + 2 * a.g[0], N1 = a.n[1] + 2 * a.g[1], N2 = a.n[2] + 2 * a.g[2];
  const long long total = (long long)nv * N0 * N1 * N2;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    int v, i, j, k;
    split4(idx, N0, N1, N2, &v, &i, &j, &k);
    bool f0, f1, f2;
    const int si = bc_source(i, a.n[0], a.g[0], a.lo[0], a.hi[0], &f0);
    const int sj = bc_source(j, a.n[1], a.g[1], a.lo[1], a.hi[1], &f1);
    const int sk = bc_source(k, a.n[2], a.g[2], a.lo[2], a.hi[2], &f2);
    double x = u[(((long long)v * a.n[0] + si) * a.n[1] + sj) * a.n[2] + sk];
    if ((f0 && v == 1) || (f1 && v == 2) || (f2 && v == 3)) x = -x;
    out[idx] = TO(x);
  }
}

__global__ void k_bc_apply_slabs(const double* __restrict__ u, double* __restrict__ out, int nv, BcArgs a,
                                 Slabs sl) {
  const int N0 = a.n[0] + 2 * a.g[0], N1 = a.n[1] + 2 * a.g[1], N2 = a.n[2] + 2 * a.g[2];
  const int bc[6] = {a.lo[0], a.hi[0], a.lo[1], a.hi[1], a.lo[2], a.hi[2]};
  const long long total = (long long)nv * N0 * N1 * N2;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    int v, i, j, k;
    split4(idx, N0, N1, N2, &v, &i, &j, &k);
    out[idx] = bc_fetch3(a.n, bc, sl, v, i - a.g[0], j - a.g[1], k - a.g[2], [&](int vv, int s0, int s1, int s2) {
      return u[(((long long)vv * a.n[0] + s0) * a.n[1] + s1) * a.n[2] + s2];
    });
  }
}

// The ghost fill and the first sweep of the padded array in one pass: out is the sweep along axis
// (1 = x or 2 = y) of the array fvhoe_bc_apply would write, each padded element fetched from u through
// the boundary index maps (the same values as the fill, rounded to T as the float32 fill rounds
// them), so the padded array is never stored.  Threads along the contiguous z, a register window of R
// outputs along the sweep axis, as k_sweep_slide (bit-identical to fill + sweep).
template <typename T, int KIND, int P, int R>
__global__ void __launch_bounds__(256) k_bc_sweep_slide(const double* __restrict__ u, T* __restrict__ out, BcArgs a,
                                                        int axis, int nblk_fast, int nb_axis) {
  constexpr int K = ntaps(P), NW = R + K - 1;
  const int N0 = a.n[0] + 2 * a.g[0], N1 = a.n[1] + 2 * a.g[1], N2 = a.n[2] + 2 * a.g[2];
  const int na = axis == 1 ? N0 : N1, ma = na - (K - 1);
  const int m0 = axis == 1 ? ma : N0, m1 = axis == 2 ? ma : N1;
  // threads over the (other axis, z) plane, z fastest; blocks: window along the axis fastest (neighbours on
  // one XCD), then the plane, then v
  const unsigned blk = (unsigned)xcd_tile((int)blockIdx.x, (int)gridDim.x);
  const int ab = (int)(blk % (unsigned)nb_axis);
  const unsigned rest = blk / (unsigned)nb_axis;
  const int no = axis == 1 ? N1 : N0;
  const int t = (int)(rest % (unsigned)nblk_fast) * 256 + (int)threadIdx.x, v = (int)(rest / (unsigned)nblk_fast);
  if (t >= no * N2) return;
  const int o = t / N2, k = t - (t / N2) * N2;
  const int a0 = ab * R;
  const int dS = axis == 1 ? 0 : 1, dO = axis == 1 ? 1 : 0;  // sweep axis and the other strided axis
  bool fk, fo;
  const int sk = bc_source(k, a.n[2], a.g[2], a.lo[2], a.hi[2], &fk);
  const int so = bc_source(o, a.n[dO], a.g[dO], a.lo[dO], a.hi[dO], &fo);
  // 32-bit offsets (the caller checks the sizes): u[v][i][j][k] = fixed + source index along the sweep
  // axis x its stride; a reflective side negates its normal momentum (variable 1 + axis)
  const int plane = a.n[1] * a.n[2], sstride = axis == 1 ? plane : a.n[2];
  const int fixed = v * a.n[0] * plane + so * (axis == 1 ? a.n[2] : plane) + sk;
  const bool flip_fixed = (fo && v == 1 + dO) || (fk && v == 3);
  T win[NW];
#pragma unroll
  for (int r = 0; r < NW; ++r) {
    win[r] = T(0);
    if (a0 + r < na) {
      bool fs;
      const int ss = bc_source(a0 + r, a.n[dS], a.g[dS], a.lo[dS], a.hi[dS], &fs);
      double x = u[fixed + ss * sstride];
      if (flip_fixed || (fs && v == 1 + dS)) x = -x;
      win[r] = T(x);
    }
  }
  const long long os = axis == 1 ? (long long)m1 * N2 : N2;
  const long long o0 = (((long long)v * m0 + (axis == 1 ? a0 : o)) * m1 + (axis == 2 ? a0 : o)) * N2 + k;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (a0 + q < ma) {
      T acc = T(0);
      static_for<K>([&](auto I) {
        constexpr double wt = weight(KIND, P, decltype(I)::value);
        if constexpr (wt != 0.0) acc = acc + T(wt) * win[q + decltype(I)::value];
      });
      out[o0 + q * os] = acc;
    }
  }
}

// dudt = 0.0 + cx*(F[i+1]-F[i]) + cy*(G[j+1]-G[j]) + cz*(H[k+1]-H[k]), c = -(1/h)  (fvhoe/solver.py:744-750)
// element strides (variable, axis 0, axis 1) of a face-flux array whose last axis is contiguous: the
// fluxes may be views into larger arrays (the trimmed fluxes of hydrofluxes)
struct Strides3 {
  long long v, a, b;
};
template <typename T>
__global__ void k_euler_eq(const T* __restrict__ F, const T* __restrict__ G, const T* __restrict__ H,
                           double* __restrict__ dudt, int nv, int n0, int n1, int n2, int active, double cx,
                           double cy, double cz, Strides3 fs, Strides3 gs, Strides3 hs) {
  const long long total = (long long)nv * n0 * n1 * n2;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    int v, i, j, k;
    split4(idx, n0, n1, n2, &v, &i, &j, &k);
    double acc = 0.0;
    if (active & 1) {
      const long long b = v * fs.v + i * fs.a + j * fs.b + k;
      acc = acc + cx * ((double)F[b + fs.a] - (double)F[b]);
    }
    if (active & 2) {
      const long long b = v * gs.v + i * gs.a + j * gs.b + k;
      acc = acc + cy * ((double)G[b + gs.b] - (double)G[b]);
    }
    if (active & 4) {
      const long long b = v * hs.v + i * hs.a + j * hs.b + k;
      acc = acc + cz * ((double)H[b + 1] - (double)H[b]);
    }
    dudt[idx] = acc;
  }
}

// The same divergence followed by the stage's RK combination in one pass (the composed stage's version
// of the fused stage's last kernel): dudt (when requested; k0 + dudt for the SSP-RK3 *_SUM stage) and
// rk.out = the combination with dudt as the stage's k, the value the separate kernels would store and
// read back.  rk.out may alias rk.u0 (pointwise).
template <typename T>
__global__ void k_euler_eq_rk(const T* __restrict__ F, const T* __restrict__ G, const T* __restrict__ H,
                              double* __restrict__ dudt, int nv, int n0, int n1, int n2, int active, double cx,
                              double cy, double cz, Strides3 fs, Strides3 gs, Strides3 hs, RkArgs rk) {
  const long long total = (long long)nv * n0 * n1 * n2;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    int v, i, j, k;
    split4(idx, n0, n1, n2, &v, &i, &j, &k);
    double acc = 0.0;
    if (active & 1) {
      const long long b = v * fs.v + i * fs.a + j * fs.b + k;
      acc = acc + cx * ((double)F[b + fs.a] - (double)F[b]);
    }
    if (active & 2) {
      const long long b = v * gs.v + i * gs.a + j * gs.b + k;
      acc = acc + cy * ((double)G[b + gs.b] - (double)G[b]);
    }
    if (active & 4) {
      const long long b = v * hs.v + i * hs.a + j * hs.b + k;
      acc = acc + cz * ((double)H[b + 1] - (double)H[b]);
    }
    if (dudt) dudt[idx] = rk.kind == FVHOE_RK_SSPRK3_STAGE3_SUM ? rk.k0[idx] + acc : acc;
    rk.out[idx] = rk_combine(rk, (int)idx, acc);
  }
}

// The divergence (and, with RK, the stage's RK combination) for grids whose contiguous axis holds at
// least a wave: blocks of 64 lanes along z x 4 waves along y, the (v, x) pair from blockIdx.z, so the
// row offsets are wave-uniform and no cell index is divided (the grid-stride kernels above split a
// linear index per cell).  The same operations in the same order as those kernels.
template <typename T, bool RK>
__global__ void __launch_bounds__(256) k_euler_eq_rows(const T* __restrict__ F, const T* __restrict__ G,
                                                       const T* __restrict__ H, double* __restrict__ dudt, int n0,
                                                       int n1, int n2, int active, double cx, double cy, double cz,
                                                       Strides3 fs, Strides3 gs, Strides3 hs, RkArgs rk) {
  const int k = (int)blockIdx.x * 64 + ((int)threadIdx.x & 63);
  const int j = (int)blockIdx.y * 4 + ((int)threadIdx.x >> 6);
  if (k >= n2 || j >= n1) return;
  const int vi = (int)blockIdx.z, v = vi / n0, i = vi - v * n0;
  const long long idx = ((long long)vi * n1 + j) * n2 + k;
  double acc = 0.0;
  if (active & 1) {
    const long long b = v * fs.v + i * fs.a + j * fs.b + k;
    acc = acc + cx * ((double)F[b + fs.a] - (double)F[b]);
  }
  if (active & 2) {
    const long long b = v * gs.v + i * gs.a + j * gs.b + k;
    acc = acc + cy * ((double)G[b + gs.b] - (double)G[b]);
  }
  if (active & 4) {
    const long long b = v * hs.v + i * hs.a + j * hs.b + k;
    acc = acc + cz * ((double)H[b + 1] - (double)H[b]);
  }
  if (RK) {
    if (dudt) dudt[idx] = rk.kind == FVHOE_RK_SSPRK3_STAGE3_SUM ? rk.k0[idx] + acc : acc;
    rk.out[idx] = rk_combine(rk, (int)idx, acc);
  } else {
    dudt[idx] = acc;
  }
}

// whether the row kernel applies: a z row of at least one wave and a grid within the launch limits
inline bool rows_grid(int nv, int n0, int n1, int n2, dim3* g) {
  if (n2 < 64 || (long long)nv * n0 > 65535 || (n1 + 3) / 4 > 65535) return false;
  *g = dim3((unsigned)((n2 + 63) / 64), (unsigned)((n1 + 3) / 4), (unsigned)(nv * n0));
  return true;
}

__global__ void k_init_minmax(unsigned long long* acc, int nv) {
  const int i = threadIdx.x;
  if (i < 2 * nv) acc[i] = (i & 1) ? 0ull : ~0ull;  // (min, max) pairs
}

__global__ void k_minmax_partial(const double* __restrict__ u, long long nc, unsigned long long* acc) {
  const int v = blockIdx.y;
  unsigned long long mx = 0ull, mn = ~0ull;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nc; i += (long long)gridDim.x * blockDim.x) {
    const unsigned long long b = ordered_bits(u[v * nc + i]);
    mx = b > mx ? b : mx;
    mn = b < mn ? b : mn;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long a = __shfl_down(mx, off, 64), c = __shfl_down(mn, off, 64);
    mx = a > mx ? a : mx;
    mn = c < mn ? c : mn;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&acc[2 * v + 1], mx);
    atomicMin(&acc[2 * v], mn);
  }
}

__global__ void k_minmax_final(const unsigned long long* acc, double* out, int nv) {
  const int i = threadIdx.x;
  if (i < 2 * nv) out[i] = from_ordered_bits(acc[i]);
}

struct DetectArgs {
  int N[3];        // padded extents
  int dims;        // bit d set: axis d active (3 ghosts)
  int mode;        // 0 global, 1 local
  int range_type;  // 0 relative, 1 absolute
  int nad_vars;    // bit mask over the 5 variables
  int pad_var;     // -1: no PAD
  int sed;
  double eps, pad_lo, pad_hi, sed_eps;
};

__device__ __forceinline__ double sed_alpha_cell(const double* __restrict__ x, long long base, const long long st[3],
                                                 int dims, double eps) {
  double alpha = 0.0;
  bool first = true;
  for (int d = 0; d < 3; ++d) {
    if (!(dims & (1 << d))) continue;
    const long long s = st[d];
    const double a = sed_alpha_1d<true>([&](int o) { return x[base + o * s]; }, eps);
    alpha = first ? a : np_min(alpha, a);
    first = false;
  }
  return alpha;
}

// NAD / PAD / SED on one interior cell (fvhoe/slope_limiting.py:125-244)
__global__ void k_detect(const double* __restrict__ w, const double* __restrict__ ws,
                         const double* __restrict__ range_mm, signed char* __restrict__ trouble,
                         double* __restrict__ nad_mag, double* __restrict__ pad_mag, DetectArgs a) {
  const int g0 = (a.dims & 1) ? 3 : 0, g1 = (a.dims & 2) ? 3 : 0, g2 = (a.dims & 4) ? 3 : 0;
  const int n0 = a.N[0] - 2 * g0, n1 = a.N[1] - 2 * g1, n2 = a.N[2] - 2 * g2;
  const long long nc = (long long)n0 * n1 * n2;
  const long long NC = (long long)a.N[0] * a.N[1] * a.N[2];
  const long long st[3] = {(long long)a.N[1] * a.N[2], a.N[2], 1};
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < nc; c += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    split3(c, n1, n2, &i, &j, &k);
    const long long base = ((long long)(i + g0) * a.N[1] + (j + g1)) * a.N[2] + (k + g2);
    bool nad = false;
    double ind_min = 0.0;
    for (int v = 0; v < 5; ++v) {
      const double* wv = w + v * NC;
      double M = wv[base], m = wv[base];
      for (int di = -(g0 > 0); di <= (g0 > 0); ++di)
        for (int dj = -(g1 > 0); dj <= (g1 > 0); ++dj)
          for (int dk = -(g2 > 0); dk <= (g2 > 0); ++dk) {
            const double x = wv[base + di * st[0] + dj * st[1] + dk * st[2]];
            M = np_max(M, x);
            m = np_min(m, x);
          }
      const double uc = ws[v * NC + base];
      double ind;
      bool tv;
      if (a.range_type == 0) {
        const double rng = a.mode == 0 ? range_mm[2 * v + 1] - range_mm[2 * v] : M - m;
        ind = np_min(uc - m, M - uc);
        tv = ind < -a.eps * rng;
      } else {
        const double lo = a.mode == 0 ? (1.0 - a.eps) * range_mm[2 * v] : (1.0 - a.eps) * m;
        const double hi = a.mode == 0 ? (1.0 + a.eps) * range_mm[2 * v + 1] : (1.0 + a.eps) * M;
        ind = np_min(uc - lo, hi - uc);
        tv = ind < 0.0;
      }
      if (tv && a.sed) tv = sed_alpha_cell(ws + v * NC, base, st, a.dims, a.sed_eps) < 1.0;
      if (!((a.nad_vars >> v) & 1)) tv = false;
      nad = nad || tv;
      ind_min = v == 0 ? ind : np_min(ind_min, ind);
    }
    nad_mag[c] = nad ? -ind_min : 0.0;
    double pind = 0.0;
    if (a.pad_var >= 0) {
      const double x = ws[a.pad_var * NC + base];
      pind = np_min(x - a.pad_lo, a.pad_hi - x);
    }
    const bool pad = pind < 0.0;
    pad_mag[c] = pad ? -pind : 0.0;
    trouble[c] = (pad || nad) ? 1 : 0;
  }
}

__global__ void k_sed(const double* __restrict__ u, double* __restrict__ alpha, int nv, int N0, int N1, int N2,
                      int dims, double eps) {
  const int g0 = (dims & 1) ? 3 : 0, g1 = (dims & 2) ? 3 : 0, g2 = (dims & 4) ? 3 : 0;
  const int n0 = N0 - 2 * g0, n1 = N1 - 2 * g1, n2 = N2 - 2 * g2;
  const long long nc = (long long)n0 * n1 * n2, NC = (long long)N0 * N1 * N2;
  const long long st[3] = {(long long)N1 * N2, N2, 1};
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < nc * nv;
       idx += (long long)gridDim.x * blockDim.x) {
    const int v = (int)(idx / nc);
    int i, j, k;
    split3(idx % nc, n1, n2, &i, &j, &k);
    const long long base = ((long long)(i + g0) * N1 + (j + g1)) * N2 + (k + g2);
    alpha[idx] = sed_alpha_cell(u + v * NC, base, st, dims, eps);
  }
}

__global__ void k_broadcast_simple(const signed char* __restrict__ t, double* __restrict__ tx, double* __restrict__ ty,
                                   double* __restrict__ tz, int n0, int n1, int n2, int dims) {
  // theta on faces = max of the flags of the cells the face touches (fvhoe/slope_limiting.py:366-380)
  const long long fx = (long long)(n0 + 1) * n1 * n2, fy = (long long)n0 * (n1 + 1) * n2,
                  fz = (long long)n0 * n1 * (n2 + 1);
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < fx + fy + fz;
       idx += (long long)gridDim.x * blockDim.x) {
    if (idx < fx) {
      int i, j, k;
      split3(idx, n1, n2, &i, &j, &k);
      double th = 0.0;
      if (dims & 1) {
        const int a = i > 0 ? t[((long long)(i - 1) * n1 + j) * n2 + k] : 0;
        const int b = i < n0 ? t[((long long)i * n1 + j) * n2 + k] : 0;
        th = (double)(a > b ? a : b);
      }
      tx[idx] = th;
    } else if (idx < fx + fy) {
      int i, j, k;
      split3(idx - fx, (n1 + 1), n2, &i, &j, &k);
      double th = 0.0;
      if (dims & 2) {
        const int a = j > 0 ? t[((long long)i * n1 + (j - 1)) * n2 + k] : 0;
        const int b = j < n1 ? t[((long long)i * n1 + j) * n2 + k] : 0;
        th = (double)(a > b ? a : b);
      }
      ty[idx - fx] = th;
    } else {
      int i, j, k;
      split3(idx - fx - fy, n1, (n2 + 1), &i, &j, &k);
      double th = 0.0;
      if (dims & 4) {
        const int a = k > 0 ? t[((long long)i * n1 + j) * n2 + (k - 1)] : 0;
        const int b = k < n2 ? t[((long long)i * n1 + j) * n2 + k] : 0;
        th = (double)(a > b ? a : b);
      }
      tz[idx - fx - fy] = th;
    }
  }
}

struct ShiftMaxArgs {
  int ext[3];                       // region extents
  long long os[3], ss[3], ts[3];    // strides of out, S1, S2
  long long ooff, soff, toff;       // element offsets of the region's origin
  double coef;
  int gt;                           // S1 term is coef*(S1 > 0) instead of coef*S1
};

// out[r] = max(coef * g(S1[r]), S2[r]) over a 3-D region
__global__ void k_shift_max(double* out, const double* __restrict__ S1, const double* __restrict__ S2, ShiftMaxArgs a) {
  const long long total = (long long)a.ext[0] * a.ext[1] * a.ext[2];
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    split3(idx, a.ext[1], a.ext[2], &i, &j, &k);
    const double s = S1[a.soff + i * a.ss[0] + j * a.ss[1] + k * a.ss[2]];
    const double x = a.coef * (a.gt ? (s > 0.0 ? 1.0 : 0.0) : s);
    const double y = S2[a.toff + i * a.ts[0] + j * a.ts[1] + k * a.ts[2]];
    out[a.ooff + i * a.os[0] + j * a.os[1] + k * a.os[2]] = np_max(x, y);
  }
}

// F = (1 - theta) * F + theta * Fl, theta broadcast over variables (fvhoe/solver.py:729-731)
__global__ void k_blend(double* __restrict__ F, const double* __restrict__ Fl, const double* __restrict__ th, int nv,
                        long long nf) {
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < nv * nf;
       idx += (long long)gridDim.x * blockDim.x) {
    const double t = th[idx % nf];
    F[idx] = (1.0 - t) * F[idx] + t * Fl[idx];
  }
}

// Runge-Kutta combinations in the reference's k-form (fvhoe/ode.py:245-298); dt read from device memory
__global__ void k_rk(int kind, double* out, const double* u, const double* __restrict__ k0,
                     const double* __restrict__ k1, const double* __restrict__ k2, const double* __restrict__ k3,
                     const double* __restrict__ dtp, long long n) {
  const double dt = dtp[0];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double r;
    if (dt == 0.0) {  // a step at its target time leaves the state exactly as it is
      out[i] = u[i];
      continue;
    }
    switch (kind) {
      case FVHOE_RK_AXPY: r = u[i] + dt * k0[i]; break;
      case FVHOE_RK_SSPRK3_STAGE3: r = u[i] + 0.25 * dt * k0[i] + 0.25 * dt * k1[i]; break;
      case FVHOE_RK_SSPRK3_FINAL: r = u[i] + (1.0 / 6.0) * dt * (k0[i] + k1[i] + 4.0 * k2[i]); break;
      case FVHOE_RK_SSPRK2_FINAL: r = 0.5 * u[i] + 0.5 * (k0[i] + dt * k1[i]); break;  // k0 holds u1
      case FVHOE_RK_HALF_AXPY: r = u[i] + 0.5 * dt * k0[i]; break;
      case FVHOE_RK_RK4_FINAL: r = u[i] + (1.0 / 6.0) * dt * (k0[i] + 2.0 * k1[i] + 2.0 * k2[i] + k3[i]); break;
      default: r = 0.0;
    }
    out[i] = r;
  }
}

// trouble log (fvhoe/solver.py:794-833): 0: dst = t, 1: dst = dst + t, 2: dst = dst / d; nothing when
// the gate (a step's dt) is zero
__global__ void k_log(double* __restrict__ dst, const signed char* __restrict__ t, long long n, int mode, double d,
                      const double* __restrict__ gate) {
  if (gate && gate[0] == 0.0) return;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    if (mode == 0) dst[i] = (double)t[i];
    else if (mode == 1) dst[i] = dst[i] + (double)t[i];
    else dst[i] = dst[i] / d;
  }
}

__global__ void k_log_count(int* __restrict__ cnt, const signed char* __restrict__ t, long long n, int mode,
                            const double* __restrict__ gate) {
  if (gate && gate[0] == 0.0) return;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    cnt[i] = mode == 0 ? (int)t[i] : cnt[i] + (int)t[i];
}

// The same over 16 cells per thread (16-byte flag words, n a multiple of 16, aligned arrays): mode 1
// touches the counts only where a flag is set (the flags are sparse), mode 0 writes them all.
__device__ __forceinline__ int flag_byte(unsigned w, int b) { return (int)(signed char)((w >> (8 * b)) & 0xffu); }
__global__ void k_log_count16(int* __restrict__ cnt, const uint4* __restrict__ t16, long long n16, int mode,
                              const double* __restrict__ gate) {
  if (gate && gate[0] == 0.0) return;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n16;
       i += (long long)gridDim.x * blockDim.x) {
    const uint4 w = t16[i];
    const unsigned ws[4] = {w.x, w.y, w.z, w.w};
    int4* c = (int4*)(cnt + 16 * i);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (mode == 0) {
        c[q] = int4{flag_byte(ws[q], 0), flag_byte(ws[q], 1), flag_byte(ws[q], 2), flag_byte(ws[q], 3)};
      } else if (ws[q]) {
        int4 v = c[q];
        v.x += flag_byte(ws[q], 0);
        v.y += flag_byte(ws[q], 1);
        v.z += flag_byte(ws[q], 2);
        v.w += flag_byte(ws[q], 3);
        c[q] = v;
      }
    }
  }
}

__global__ void k_log_mean(double* __restrict__ dst, const int* __restrict__ cnt, long long n, double d,
                           const double* __restrict__ gate) {
  if (gate && gate[0] == 0.0) return;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = (double)cnt[i] / d;
}

__global__ void k_set_scalar(double* p, double v) { p[0] = v; }
__global__ void k_set_scalar_gated(double* p, double v, const double* __restrict__ gate) {
  if (!(gate && gate[0] == 0.0)) p[0] = v;
}

// dense violation-magnitude logs from the fused stage's slots: out = base (or 0), then one slot after the
// other out[cell] = out[cell] + magnitude (a cell appears at most once per slot), then out / divisor
__global__ void k_mag_base(double* __restrict__ nad_out, double* __restrict__ pad_out,
                           const double* __restrict__ nad_base, const double* __restrict__ pad_base, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    nad_out[i] = nad_base ? nad_base[i] : 0.0;
    pad_out[i] = pad_base ? pad_base[i] : 0.0;
  }
}
__global__ void k_mag_scatter(double* __restrict__ nad_out, double* __restrict__ pad_out,
                              const int* __restrict__ count, const int* __restrict__ idx,
                              const double* __restrict__ nad, const double* __restrict__ pad, long long n,
                              long long cap) {
  const int m = (long long)count[0] < cap ? count[0] : (int)cap;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    const int c = idx[e];
    if (c < 0 || c >= n) continue;  // entries always name interior cells; a bad slot drops out
    nad_out[c] = nad_out[c] + nad[e];
    pad_out[c] = pad_out[c] + pad[e];
  }
}
__global__ void k_mag_divide(double* __restrict__ nad_out, double* __restrict__ pad_out, long long n,
                             const double* __restrict__ divisor) {
  const double d = divisor[0];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    nad_out[i] = nad_out[i] / d;
    pad_out[i] = pad_out[i] / d;
  }
}

__global__ void k_hit_target(double* dt_step, const double* dt_stage, const double* t, double target, int has_target) {
  const double dt = dt_stage[0];
  dt_step[0] = (has_target && target - t[0] < dt) ? target - t[0] : dt;
}

__global__ void k_advance_time(double* t, const double* dt) { t[0] = t[0] + dt[0]; }

// t += dt for a step that moved (dt != 0): count it and record its end time while there is room
__global__ void k_advance_time_counted(double* t, const double* dt, long long* steps, double* times, long long cap) {
  if (dt[0] == 0.0) return;
  t[0] = t[0] + dt[0];
  const long long k = steps[0];
  if (k < cap) times[k] = t[0];
  steps[0] = k + 1;
}

// entry bodies shared by the float64 entries and their float32 (_f32) twins
template <typename T>
int stencil_sweep_impl(const T* u, T* out, int nv, int n0, int n1, int n2, int kind, int p, int axis, void* stream) {
  if (!u || !out || nv < 1 || axis < 1 || axis > 3 || kind < 0 || kind > 3) return FVHOE_ERR_ARG;
  if (p < 0 || p > 8) return FVHOE_ERR_UNSUPPORTED;
  Weights9 w{};
  w.k = ntaps(p);
  const int na = axis == 1 ? n0 : axis == 2 ? n1 : n2;
  if (na < w.k) return FVHOE_ERR_ARG;
  for (int i = 0; i < w.k; ++i) w.w[i] = weight(kind, p, i);
  const long long total = (long long)nv * n0 * n1 * n2;
  const int m0 = axis == 1 ? n0 - (w.k - 1) : n0, m1 = axis == 2 ? n1 - (w.k - 1) : n1,
            m2 = axis == 3 ? n2 - (w.k - 1) : n2;
  // large grids: the sliding-window kernel (32-bit element offsets); small ones keep one output per thread
  if (total >= (1ll << 16) && total < (1ll << 31)) {
    const hipStream_t st = (hipStream_t)stream;
    int r = -1;
    switch (kind) {
      case 0: r = launch_slide_p<T, 0>(p, u, out, nv, n0, n1, n2, axis, st); break;
      case 1: r = launch_slide_p<T, 1>(p, u, out, nv, n0, n1, n2, axis, st); break;
      case 2: r = launch_slide_p<T, 2>(p, u, out, nv, n0, n1, n2, axis, st); break;
      case 3: r = launch_slide_p<T, 3>(p, u, out, nv, n0, n1, n2, axis, st); break;
    }
    if (r == 0) return launch_status();
  }
  if ((long long)m1 * m2 >= 64 && (long long)m1 * m2 < (1ll << 30) && (long long)nv * m0 <= 65535) {
    const dim3 grid((unsigned)((m1 * m2 + TPB - 1) / TPB), (unsigned)(nv * m0));
    k_sweep2<T><<<grid, TPB, 0, (hipStream_t)stream>>>(u, out, n0, n1, n2, axis, w, m0, m1, m2);
  } else {
    k_sweep<T><<<grid_for(total), TPB, 0, (hipStream_t)stream>>>(u, out, nv, n0, n1, n2, axis, w);
  }
  return launch_status();
}

template <typename T>
int compute_primitives_impl(const T* u, T* w, int nv, long long ncell, double gamma, void* stream) {
  if (!u || !w || nv < 5 || ncell < 0) return FVHOE_ERR_ARG;
  if (ncell == 0) return FVHOE_OK;
  k_prims<T><<<grid_for(ncell), TPB, 0, (hipStream_t)stream>>>(u, w, nv, ncell, gamma);
  return launch_status();
}

template <typename T>
int hydro_dt_impl(const T* w, long long ncell, double h, int ndim, double CFL, double gamma, double csq_floor,
                  double* dt_out, unsigned long long* scratch, void* stream) {
  if (!w || !dt_out || !scratch || ncell <= 0) return FVHOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  k_fill_u64<<<1, 64, 0, s>>>(scratch, 1, 0ull);
  k_dt_partial<T><<<grid_for(ncell) < 2048 ? grid_for(ncell) : 2048, TPB, 0, s>>>(w, ncell, ndim, gamma, csq_floor,
                                                                                 scratch);
  k_dt_final<<<1, 1, 0, s>>>(scratch, CFL * h, dt_out);
  return launch_status();
}

template <typename T>
int zpair_impl(const T* u, T* out, int nv, int n0, int n1, int n2, int kind, int p, int axis, int zfirst,
               void* stream) {
  if (!u || !out || nv < 1 || (axis != 1 && axis != 2) || (kind != FVHOE_STENCIL_C && kind != FVHOE_STENCIL_T))
    return FVHOE_ERR_ARG;
  if (p < 1 || p > 8) return FVHOE_ERR_UNSUPPORTED;
  const int k = ntaps(p), ns = axis == 1 ? n0 : n1;
  if (ns < k || n2 < 64) return FVHOE_ERR_UNSUPPORTED;  // a z window spans a whole wave
  if ((long long)nv * n0 * n1 * n2 >= (1ll << 31)) return FVHOE_ERR_UNSUPPORTED;
  const hipStream_t st = (hipStream_t)stream;
  int r;
  if (kind == FVHOE_STENCIL_C)
    r = zfirst ? launch_zpair_p<T, KIND_C, true>(p, u, out, nv, n0, n1, n2, axis, st)
               : launch_zpair_p<T, KIND_C, false>(p, u, out, nv, n0, n1, n2, axis, st);
  else
    r = zfirst ? launch_zpair_p<T, KIND_T, true>(p, u, out, nv, n0, n1, n2, axis, st)
               : launch_zpair_p<T, KIND_T, false>(p, u, out, nv, n0, n1, n2, axis, st);
  if (r != 0) return FVHOE_ERR_UNSUPPORTED;
  return launch_status();
}

template <typename T>
int xypair_impl(const T* u, T* out, int nv, int n0, int n1, int n2, int kind, int p, int afirst, void* stream) {
  if (!u || !out || nv < 1 || (afirst != 1 && afirst != 2) || (kind != FVHOE_STENCIL_C && kind != FVHOE_STENCIL_T))
    return FVHOE_ERR_ARG;
  if (p < 1 || p > 8) return FVHOE_ERR_UNSUPPORTED;
  const int k = ntaps(p);
  if (n0 < k || n1 < k) return FVHOE_ERR_ARG;
  const long long total = (long long)nv * n0 * n1 * n2;
  if (total < (1ll << 16) || total >= (1ll << 31)) return FVHOE_ERR_UNSUPPORTED;  // small grids: two sweeps
  const hipStream_t st = (hipStream_t)stream;
  const int r = kind == FVHOE_STENCIL_C ? launch_xypair_p<T, KIND_C>(p, u, out, nv, n0, n1, n2, afirst, st)
                                        : launch_xypair_p<T, KIND_T>(p, u, out, nv, n0, n1, n2, afirst, st);
  if (r != 0) return FVHOE_ERR_UNSUPPORTED;
  return launch_status();
}

template <typename T>
int face_fluxes_impl(const T* w, T* out, int n0, int n1, int n2, int axis, int p, int solver, double gamma,
                     double csq_floor, void* stream) {
  if (!w || !out || axis < 1 || axis > 3 || solver < 0 || solver > 2 || n0 < 1 || n1 < 1 || n2 < 1)
    return FVHOE_ERR_ARG;
  if (p < 0 || p > 8) return FVHOE_ERR_UNSUPPORTED;
  const int na = axis == 1 ? n0 : axis == 2 ? n1 : n2;
  if (na < ntaps(p) + 1) return FVHOE_ERR_ARG;
  if ((long long)5 * n0 * n1 * n2 >= (1ll << 31)) return FVHOE_ERR_UNSUPPORTED;
  const hipStream_t st = (hipStream_t)stream;
  int r = -1;
  if (axis == 1) r = launch_face_flux_p<T, 0>(p, w, out, n0, n1, n2, solver, gamma, csq_floor, st);
  else if (axis == 2) r = launch_face_flux_p<T, 1>(p, w, out, n0, n1, n2, solver, gamma, csq_floor, st);
  else r = launch_face_flux_p<T, 2>(p, w, out, n0, n1, n2, solver, gamma, csq_floor, st);
  if (r != 0) return FVHOE_ERR_UNSUPPORTED;
  return launch_status();
}

template <typename TO>
int bc_apply_impl(const double* u, TO* out, int nv, int n0, int n1, int n2, int g0, int g1, int g2, const int* bc6,
                  void* stream) {
  if (!u || !out || !bc6 || nv < 1) return FVHOE_ERR_ARG;
  BcArgs a{{n0, n1, n2}, {g0, g1, g2}, {bc6[0], bc6[2], bc6[4]}, {bc6[1], bc6[3], bc6[5]}};
  for (int d = 0; d < 3; ++d) {
    if (a.g[d] < 0 || a.n[d] < 1) return FVHOE_ERR_ARG;
    if (a.g[d] > 0 && a.g[d] > a.n[d]) return FVHOE_ERR_ARG;  // slabs must come from the interior
    if (a.lo[d] < 0 || a.lo[d] > 3 || a.hi[d] < 0 || a.hi[d] > 3) return FVHOE_ERR_UNSUPPORTED;
  }
  const long long total = (long long)nv * (n0 + 2 * g0) * (n1 + 2 * g1) * (n2 + 2 * g2);
  k_bc_apply<TO><<<grid_for(total), TPB, 0, (hipStream_t)stream>>>(u, out, nv, a);
  return launch_status();
}

template <typename T>
int euler_equation_impl(const T* F, const long long* fst, const T* G, const long long* gst, const T* H,
                        const long long* hst, double* dudt, int nv, int n0, int n1, int n2, int active, double cx,
                        double cy, double cz, void* stream) {
  if (!dudt || nv < 1 || n0 < 1 || n1 < 1 || n2 < 1) return FVHOE_ERR_ARG;
  if (((active & 1) && !F) || ((active & 2) && !G) || ((active & 4) && !H)) return FVHOE_ERR_ARG;
  const long long N0 = n0, N1 = n1, N2 = n2;
  // contiguous (nv, n0+1, n1, n2), (nv, n0, n1+1, n2), (nv, n0, n1, n2+1) unless strides are given
  Strides3 fs{(N0 + 1) * N1 * N2, N1 * N2, N2}, gs{N0 * (N1 + 1) * N2, (N1 + 1) * N2, N2},
      hs{N0 * N1 * (N2 + 1), N1 * (N2 + 1), N2 + 1};
  if (fst) fs = Strides3{fst[0], fst[1], fst[2]};
  if (gst) gs = Strides3{gst[0], gst[1], gst[2]};
  if (hst) hs = Strides3{hst[0], hst[1], hst[2]};
  const long long total = (long long)nv * n0 * n1 * n2;
  dim3 rg;
  if (rows_grid(nv, n0, n1, n2, &rg))
    k_euler_eq_rows<T, false><<<rg, 256, 0, (hipStream_t)stream>>>(F, G, H, dudt, n0, n1, n2, active, cx, cy, cz, fs,
                                                                   gs, hs, RkArgs{});
  else
    k_euler_eq<T><<<grid_for(total), TPB, 0, (hipStream_t)stream>>>(F, G, H, dudt, nv, n0, n1, n2, active, cx, cy,
                                                                    cz, fs, gs, hs);
  return launch_status();
}

template <typename T, int KIND, int P>
int launch_bc_slide(const double* u, T* out, const BcArgs& a, int nv, int axis, hipStream_t st) {
  constexpr int K = ntaps(P), R = FVHOE_SLIDE_R;
  const int N0 = a.n[0] + 2 * a.g[0], N1 = a.n[1] + 2 * a.g[1], N2 = a.n[2] + 2 * a.g[2];
  const int na = axis == 1 ? N0 : N1, ma = na - (K - 1);
  const long long plane = (long long)(axis == 1 ? N1 : N0) * N2;
  const int nbp = (int)((plane + 255) / 256), nb = (ma + R - 1) / R;
  const long long blocks = (long long)nbp * nb * nv;
  if (blocks >= (1ll << 31) || plane >= (1ll << 31)) return -1;
  k_bc_sweep_slide<T, KIND, P, R><<<(unsigned)blocks, 256, 0, st>>>(u, out, a, axis, nbp, nb);
  return 0;
}

template <typename T, int KIND>
int launch_bc_slide_p(int p, const double* u, T* out, const BcArgs& a, int nv, int axis, hipStream_t st) {
  switch (p) {
    case 1: return launch_bc_slide<T, KIND, 1>(u, out, a, nv, axis, st);
    case 2: return launch_bc_slide<T, KIND, 2>(u, out, a, nv, axis, st);
    case 3: return launch_bc_slide<T, KIND, 3>(u, out, a, nv, axis, st);
    case 4: return launch_bc_slide<T, KIND, 4>(u, out, a, nv, axis, st);
    case 5: return launch_bc_slide<T, KIND, 5>(u, out, a, nv, axis, st);
    case 6: return launch_bc_slide<T, KIND, 6>(u, out, a, nv, axis, st);
    case 7: return launch_bc_slide<T, KIND, 7>(u, out, a, nv, axis, st);
    case 8: return launch_bc_slide<T, KIND, 8>(u, out, a, nv, axis, st);
  }
  return -1;
}

template <typename T>
int bc_stencil_sweep_impl(const double* u, T* out, int nv, int n0, int n1, int n2, int g0, int g1, int g2,
                          const int* bc6, int kind, int p, int axis, void* stream) {
  if (!u || !out || !bc6 || nv < 1 || (axis != 1 && axis != 2) || (kind != FVHOE_STENCIL_C && kind != FVHOE_STENCIL_T))
    return FVHOE_ERR_ARG;
  BcArgs a{{n0, n1, n2}, {g0, g1, g2}, {bc6[0], bc6[2], bc6[4]}, {bc6[1], bc6[3], bc6[5]}};
  for (int d = 0; d < 3; ++d) {
    if (a.g[d] < 0 || a.n[d] < 1) return FVHOE_ERR_ARG;
    if (a.g[d] > 0 && a.g[d] > a.n[d]) return FVHOE_ERR_ARG;
    if (a.lo[d] < 0 || a.lo[d] > 3 || a.hi[d] < 0 || a.hi[d] > 3) return FVHOE_ERR_UNSUPPORTED;  // no slab kinds
  }
  if (p < 1 || p > 8) return FVHOE_ERR_UNSUPPORTED;
  const long long N0 = n0 + 2 * g0, N1 = n1 + 2 * g1, N2 = n2 + 2 * g2;
  if ((axis == 1 ? N0 : N1) < ntaps(p)) return FVHOE_ERR_ARG;
  const long long total = (long long)nv * N0 * N1 * N2;
  if (total < (1ll << 16) || total >= (1ll << 31)) return FVHOE_ERR_UNSUPPORTED;  // small grids: fill + sweep
  const hipStream_t st = (hipStream_t)stream;
  const int r = kind == FVHOE_STENCIL_C ? launch_bc_slide_p<T, KIND_C>(p, u, out, a, nv, axis, st)
                                        : launch_bc_slide_p<T, KIND_T>(p, u, out, a, nv, axis, st);
  if (r != 0) return FVHOE_ERR_UNSUPPORTED;
  return launch_status();
}

template <typename T>
int euler_equation_rk_impl(const T* F, const long long* fst, const T* G, const long long* gst, const T* H,
                           const long long* hst, double* dudt, int nv, int n0, int n1, int n2, int active, double cx,
                           double cy, double cz, const FvhoeRkFuse* rkf, void* stream) {
  if (!rkf || nv < 1 || n0 < 1 || n1 < 1 || n2 < 1) return FVHOE_ERR_ARG;
  if (((active & 1) && !F) || ((active & 2) && !G) || ((active & 4) && !H)) return FVHOE_ERR_ARG;
  const int kind = rkf->kind;
  if (kind < 0 || kind > FVHOE_RK_SSPRK3_FINAL_SUM || !rkf->out || !rkf->u0 || !rkf->dt_step) return FVHOE_ERR_ARG;
  const bool need_k0 = kind == FVHOE_RK_SSPRK3_STAGE3 || kind == FVHOE_RK_SSPRK3_FINAL ||
                       kind == FVHOE_RK_SSPRK2_FINAL || kind == FVHOE_RK_RK4_FINAL ||
                       kind == FVHOE_RK_SSPRK3_STAGE3_SUM || kind == FVHOE_RK_SSPRK3_FINAL_SUM;
  if (need_k0 && !rkf->k0) return FVHOE_ERR_ARG;
  if ((kind == FVHOE_RK_SSPRK3_FINAL || kind == FVHOE_RK_RK4_FINAL) && !rkf->k1) return FVHOE_ERR_ARG;
  if (kind == FVHOE_RK_RK4_FINAL && !rkf->k2) return FVHOE_ERR_ARG;
  if (kind == FVHOE_RK_SSPRK3_STAGE3_SUM && !dudt) return FVHOE_ERR_ARG;
  const long long total = (long long)nv * n0 * n1 * n2;
  if (total >= (1ll << 31)) return FVHOE_ERR_UNSUPPORTED;  // rk_combine takes 32-bit indices
  const long long N0 = n0, N1 = n1, N2 = n2;
  Strides3 fs{(N0 + 1) * N1 * N2, N1 * N2, N2}, gs{N0 * (N1 + 1) * N2, (N1 + 1) * N2, N2},
      hs{N0 * N1 * (N2 + 1), N1 * (N2 + 1), N2 + 1};
  if (fst) fs = Strides3{fst[0], fst[1], fst[2]};
  if (gst) gs = Strides3{gst[0], gst[1], gst[2]};
  if (hst) hs = Strides3{hst[0], hst[1], hst[2]};
  const RkArgs rk{kind, rkf->out, rkf->u0, rkf->k0, rkf->k1, rkf->k2, rkf->dt_step};
  dim3 rg;
  if (rows_grid(nv, n0, n1, n2, &rg))
    k_euler_eq_rows<T, true><<<rg, 256, 0, (hipStream_t)stream>>>(F, G, H, dudt, n0, n1, n2, active, cx, cy, cz, fs,
                                                                  gs, hs, rk);
  else
    k_euler_eq_rk<T><<<grid_for(total), TPB, 0, (hipStream_t)stream>>>(F, G, H, dudt, nv, n0, n1, n2, active, cx,
                                                                       cy, cz, fs, gs, hs, rk);
  return launch_status();
}

}  // namespace

// ------------------------------------------------------------------- C-ABI

extern "C" {

int fvhoe_version(void) { return FVHOE_ABI_VERSION; }

int fvhoe_stencil_weights(int p, int kind, double* out, int* n) {
  if (p < 0 || p > 8 || kind < 0 || kind > 3 || !out || !n) return FVHOE_ERR_ARG;
  *n = ntaps(p);
  for (int i = 0; i < *n; ++i) out[i] = weight(kind, p, i);
  return FVHOE_OK;
}

int fvhoe_stencil_sweep(const double* u, double* out, int nv, int n0, int n1, int n2, int kind, int p, int axis,
                        void* stream) {
  return stencil_sweep_impl<double>(u, out, nv, n0, n1, n2, kind, p, axis, stream);
}

int fvhoe_stencil_sweep_f32(const float* u, float* out, int nv, int n0, int n1, int n2, int kind, int p, int axis,
                            void* stream) {
  return stencil_sweep_impl<float>(u, out, nv, n0, n1, n2, kind, p, axis, stream);
}

int fvhoe_compute_primitives(const double* u, double* w, int nv, long long ncell, double gamma, void* stream) {
  return compute_primitives_impl<double>(u, w, nv, ncell, gamma, stream);
}

int fvhoe_compute_primitives_f32(const float* u, float* w, int nv, long long ncell, double gamma, void* stream) {
  return compute_primitives_impl<float>(u, w, nv, ncell, gamma, stream);
}

int fvhoe_compute_conservatives(const double* w, double* u, int nv, long long ncell, double gamma, void* stream) {
  if (!u || !w || nv < 5 || ncell < 0) return FVHOE_ERR_ARG;
  if (ncell == 0) return FVHOE_OK;
  k_cons<<<grid_for(ncell), TPB, 0, (hipStream_t)stream>>>(w, u, nv, ncell, gamma);
  return launch_status();
}

int fvhoe_compute_sound_speed(const double* w, double* c, long long ncell, double gamma, double csq_floor,
                              void* stream) {
  if (!w || !c || ncell < 0) return FVHOE_ERR_ARG;
  if (ncell == 0) return FVHOE_OK;
  k_sound<<<grid_for(ncell), TPB, 0, (hipStream_t)stream>>>(w, c, ncell, gamma, csq_floor);
  return launch_status();
}

int fvhoe_hydro_dt(const double* w, long long ncell, double h, int ndim, double CFL, double gamma, double csq_floor,
                   double* dt_out, unsigned long long* scratch, void* stream) {
  return hydro_dt_impl<double>(w, ncell, h, ndim, CFL, gamma, csq_floor, dt_out, scratch, stream);
}

int fvhoe_hydro_dt_f32(const float* w, long long ncell, double h, int ndim, double CFL, double gamma,
                       double csq_floor, double* dt_out, unsigned long long* scratch, void* stream) {
  return hydro_dt_impl<float>(w, ncell, h, ndim, CFL, gamma, csq_floor, dt_out, scratch, stream);
}

int fvhoe_riemann(const double* wl, const double* wr, double* out, int nv, long long ncell, int solver, int dim,
                  double gamma, double csq_floor, void* stream) {
  if (!wl || !wr || !out || nv < 5 || dim < 0 || dim > 2 || solver < 0 || solver > 2) return FVHOE_ERR_ARG;
  if (ncell == 0) return FVHOE_OK;
  hipStream_t s = (hipStream_t)stream;
  const int g = grid_for(ncell);
  if (dim == 0) k_riemann_pt<0><<<g, TPB, 0, s>>>(wl, wr, out, nv, ncell, solver, gamma, csq_floor);
  else if (dim == 1) k_riemann_pt<1><<<g, TPB, 0, s>>>(wl, wr, out, nv, ncell, solver, gamma, csq_floor);
  else k_riemann_pt<2><<<g, TPB, 0, s>>>(wl, wr, out, nv, ncell, solver, gamma, csq_floor);
  return launch_status();
}

int fvhoe_riemann_faces(const double* wr_face, const double* wl_face, double* out, int n0, int n1, int n2, int axis,
                        int solver, double gamma, double csq_floor, void* stream) {
  if (!wr_face || !wl_face || !out || axis < 1 || axis > 3 || solver < 0 || solver > 2) return FVHOE_ERR_ARG;
  const long long nout = (long long)(axis == 1 ? n0 - 1 : n0) * (axis == 2 ? n1 - 1 : n1) * (axis == 3 ? n2 - 1 : n2);
  if (nout <= 0) return nout == 0 ? FVHOE_OK : FVHOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int g = grid_for(nout);
  if (axis == 1) k_riemann_faces<0><<<g, TPB, 0, s>>>(wr_face, wl_face, out, n0, n1, n2, axis, solver, gamma, csq_floor);
  else if (axis == 2) k_riemann_faces<1><<<g, TPB, 0, s>>>(wr_face, wl_face, out, n0, n1, n2, axis, solver, gamma, csq_floor);
  else k_riemann_faces<2><<<g, TPB, 0, s>>>(wr_face, wl_face, out, n0, n1, n2, axis, solver, gamma, csq_floor);
  return launch_status();
}

int fvhoe_stencil_sweep_zpair(const double* u, double* out, int nv, int n0, int n1, int n2, int kind, int p, int axis,
                              int zfirst, void* stream) {
  return zpair_impl<double>(u, out, nv, n0, n1, n2, kind, p, axis, zfirst, stream);
}

int fvhoe_stencil_sweep_zpair_f32(const float* u, float* out, int nv, int n0, int n1, int n2, int kind, int p,
                                  int axis, int zfirst, void* stream) {
  return zpair_impl<float>(u, out, nv, n0, n1, n2, kind, p, axis, zfirst, stream);
}

int fvhoe_stencil_sweep_xypair(const double* u, double* out, int nv, int n0, int n1, int n2, int kind, int p,
                               int first_axis, void* stream) {
  return xypair_impl<double>(u, out, nv, n0, n1, n2, kind, p, first_axis, stream);
}

int fvhoe_stencil_sweep_xypair_f32(const float* u, float* out, int nv, int n0, int n1, int n2, int kind, int p,
                                   int first_axis, void* stream) {
  return xypair_impl<float>(u, out, nv, n0, n1, n2, kind, p, first_axis, stream);
}

int fvhoe_face_fluxes(const double* w, double* out, int n0, int n1, int n2, int axis, int p, int solver, double gamma,
                      double csq_floor, void* stream) {
  return face_fluxes_impl<double>(w, out, n0, n1, n2, axis, p, solver, gamma, csq_floor, stream);
}

int fvhoe_face_fluxes_f32(const float* w, float* out, int n0, int n1, int n2, int axis, int p, int solver,
                          double gamma, double csq_floor, void* stream) {
  return face_fluxes_impl<float>(w, out, n0, n1, n2, axis, p, solver, gamma, csq_floor, stream);
}

int fvhoe_slope_limit(int limiter, const double* du_left, const double* du_right, double* out, long long n,
                      void* stream) {
  if (!du_left || !du_right || !out || n < 0 || (limiter != LIM_MINMOD && limiter != LIM_MONCEN))
    return FVHOE_ERR_ARG;
  if (n == 0) return FVHOE_OK;
  k_limit<<<grid_for(n), TPB, 0, (hipStream_t)stream>>>(limiter, du_left, du_right, out, n);
  return launch_status();
}

int fvhoe_central_difference(const double* u, double* out, long long outer, long long len, long long inner,
                             double h, int has_h, void* stream) {
  if (!u || !out || outer < 0 || len < 2 || inner < 0) return FVHOE_ERR_ARG;
  if (outer * (len - 2) * inner == 0) return FVHOE_OK;
  k_central_diff<<<grid_for(outer * (len - 2) * inner), TPB, 0, (hipStream_t)stream>>>(u, out, outer, len, inner, h,
                                                                                        has_h);
  return launch_status();
}

int fvhoe_compute_fluxes(const double* w, const double* u, double* out, int nv, long long ncell, int dim,
                         int include_pressure, void* stream) {
  if (!w || !u || !out || nv < 5 || ncell < 0 || dim < 0 || dim > 2) return FVHOE_ERR_ARG;
  if (ncell == 0) return FVHOE_OK;
  k_fluxes<<<grid_for(ncell), TPB, 0, (hipStream_t)stream>>>(w, u, out, nv, ncell, dim, include_pressure);
  return launch_status();
}

int fvhoe_muscl(const double* u, double* left, double* right, int nv, int n0, int n1, int n2, int axis, int limiter,
                void* stream) {
  if (!u || !left || !right || axis < 1 || axis > 3 || limiter < 0 || limiter > 2) return FVHOE_ERR_ARG;
  const int na = axis == 1 ? n0 : axis == 2 ? n1 : n2;
  if (na < 3) return FVHOE_ERR_ARG;
  const long long total = (long long)nv * n0 * n1 * n2;
  k_muscl<<<grid_for(total), TPB, 0, (hipStream_t)stream>>>(u, left, right, nv, n0, n1, n2, axis, limiter);
  return launch_status();
}

int fvhoe_bc_apply(const double* u, double* out, int nv, int n0, int n1, int n2, int g0, int g1, int g2,
                   const int* bc6, void* stream) {
  return bc_apply_impl<double>(u, out, nv, n0, n1, n2, g0, g1, g2, bc6, stream);
}

int fvhoe_bc_apply_f32(const double* u, float* out, int nv, int n0, int n1, int n2, int g0, int g1, int g2,
                       const int* bc6, void* stream) {
  return bc_apply_impl<float>(u, out, nv, n0, n1, n2, g0, g1, g2, bc6, stream);
}

int fvhoe_bc_stencil_sweep(const double* u, double* out, int nv, int n0, int n1, int n2, int g0, int g1, int g2,
                           const int* bc6, int kind, int p, int axis, void* stream) {
  return bc_stencil_sweep_impl<double>(u, out, nv, n0, n1, n2, g0, g1, g2, bc6, kind, p, axis, stream);
}

int fvhoe_bc_stencil_sweep_f32(const double* u, float* out, int nv, int n0, int n1, int n2, int g0, int g1, int g2,
                               const int* bc6, int kind, int p, int axis, void* stream) {
  return bc_stencil_sweep_impl<float>(u, out, nv, n0, n1, n2, g0, g1, g2, bc6, kind, p, axis, stream);
}

int fvhoe_bc_apply_dirichlet(const double* u, double* out, int nv, int n0, int n1, int n2, int g0, int g1, int g2,
                             const int* bc6, const double* const* slabs6, void* stream) {
  if (!u || !out || !bc6 || !slabs6 || nv < 1) return FVHOE_ERR_ARG;
  BcArgs a{{n0, n1, n2}, {g0, g1, g2}, {bc6[0], bc6[2], bc6[4]}, {bc6[1], bc6[3], bc6[5]}};
  Slabs sl{};
  for (int d = 0; d < 3; ++d) {
    if (a.g[d] < 0 || a.n[d] < 1) return FVHOE_ERR_ARG;
    sl.g[d] = sl.pad[d] = a.g[d];  // the slabs are the ghost regions of out, corners included
  }
  for (int side = 0; side < 6; ++side) {
    const int d = side / 2, kind = bc6[side];
    if (kind < 0 || kind > BC_DIRICHLET) return FVHOE_ERR_UNSUPPORTED;
    sl.s[side] = nullptr;
    if (a.g[d] == 0) continue;
    if (kind == BC_DIRICHLET) {
      if (!slabs6[side]) return FVHOE_ERR_ARG;
      sl.s[side] = slabs6[side];
    } else if (a.g[d] > a.n[d]) {
      return FVHOE_ERR_ARG;  // slabs of the other kinds come from the interior
    }
  }
  const long long total = (long long)nv * (n0 + 2 * g0) * (n1 + 2 * g1) * (n2 + 2 * g2);
  k_bc_apply_slabs<<<grid_for(total), TPB, 0, (hipStream_t)stream>>>(u, out, nv, a, sl);
  return launch_status();
}

int fvhoe_euler_equation(const double* F, const double* G, const double* H, double* dudt, int nv, int n0, int n1,
                         int n2, int active, double cx, double cy, double cz, void* stream) {
  return euler_equation_impl<double>(F, nullptr, G, nullptr, H, nullptr, dudt, nv, n0, n1, n2, active, cx, cy, cz,
                                     stream);
}

int fvhoe_euler_equation_strided(const double* F, const long long* fst, const double* G, const long long* gst,
                                 const double* H, const long long* hst, double* dudt, int nv, int n0, int n1, int n2,
                                 int active, double cx, double cy, double cz, void* stream) {
  return euler_equation_impl<double>(F, fst, G, gst, H, hst, dudt, nv, n0, n1, n2, active, cx, cy, cz, stream);
}

int fvhoe_euler_equation_rk(const double* F, const long long* fst, const double* G, const long long* gst,
                            const double* H, const long long* hst, double* dudt, int nv, int n0, int n1, int n2,
                            int active, double cx, double cy, double cz, const FvhoeRkFuse* rk, void* stream) {
  return euler_equation_rk_impl<double>(F, fst, G, gst, H, hst, dudt, nv, n0, n1, n2, active, cx, cy, cz, rk, stream);
}

int fvhoe_euler_equation_rk_f32(const float* F, const long long* fst, const float* G, const long long* gst,
                                const float* H, const long long* hst, double* dudt, int nv, int n0, int n1, int n2,
                                int active, double cx, double cy, double cz, const FvhoeRkFuse* rk, void* stream) {
  return euler_equation_rk_impl<float>(F, fst, G, gst, H, hst, dudt, nv, n0, n1, n2, active, cx, cy, cz, rk, stream);
}

int fvhoe_euler_equation_strided_f32(const float* F, const long long* fst, const float* G, const long long* gst,
                                     const float* H, const long long* hst, double* dudt, int nv, int n0, int n1,
                                     int n2, int active, double cx, double cy, double cz, void* stream) {
  return euler_equation_impl<float>(F, fst, G, gst, H, hst, dudt, nv, n0, n1, n2, active, cx, cy, cz, stream);
}

int fvhoe_euler_equation_f32(const float* F, const float* G, const float* H, double* dudt, int nv, int n0, int n1,
                             int n2, int active, double cx, double cy, double cz, void* stream) {
  return euler_equation_impl<float>(F, nullptr, G, nullptr, H, nullptr, dudt, nv, n0, n1, n2, active, cx, cy, cz,
                                    stream);
}

int fvhoe_minmax(const double* u, int nv, long long ncell, double* out, unsigned long long* scratch, void* stream) {
  if (!u || !out || !scratch || nv < 1 || nv > 64 || ncell <= 0) return FVHOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  k_init_minmax<<<1, 128, 0, s>>>(scratch, nv);
  dim3 grid(grid_for(ncell) < 1024 ? grid_for(ncell) : 1024, nv);
  k_minmax_partial<<<grid, TPB, 0, s>>>(u, ncell, scratch);
  k_minmax_final<<<1, 128, 0, s>>>(scratch, out, nv);
  return launch_status();
}

int fvhoe_detect_troubled_cells(const double* w, const double* ws, int N0, int N1, int N2, int dims,
                                const double* range_minmax, int mode, int range_type, double eps, int nad_vars,
                                int pad_var, double pad_lo, double pad_hi, int sed, double sed_eps,
                                signed char* trouble, double* nad_mag, double* pad_mag, void* stream) {
  if (!w || !ws || !trouble || !nad_mag || !pad_mag || dims < 1 || dims > 7) return FVHOE_ERR_ARG;
  if (mode < 0 || mode > 1 || range_type < 0 || range_type > 1 || pad_var > 4) return FVHOE_ERR_ARG;
  if (mode == 0 && !range_minmax) return FVHOE_ERR_ARG;
  DetectArgs a{{N0, N1, N2}, dims, mode, range_type, nad_vars, pad_var, sed, eps, pad_lo, pad_hi, sed_eps};
  const long long nc = (long long)(N0 - ((dims & 1) ? 6 : 0)) * (N1 - ((dims & 2) ? 6 : 0)) * (N2 - ((dims & 4) ? 6 : 0));
  if (nc < 0) return FVHOE_ERR_ARG;
  if (nc == 0) return FVHOE_OK;
  k_detect<<<grid_for(nc), TPB, 0, (hipStream_t)stream>>>(w, ws, range_minmax, trouble, nad_mag, pad_mag, a);
  return launch_status();
}

int fvhoe_smooth_extrema_detector(const double* u, double* alpha, int nv, int N0, int N1, int N2, int dims,
                                  double eps, void* stream) {
  if (!u || !alpha || nv < 1 || dims < 1 || dims > 7) return FVHOE_ERR_ARG;
  const long long nc = (long long)(N0 - ((dims & 1) ? 6 : 0)) * (N1 - ((dims & 2) ? 6 : 0)) * (N2 - ((dims & 4) ? 6 : 0));
  if (nc < 0) return FVHOE_ERR_ARG;
  if (nc == 0) return FVHOE_OK;
  k_sed<<<grid_for(nc * nv), TPB, 0, (hipStream_t)stream>>>(u, alpha, nv, N0, N1, N2, dims, eps);
  return launch_status();
}

int fvhoe_broadcast_simple(const signed char* trouble, double* tx, double* ty, double* tz, int n0, int n1, int n2,
                           int dims, void* stream) {
  if (!trouble || !tx || !ty || !tz) return FVHOE_ERR_ARG;
  const long long total = (long long)(n0 + 1) * n1 * n2 + (long long)n0 * (n1 + 1) * n2 + (long long)n0 * n1 * (n2 + 1);
  k_broadcast_simple<<<grid_for(total), TPB, 0, (hipStream_t)stream>>>(trouble, tx, ty, tz, n0, n1, n2, dims);
  return launch_status();
}

int fvhoe_shift_max(double* out, const long long* out_strides, long long out_offset, const double* s1,
                    const long long* s1_strides, long long s1_offset, double coef, int gt, const double* s2,
                    const long long* s2_strides, long long s2_offset, const int* extent, void* stream) {
  if (!out || !s1 || !s2 || !out_strides || !s1_strides || !s2_strides || !extent) return FVHOE_ERR_ARG;
  ShiftMaxArgs a{};
  for (int d = 0; d < 3; ++d) {
    a.ext[d] = extent[d];
    a.os[d] = out_strides[d];
    a.ss[d] = s1_strides[d];
    a.ts[d] = s2_strides[d];
    if (extent[d] < 0) return FVHOE_ERR_ARG;
  }
  a.ooff = out_offset;
  a.soff = s1_offset;
  a.toff = s2_offset;
  a.coef = coef;
  a.gt = gt;
  const long long total = (long long)a.ext[0] * a.ext[1] * a.ext[2];
  if (total == 0) return FVHOE_OK;
  k_shift_max<<<grid_for(total), TPB, 0, (hipStream_t)stream>>>(out, s1, s2, a);
  return launch_status();
}

int fvhoe_blend(double* F, const double* Fl, const double* theta, int nv, long long nface, void* stream) {
  if (!F || !Fl || !theta || nv < 1 || nface < 0) return FVHOE_ERR_ARG;
  if (nface == 0) return FVHOE_OK;
  k_blend<<<grid_for(nv * nface), TPB, 0, (hipStream_t)stream>>>(F, Fl, theta, nv, nface);
  return launch_status();
}

int fvhoe_rk_combine(int kind, double* out, const double* u, const double* k0, const double* k1, const double* k2,
                     const double* k3, const double* dt, long long n, void* stream) {
  if (!out || !u || !k0 || !dt || n < 0 || kind < 0 || kind > FVHOE_RK_RK4_FINAL) return FVHOE_ERR_ARG;
  if ((kind == FVHOE_RK_SSPRK3_STAGE3 || kind == FVHOE_RK_SSPRK2_FINAL) && !k1) return FVHOE_ERR_ARG;
  if (kind == FVHOE_RK_SSPRK3_FINAL && (!k1 || !k2)) return FVHOE_ERR_ARG;
  if (kind == FVHOE_RK_RK4_FINAL && (!k1 || !k2 || !k3)) return FVHOE_ERR_ARG;
  if (n == 0) return FVHOE_OK;
  k_rk<<<grid_for(n), TPB, 0, (hipStream_t)stream>>>(kind, out, u, k0, k1, k2, k3, dt, n);
  return launch_status();
}

int fvhoe_log_troubles(double* mean, const signed char* trouble, long long n, int mode, double divisor,
                       void* stream) {
  if (!mean || n < 0 || mode < 0 || mode > 2 || (mode < 2 && !trouble)) return FVHOE_ERR_ARG;
  if (n == 0) return FVHOE_OK;
  k_log<<<grid_for(n), TPB, 0, (hipStream_t)stream>>>(mean, trouble, n, mode, divisor, nullptr);
  return launch_status();
}

int fvhoe_log_troubles_gated(double* mean, const signed char* trouble, long long n, int mode, double divisor,
                             const double* gate, void* stream) {
  if (!mean || !gate || n < 0 || mode < 0 || mode > 2 || (mode < 2 && !trouble)) return FVHOE_ERR_ARG;
  if (n == 0) return FVHOE_OK;
  k_log<<<grid_for(n), TPB, 0, (hipStream_t)stream>>>(mean, trouble, n, mode, divisor, gate);
  return launch_status();
}

int fvhoe_log_count(int* count, const signed char* trouble, long long n, int mode, const double* gate, void* stream) {
  if (!count || !trouble || n < 0 || mode < 0 || mode > 1) return FVHOE_ERR_ARG;
  if (n == 0) return FVHOE_OK;
  if (n % 16 == 0 && ((uintptr_t)count % 16) == 0 && ((uintptr_t)trouble % 16) == 0) {
    k_log_count16<<<grid_for(n / 16), TPB, 0, (hipStream_t)stream>>>(count, (const uint4*)trouble, n / 16, mode,
                                                                      gate);
  } else {
    k_log_count<<<grid_for(n), TPB, 0, (hipStream_t)stream>>>(count, trouble, n, mode, gate);
  }
  return launch_status();
}

int fvhoe_log_mean(double* mean, const int* count, long long n, double divisor, const double* gate, void* stream) {
  if (!mean || !count || n < 0) return FVHOE_ERR_ARG;
  if (n == 0) return FVHOE_OK;
  k_log_mean<<<grid_for(n), TPB, 0, (hipStream_t)stream>>>(mean, count, n, divisor, gate);
  return launch_status();
}

int fvhoe_log_magnitudes(double* nad_out, double* pad_out, const double* nad_base, const double* pad_base,
                         long long n, const int* count, const int* idx, const double* nad, const double* pad,
                         int nslots, long long cap, const double* divisor, void* stream) {
  if (!nad_out || !pad_out || n < 0 || nslots < 0 || cap < 0 || (nslots > 0 && (!count || !idx || !nad || !pad)))
    return FVHOE_ERR_ARG;
  if ((nad_base == nullptr) != (pad_base == nullptr) || n > 2147483647LL || cap > 2147483647LL) return FVHOE_ERR_ARG;
  if (n == 0) return FVHOE_OK;
  hipStream_t s = (hipStream_t)stream;
  k_mag_base<<<grid_for(n), TPB, 0, s>>>(nad_out, pad_out, nad_base, pad_base, n);
  const int gs = grid_for(cap > 0 ? cap : 1);
  for (int k = 0; k < nslots; ++k)
    k_mag_scatter<<<gs, TPB, 0, s>>>(nad_out, pad_out, count + k, idx + k * cap, nad + k * cap, pad + k * cap, n,
                                     cap);
  if (divisor) k_mag_divide<<<grid_for(n), TPB, 0, s>>>(nad_out, pad_out, n, divisor);
  return launch_status();
}

int fvhoe_set_scalar_gated(double* p, double v, const double* gate, void* stream) {
  if (!p) return FVHOE_ERR_ARG;
  k_set_scalar_gated<<<1, 1, 0, (hipStream_t)stream>>>(p, v, gate);
  return launch_status();
}

int fvhoe_set_scalar(double* p, double v, void* stream) {
  if (!p) return FVHOE_ERR_ARG;
  k_set_scalar<<<1, 1, 0, (hipStream_t)stream>>>(p, v);
  return launch_status();
}

int fvhoe_hit_target(double* dt_step, const double* dt_stage, const double* t, double target, int has_target,
                     void* stream) {
  if (!dt_step || !dt_stage || !t) return FVHOE_ERR_ARG;
  k_hit_target<<<1, 1, 0, (hipStream_t)stream>>>(dt_step, dt_stage, t, target, has_target);
  return launch_status();
}

int fvhoe_advance_time(double* t, const double* dt, void* stream) {
  if (!t || !dt) return FVHOE_ERR_ARG;
  k_advance_time<<<1, 1, 0, (hipStream_t)stream>>>(t, dt);
  return launch_status();
}

int fvhoe_advance_time_counted(double* t, const double* dt, long long* steps, double* times, long long cap,
                               void* stream) {
  if (!t || !dt || !steps || (cap > 0 && !times) || cap < 0) return FVHOE_ERR_ARG;
  k_advance_time_counted<<<1, 1, 0, (hipStream_t)stream>>>(t, dt, steps, times, cap);
  return launch_status();
}

}  // extern "C"
