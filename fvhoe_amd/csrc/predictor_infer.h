// Fused inference kernel of the troubled-cell predictor (included by the
// library sources that launch it).
//
// Persistent blocks of 256 threads (4 waves) walk over 16x16 output tiles.
// Per tile: the 28x28x4 feature patch is staged in LDS as bfloat16 (channel
// last), layer 1 (24x24 pixels) and layer 2 (20x20 pixels) run on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation (A = 16 pixels x 32 of K,
// B = weights held in registers for the whole kernel, D = 16 pixels x 16
// channels), activations go back to LDS as bfloat16.  Layer 3 (16 -> 1
// channel) is split as G[q][t] = sum_c a2[q][c] w3[c][t] for every layer-2
// pixel q and tap t (v_mfma_f32_16x16x16_bf16, taps as the N dimension, in two
// halves of 13 and 12 taps kept in FP32 in LDS) and z[p] = b3 + sum_t
// G[p + offset(t)][t], one pixel per thread.  Lane l of a wave holds A row
// l%16 and B column l%16 for the K slice 8*(l/16)..+7, and D rows
// 4*(l/16)..+3 of column l%16.
#pragma once
#include <hip/hip_runtime.h>

#include "fvhoe_hip.h"

namespace fvhoe {
namespace pred {

constexpr int KS = 5, KK = 25;
constexpr int C0 = 4, C1 = 16, C2 = 16;
constexpr int HALO = 6;
constexpr int PT = 16;  // output tile edge
constexpr int L0 = PT + 12, L1 = PT + 8, L2 = PT + 4;
constexpr int K1STEPS = 4;   // layer 1: K = 25 taps x 4 channels in 4 steps of 32 (tap = 8 step + k/4, channel k%4)
constexpr int K2STEPS = 13;  // layer 2: K = 25 taps x 16 channels in 13 steps of 32 (tap = 2 step + k/16, channel k%16)

// packed inference weights (bytes); taps beyond 24 are zero
constexpr int PW1 = 0;                            // bf16 [4][16][32]  (step, n, k)
constexpr int PW2 = PW1 + K1STEPS * 16 * 32 * 2;  // bf16 [13][16][32] (step, n, k)
constexpr int PW3 = PW2 + K2STEPS * 16 * 32 * 2;  // f32 [16][25]      (bf16-rounded values)
constexpr int PB1 = PW3 + C2 * KK * 4, PB2 = PB1 + 64, PB3 = PB2 + 64, PACKED = PB3 + 64;

// LDS carve-up of the inference kernel; the layer-3 tap sums G alias xin and a1
constexpr int GT = 13;  // taps per layer-3 half
constexpr int SM_XIN = 0, SM_A1 = SM_XIN + L0 * L0 * C0 * 2, SM_A2 = SM_A1 + L1 * L1 * C1 * 2,
              SM_TOTAL = SM_A2 + L2 * L2 * C2 * 2;
static_assert(L2 * L2 * GT * 4 <= SM_A2, "layer-3 tap sums fit over xin and a1");

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ unsigned short f2bf(float f) {  // round to nearest even
#ifdef __HIP_DEVICE_COMPILE__
  unsigned int b = __float_as_uint(f);
#else
  unsigned int b;
  __builtin_memcpy(&b, &f, 4);
#endif
  b += 0x7FFFu + ((b >> 16) & 1u);
  return (unsigned short)(b >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }
__device__ __forceinline__ float bf_lo(unsigned int w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(unsigned int w) { return __uint_as_float(w & 0xFFFF0000u); }

// Features from planar float X (4, n0+12, n1+12).
struct FeatPlanar {
  const float* __restrict__ X;
  int HX, WX;
  __device__ __forceinline__ void init(double*) const {}
  __device__ __forceinline__ void load(int gy, int gx, float f[4], const double*) const {
    const bool in = gy < HX && gx < WX;
#pragma unroll
    for (int c = 0; c < 4; ++c) f[c] = in ? X[((long long)c * HX + gy) * WX + gx] : 0.0f;
  }
};

// The four features of a pixel as bfloat16, the form the network stages them in.
template <typename Feat>
__device__ __forceinline__ bf16x4v load_bf16(const Feat& feat, int gy, int gx, const double* sh) {
  float f[4];
  feat.load(gy, gx, f, sh);
  return bf16x4v{(__bf16)f[0], (__bf16)f[1], (__bf16)f[2], (__bf16)f[3]};
}

// Features already rounded to bfloat16, channel last: XB[(gy * WX + gx) * 4 + c].
struct FeatBF16 {
  const bf16x4v* __restrict__ XB;
  int HX, WX;
  __device__ __forceinline__ void init(double*) const {}
};
template <>
__device__ __forceinline__ bf16x4v load_bf16<FeatBF16>(const FeatBF16& feat, int gy, int gx, const double*) {
  if (gy < feat.HX && gx < feat.WX) return feat.XB[(long long)gy * feat.WX + gx];
  return bf16x4v{(__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f};
}

constexpr int INFER_THREADS = 256;

// Activations are stored 16 channels (32 bytes) per pixel.  Element index of
// channel quad k (channels 4k..4k+3) of pixel q:
__device__ __forceinline__ int swz(int q, int k) { return q * 16 + 4 * k; }

template <typename Feat>
__global__ void __launch_bounds__(INFER_THREADS, 3) k_infer(Feat feat, int n0, int n1,
                                                        const unsigned char* __restrict__ pw, float thr,
                                                        signed char* __restrict__ flags,
                                                        float* __restrict__ logits, int* __restrict__ count,
                                                        int* __restrict__ list, int tiles1, int ntiles) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[SM_TOTAL];
  unsigned short* xin = (unsigned short*)(smem + SM_XIN);
  unsigned short* a1 = (unsigned short*)(smem + SM_A1);
  unsigned short* a2 = (unsigned short*)(smem + SM_A2);
  float* G = (float*)smem;
  __shared__ int wbase[4];
  __shared__ int bbase;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int mrow = lane & 15, kb = lane >> 4;
  __shared__ double fconst[8];
  feat.init(fconst);  // per-block constants of the feature map (visible after the first barrier)
  // B fragments for the whole kernel: lane (n = mrow, kb) holds B[8kb..8kb+7][n]
  bf16x8 wb1[K1STEPS], wb2[K2STEPS];
#pragma unroll
  for (int s = 0; s < K1STEPS; ++s) wb1[s] = *(const bf16x8*)(pw + PW1 + 2 * ((s * 16 + mrow) * 32 + kb * 8));
#pragma unroll
  for (int s = 0; s < K2STEPS; ++s) wb2[s] = *(const bf16x8*)(pw + PW2 + 2 * ((s * 16 + mrow) * 32 + kb * 8));
  // the products are formed as D = W x pixels, so lane (mrow, kb) receives channels 4kb..4kb+3 of pixel mrow
  float bias1[4], bias2[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    bias1[r] = ((const float*)(pw + PB1))[4 * kb + r];
    bias2[r] = ((const float*)(pw + PB2))[4 * kb + r];
  }
  const float bias3 = ((const float*)(pw + PB3))[0];
  // layer-3 B fragments: lane (n = mrow, kb) holds w3[4kb..4kb+3][tap 13h + n] for half h
  bf16x4 wb3[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int t = GT * h + mrow;
    const bool ok = mrow < (h ? KK - GT : GT);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      wb3[h][i] = ok ? (short)f2bf(((const float*)(pw + PW3))[(4 * kb + i) * KK + t]) : (short)0;
  }

  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int oy0 = (tile / tiles1) * PT, ox0 = (tile % tiles1) * PT;
    __syncthreads();  // the previous tile is done with xin / a2 / wbase
    for (int p = tid; p < L0 * L0; p += INFER_THREADS) {
      const int y = p / L0, x = p % L0;
      *(bf16x4v*)&xin[p * C0] = load_bf16(feat, oy0 + y, ox0 + x, fconst);
    }
    __syncthreads();
    // layer 1: 24x24 pixels = 36 groups of 16; lane slice = taps 8s+2kb and 8s+2kb+1, 4 channels each
    for (int grp = wave; grp < (L1 * L1) / 16; grp += 4) {
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
      const int p = grp * 16 + mrow, py = p / L1, px = p % L1;
      const unsigned short* base = &xin[(py * L0 + px) * C0];
#pragma unroll
      for (int s = 0; s < K1STEPS; ++s) {
        const int t0 = 8 * s + 2 * kb, t1 = t0 + 1;
        u16x4 lo = {0, 0, 0, 0}, hi = {0, 0, 0, 0};
        if (t0 < KK) lo = *(const u16x4*)(base + ((t0 / KS) * L0 + t0 % KS) * C0);
        if (t1 < KK) hi = *(const u16x4*)(base + ((t1 / KS) * L0 + t1 % KS) * C0);
        const u16x8 a = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wb1[s], __builtin_bit_cast(bf16x8, a), acc, 0, 0, 0);
      }
      // acc[j] = D[4kb + j][mrow]: channel 4kb + j of pixel grp*16 + mrow
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] = fmaxf(acc[j] + bias1[j], 0.f);
      const bf16x4v v = {(__bf16)acc[0], (__bf16)acc[1], (__bf16)acc[2], (__bf16)acc[3]};
      *(bf16x4v*)&a1[swz(grp * 16 + mrow, kb)] = v;
    }
    __syncthreads();
    // layer 2: 20x20 pixels = 25 groups of 16; lane slice = tap 2s + kb/2, channels 8(kb%2)..+7
    for (int grp = wave; grp < (L2 * L2) / 16; grp += 4) {
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
      const int p = grp * 16 + mrow, py = p / L2, px = p % L2;
      const int q0 = py * L1 + px, hb = kb >> 1;
#pragma unroll
      for (int s = 0; s < K2STEPS; ++s) {
        const int ta = 2 * s, tb = 2 * s + 1;
        const int q = q0 + (hb ? (tb / KS) * L1 + tb % KS : (ta / KS) * L1 + ta % KS);
        u16x8 a = {0, 0, 0, 0, 0, 0, 0, 0};
        if (tb < KK || !hb) a = *(const u16x8*)&a1[swz(q, 2 * (kb & 1))];
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wb2[s], __builtin_bit_cast(bf16x8, a), acc, 0, 0, 0);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] = fmaxf(acc[j] + bias2[j], 0.f);
      const bf16x4v v = {(__bf16)acc[0], (__bf16)acc[1], (__bf16)acc[2], (__bf16)acc[3]};
      *(bf16x4v*)&a2[swz(grp * 16 + mrow, kb)] = v;
    }
    __syncthreads();
    // layer 3 in two halves of taps: tap sums G on MFMA, then one gather-sum per output pixel
    const int y = tid / PT, x = tid % PT;
    float z = bias3;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int tb = GT * h, nt = h ? KK - GT : GT;
      for (int grp = wave; grp < (L2 * L2) / 16; grp += 4) {
        const bf16x4 a = *(const bf16x4*)&a2[swz(grp * 16 + mrow, kb)];
        floatx4 g = {0.f, 0.f, 0.f, 0.f};
        g = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wb3[h], a, g, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j)  // lane (mrow, kb): taps 4kb+j of pixel grp*16 + mrow
          if (4 * kb + j < nt) G[(grp * 16 + mrow) * GT + 4 * kb + j] = g[j];
      }
      __syncthreads();
#pragma unroll
      for (int n = 0; n < nt; ++n) {
        const int t = tb + n;
        z += G[((y + t / KS) * L2 + x + t % KS) * GT + n];
      }
      __syncthreads();
    }
    const int gy = oy0 + y, gx = ox0 + x;
    const bool inside = gy < n0 && gx < n1;
    const int hit = (inside && z > thr) ? 1 : 0;
    if (inside) {
      const long long o = (long long)gy * n1 + gx;
      if (flags) flags[o] = (signed char)hit;
      if (logits) logits[o] = z;
    }
    if (list) {  // compaction of the flagged cells: wave ballot, block prefix, one atomic per block
      const unsigned long long ball = __ballot(hit);
      if (lane == 0) wbase[wave] = __popcll(ball);
      __syncthreads();
      if (tid == 0) {
        int tot = 0;
        for (int w = 0; w < 4; ++w) {
          const int c = wbase[w];
          wbase[w] = tot;
          tot += c;
        }
        bbase = tot ? atomicAdd(count, tot) : 0;
      }
      __syncthreads();
      if (hit) list[bbase + wbase[wave] + __popcll(ball & ((1ull << lane) - 1ull))] = gy * n1 + gx;
    }
  }
}

// Grid of persistent blocks: up to 4 per CU on 256 CUs.
inline int infer_grid(int ntiles) { return ntiles < 1024 ? ntiles : 1024; }

template <typename Feat>
inline int launch_infer(const Feat& f, int n0, int n1, const void* packed, float thr, signed char* flags,
                        float* logits, int* count, int* list, hipStream_t s) {
  const int t0 = (n0 + PT - 1) / PT, t1 = (n1 + PT - 1) / PT, nt = t0 * t1;
  k_infer<Feat><<<infer_grid(nt), INFER_THREADS, 0, s>>>(f, n0, n1, (const unsigned char*)packed, thr, flags,
                                                         logits, count, list, t1, nt);
  return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
}

}  // namespace pred
}  // namespace fvhoe
