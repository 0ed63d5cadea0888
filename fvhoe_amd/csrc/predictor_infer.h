// Fused inference kernel of the troubled-cell predictor (included by the
// library sources that launch it).
//
// Persistent blocks of 256 threads (4 waves) walk over output tiles of 16
// rows x 28 columns (x = the contiguous axis).  Per tile the 28x40x4 feature
// patch is staged in LDS as bfloat16 (channel last); layer 1 (24x36 pixels)
// and layer 2 (20x32 pixels) run on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation, A = weights (16 output channels x 32 of K, held in registers
// for the whole kernel), B = 16 consecutive pixels of one row x 32 of K,
// D = 16 channels x 16 pixels.  Each wave owns a column group of 16 pixels
// and a run of output rows, and slides down the input rows: an input row's
// B operand (taps dx of one row dy offset) is loaded from LDS once and used
// by the 5 output rows it feeds (one accumulator per output row), so each
// activation is read from LDS 5 (rows + 4) / rows times (once per tap dx)
// instead of 25 times (once per tap).  K of layer 1: taps dx = 0..4 of one row x 4 channels (20 of 32, the
// rest zero); K of layer 2: a pair of taps dx x 16 channels (dx = 5 zero).
// Layer 3 (16 -> 1 channel) is split as G[q][t] = sum_c a2[q][c] w3[c][t]
// for every layer-2 pixel q and tap t (v_mfma_f32_16x16x16_bf16, taps as the
// M dimension, in two halves of 13 and 12 taps kept in FP32 in LDS) and
// z[p] = b3 + sum_t G[p + offset(t)][t].  Lane l of a wave holds A row l%16
// and B column l%16 for the K slice 8*(l/16)..+7, and D rows 4*(l/16)..+3 of
// column l%16.
#pragma once
#include <hip/hip_runtime.h>

#include "fvhoe_hip.h"

namespace fvhoe {
namespace pred {

constexpr int KS = 5, KK = 25;
constexpr int C0 = 4, C1 = 16, C2 = 16;
constexpr int HALO = 6;
constexpr int PTY = 16, PTX = 28;                 // output tile: rows x columns
constexpr int L2Y = PTY + 4, L2X = PTX + 4;       // layer-2 pixels (32 columns: 2 groups of 16)
constexpr int L1Y = PTY + 8, L1X = PTX + 8;       // layer-1 pixels (36 columns: groups at 0, 16, 20)
constexpr int L0Y = PTY + 12, L0X = PTX + 12;     // feature pixels
constexpr int R1 = 6, R2 = 10;                    // output rows per wave unit in layers 1 and 2
static_assert(L2X == 32 && L1Y % R1 == 0 && L2Y == 2 * R2, "tile geometry");

// packed inference weights (bytes)
constexpr int PW1 = 0;                        // bf16 [5 dy][16 n][32 k]: dx = k/4 (k < 20), channel k%4
constexpr int PW2 = PW1 + 5 * 16 * 32 * 2;    // bf16 [5 dy][3 pair][16 n][32 k]: dx = 2 pair + k/16, channel k%16
constexpr int PW3 = PW2 + 5 * 3 * 16 * 32 * 2;  // f32 [16][25] (bf16-rounded values)
constexpr int PB1 = PW3 + C2 * KK * 4, PB2 = PB1 + 64, PB3 = PB2 + 64, PACKED = PB3 + 64;

// packed weights of the planar network (the network on 1-D runs seen as 2-D states constant along y,
// where each 5x5 kernel acts as its sum over the y taps): bf16 [16 n][32 k] layer 1 (tap k/4, channel k%4,
// k < 20), bf16 [3 pair][16 n][32 k] layer 2 (tap 2 pair + k/16, channel k%16, tap 5 zero), f32 [16][5]
// layer 3, then the biases
constexpr int QW1 = 0, QW2 = QW1 + 16 * 32 * 2, QW3 = QW2 + 3 * 16 * 32 * 2, QB1 = QW3 + C2 * KS * 4,
              QB2 = QB1 + 64, QB3 = QB2 + 64, QPACKED = QB3 + 64;

// LDS carve-up of the inference kernel; the layer-3 tap sums G alias xin and a1
constexpr int GT = 13;  // taps per layer-3 half
constexpr int SM_XIN = 0, SM_A1 = SM_XIN + L0Y * L0X * C0 * 2, SM_A2 = SM_A1 + L1Y * L1X * C1 * 2,
              SM_TOTAL = SM_A2 + L2Y * L2X * C2 * 2;
static_assert(L2Y * L2X * GT * 4 <= SM_A2, "layer-3 tap sums fit over xin and a1");

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
__global__ void __launch_bounds__(INFER_THREADS, 2) k_infer(Feat feat, int n0, int n1,
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
  // A fragments for the whole kernel: lane (n = mrow, kb) holds A[n][8kb..8kb+7]
  bf16x8 wa1[KS], wa2[KS][3];
#pragma unroll
  for (int dy = 0; dy < KS; ++dy) {
    wa1[dy] = *(const bf16x8*)(pw + PW1 + 2 * ((dy * 16 + mrow) * 32 + kb * 8));
#pragma unroll
    for (int k = 0; k < 3; ++k)
      wa2[dy][k] = *(const bf16x8*)(pw + PW2 + 2 * (((dy * 3 + k) * 16 + mrow) * 32 + kb * 8));
  }
  // the products are formed as D = W x pixels, so lane (mrow, kb) receives channels 4kb..4kb+3 of pixel mrow
  float bias1[4], bias2[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    bias1[r] = ((const float*)(pw + PB1))[4 * kb + r];
    bias2[r] = ((const float*)(pw + PB2))[4 * kb + r];
  }
  const float bias3 = ((const float*)(pw + PB3))[0];
  // layer-3 A fragments: lane (n = mrow, kb) holds w3[4kb..4kb+3][tap 13h + n] for half h
  bf16x4 wb3[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int t = GT * h + mrow;
    const bool ok = mrow < (h ? KK - GT : GT);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      wb3[h][i] = ok ? (short)f2bf(((const float*)(pw + PW3))[(4 * kb + i) * KK + t]) : (short)0;
  }
  const u16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};

  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int oy0 = (tile / tiles1) * PTY, ox0 = (tile % tiles1) * PTX;
    __syncthreads();  // the previous tile is done with xin / a2 / G / wbase
    for (int p = tid; p < L0Y * L0X; p += INFER_THREADS) {
      const int y = p / L0X, x = p - (p / L0X) * L0X;
      *(bf16x4v*)&xin[p * C0] = load_bf16(feat, oy0 + y, ox0 + x, fconst);
    }
    __syncthreads();
    // layer 1: 12 units = 3 column groups (x0 = 0, 16, 20; 20..31 are computed twice, identically) x 4 runs
    // of R1 rows; lane slice kb: pixels x0 + mrow + 2kb, +1 (taps dx = 2kb, 2kb+1), 4 channels each
    for (int u = wave; u < 12; u += 4) {
      const int g = u % 3, y0 = (u / 3) * R1, x0 = g == 0 ? 0 : (g == 1 ? 16 : 20);
      floatx4 acc[R1];
#pragma unroll
      for (int j = 0; j < R1; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < R1 + 4; ++r) {
        const unsigned short* px = &xin[((y0 + r) * L0X + x0 + mrow + 2 * kb) * C0];
        // two pixels (dx = 2kb, 2kb + 1), channels 0..3 each, as two 8-byte reads (a pixel is 8-byte
        // aligned); for kb = 2 only dx = 4 (the K slots of dx = 5 stay zero), nothing for kb = 3
        u16x4 lo = {0, 0, 0, 0}, hi = {0, 0, 0, 0};
        if (kb < 3) lo = *(const u16x4*)px;
        if (kb < 2) hi = *(const u16x4*)(px + C0);
        const u16x8 b = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        const bf16x8 bb = __builtin_bit_cast(bf16x8, b);
#pragma unroll
        for (int dy = 0; dy < KS; ++dy) {
          const int j = r - dy;  // output row fed by input row r through tap row dy
          if (j >= 0 && j < R1) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa1[dy], bb, acc[j], 0, 0, 0);
        }
        const int j = r - (KS - 1);  // row j is complete: at most 5 accumulators are live at a time
        if (j >= 0) {  // acc[j][i] = channel 4kb + i of pixel (y0 + j, x0 + mrow)
          const bf16x4v v = {(__bf16)fmaxf(acc[j][0] + bias1[0], 0.f), (__bf16)fmaxf(acc[j][1] + bias1[1], 0.f),
                             (__bf16)fmaxf(acc[j][2] + bias1[2], 0.f), (__bf16)fmaxf(acc[j][3] + bias1[3], 0.f)};
          *(bf16x4v*)&a1[swz((y0 + j) * L1X + x0 + mrow, kb)] = v;
        }
        __builtin_amdgcn_sched_barrier(0);  // keep each row's loads with its products (register pressure)
      }
    }
    __syncthreads();
    // layer 2: 4 units = 2 column groups (x0 = 0, 16) x 2 runs of R2 rows, one per wave; K step k: taps
    // dx = 2k + kb/2 of the row, channels 8(kb%2)..+7 (dx = 5 is zero)
    {
      const int x0 = 16 * (wave & 1), y0 = R2 * (wave >> 1);
      floatx4 acc[R2];
#pragma unroll
      for (int j = 0; j < R2; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < R2 + 4; ++r) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int dx = 2 * k + (kb >> 1);
          u16x8 b = zero8;
          if (dx < KS) b = *(const u16x8*)&a1[swz((y0 + r) * L1X + x0 + mrow + dx, 2 * (kb & 1))];
          const bf16x8 bb = __builtin_bit_cast(bf16x8, b);
#pragma unroll
          for (int dy = 0; dy < KS; ++dy) {
            const int j = r - dy;
            if (j >= 0 && j < R2) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa2[dy][k], bb, acc[j], 0, 0, 0);
          }
        }
        const int j = r - (KS - 1);  // row j is complete
        if (j >= 0) {
          const bf16x4v v = {(__bf16)fmaxf(acc[j][0] + bias2[0], 0.f), (__bf16)fmaxf(acc[j][1] + bias2[1], 0.f),
                             (__bf16)fmaxf(acc[j][2] + bias2[2], 0.f), (__bf16)fmaxf(acc[j][3] + bias2[3], 0.f)};
          *(bf16x4v*)&a2[swz((y0 + j) * L2X + x0 + mrow, kb)] = v;
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __syncthreads();
    // layer 3 in two halves of taps: tap sums G on MFMA, then one gather-sum per output pixel; each thread
    // owns output pixels tid and tid + 256 (of 16 x 28 = 448)
    constexpr int NOUT = PTY * PTX;
    float z[2] = {bias3, bias3};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int tb = GT * h, nt = h ? KK - GT : GT;
      for (int grp = wave; grp < (L2Y * L2X) / 16; grp += 4) {
        const bf16x4 a = *(const bf16x4*)&a2[swz(grp * 16 + mrow, kb)];
        floatx4 gsum = {0.f, 0.f, 0.f, 0.f};
        gsum = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wb3[h], a, gsum, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j)  // lane (mrow, kb): taps 4kb+j of pixel grp*16 + mrow
          if (4 * kb + j < nt) G[(grp * 16 + mrow) * GT + 4 * kb + j] = gsum[j];
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int p = tid + s * INFER_THREADS;
        if (p < NOUT) {
          const int y = p / PTX, x = p - (p / PTX) * PTX;
#pragma unroll
          for (int n = 0; n < nt; ++n) {
            const int t = tb + n;
            z[s] += G[((y + t / KS) * L2X + x + t % KS) * GT + n];
          }
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int p = tid + s * INFER_THREADS;
      const int y = p / PTX, x = p - (p / PTX) * PTX;
      const int gy = oy0 + y, gx = ox0 + x;
      const bool inside = p < NOUT && gy < n0 && gx < n1;
      const int hit = (inside && z[s] > thr) ? 1 : 0;
      if (inside) {
        const long long o = (long long)gy * n1 + gx;
        if (flags) flags[o] = (signed char)hit;
        if (logits) logits[o] = z[s];
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
        __syncthreads();  // wbase / bbase are reused by the next slot
      }
    }
  }
}

// Persistent blocks: up to 2 per CU on 256 CUs (57 KB of LDS each) when the network runs alone;
// fewer when it runs beside a stage's kernels, which need LDS and CUs of their own.
#ifndef FVHOE_INFER_BLOCKS
#define FVHOE_INFER_BLOCKS 512
#endif
#ifndef FVHOE_INFER_BLOCKS_BESIDE
#define FVHOE_INFER_BLOCKS_BESIDE 256
#endif

template <typename Feat>
inline int launch_infer(const Feat& f, int n0, int n1, const void* packed, float thr, signed char* flags,
                        float* logits, int* count, int* list, hipStream_t s, bool beside = false) {
  const int t0 = (n0 + PTY - 1) / PTY, t1 = (n1 + PTX - 1) / PTX, nt = t0 * t1;
  const int cap = beside ? FVHOE_INFER_BLOCKS_BESIDE : FVHOE_INFER_BLOCKS;
  k_infer<Feat><<<nt < cap ? nt : cap, INFER_THREADS, 0, s>>>(f, n0, n1, (const unsigned char*)packed, thr, flags,
                                                              logits, count, list, t1, nt);
  return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
}

}  // namespace pred
}  // namespace fvhoe
