// Learned troubled-cell predictor (no counterpart in FVHOE): a three-layer
// 5x5 convolutional network 4 -> 16 -> 16 -> 1 channels on normalized
// primitive fv averages (rho, vx, vy, P), predicting the a-posteriori flags.
//
// Inference: one fused kernel per 16x16 tile of cells, layers 1 and 2 on
// BF16 MFMA (v_mfma_f32_16x16x16_bf16, FP32 accumulation), layer 3 in FP32
// SIMT; inputs and activations are BF16 in LDS.
// Training: FP32 SIMT forward / backward / Adam kernels with the weights
// rounded to BF16 in the forward pass (the gradient is applied to the FP32
// master weights).
#include "fvhoe_common.h"
#include "fvhoe_hip.h"
#include "predictor.h"

using namespace fvhoe;

namespace {

constexpr int KS = 5, KK = 25;
constexpr int C0 = 4, C1 = 16, C2 = 16;
constexpr int HALO = 6;
constexpr int PT = 16;             // output tile edge
constexpr int L0 = PT + 12, L1 = PT + 8, L2 = PT + 4;
constexpr int K1STEPS = 7;         // ceil(25 taps * 4 channels / 16)

// parameter vector layout (float32): w1 (16,4,5,5), b1 (16), w2 (16,16,5,5), b2 (16), w3 (1,16,5,5), b3 (1)
constexpr int OW1 = 0, OB1 = OW1 + C1 * C0 * KK, OW2 = OB1 + C1, OB2 = OW2 + C2 * C1 * KK, OW3 = OB2 + C2,
              OB3 = OW3 + C2 * KK, NPARAM = OB3 + 1;
static_assert(NPARAM == FVHOE_PRED_NPARAMS, "parameter count");

// packed inference weights (bytes)
constexpr int PW1 = 0;                                   // bf16 [7][16][16]  (k-step, n, k)
constexpr int PW2 = PW1 + K1STEPS * 16 * 16 * 2;         // bf16 [25][16][16] (tap, n, cin)
constexpr int PW3 = PW2 + KK * 16 * 16 * 2;              // f32 [16][25]      (bf16-rounded values)
constexpr int PB1 = PW3 + C2 * KK * 4, PB2 = PB1 + 64, PB3 = PB2 + 64, PACKED = PB3 + 64;
static_assert(PACKED <= FVHOE_PRED_PACKED_BYTES, "packed size");

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
__device__ __forceinline__ float bfround(float f) { return bf2f(f2bf(f)); }

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// weight packing for inference
// ---------------------------------------------------------------------------
__global__ void k_pack(const float* __restrict__ th, unsigned char* __restrict__ out) {
  unsigned short* w1s = (unsigned short*)(out + PW1);
  unsigned short* w2s = (unsigned short*)(out + PW2);
  float* w3s = (float*)(out + PW3);
  float* b1 = (float*)(out + PB1);
  float* b2 = (float*)(out + PB2);
  float* b3 = (float*)(out + PB3);
  for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < K1STEPS * 256 + KK * 256 + C2 * KK + 33;
       i += blockDim.x * gridDim.x) {
    if (i < K1STEPS * 256) {
      const int s = i / 256, n = (i / 16) % 16, k = i % 16;
      const int t = 4 * s + k / 4, c = k % 4;
      w1s[i] = t < KK ? f2bf(th[OW1 + (n * C0 + c) * KK + t]) : (unsigned short)0;
    } else if (i < K1STEPS * 256 + KK * 256) {
      const int j = i - K1STEPS * 256;
      const int t = j / 256, n = (j / 16) % 16, c = j % 16;
      w2s[j] = f2bf(th[OW2 + (n * C1 + c) * KK + t]);
    } else if (i < K1STEPS * 256 + KK * 256 + C2 * KK) {
      const int j = i - K1STEPS * 256 - KK * 256;
      w3s[j] = bf2f(f2bf(th[OW3 + j]));
    } else {
      const int j = i - K1STEPS * 256 - KK * 256 - C2 * KK;
      if (j < 16) b1[j] = th[OB1 + j];
      else if (j < 32) b2[j - 16] = th[OB2 + j - 16];
      else b3[0] = th[OB3];
    }
  }
}

// ---------------------------------------------------------------------------
// fused inference: X (4, n0+12, n1+12) -> logits / flags (n0, n1)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_infer(const float* __restrict__ X, int n0, int n1,
                                              const unsigned char* __restrict__ pw, float thr,
                                              signed char* __restrict__ flags, float* __restrict__ logits,
                                              int* __restrict__ count, int* __restrict__ list, int tiles1) {
  __shared__ __attribute__((aligned(16))) unsigned short xin[L0 * L0 * C0];
  __shared__ __attribute__((aligned(16))) unsigned short a1[L1 * L1 * C1];
  __shared__ __attribute__((aligned(16))) unsigned short a2[L2 * L2 * C2];
  __shared__ __attribute__((aligned(16))) unsigned short w1s[K1STEPS * 256];
  __shared__ __attribute__((aligned(16))) unsigned short w2s[KK * 256];
  __shared__ float w3s[C2 * KK], b1[16], b2[16], b3;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int oy0 = (blockIdx.x / tiles1) * PT, ox0 = (blockIdx.x % tiles1) * PT;
  const int HX = n0 + 2 * HALO, WX = n1 + 2 * HALO;
  // weights
  for (int i = tid; i < K1STEPS * 256; i += 256) w1s[i] = ((const unsigned short*)(pw + PW1))[i];
  for (int i = tid; i < KK * 256; i += 256) w2s[i] = ((const unsigned short*)(pw + PW2))[i];
  for (int i = tid; i < C2 * KK; i += 256) w3s[i] = ((const float*)(pw + PW3))[i];
  if (tid < 16) {
    b1[tid] = ((const float*)(pw + PB1))[tid];
    b2[tid] = ((const float*)(pw + PB2))[tid];
  }
  if (tid == 0) b3 = ((const float*)(pw + PB3))[0];
  // input tile, channel-last bf16
  for (int i = tid; i < L0 * L0 * C0; i += 256) {
    const int c = i / (L0 * L0), p = i % (L0 * L0), y = p / L0, x = p % L0;
    const int gy = oy0 + y, gx = ox0 + x;
    const float v = (gy < HX && gx < WX) ? X[((long long)c * HX + gy) * WX + gx] : 0.0f;
    xin[p * C0 + c] = f2bf(v);
  }
  __syncthreads();
  const int mrow = lane & 15, kb = lane >> 4;
  // layer 1: 24x24 pixels = 36 groups of 16, K = 25 taps x 4 channels in 7 steps of 16
  for (int grp = wave; grp < (L1 * L1) / 16; grp += 4) {
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    const int p = grp * 16 + mrow, py = p / L1, px = p % L1;
#pragma unroll
    for (int s = 0; s < K1STEPS; ++s) {
      const int t = 4 * s + kb;
      bf16x4 a = {0, 0, 0, 0};
      if (t < KK) {
        const int dy = t / KS, dx = t % KS;
        a = *(const bf16x4*)&xin[((py + dy) * L0 + (px + dx)) * C0];
      }
      const bf16x4 b = *(const bf16x4*)&w1s[(s * 16 + mrow) * 16 + kb * 4];
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    // acc[j] = D[row 4*kb + j][col mrow]: pixel grp*16 + 4*kb + j, channel mrow
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int q = grp * 16 + 4 * kb + j;
      const float v = acc[j] + b1[mrow];
      a1[q * C1 + mrow] = f2bf(v > 0.f ? v : 0.f);
    }
  }
  __syncthreads();
  // layer 2: 20x20 pixels = 25 groups of 16, K = 16 channels per tap, 25 taps
  for (int grp = wave; grp < (L2 * L2) / 16; grp += 4) {
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    const int p = grp * 16 + mrow, py = p / L2, px = p % L2;
#pragma unroll 5
    for (int t = 0; t < KK; ++t) {
      const int dy = t / KS, dx = t % KS;
      const bf16x4 a = *(const bf16x4*)&a1[((py + dy) * L1 + (px + dx)) * C1 + kb * 4];
      const bf16x4 b = *(const bf16x4*)&w2s[(t * 16 + mrow) * 16 + kb * 4];
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int q = grp * 16 + 4 * kb + j;
      const float v = acc[j] + b2[mrow];
      a2[q * C2 + mrow] = f2bf(v > 0.f ? v : 0.f);
    }
  }
  __syncthreads();
  // layer 3: one output pixel per thread, FP32
  const int y = tid / PT, x = tid % PT;
  float z = b3;
  for (int t = 0; t < KK; ++t) {
    const int dy = t / KS, dx = t % KS;
    const unsigned short* ap = &a2[((y + dy) * L2 + (x + dx)) * C2];
#pragma unroll
    for (int c = 0; c < C2; ++c) z += w3s[c * KK + t] * bf2f(ap[c]);
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
    __shared__ int wbase[4];
    __shared__ int bbase;
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

// ---------------------------------------------------------------------------
// training kernels (FP32)
// ---------------------------------------------------------------------------
// out[b,o,y,x] = bias[o] + sum_{c,dy,dx} bf16(w[o,c,dy,dx]) in[b,c,y+dy,x+dx]; optional ReLU
__global__ void k_conv_fwd(const float* __restrict__ in, const float* __restrict__ w, const float* __restrict__ bias,
                           float* __restrict__ out, int B, int Cin, int Cout, int H, int W, int relu) {
  const int Ho = H - 4, Wo = W - 4;
  const long long total = (long long)B * Cout * Ho * Wo;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int x = i % Wo, y = (i / Wo) % Ho, o = (i / ((long long)Wo * Ho)) % Cout;
    const int b = (int)(i / ((long long)Wo * Ho * Cout));
    float acc = bias[o];
    for (int c = 0; c < Cin; ++c) {
      const float* ip = in + (((long long)b * Cin + c) * H + y) * W + x;
      const float* wp = w + (o * Cin + c) * KK;
#pragma unroll
      for (int dy = 0; dy < KS; ++dy)
#pragma unroll
        for (int dx = 0; dx < KS; ++dx) acc += bfround(wp[dy * KS + dx]) * ip[dy * W + dx];
    }
    out[i] = (relu && acc < 0.f) ? 0.f : acc;
  }
}

// din[b,c,y,x] = mask(a_in) * sum_{o,dy,dx} dout[b,o,y-dy,x-dx] bf16(w[o,c,dy,dx])
__global__ void k_conv_bwd_in(const float* __restrict__ dout, const float* __restrict__ w,
                              const float* __restrict__ act_in, float* __restrict__ din, int B, int Cin, int Cout,
                              int H, int W) {
  const int Ho = H - 4, Wo = W - 4;
  const long long total = (long long)B * Cin * H * W;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int x = i % W, y = (i / W) % H, c = (i / ((long long)W * H)) % Cin;
    const int b = (int)(i / ((long long)W * H * Cin));
    float acc = 0.f;
    if (act_in[i] > 0.f) {
      for (int o = 0; o < Cout; ++o) {
        const float* dp = dout + ((long long)b * Cout + o) * Ho * Wo;
        const float* wp = w + (o * Cin + c) * KK;
        for (int dy = 0; dy < KS; ++dy) {
          const int yo = y - dy;
          if (yo < 0 || yo >= Ho) continue;
          for (int dx = 0; dx < KS; ++dx) {
            const int xo = x - dx;
            if (xo < 0 || xo >= Wo) continue;
            acc += dp[yo * Wo + xo] * bfround(wp[dy * KS + dx]);
          }
        }
      }
    }
    din[i] = acc;
  }
}

// dw[o,c,:,:] and db[o]: one block per (o, c); 25 tap sums reduced over the block
__global__ void __launch_bounds__(256) k_conv_bwd_w(const float* __restrict__ in, const float* __restrict__ dout,
                                                   float* __restrict__ dw, float* __restrict__ db, int B, int Cin,
                                                   int Cout, int H, int W) {
  const int o = blockIdx.x / Cin, c = blockIdx.x % Cin;
  const int Ho = H - 4, Wo = W - 4;
  float acc[KK];
#pragma unroll
  for (int t = 0; t < KK; ++t) acc[t] = 0.f;
  float accb = 0.f;
  const long long npix = (long long)B * Ho * Wo;
  for (long long q = threadIdx.x; q < npix; q += blockDim.x) {
    const int x = q % Wo, y = (q / Wo) % Ho, b = (int)(q / ((long long)Wo * Ho));
    const float g = dout[((long long)b * Cout + o) * Ho * Wo + (long long)y * Wo + x];
    if (c == 0) accb += g;
    const float* ip = in + (((long long)b * Cin + c) * H + y) * W + x;
#pragma unroll
    for (int dy = 0; dy < KS; ++dy)
#pragma unroll
      for (int dx = 0; dx < KS; ++dx) acc[dy * KS + dx] += g * ip[dy * W + dx];
  }
  __shared__ float red[256];
#pragma unroll
  for (int t = 0; t <= KK; ++t) {
    red[threadIdx.x] = t < KK ? acc[t] : accb;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (t < KK) dw[(o * Cin + c) * KK + t] = red[0];
      else if (c == 0) db[o] = red[0];
    }
    __syncthreads();
  }
}

// dz = (pw*y*(sigmoid(z)-1) + (1-y)*sigmoid(z)) / N; loss partial sums per block into lsum
__global__ void k_loss(const float* __restrict__ z, const signed char* __restrict__ y, float pos_weight, long long n,
                       float* __restrict__ dz, float* __restrict__ lsum) {
  float part = 0.f;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float zz = z[i], yy = (float)y[i];
    const float s = 1.f / (1.f + expf(-zz));
    const float sp = zz > 0.f ? log1pf(expf(-zz)) : -zz + log1pf(expf(zz));  // softplus(-z)
    part += pos_weight * yy * sp + (1.f - yy) * (zz + sp);
    dz[i] = (pos_weight * yy * (s - 1.f) + (1.f - yy) * s) / (float)n;
  }
  for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(lsum, part / (float)n);
}

__global__ void k_adam(float* __restrict__ th, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  th[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

inline int grid_for(long long n) {
  long long b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

}  // namespace

namespace fvhoe {

int predictor_infer(const float* X, int n0, int n1, const void* packed, float thr, signed char* flags,
                    float* logits, int* count, int* list, hipStream_t s) {
  const int t0 = (n0 + PT - 1) / PT, t1 = (n1 + PT - 1) / PT;
  k_infer<<<t0 * t1, 256, 0, s>>>(X, n0, n1, (const unsigned char*)packed, thr, flags, logits, count, list, t1);
  return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
}

}  // namespace fvhoe

extern "C" {

int fvhoe_pred_pack(const float* theta, void* packed, void* stream) {
  if (!theta || !packed) return FVHOE_ERR_ARG;
  k_pack<<<40, 256, 0, (hipStream_t)stream>>>(theta, (unsigned char*)packed);
  return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
}

int fvhoe_pred_infer(const float* X, int n0, int n1, const void* packed, float logit_threshold,
                     signed char* flags, float* logits, void* stream) {
  if (!X || !packed || n0 < 1 || n1 < 1 || (!flags && !logits)) return FVHOE_ERR_ARG;
  return fvhoe::predictor_infer(X, n0, n1, packed, logit_threshold, flags, logits, nullptr, nullptr,
                                 (hipStream_t)stream);
}

int fvhoe_pred_conv_forward(const float* in, const float* w, const float* bias, float* out, int B, int Cin, int Cout,
                            int H, int W, int relu, void* stream) {
  if (!in || !w || !bias || !out || B < 1 || H < 5 || W < 5) return FVHOE_ERR_ARG;
  k_conv_fwd<<<grid_for((long long)B * Cout * (H - 4) * (W - 4)), 256, 0, (hipStream_t)stream>>>(
      in, w, bias, out, B, Cin, Cout, H, W, relu);
  return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
}

int fvhoe_pred_conv_backward(const float* in, const float* dout, const float* w, float* din, float* dw, float* db,
                             int B, int Cin, int Cout, int H, int W, void* stream) {
  if (!in || !dout || !w || !dw || !db || B < 1 || H < 5 || W < 5) return FVHOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  k_conv_bwd_w<<<Cout * Cin, 256, 0, s>>>(in, dout, dw, db, B, Cin, Cout, H, W);
  if (din) k_conv_bwd_in<<<grid_for((long long)B * Cin * H * W), 256, 0, s>>>(dout, w, in, din, B, Cin, Cout, H, W);
  return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
}

int fvhoe_pred_loss(const float* logits, const signed char* labels, float pos_weight, long long n, float* dlogits,
                    float* loss, void* stream) {
  if (!logits || !labels || !dlogits || !loss || n < 1) return FVHOE_ERR_ARG;
  k_loss<<<grid_for(n) < 1024 ? grid_for(n) : 1024, 256, 0, (hipStream_t)stream>>>(logits, labels, pos_weight, n,
                                                                                    dlogits, loss);
  return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
}

int fvhoe_pred_adam(float* theta, const float* grad, float* m, float* v, int n, float lr, float beta1, float beta2,
                    float eps, int step, void* stream) {
  if (!theta || !grad || !m || !v || n < 1 || step < 1) return FVHOE_ERR_ARG;
  float bc1 = 1.f, bc2 = 1.f;
  for (int i = 0; i < step; ++i) {
    bc1 *= beta1;
    bc2 *= beta2;
  }
  k_adam<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(theta, grad, m, v, n, lr, beta1, beta2, eps, 1.f - bc1,
                                                           1.f - bc2);
  return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
}

}  // extern "C"
