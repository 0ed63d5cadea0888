// Learned troubled-cell predictor (no counterpart in FVHOE): a three-layer
// 5x5 convolutional network 4 -> 16 -> 16 -> 1 channels on normalized
// primitive fv averages (rho, vx, vy, P), predicting the a-posteriori flags.
//
// Inference: the fused kernel of predictor_infer.h (layers 1 and 2 on BF16
// MFMA with FP32 accumulation, layer 3 in FP32, activations in LDS as BF16)
// on planar float features.
// Training: FP32 SIMT forward / backward / Adam kernels with the weights
// rounded to BF16 in the forward pass (the gradient is applied to the FP32
// master weights).
#include "fvhoe_common.h"
#include "fvhoe_hip.h"
#include "predictor.h"
#include "predictor_infer.h"

using namespace fvhoe;
using namespace fvhoe::pred;

namespace {

constexpr int OW1 = 0, OB1 = OW1 + C1 * C0 * KK, OW2 = OB1 + C1, OB2 = OW2 + C2 * C1 * KK, OW3 = OB2 + C2,
              OB3 = OW3 + C2 * KK, NPARAM = OB3 + 1;
static_assert(NPARAM == FVHOE_PRED_NPARAMS, "parameter count");
static_assert(PACKED <= FVHOE_PRED_PACKED_BYTES, "packed size");
static_assert(QPACKED == FVHOE_PRED_PLANAR_BYTES, "planar packed size");

__device__ __forceinline__ float bfround(float f) { return bf2f(f2bf(f)); }

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
  constexpr int N1 = KS * 16 * 32, N2 = KS * 3 * 16 * 32, N3 = C2 * KK;
  for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < N1 + N2 + N3 + 33; i += blockDim.x * gridDim.x) {
    if (i < N1) {  // [dy][n][k]: tap (dy, k/4) for k < 20, channel k%4
      const int dy = i / 512, n = (i / 32) % 16, k = i % 32;
      const int dx = k / 4, c = k % 4;
      w1s[i] = dx < KS ? f2bf(th[OW1 + (n * C0 + c) * KK + dy * KS + dx]) : (unsigned short)0;
    } else if (i < N1 + N2) {  // [dy][pair][n][k]: tap (dy, 2 pair + k/16), channel k%16
      const int j = i - N1;
      const int dy = j / 1536, pair = (j / 512) % 3, n = (j / 32) % 16, k = j % 32;
      const int dx = 2 * pair + k / 16, c = k % 16;
      w2s[j] = dx < KS ? f2bf(th[OW2 + (n * C1 + c) * KK + dy * KS + dx]) : (unsigned short)0;
    } else if (i < N1 + N2 + N3) {
      const int j = i - N1 - N2;
      w3s[j] = bf2f(f2bf(th[OW3 + j]));
    } else {
      const int j = i - N1 - N2 - N3;
      if (j < 16) b1[j] = th[OB1 + j];
      else if (j < 32) b2[j - 16] = th[OB2 + j - 16];
      else b3[0] = th[OB3];
    }
  }
}

// The planar network's weights: each kernel summed over its y taps (dx, the second image axis) in float32
// in tap order from the bfloat16-rounded weights, rounded to bfloat16 again for the MFMA layers 1 and 2,
// kept in float32 for layer 3.
__global__ void k_pack_planar(const float* __restrict__ th, unsigned char* __restrict__ out) {
  unsigned short* w1s = (unsigned short*)(out + QW1);
  unsigned short* w2s = (unsigned short*)(out + QW2);
  float* w3s = (float*)(out + QW3);
  float* b1 = (float*)(out + QB1);
  float* b2 = (float*)(out + QB2);
  float* b3 = (float*)(out + QB3);
  auto ysum = [&](int base) {  // base: the weight's (n, c, t, dx = 0) offset
    float acc = 0.0f;
    for (int dx = 0; dx < KS; ++dx) acc = acc + bf2f(f2bf(th[base + dx]));
    return acc;
  };
  constexpr int N1 = 16 * 32, N2 = 3 * 16 * 32, N3 = C2 * KS;
  for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < N1 + N2 + N3 + 33; i += blockDim.x * gridDim.x) {
    if (i < N1) {  // [n][k]: tap k/4, channel k%4
      const int n = i / 32, k = i % 32, t = k / 4, c = k % 4;
      w1s[i] = t < KS ? f2bf(ysum(OW1 + (n * C0 + c) * KK + t * KS)) : (unsigned short)0;
    } else if (i < N1 + N2) {  // [pair][n][k]: tap 2 pair + k/16, channel k%16
      const int j = i - N1, pair = j / 512, n = (j / 32) % 16, k = j % 32;
      const int t = 2 * pair + k / 16, c = k % 16;
      w2s[j] = t < KS ? f2bf(ysum(OW2 + (n * C1 + c) * KK + t * KS)) : (unsigned short)0;
    } else if (i < N1 + N2 + N3) {  // [c][t]
      const int j = i - N1 - N2, c = j / KS, t = j % KS;
      w3s[j] = ysum(OW3 + c * KK + t * KS);
    } else {
      const int j = i - N1 - N2 - N3;
      if (j < 16) b1[j] = th[OB1 + j];
      else if (j < 32) b2[j - 16] = th[OB2 + j - 16];
      else b3[0] = th[OB3];
    }
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
  return launch_infer(FeatPlanar{X, n0 + 2 * HALO, n1 + 2 * HALO}, n0, n1, packed, thr, flags, logits, count, list, s);
}

}  // namespace fvhoe

extern "C" {

int fvhoe_pred_pack(const float* theta, void* packed, void* stream) {
  if (!theta || !packed) return FVHOE_ERR_ARG;
  k_pack<<<40, 256, 0, (hipStream_t)stream>>>(theta, (unsigned char*)packed);
  return hipGetLastError() == hipSuccess ? FVHOE_OK : FVHOE_ERR_LAUNCH;
}

int fvhoe_pred_pack_planar(const float* theta, void* packed, void* stream) {
  if (!theta || !packed) return FVHOE_ERR_ARG;
  k_pack_planar<<<8, 256, 0, (hipStream_t)stream>>>(theta, (unsigned char*)packed);
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
