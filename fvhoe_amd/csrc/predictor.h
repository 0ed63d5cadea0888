// Internal interface between the fused stage and the predictor kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace fvhoe {

// Network logits on the features X (4, n0+12, n1+12) float, BF16 MFMA.
// flags (n0*n1 int8, may be null) = logit > thr; logits (may be null);
// with count and list, the flagged cells are appended to list (count += n).
int predictor_infer(const float* X, int n0, int n1, const void* packed, float thr, signed char* flags, float* logits,
                    int* count, int* list, hipStream_t s);

}  // namespace fvhoe
