// Grouped/batched GEMM descriptors shared by the host engine and the kernels.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// C[b] (M x N) = alpha * A[b] (M x K) * B[b] (K x N) + beta * C[b]  (+ diag on C's diagonal)
// Element (m,k) of A lives at A + b*sAb + m*sAm + k*sAk (likewise B, C).
// For the fp32/bf16/fp8 kernels exactly one of sAm/sAk must be 1 (and of sBk/sBn);
// the contiguous extent must be readable up to a multiple of 16 bytes.
typedef struct FhGemmDesc {
  const void* A;
  const void* B;
  void* C;
  int64_t M, N, K;
  int64_t sAm, sAk, sBk, sBn, sCm, sCn;
  int64_t sAb, sBb, sCb;
  int64_t sCk;            // FH_GEMM_SPLIT_PARTIAL: stride between the k-split partial outputs
  int32_t batch;
  int32_t flags;          // FH_GEMM_ATOMIC: C += alpha*AB with atomics (beta ignored)
  double alpha, beta, diag;
  const float* scale_a;   // fp8: per-batch dequant scale of A (nullable)
  const float* scale_b;   // fp8: per-batch dequant scale of B (nullable)
  const int* skip;        // nullable device flag: when *skip != 0 the launch is a no-op
  int64_t tile_begin;     // filled by fh_gemm_prepare
  int32_t tiles_m, tiles_n;
} FhGemmDesc;

enum { FH_GEMM_ATOMIC = 1, FH_GEMM_OUT_BF16 = 2, FH_GEMM_SPLIT_PARTIAL = 4 };
// bits 8.. of flags: k-split count (0 or 1 = no split; >1 requires FH_GEMM_ATOMIC or
// FH_GEMM_SPLIT_PARTIAL, which writes split ks to C + ks*sCk for a fixed-order reduction)

#ifdef __cplusplus
}
#endif
