// Backward of the PARAFAC2 reconstruction loss
//   L = sum_c sum_i || X_i - P_i B_c diag(A_ci) C_c^T ||^2
// with respect to the per-bin weights A and the chromosome factor B.  With
// F_i = X_i C - Q_i C^T C (Q_i = P_i B diag(A_i)) and H_i = P_i^T F_i (s x s per bin, one batched
// GEMM before this kernel):
//   dL/dA[i, r] = -2 sum_p B[p, r] H_i[p, r]
//   dL/dB[p, r] = -2 sum_i H_i[p, r] A[i, r]
// The A part reduces over p for every (bin, column); the B part reduces over the bins of one
// chunk of FH_LOSS_CHUNK bins into a partial buffer (summed afterwards in fixed order).
// Accumulation is fp64; every reduction has a fixed order.
#include "fh_common.h"

namespace {
constexpr int NT = 256;

__global__ __launch_bounds__(NT) void loss_bwd_kernel(const FhLossChrom* __restrict__ tab, int n_tab,
                                                      float scale) {
  const int64_t blk = blockIdx.x;
  int lo = 0, hi = n_tab - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tab[mid].blk_begin <= blk) lo = mid; else hi = mid - 1;
  }
  const FhLossChrom q = tab[lo];
  const int64_t local = blk - q.blk_begin;
  const int64_t n = q.n, s = q.s, sp = q.sp;
  if (local < q.blk_a) {
    const int64_t e = local * NT + threadIdx.x;
    if (e >= n * sp) return;
    const int64_t i = e / sp, r = e - i * sp;
    double v = 0.0;
    if (r < s) {
      const float* h = q.H + i * s * sp + r;
      const float* bb = q.B + r;
      for (int64_t p = 0; p < s; ++p) v = fma((double)bb[p * sp], (double)h[p * sp], v);
    }
    q.gA[e] = (float)(scale * v);
    return;
  }
  const int64_t l2 = local - q.blk_a;
  const int64_t chunk = l2 / q.blk_b;
  if (chunk >= q.n_chunks) return;
  const int64_t e = (l2 - chunk * q.blk_b) * NT + threadIdx.x;
  if (e >= s * sp) return;
  const int64_t p = e / sp, r = e - p * sp;
  double v = 0.0;
  if (r < s) {
    const int64_t i0 = chunk * FH_LOSS_CHUNK;
    const int64_t i1 = i0 + FH_LOSS_CHUNK < n ? i0 + FH_LOSS_CHUNK : n;
    for (int64_t i = i0; i < i1; ++i)
      v = fma((double)q.H[(i * s + p) * sp + r], (double)q.A[i * sp + r], v);
  }
  q.gBpart[chunk * q.chunk_stride + e] = (float)(scale * v);
}

}  // namespace

extern "C" int64_t fh_loss_prepare(FhLossChrom* tab_host, int n) {
  int64_t total = 0;
  for (int c = 0; c < n; ++c) {
    FhLossChrom& q = tab_host[c];
    q.blk_a = (int32_t)fh_cdiv(q.n * q.sp, NT);
    q.n_chunks = (int32_t)fh_cdiv(q.n, FH_LOSS_CHUNK);
    q.blk_b = (int32_t)fh_cdiv(q.s * q.sp, NT);
    q.blk_begin = total;
    total += q.blk_a + (int64_t)q.n_chunks * q.blk_b;
  }
  return total;
}

extern "C" int fh_loss_bwd(const FhLossChrom* tab_dev, int n_tab, int64_t total_blocks, float scale,
                           hipStream_t stream) {
  if (total_blocks <= 0) return FH_OK;
  FH_REQUIRE(tab_dev != nullptr && n_tab > 0, "loss_bwd: empty table");
  FH_REQUIRE(total_blocks < (int64_t)1 << 31, "loss_bwd: too many blocks");
  hipLaunchKernelGGL(loss_bwd_kernel, dim3((unsigned)total_blocks), dim3(NT), 0, stream, tab_dev, n_tab, scale);
  FH_CHECK_LAUNCH();
  return FH_OK;
}
