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

// Forward, row-stationary: a workgroup keeps a 64-row block of Q (k = s contiguous) in LDS and
// walks the cells in tiles of 32: per tile the s x 32 slice of C^T goes through LDS (the next
// slice and the next X tile are loaded into registers while the current one is multiplied),
// 4 waves each compute a 16 x 32 block of Q C^T with MFMA 16x16x4 f32, and subtract it from X.
// One fp64 partial sum of squares per workgroup.  Q is read once; C^T, which every row block
// reads, stays in L2.
constexpr int RB = 64, CB = 32;
template <int SMAX>
struct RowsCfg {
  static constexpr int QLD = SMAX + 4;   // Q rows: lane reads (row l % 16, k l / 16) hit 64 banks
  static constexpr int CLD = 48;         // C^T rows: 4 k-rows x 16 columns hit 64 banks
  static constexpr int CQ = SMAX / 32;   // float4 loads of a C^T slice per thread
};

template <int SMAX>
__global__ __launch_bounds__(NT) void loss_fwd_rows_kernel(const FhLossRows* __restrict__ tab, int n_tab,
                                                           double* __restrict__ part) {
  using Cfg = RowsCfg<SMAX>;
  __shared__ __attribute__((aligned(16))) float Qs[RB * Cfg::QLD];
  __shared__ __attribute__((aligned(16))) float Cs[SMAX * Cfg::CLD];
  __shared__ double red[NT / 64];
  const int64_t blk = blockIdx.x;
  int lo = 0, hi = n_tab - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tab[mid].blk_begin <= blk) lo = mid; else hi = mid - 1;
  }
  const FhLossRows d = tab[lo];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t m0 = (blk - d.blk_begin) * RB;
  const int kpad = (int)((d.s + 3) & ~3);
  // Q block -> LDS (rows past W are zero)
  const int q4 = kpad / 4;
  for (int e = tid; e < RB * q4; e += NT) {
    const int r = e / q4, q = e - r * q4;
    floatx4 v = {0.f, 0.f, 0.f, 0.f};
    if (m0 + r < d.W) v = *(const floatx4*)(d.Q + (m0 + r) * d.ldq + 4 * q);
    *(floatx4*)(Qs + r * Cfg::QLD + 4 * q) = v;
  }
  // C^T slice: thread t holds rows t / 8 + 32 p, columns 4 (t % 8) .. + 3
  const int crow = tid >> 3, cq = tid & 7;
  floatx4 cr[Cfg::CQ];
  auto load_c = [&](int64_t n0) {
#pragma unroll
    for (int p = 0; p < Cfg::CQ; ++p) {
      const int k = crow + 32 * p;
      cr[p] = k < kpad ? *(const floatx4*)(d.Ct + (int64_t)k * d.ldc + n0 + 4 * cq) : floatx4{0.f, 0.f, 0.f, 0.f};
    }
  };
  // X values of the lane's outputs: rows 16 wave + 4 (l / 16) + i, columns 16 j + l % 16
  float xr[2][4];
  auto load_x = [&](int64_t n0) {
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t row = m0 + 16 * wave + 4 * (lane >> 4) + i;
        xr[j][i] = row < d.W ? d.X[row * d.ldx + n0 + 16 * j + (lane & 15)] : 0.f;
      }
  };
  load_c(0);
  load_x(0);
  double ss = 0.0;
  for (int64_t n0 = 0; n0 < d.N; n0 += CB) {
#pragma unroll
    for (int p = 0; p < Cfg::CQ; ++p) {
      const int k = crow + 32 * p;
      if (k < kpad) *(floatx4*)(Cs + k * Cfg::CLD + 4 * cq) = cr[p];
    }
    __syncthreads();
    float xc[2][4];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) xc[j][i] = xr[j][i];
    if (n0 + CB < d.N) {
      load_c(n0 + CB);
      load_x(n0 + CB);
    }
    floatx4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
    const float* qa = Qs + (16 * wave + (lane & 15)) * Cfg::QLD + (lane >> 4);
    const float* cb = Cs + (lane >> 4) * Cfg::CLD + (lane & 15);
    for (int k0 = 0; k0 < kpad; k0 += 4) {
      const float a = qa[k0];
      const float b0 = cb[k0 * Cfg::CLD], b1 = cb[k0 * Cfg::CLD + 16];
      acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b0, acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b1, acc[1], 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t row = m0 + 16 * wave + 4 * (lane >> 4) + i;
        const int64_t col = n0 + 16 * j + (lane & 15);
        if (row < d.W && col < d.N) {
          const float e = xc[j][i] - acc[j][i];
          ss = fma((double)e, (double)e, ss);
        }
      }
    __syncthreads();
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off);
  if (lane == 0) red[wave] = ss;
  __syncthreads();
  if (tid == 0) part[blk] = (red[0] + red[1]) + (red[2] + red[3]);
}

}  // namespace

extern "C" int64_t fh_loss_rows_prepare(FhLossRows* tab_host, int n) {
  int64_t total = 0;
  for (int c = 0; c < n; ++c) {
    tab_host[c].blk_begin = total;
    total += fh_cdiv(tab_host[c].W, RB);
  }
  return total;
}

extern "C" int fh_loss_fwd_rows(const FhLossRows* tab_dev, int n_tab, int64_t total_blocks, int smax, double* part,
                                hipStream_t stream) {
  if (total_blocks <= 0) return FH_OK;
  FH_REQUIRE(tab_dev != nullptr && n_tab > 0 && part != nullptr, "loss_fwd_rows: empty table or partials");
  FH_REQUIRE(total_blocks < (int64_t)1 << 31, "loss_fwd_rows: too many blocks");
  if (smax == 128)
    hipLaunchKernelGGL(loss_fwd_rows_kernel<128>, dim3((unsigned)total_blocks), dim3(NT), 0, stream, tab_dev, n_tab,
                       part);
  else if (smax == 256)
    hipLaunchKernelGGL(loss_fwd_rows_kernel<256>, dim3((unsigned)total_blocks), dim3(NT), 0, stream, tab_dev, n_tab,
                       part);
  else {
    fh_set_error("loss_fwd_rows: smax %d (128 or 256)", smax);
    return FH_ERR_ARG;
  }
  FH_CHECK_LAUNCH();
  return FH_OK;
}

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
