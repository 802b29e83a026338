// Partial random walk with restart on band tiles (partial_rwr.py:45-175).
//
// Per cell, with A the (rows x w) tile and [ls, ls + rows) its diagonal block columns:
//   S2 = A A^T (GEMM), diag(S2) = 0
//   local = 0.75 colnorm(A[:, ls:ls+rows]) + 0.25 colnorm(S2);  zero columns get a 1 on the diagonal
//   P = colnorm(local);  Q <- 0.5 Q P + 0.5 I  (GEMM with diag epilogue), n epochs
//   [do_col]  Q <- rownorm(max((Q + Q^T)/2, 0)),  A <- A / bin_cov
//   out = Q A (GEMM)
// The GEMMs run through the grouped MFMA kernel; this file holds the per-cell
// normalisations, the auto-stop delta, do_col, and the init-time pooling / coverage sums.
#include "fh_common.h"

namespace {
constexpr int NT = 256;

// one block per cell; thread j owns column j (rows <= 256)
__global__ __launch_bounds__(NT) void rwr_prep_kernel(const float* __restrict__ S2, int64_t ldS, int64_t strS,
                                                      const float* __restrict__ X, int64_t ldX, int64_t strX,
                                                      int64_t ls, int64_t rows, float* __restrict__ P,
                                                      int64_t ldP, int64_t strP) {
  const int64_t c = blockIdx.x;
  const int j = threadIdx.x;
  if (j >= rows) return;
  const float* s2 = S2 + c * strS;
  const float* x = X + c * strX;
  float* p = P + c * strP;
  float cs2 = 0.f, cs1 = 0.f;
  for (int64_t i = 0; i < rows; ++i) {
    if (i != j) cs2 += s2[i * ldS + j];
    cs1 += x[i * ldX + ls + j];
  }
  const float d2 = cs2 + 1e-15f, d1 = cs1 + 1e-15f;
  float lsum = 0.f;
  for (int64_t i = 0; i < rows; ++i) {
    float v2 = (i == j) ? 0.f : s2[i * ldS + j];
    float v = (x[i * ldX + ls + j] / d1) * 0.75f + (v2 / d2) * 0.25f;
    p[i * ldP + j] = v;
    lsum += v;
  }
  float fill = (lsum == 0.f) ? 1.f : 0.f;
  if (fill != 0.f) {
    p[(int64_t)j * ldP + j] += fill;
    lsum = 0.f;
    for (int64_t i = 0; i < rows; ++i) lsum += p[i * ldP + j];
  }
  const float dn = lsum + 1e-15f;
  for (int64_t i = 0; i < rows; ++i) p[i * ldP + j] = p[i * ldP + j] / dn;
}

// delta[c] = ||Q_old - Q_new||_F  over rows x rows
__global__ __launch_bounds__(NT) void rwr_delta_kernel(const float* __restrict__ Qa, const float* __restrict__ Qb,
                                                       int64_t ld, int64_t str, int64_t rows,
                                                       float* __restrict__ delta) {
  __shared__ float red[NT / 64];
  const int64_t c = blockIdx.x;
  float p = 0.f;
  for (int64_t e = threadIdx.x; e < rows * rows; e += NT) {
    int64_t i = e / rows, j = e % rows;
    float d = Qa[c * str + i * ld + j] - Qb[c * str + i * ld + j];
    p += d * d;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = p;
  __syncthreads();
  if (threadIdx.x == 0) delta[c] = sqrtf(red[0] + red[1] + red[2] + red[3]);
}

__global__ void max_kernel(const float* __restrict__ v, int64_t n, float* __restrict__ out) {
  __shared__ float red[NT];
  float m = -1.f;
  for (int64_t i = threadIdx.x; i < n; i += NT) m = fmaxf(m, v[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int off = NT / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// do_col, part 1: Qo = rownorm(max((Q + Q^T)/2, 0)); one block per cell, thread i owns row i
__global__ __launch_bounds__(NT) void rwr_symnorm_kernel(const float* __restrict__ Q, float* __restrict__ Qo,
                                                         int64_t ld, int64_t str, int64_t rows) {
  const int64_t c = blockIdx.x;
  const int i = threadIdx.x;
  if (i >= rows) return;
  const float* q = Q + c * str;
  float* o = Qo + c * str;
  float sum = 0.f;
  for (int64_t j = 0; j < rows; ++j) {
    float v = fmaxf((q[i * ld + j] + q[j * ld + i]) * 0.5f, 0.f);
    o[i * ld + j] = v;
    sum += v;
  }
  const float dn = sum + 1e-15f;
  for (int64_t j = 0; j < rows; ++j) o[i * ld + j] = o[i * ld + j] / dn;
}

// do_col, part 2: X[c][i][k] /= cov[c * ldc + k]  (k < w)
__global__ void div_cols_kernel(float* __restrict__ X, int64_t ldX, int64_t strX, int64_t rows, int64_t w,
                                const float* __restrict__ cov, int64_t ldc, int64_t ncell) {
  const int64_t tot = ncell * rows * w;
  for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < tot; e += (int64_t)gridDim.x * NT) {
    int64_t c = e / (rows * w), q = e % (rows * w), i = q / w, k = q % w;
    X[c * strX + i * ldX + k] = X[c * strX + i * ldX + k] / cov[c * ldc + k];
  }
}

// avg_pool2d(kernel = stride = ll, floor mode) of each cell's (rows x w) tile written as a flat
// feature row: F[c * ldf + p * (w / ll) + q]
__global__ void pool_kernel(const float* __restrict__ X, int64_t ldX, int64_t strX, int64_t rows, int64_t w,
                            int64_t ll, int64_t ncell, float* __restrict__ F, int64_t ldf) {
  const int64_t pr = rows / ll, pc = w / ll;
  const int64_t tot = ncell * pr * pc;
  const float inv = 1.f / (float)(ll * ll);
  for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < tot; e += (int64_t)gridDim.x * NT) {
    int64_t c = e / (pr * pc), q = e % (pr * pc), p = q / pc, qq = q % pc;
    float s = 0.f;
    for (int64_t a = 0; a < ll; ++a)
      for (int64_t b = 0; b < ll; ++b) s += X[c * strX + (p * ll + a) * ldX + qq * ll + b];
    F[c * ldf + q] = s * inv;
  }
}

// cov[c * ldc + k] += sum_i X[c][i][k]   (k < w)
__global__ void colsum_add_kernel(const float* __restrict__ X, int64_t ldX, int64_t strX, int64_t rows, int64_t w,
                                  int64_t ncell, float* __restrict__ cov, int64_t ldc) {
  const int64_t tot = ncell * w;
  for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < tot; e += (int64_t)gridDim.x * NT) {
    int64_t c = e / w, k = e % w;
    float s = 0.f;
    for (int64_t i = 0; i < rows; ++i) s += X[c * strX + i * ldX + k];
    cov[c * ldc + k] += s;
  }
}

// F.avg_pool2d(x, 3, stride 1, padding 1, ceil_mode=True) with count_include_pad, then
// clamp(min=1e-8) (partial_rwr.py:77-81): out[c][i][k] = max(sum_{3x3} x / 9, 1e-8)
__global__ void avgpool3_kernel(const float* __restrict__ X, int64_t ldX, int64_t strX, int64_t rows, int64_t w,
                                int64_t ncell, float* __restrict__ out, int64_t ldO, int64_t strO) {
  const int64_t tot = ncell * rows * w;
  for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < tot; e += (int64_t)gridDim.x * NT) {
    int64_t c = e / (rows * w), q = e % (rows * w), i = q / w, k = q % w;
    float s = 0.f;
    for (int di = -1; di <= 1; ++di)
      for (int dk = -1; dk <= 1; ++dk) {
        int64_t ii = i + di, kk = k + dk;
        if (ii >= 0 && ii < rows && kk >= 0 && kk < w) s += X[c * strX + ii * ldX + kk];
      }
    out[c * strO + i * ldO + k] = fmaxf(s / 9.f, 1e-8f);
  }
}

inline unsigned gsz(int64_t n) {
  int64_t g = (n + NT - 1) / NT;
  return (unsigned)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}
}  // namespace

extern "C" {

int fh_rwr_prep(const float* S2, int64_t ldS, int64_t strS, const float* X, int64_t ldX, int64_t strX, int64_t ls,
                int64_t rows, float* P, int64_t ldP, int64_t strP, int64_t ncell, hipStream_t st) {
  FH_REQUIRE(rows >= 1 && rows <= NT, "rwr_prep: rows=%lld out of range (max %d)", (long long)rows, NT);
  if (ncell <= 0) return FH_OK;
  hipLaunchKernelGGL(rwr_prep_kernel, dim3((unsigned)ncell), dim3(NT), 0, st, S2, ldS, strS, X, ldX, strX, ls, rows,
                     P, ldP, strP);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

// delta per cell and their max (out_max: one float)
int fh_rwr_delta(const float* Qa, const float* Qb, int64_t ld, int64_t str, int64_t rows, int64_t ncell,
                 float* delta, float* out_max, hipStream_t st) {
  if (ncell <= 0) return FH_OK;
  hipLaunchKernelGGL(rwr_delta_kernel, dim3((unsigned)ncell), dim3(NT), 0, st, Qa, Qb, ld, str, rows, delta);
  FH_CHECK_LAUNCH();
  hipLaunchKernelGGL(max_kernel, dim3(1), dim3(NT), 0, st, delta, ncell, out_max);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

int fh_rwr_symnorm(const float* Q, float* Qo, int64_t ld, int64_t str, int64_t rows, int64_t ncell, hipStream_t st) {
  FH_REQUIRE(rows >= 1 && rows <= NT, "rwr_symnorm: rows out of range");
  if (ncell <= 0) return FH_OK;
  hipLaunchKernelGGL(rwr_symnorm_kernel, dim3((unsigned)ncell), dim3(NT), 0, st, Q, Qo, ld, str, rows);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

int fh_div_cols(float* X, int64_t ldX, int64_t strX, int64_t rows, int64_t w, const float* cov, int64_t ldc,
                int64_t ncell, hipStream_t st) {
  hipLaunchKernelGGL(div_cols_kernel, dim3(gsz(ncell * rows * w)), dim3(NT), 0, st, X, ldX, strX, rows, w, cov, ldc,
                     ncell);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

int fh_pool(const float* X, int64_t ldX, int64_t strX, int64_t rows, int64_t w, int64_t ll, int64_t ncell, float* F,
            int64_t ldf, hipStream_t st) {
  FH_REQUIRE(ll >= 1, "pool: ll");
  hipLaunchKernelGGL(pool_kernel, dim3(gsz(ncell * (rows / ll) * (w / ll))), dim3(NT), 0, st, X, ldX, strX, rows, w,
                     ll, ncell, F, ldf);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

int fh_avgpool3(const float* X, int64_t ldX, int64_t strX, int64_t rows, int64_t w, int64_t ncell, float* out,
                int64_t ldO, int64_t strO, hipStream_t st) {
  if (ncell * rows * w <= 0) return FH_OK;
  hipLaunchKernelGGL(avgpool3_kernel, dim3(gsz(ncell * rows * w)), dim3(NT), 0, st, X, ldX, strX, rows, w, ncell, out,
                     ldO, strO);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

int fh_colsum_add(const float* X, int64_t ldX, int64_t strX, int64_t rows, int64_t w, int64_t ncell, float* cov,
                  int64_t ldc, hipStream_t st) {
  hipLaunchKernelGGL(colsum_add_kernel, dim3(gsz(ncell * w)), dim3(NT), 0, st, X, ldX, strX, rows, w, ncell, cov,
                     ldc);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

}  // extern "C"
