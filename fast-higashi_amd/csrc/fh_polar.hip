// Batched FP64 polar factors and SPD solves.
//
// Polar factor of Y (m x s, m >= s):  polar(Y) = Y G^{-1/2},  G = Y^T Y (FP64).
// G^{-1/2} comes from the coupled Newton-Schulz iteration on Y_0 = G/c, Z_0 = I
// (c = ||G||_F >= lambda_max):
//     T_k = 1.5 I - 0.5 Z_k Y_k,   Y_{k+1} = Y_k T_k,   Z_{k+1} = T_k Z_k
// Y_k -> (G/c)^{1/2}, Z_k -> (G/c)^{-1/2}.  Every matrix of the batch iterates until
// ||T_k - I||_F^2 < tol, then is frozen (copied through) so one fixed launch
// sequence serves all of them.  Nuclear norm: trace(G^{1/2}) = sqrt(c) trace(Y).
//
// The iterations are s x s FP64 GEMMs on v_mfma_f64_16x16x4_f64, 64 x 64 tiles.
#include "fh_common.h"

namespace {

constexpr int NT = 256;


}  // namespace

extern "C" {
typedef struct FhNsGroup {
  double* Y0;
  double* Z0;
  double* Y1;
  double* Z1;
  double* T;
  double* scale;
  double* res;
  int64_t s, ld, stride, count;
  int64_t tile_begin;
  int32_t tiles, tm;
} FhNsGroup;
}

namespace {

__device__ __forceinline__ int find_group(const FhNsGroup* g, int n, int64_t t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (g[mid].tile_begin <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// 64x64 output tile of  alpha * L * R  (+ diag on the diagonal),  L, R row-major s x s.
// Writes out and, if res != null, accumulates sum((out - I)^2) over the tile.
__device__ __forceinline__ void tile_gemm_f64(const double* __restrict__ L, const double* __restrict__ R,
                                              double* __restrict__ out, int64_t s, int64_t ld,
                                              int64_t m0, int64_t n0, double alpha, double diag,
                                              double* res, double* smem) {
  constexpr int BKD = 16, LR = BKD + 1;
  double* As = smem;
  double* Bs = smem + 64 * LR;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  doublex4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = doublex4{0., 0., 0., 0.};
  for (int64_t k0 = 0; k0 < s; k0 += BKD) {
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      int e = tid + p * NT;
      int r = e >> 4, k = e & 15;
      // L[m][k] (row-major, k contiguous)
      As[r * LR + k] = (m0 + r < s && k0 + k < s) ? L[(m0 + r) * ld + k0 + k] : 0.;
      // R[k][n] (row-major, n contiguous).  The iterates are symmetric only up to
      // rounding; using R^T here instead makes the coupled iteration diverge.
      int n = e & 63, kk = e >> 6;
      Bs[n * LR + kk] = (n0 + n < s && k0 + kk < s) ? R[(k0 + kk) * ld + n0 + n] : 0.;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BKD; kk += 4) {
      double a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[(wm * 32 + i * 16 + (lane & 15)) * LR + kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[(wn * 32 + j * 16 + (lane & 15)) * LR + kk + (lane >> 4)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
  double part = 0.;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int64_t row = m0 + wm * 32 + i * 16 + (lane >> 4) + 4 * r;  // f64 MFMA: item r of lane l is row 4r + l/16
        int64_t col = n0 + wn * 32 + j * 16 + (lane & 15);
        if (row < s && col < s) {
          double v = alpha * acc[i][j][r] + (row == col ? diag : 0.);
          out[row * ld + col] = v;
          double dv = v - (row == col ? 1. : 0.);
          part += dv * dv;
        }
      }
  if (res) {
    // wave reduce then one atomic per wave
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
    if (lane == 0) atomicAdd(res, part);
  }
}

__global__ __launch_bounds__(NT) void ns_init_kernel(const FhNsGroup* __restrict__ groups, int n_groups,
                                                     const int64_t* __restrict__ mat_begin, int n_iter) {
  // one workgroup per matrix; blockIdx.x = global matrix index
  const int64_t gm = blockIdx.x;
  int lo = 0, hi = n_groups - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (mat_begin[mid] <= gm) lo = mid; else hi = mid - 1;
  }
  const FhNsGroup g = groups[lo];
  const int64_t b = gm - mat_begin[lo];
  double* Y = g.Y0 + b * g.stride;
  double* Z = g.Z0 + b * g.stride;
  __shared__ double red[NT / 64];
  double part = 0.;
  for (int64_t e = threadIdx.x; e < g.s * g.s; e += NT) {
    double v = Y[(e / g.s) * g.ld + e % g.s];
    part += v * v;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  double c = 0.;
  for (int w = 0; w < NT / 64; ++w) c += red[w];
  c = sqrt(c);
  if (!(c > 0.)) c = 1.;
  const double inv = 1. / c;
  for (int64_t e = threadIdx.x; e < g.s * g.s; e += NT) {
    int64_t r = e / g.s, col = e % g.s;
    Y[r * g.ld + col] *= inv;
    Z[r * g.ld + col] = (r == col) ? 1. : 0.;
  }
  if (threadIdx.x == 0) g.scale[b] = c;
  for (int it = threadIdx.x; it < n_iter; it += NT) g.res[(int64_t)it * g.count + b] = 0.;
}

// T = 1.5 I - 0.5 Z_k Y_k  for matrices still iterating.
__global__ __launch_bounds__(NT) void ns_t_kernel(const FhNsGroup* __restrict__ groups, int n_groups,
                                                  int iter, double tol) {
  __shared__ double smem[2 * 64 * 17];
  const int64_t t = blockIdx.x;
  const FhNsGroup g = groups[find_group(groups, n_groups, t)];
  int64_t local = t - g.tile_begin;
  const int64_t b = local / g.tiles;
  if (b >= g.count) return;
  local -= b * g.tiles;
  const int64_t tm = local / g.tm, tn = local % g.tm;
  double* res_now = g.res + (int64_t)iter * g.count + b;
  if (iter > 0) {
    double prev = g.res[(int64_t)(iter - 1) * g.count + b];
    if (prev < tol) {              // converged: carry the residual forward
      if (tm == 0 && tn == 0 && threadIdx.x == 0) *res_now = prev;
      return;
    }
  }
  const int cur = iter & 1;
  const double* Y = (cur ? g.Y1 : g.Y0) + b * g.stride;
  const double* Z = (cur ? g.Z1 : g.Z0) + b * g.stride;
  double* T = g.T + b * g.stride;
  tile_gemm_f64(Z, Y, T, g.s, g.ld, tm * 64, tn * 64, -0.5, 1.5, res_now, smem);
}

// Y_{k+1} = Y_k T, Z_{k+1} = T Z_k  (grid: 2 x tiles); converged matrices are copied.
__global__ __launch_bounds__(NT) void ns_yz_kernel(const FhNsGroup* __restrict__ groups, int n_groups,
                                                   int iter, double tol) {
  __shared__ double smem[2 * 64 * 17];
  const int64_t t = blockIdx.x >> 1;
  const int which = blockIdx.x & 1;
  const FhNsGroup g = groups[find_group(groups, n_groups, t)];
  int64_t local = t - g.tile_begin;
  const int64_t b = local / g.tiles;
  if (b >= g.count) return;
  local -= b * g.tiles;
  const int64_t tm = local / g.tm, tn = local % g.tm;
  const int cur = iter & 1;
  const double* Yc = (cur ? g.Y1 : g.Y0) + b * g.stride;
  const double* Zc = (cur ? g.Z1 : g.Z0) + b * g.stride;
  double* Yn = (cur ? g.Y0 : g.Y1) + b * g.stride;
  double* Zn = (cur ? g.Z0 : g.Z1) + b * g.stride;
  const double r = g.res[(int64_t)iter * g.count + b];
  const int64_t m0 = tm * 64, n0 = tn * 64;
  if (r < tol) {
    const double* src = which ? Zc : Yc;
    double* dst = which ? Zn : Yn;
    for (int e = threadIdx.x; e < 64 * 64; e += NT) {
      int64_t row = m0 + (e >> 6), col = n0 + (e & 63);
      if (row < g.s && col < g.s) dst[row * g.ld + col] = src[row * g.ld + col];
    }
    return;
  }
  const double* T = g.T + b * g.stride;
  if (which == 0) tile_gemm_f64(Yc, T, Yn, g.s, g.ld, m0, n0, 1.0, 0.0, nullptr, smem);
  else tile_gemm_f64(T, Zc, Zn, g.s, g.ld, m0, n0, 1.0, 0.0, nullptr, smem);
}

// After the last iteration (result in buffer `fin`): Zf = Z / sqrt(c) (= G^{-1/2}),
// nuc[b] = sqrt(c) * trace(Y)  (= trace(G^{1/2}) = sum of singular values).
__global__ __launch_bounds__(NT) void ns_final_kernel(const FhNsGroup* __restrict__ groups, int n_groups,
                                                      const int64_t* __restrict__ mat_begin, int fin,
                                                      double* __restrict__ zout_unused,
                                                      double* __restrict__ nuc) {
  const int64_t gm = blockIdx.x;
  int lo = 0, hi = n_groups - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (mat_begin[mid] <= gm) lo = mid; else hi = mid - 1;
  }
  const FhNsGroup g = groups[lo];
  const int64_t b = gm - mat_begin[lo];
  double* Y = (fin ? g.Y1 : g.Y0) + b * g.stride;
  double* Z = (fin ? g.Z1 : g.Z0) + b * g.stride;
  const double c = g.scale[b];
  const double rs = sqrt(c), irs = 1. / rs;
  __shared__ double red[NT / 64];
  double part = 0.;
  for (int64_t i = threadIdx.x; i < g.s; i += NT) part += Y[i * g.ld + i];
  for (int64_t e = threadIdx.x; e < g.s * g.s; e += NT) {
    int64_t r = e / g.s, col = e % g.s;
    Z[r * g.ld + col] *= irs;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tr = 0.;
    for (int w = 0; w < NT / 64; ++w) tr += red[w];
    nuc[gm] = rs * tr;
  }
}

// ---------------------------------------------------------------------------
// Batched SPD inverse  Ginv = (G + eps I)^{-1}  (G symmetric s x s FP64).
// One workgroup per system: Cholesky with the packed lower factor in LDS, then
// every thread solves L L^T x = e_j for its own columns j, in place in the output.
// ---------------------------------------------------------------------------
constexpr int CHOL_MAX = 256;
constexpr int CHOL_LDS = 160 * 1024 - 4096;   // bytes available for the packed factor

__device__ __forceinline__ int64_t pk(int64_t i, int64_t j) { return i * (i + 1) / 2 + j; }  // i >= j

}  // namespace
extern "C" {
typedef struct FhSpdSys {
  const double* G;
  double* out;
  int64_t s, ldg, ldo;
  double eps;
} FhSpdSys;
}
namespace {

__global__ __launch_bounds__(NT) void spd_inverse_kernel(const FhSpdSys* __restrict__ sys, int* __restrict__ status) {
  extern __shared__ double Lp[];
  const FhSpdSys q = sys[blockIdx.x];
  const int64_t s = q.s;
  const int tid = threadIdx.x;
  for (int64_t i = tid; i < s; i += NT)
    for (int64_t j = 0; j <= i; ++j) Lp[pk(i, j)] = q.G[i * q.ldg + j] + (i == j ? q.eps : 0.);
  __syncthreads();
  __shared__ double dkk;
  for (int64_t k = 0; k < s; ++k) {
    if (tid == 0) {
      double v = Lp[pk(k, k)];
      if (!(v > 0.)) { if (status) atomicOr(status, 1); v = 1e-300; }
      dkk = sqrt(v);
      Lp[pk(k, k)] = dkk;
    }
    __syncthreads();
    const double inv = 1. / dkk;
    for (int64_t i = k + 1 + tid; i < s; i += NT) Lp[pk(i, k)] *= inv;
    __syncthreads();
    for (int64_t i = k + 1 + tid; i < s; i += NT) {
      const double lik = Lp[pk(i, k)];
      for (int64_t j = k + 1; j <= i; ++j) Lp[pk(i, j)] -= lik * Lp[pk(j, k)];
    }
    __syncthreads();
  }
  for (int64_t j = tid; j < s; j += NT) {
    double* x = q.out + j;                 // column j, stride ldo
    for (int64_t i = 0; i < j; ++i) x[i * q.ldo] = 0.;
    for (int64_t i = j; i < s; ++i) {      // L y = e_j
      double v = (i == j) ? 1. : 0.;
      for (int64_t m = j; m < i; ++m) v -= Lp[pk(i, m)] * x[m * q.ldo];
      x[i * q.ldo] = v / Lp[pk(i, i)];
    }
    for (int64_t i = s - 1; i >= 0; --i) { // L^T x = y
      double v = x[i * q.ldo];
      for (int64_t m = i + 1; m < s; ++m) v -= Lp[pk(m, i)] * x[m * q.ldo];
      x[i * q.ldo] = v / Lp[pk(i, i)];
    }
  }
}

}  // namespace

extern "C" int64_t fh_ns_prepare(FhNsGroup* groups, int n) {
  int64_t total = 0;
  for (int i = 0; i < n; ++i) {
    FhNsGroup& g = groups[i];
    g.tm = (int32_t)fh_cdiv(g.s, 64);
    g.tiles = g.tm * g.tm;
    g.tile_begin = total;
    total += (int64_t)g.tiles * g.count;
  }
  return total;
}

// Full Newton-Schulz run: G (in Y0 of each group) -> Z0/Z1 holds G^{-1/2}, nuc = trace(G^{1/2}).
// groups_dev / mat_begin_dev: device copies prepared by the host; total_tiles from fh_ns_prepare;
// total_mats = sum of counts.  Returns the buffer index (0/1) holding the result.
extern "C" int fh_ns_run(const FhNsGroup* groups_dev, int n_groups, const int64_t* mat_begin_dev,
                         int64_t total_tiles, int64_t total_mats, int n_iter, double tol,
                         double* nuc_dev, int* result_buffer, hipStream_t stream) {
  FH_REQUIRE(n_iter >= 1 && n_iter <= 200, "ns: n_iter out of range");
  if (total_mats == 0) { *result_buffer = 0; return FH_OK; }
  hipLaunchKernelGGL(ns_init_kernel, dim3((unsigned)total_mats), dim3(NT), 0, stream, groups_dev, n_groups,
                     mat_begin_dev, n_iter);
  FH_CHECK_LAUNCH();
  for (int it = 0; it < n_iter; ++it) {
    hipLaunchKernelGGL(ns_t_kernel, dim3((unsigned)total_tiles), dim3(NT), 0, stream, groups_dev, n_groups, it, tol);
    FH_CHECK_LAUNCH();
    hipLaunchKernelGGL(ns_yz_kernel, dim3((unsigned)(2 * total_tiles)), dim3(NT), 0, stream, groups_dev, n_groups, it, tol);
    FH_CHECK_LAUNCH();
  }
  const int fin = n_iter & 1;
  hipLaunchKernelGGL(ns_final_kernel, dim3((unsigned)total_mats), dim3(NT), 0, stream, groups_dev, n_groups,
                     mat_begin_dev, fin, nullptr, nuc_dev);
  FH_CHECK_LAUNCH();
  *result_buffer = fin;
  return FH_OK;
}

// Batched SPD inverses; sys_dev: device table of n systems, max_s the largest s.
extern "C" int fh_spd_inverse(const FhSpdSys* sys_dev, int n, int64_t max_s, int* status_dev,
                              hipStream_t stream) {
  if (n <= 0) return FH_OK;
  FH_REQUIRE(max_s >= 1 && max_s <= CHOL_MAX, "spd_inverse: s=%lld out of range (max %d)", (long long)max_s, CHOL_MAX);
  size_t lds = (size_t)(max_s * (max_s + 1) / 2) * sizeof(double);
  FH_REQUIRE(lds <= (size_t)CHOL_LDS, "spd_inverse: s=%lld exceeds LDS", (long long)max_s);
  hipLaunchKernelGGL(spd_inverse_kernel, dim3(n), dim3(NT), lds, stream, sys_dev, status_dev);
  FH_CHECK_LAUNCH();
  return FH_OK;
}
