// Batched FP64 polar factors and SPD solves.
//
// Polar factor of Y (m x s, m >= s):  polar(Y) = Y G^{-1/2},  G = Y^T Y (FP64).
// Two solvers for G^{-1/2}, both one workgroup per matrix:
//  * scaled Newton (fh_polar_newton, s <= 160): H <- (z^2 H + 2 I + z^-2 H^{-1}) / 4 with an
//    in-register Gauss-Jordan inverse per step, K <- K (z I + z^-1 H^{-1}) / 2 on the matrix
//    cores; about 8 steps for any condition number;
//  * coupled Newton-Schulz on Y_0 = G/c, Z_0 = I (c = ||G||_F >= lambda_max):
//        T_k = 1.5 I - 0.5 Z_k Y_k,   Y_{k+1} = Y_k T_k,   Z_{k+1} = T_k Z_k
//    Y_k -> (G/c)^{1/2}, Z_k -> (G/c)^{-1/2}, until ||T_k - I||_F^2 < tol; only products
//    (v_mfma_f64_16x16x4_f64), O(log cond G) steps.  Used above s = 160 and by fh_ns_run.
// Nuclear norm: trace(G^{1/2}).
#include "fh_common.h"

namespace {

constexpr int NT = 256;


}  // namespace


namespace {

__device__ __forceinline__ int find_group(const FhNsGroup* g, int n, int64_t t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (g[mid].tile_begin <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// 64x64 output tile of  alpha * L * R  (+ diag on the diagonal),  L, R row-major s x s, in
// k-steps of BKD (smem: 2 * 64 * (BKD + 1) doubles).  The MFMA sequence over k does not depend
// on BKD.  Writes out and, if res != null, accumulates sum((out - I)^2) over the tile.
template <int BKD = 16>
__device__ __forceinline__ void tile_gemm_f64(const double* __restrict__ L, const double* __restrict__ R,
                                              double* __restrict__ out, int64_t s, int64_t ld,
                                              int64_t m0, int64_t n0, double alpha, double diag,
                                              double* res, double* smem) {
  constexpr int LR = BKD + 1, PER = 64 * BKD / NT;
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
    for (int p = 0; p < PER; ++p) {
      int e = tid + p * NT;
      int r = e / BKD, k = e % BKD;
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

// k-step of the per-iteration Newton-Schulz products: four k-steps of 64 at s = 256 (one
// global-load round trip each) instead of sixteen of 16
constexpr int kNsBK = 64;

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

// T = diag I + alpha Z_k Y_k for matrices still iterating (1.5 I - 0.5 Z_k Y_k unscaled).
__global__ __launch_bounds__(NT) void ns_t_kernel(const FhNsGroup* __restrict__ groups, int n_groups,
                                                  int iter, double tol, double alpha, double diag) {
  __shared__ double smem[2 * 64 * (kNsBK + 1)];
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
  tile_gemm_f64<kNsBK>(Z, Y, T, g.s, g.ld, tm * 64, tn * 64, alpha, diag, res_now, smem);
}

// Y_{k+1} = Y_k T, Z_{k+1} = T Z_k  (grid: 2 x tiles); converged matrices are copied.
__global__ __launch_bounds__(NT) void ns_yz_kernel(const FhNsGroup* __restrict__ groups, int n_groups,
                                                   int iter, double tol) {
  __shared__ double smem[2 * 64 * (kNsBK + 1)];
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
  if (which == 0) tile_gemm_f64<kNsBK>(Yc, T, Yn, g.s, g.ld, m0, n0, 1.0, 0.0, nullptr, smem);
  else tile_gemm_f64<kNsBK>(T, Zc, Zn, g.s, g.ld, m0, n0, 1.0, 0.0, nullptr, smem);
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
// s <= 160: register Gauss-Jordan (spd_gj_kernel below).  160 < s <= 256: one workgroup per
// system, Cholesky with the packed lower factor in LDS, then every thread solves
// L L^T x = e_j for its own columns j, in place in the output.
// ---------------------------------------------------------------------------
constexpr int CHOL_MAX = 256;
constexpr int CHOL_LDS = 160 * 1024 - 4096;   // bytes available for the packed factor

__device__ __forceinline__ int64_t pk(int64_t i, int64_t j) { return i * (i + 1) / 2 + j; }  // i >= j

}  // namespace
namespace {

__global__ __launch_bounds__(NT) void spd_inverse_kernel(const FhSpdSys* __restrict__ sys, int* __restrict__ status) {
  extern __shared__ double Lp[];
  const FhSpdSys q = sys[blockIdx.x];
  const int64_t s = q.s;
  if (s <= 160) return;                  // the register Gauss-Jordan kernels take these
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


// Full s x s product  out = alpha * L * R (+ diag I) by one workgroup of GM x GN waves,
// each owning TM x TN MFMA f64 16x16x4 tiles (s <= 16 TM GM = 16 TN GN), so every operand
// element is loaded once per product.  Accumulates sum((out - I)^2) into *res if set.
template <int GM, int GN, int TM, int TN>
__device__ __forceinline__ void full_gemm_f64_body(const double* __restrict__ L, const double* __restrict__ R,
                                                   double* __restrict__ out, int64_t s, int64_t ld, double alpha,
                                                   double diag, double* res, double* smem) {
  constexpr int TILE = 16 * TM * GM, BKD = 16, LR = BKD + 1, NTH = 64 * GM * GN;
  static_assert(16 * TN * GN == TILE, "square tile");
  double* As = smem;
  double* Bs = smem + TILE * LR;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % GM, wn = wave / GM;
  doublex4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = doublex4{0., 0., 0., 0.};
  constexpr int PER = TILE * BKD / NTH;
  static_assert(PER * NTH == TILE * BKD, "panel must split evenly over the threads");
  double ra[PER], rb[PER];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int e = tid + p * NTH;
      const int r = e >> 4, k = e & 15;
      ra[p] = (r < s && k0 + k < s) ? L[r * ld + k0 + k] : 0.;
      const int n = e % TILE, kk = e / TILE;
      rb[p] = (n < s && k0 + kk < s) ? R[(k0 + kk) * ld + n] : 0.;
    }
  };
  load(0);
  for (int64_t k0 = 0; k0 < s; k0 += BKD) {
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int e = tid + p * NTH;
      As[(e >> 4) * LR + (e & 15)] = ra[p];
      Bs[(e % TILE) * LR + e / TILE] = rb[p];
    }
    __syncthreads();
    if (k0 + BKD < s) load(k0 + BKD);     // next panel in flight during the MFMAs
#pragma unroll 1     // unrolling hoists every LDS fragment load and spills the accumulators
    for (int kk = 0; kk < BKD; kk += 4) {
      double a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = As[(wm * 16 * TM + i * 16 + (lane & 15)) * LR + kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = Bs[(wn * 16 * TN + j * 16 + (lane & 15)) * LR + kk + (lane >> 4)];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
  double part = 0.;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int64_t row = wm * 16 * TM + i * 16 + (lane >> 4) + 4 * r;   // f64 MFMA output row mapping
        int64_t col = wn * 16 * TN + j * 16 + (lane & 15);
        if (row < s && col < s) {
          double v = alpha * acc[i][j][r] + (row == col ? diag : 0.);
          out[row * ld + col] = v;
          double dv = v - (row == col ? 1. : 0.);
          part += dv * dv;
        }
      }
  if (res) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
    if (lane == 0) atomicAdd(res, part);
  }
}

// Out-of-line copy for callers with many call sites (keeps code size and register
// pressure of the Newton-Schulz loop down).
template <int GM, int GN, int TM, int TN>
__device__ __attribute__((noinline)) void full_gemm_f64(const double* __restrict__ L, const double* __restrict__ R,
                                              double* __restrict__ out, int64_t s, int64_t ld, double alpha,
                                              double diag, double* res, double* smem) {
  full_gemm_f64_body<GM, GN, TM, TN>(L, R, out, s, ld, alpha, diag, res, smem);
}

// wave grids of the fused solve per size class
template <int C> struct NsCfg;
template <> struct NsCfg<2> { static constexpr int GM = 2, GN = 2, TM = 2, TN = 2; };   //  64 x  64,  4 waves
template <> struct NsCfg<4> { static constexpr int GM = 2, GN = 4, TM = 4, TN = 2; };   // 128 x 128,  8 waves
template <> struct NsCfg<5> { static constexpr int GM = 2, GN = 5, TM = 5, TN = 2; };   // 160 x 160, 10 waves
template <> struct NsCfg<0> { static constexpr int GM = 2, GN = 2, TM = 2, TN = 2; };   // tiled 64 x 64

// size class of a matrix for the fused solve: full-matrix products for s <= 160
__device__ __forceinline__ int ns_class(int64_t s) { return s <= 64 ? 2 : (s <= 128 ? 4 : (s <= 160 ? 5 : 0)); }

// One workgroup per matrix runs the whole Newton-Schulz solve: scale, iterate until
// ||T - I||_F^2 < tol (or max_iter), then Z <- Z / sqrt(c) into Z0 and the nuclear
// norm.  Y/Z/T live in global memory (L2 / infinity cache); every phase is a sequence
// of 64x64 output tiles computed by the four waves together.
template <int WT>
__global__ __launch_bounds__(64 * NsCfg<WT>::GM * NsCfg<WT>::GN) void ns_fused_kernel(
    const FhNsGroup* __restrict__ groups, int n_groups, const int64_t* __restrict__ mat_begin, int max_iter,
    double tol, double* __restrict__ nuc, int* __restrict__ iters) {
  using Cfg = NsCfg<WT>;
  constexpr int NTH = 64 * Cfg::GM * Cfg::GN;
  __shared__ double smem[2 * 16 * Cfg::TM * Cfg::GM * 17];
  __shared__ double red[16];
  __shared__ double res_sh;
  const int64_t gm = blockIdx.x;
  int lo = 0, hi = n_groups - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (mat_begin[mid] <= gm) lo = mid; else hi = mid - 1;
  }
  const FhNsGroup g = groups[lo];
  const int64_t b = gm - mat_begin[lo];
  const int64_t s = g.s, ld = g.ld;
  if (ns_class(s) != WT) return;          // another launch of this solve handles this size
  double* const Y0 = g.Y0 + b * g.stride;
  double* const Y1 = g.Y1 + b * g.stride;
  double* const Z0 = g.Z0 + b * g.stride;
  double* const Z1 = g.Z1 + b * g.stride;
#define YB(i) ((i) ? Y1 : Y0)
#define ZB(i) ((i) ? Z1 : Z0)
  double* T = g.T + b * g.stride;
  const int tid = threadIdx.x;
  // scale: c = ||G||_F (G is in Y0)
  double part = 0.;
  for (int64_t e = tid; e < s * s; e += NTH) {
    double v = Y0[(e / s) * ld + e % s];
    part += v * v;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
  if ((tid & 63) == 0) red[tid >> 6] = part;
  __syncthreads();
  double c = 0.;
  for (int w = 0; w < NTH / 64; ++w) c += red[w];
  c = sqrt(c);
  if (!(c > 0.)) c = 1.;
  const double inv = 1. / c;
  for (int64_t e = tid; e < s * s; e += NTH) {
    int64_t r = e / s, col = e % s;
    Y0[r * ld + col] *= inv;
    Z0[r * ld + col] = (r == col) ? 1. : 0.;
  }
  __threadfence_block();
  __syncthreads();
  const int tm = (int)((s + 63) / 64);
  int cur = 0, it = 0;
  for (; it < max_iter; ++it) {
    if (tid == 0) res_sh = 0.;
    __syncthreads();
    // T = 1.5 I - 0.5 Z Y
    if constexpr (WT > 0) {
      full_gemm_f64<Cfg::GM, Cfg::GN, Cfg::TM, Cfg::TN>(ZB(cur), YB(cur), T, s, ld, -0.5, 1.5, &res_sh, smem);
    } else {
      for (int t = 0; t < tm * tm; ++t)
        tile_gemm_f64(ZB(cur), YB(cur), T, s, ld, (t / tm) * 64, (t % tm) * 64, -0.5, 1.5, &res_sh, smem);
    }
    __threadfence_block();
    __syncthreads();
    const double r = res_sh;
    if (r < tol) break;
    // Y' = Y T, Z' = T Z
    if constexpr (WT > 0) {
      full_gemm_f64<Cfg::GM, Cfg::GN, Cfg::TM, Cfg::TN>(YB(cur), T, YB(cur ^ 1), s, ld, 1.0, 0.0, nullptr, smem);
      full_gemm_f64<Cfg::GM, Cfg::GN, Cfg::TM, Cfg::TN>(T, ZB(cur), ZB(cur ^ 1), s, ld, 1.0, 0.0, nullptr, smem);
    } else {
      for (int t = 0; t < tm * tm; ++t) {
        tile_gemm_f64(YB(cur), T, YB(cur ^ 1), s, ld, (t / tm) * 64, (t % tm) * 64, 1.0, 0.0, nullptr, smem);
        tile_gemm_f64(T, ZB(cur), ZB(cur ^ 1), s, ld, (t / tm) * 64, (t % tm) * 64, 1.0, 0.0, nullptr, smem);
      }
    }
    __threadfence_block();
    __syncthreads();
    cur ^= 1;
  }
  // finalize into Z0: G^{-1/2} = Z / sqrt(c); nuclear norm = sqrt(c) trace(Y)
  const double rs = sqrt(c), irs = 1. / rs;
  part = 0.;
  for (int64_t i = tid; i < s; i += NTH) part += YB(cur)[i * ld + i];
  for (int64_t e = tid; e < s * s; e += NTH) {
    int64_t rr = e / s, col = e % s;
    Z0[rr * ld + col] = ZB(cur)[rr * ld + col] * irs;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = part;
  __syncthreads();
  if (tid == 0) {
    double tr = 0.;
    for (int w = 0; w < NTH / 64; ++w) tr += red[w];
    nuc[gm] = rs * tr;
    g.scale[b] = c;
    if (iters) iters[gm] = it;
  }
#undef YB
#undef ZB
}

// ---------------------------------------------------------------------------
// Scaled Newton iteration for G^{-1/2} (s <= 160).
//
// With X = Y K the Newton step for the polar factor X <- (z X + z^-1 X^{-T}) / 2 only
// needs H = X^T X:   W = (z I + z^-1 H^{-1}) / 2,  K <- K W,  H <- (z^2 H + 2 I + z^-2 H^{-1}) / 4,
// with the Frobenius-norm scaling z = (tr H^{-1} / tr H)^{1/4} until ||H - I||_F < 0.1.
// Starting from H = G / ||G||_F it converges in about log2(log(cond G)) + 3 steps.  Each
// step is two launches over the batch: the Gauss-Jordan inverse (vector FP64, the matrix
// cyclically distributed over the workgroup's registers, one barrier per pivot) with the
// H / W update, then K W on the matrix cores.  A final Newton-Schulz correction
// K <- K (3 I - K H0 K) / 2 on the true (unshifted) residual removes the drift of the H
// recurrence and the shift below.  Output: G^{-1/2} in Z0, tr(G^{1/2}) in nuc.
//
// Per-matrix state between launches lives in iters[]: bits 0-7 step count, bit 8 done,
// bit 9 scaling off, bit 10 buffer (Z1 / T) holding K.
// ---------------------------------------------------------------------------
// The iterate starts from G / ||G||_F + kNewtonShift I: rank-deficient Grams (which the
// Newton-Schulz iteration tolerates) stay invertible; the final correction on the
// unshifted G removes the shift's first-order effect on every direction with
// eigenvalue >> kNewtonShift.
constexpr double kNewtonShift = 1e-13;
constexpr int kNtDone = 1 << 8, kNtNoScale = 1 << 9, kNtKbuf = 1 << 10, kNtCount = 0xff;

template <int NTH>
__device__ __forceinline__ void block_sum2(double& a, double& b, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
  __syncthreads();                       // red may still be read by the previous reduction
  if ((threadIdx.x & 63) == 0) {
    red[2 * (threadIdx.x >> 6)] = a;
    red[2 * (threadIdx.x >> 6) + 1] = b;
  }
  __syncthreads();
  a = 0.;
  b = 0.;
#pragma unroll
  for (int w = 0; w < NTH / 64; ++w) {
    a += red[2 * w];
    b += red[2 * w + 1];
  }
}

struct MatRef {
  FhNsGroup g;
  int64_t b;
};
__device__ __forceinline__ MatRef find_mat(const FhNsGroup* groups, int n_groups, const int64_t* mat_begin,
                                           int64_t gm) {
  int lo = 0, hi = n_groups - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (mat_begin[mid] <= gm) lo = mid; else hi = mid - 1;
  }
  return MatRef{groups[lo], gm - mat_begin[lo]};
}

// c = ||G||_F;  Y0 <- G / c;  Z0 <- G / c + shift I;  state cleared
__global__ __launch_bounds__(256) void newton_init_kernel(const FhNsGroup* __restrict__ groups, int n_groups,
                                                          const int64_t* __restrict__ mat_begin,
                                                          int* __restrict__ state) {
  __shared__ double red[8];
  const MatRef m = find_mat(groups, n_groups, mat_begin, blockIdx.x);
  const int64_t s = m.g.s, ld = m.g.ld;
  if (ns_class(s) == 0) return;
  double* H0 = m.g.Y0 + m.b * m.g.stride;
  double* Hk = m.g.Z0 + m.b * m.g.stride;
  double p0 = 0., p1 = 0.;
  for (int64_t e = threadIdx.x; e < s * s; e += 256) {
    const double v = H0[(e / s) * ld + e % s];
    p0 += v * v;
  }
  block_sum2<256>(p0, p1, red);
  double c = sqrt(p0);
  if (!(c > 0.)) c = 1.;
  const double inv_c = 1. / c;
  for (int64_t e = threadIdx.x; e < s * s; e += 256) {
    const int64_t r = e / s, q = e % s;
    const double v = H0[r * ld + q] * inv_c;
    H0[r * ld + q] = v;
    Hk[r * ld + q] = v + (r == q ? kNewtonShift : 0.);
  }
  if (threadIdx.x == 0) {
    m.g.scale[m.b] = c;
    state[blockIdx.x] = 0;
  }
}

// register grid of the in-place inverse per size class: RG x CG threads, SMAX = matrix bound
template <int C> struct GjCfg;
// WPE: waves per SIMD the register budget is sized for (4: two 8-wave workgroups per CU)
template <> struct GjCfg<2> { static constexpr int RG = 16, CG = 16, SMAX = 64, WPE = 4; };
template <> struct GjCfg<4> { static constexpr int RG = 16, CG = 32, SMAX = 128, WPE = 4; };
template <> struct GjCfg<5> { static constexpr int RG = 16, CG = 32, SMAX = 160, WPE = 2; };

// LDS of the pivot broadcast: row k and column k stored thread-major (the CPT row values a
// thread needs, then the next thread's), each thread's run padded to an even count, so a
// thread reads its values with 16-byte loads.
template <int RG, int CG, int SMAX>
struct GjBuf {
  static constexpr int RPT = SMAX / RG, CPT = SMAX / CG;
  static constexpr int RPTP = (RPT + 1) & ~1, CPTP = (CPT + 1) & ~1;
  static constexpr int ROW = CG * CPTP, COL = RG * RPTP;   // doubles per buffer
};

// In-place Gauss-Jordan inverse of an SPD s x s matrix (s <= SMAX) held by the workgroup's
// RG x CG threads: element (i, j) in a[i / RG][j / CG] of thread (i % RG, j % CG), padding
// rows / columns set to the identity.  One barrier per pivot.
//
// Pivot k = kr RG + ky: the outer loop over kr is unrolled, so the owners' register slots
// (row slot kr, column slot kq = kr RG / CG) are compile-time: publishing row / column k and
// rewriting them after the rank-1 update a <- a - u v^T (u = column k, v = row k / p) are
// plain masked register moves, with no selects over slots.  The pivot owner broadcasts 1 / p.
// The row is scaled by -1 / p, so the update is fma(u, -v, a) (v_fmac, no sign flips of u).
template <int RG, int CG, int SMAX>
__device__ __forceinline__ void gj_invert(double (&a)[SMAX / RG][SMAX / CG], int s, double* rowb, double* colb,
                                          double* ipb) {
  using Buf = GjBuf<RG, CG, SMAX>;
  constexpr int RPT = Buf::RPT, CPT = Buf::CPT, RPTP = Buf::RPTP, CPTP = Buf::CPTP;
  typedef double doublex2 __attribute__((ext_vector_type(2)));
  static_assert(CG % RG == 0 && RG % 2 == 0, "column slot of a pivot must follow from its row slot");
  const int tid = threadIdx.x, ty = tid / CG, tx = tid % CG;
#pragma unroll
  for (int kr = 0; kr < RPT; ++kr) {
    const int kq = kr * RG / CG;                     // compile-time in the unrolled loop
    const int kend = min(RG, s - kr * RG);
#pragma unroll 1
    for (int ky = 0; ky < kend; ++ky) {
      const int bf = ky & 1;
      double* const rb = rowb + bf * Buf::ROW;
      double* const cb = colb + bf * Buf::COL;
      const int kx = kr * RG - kq * CG + ky;
      if (ty == ky) {
#pragma unroll
        for (int q = 0; q < CPT; ++q) rb[tx * CPTP + q] = a[kr][q];
      }
      if (tx == kx) {
#pragma unroll
        for (int r = 0; r < RPT; ++r) cb[ty * RPTP + r] = a[r][kq];
        if (ty == ky) {
          double p = a[kr][kq];
          p = p > 1e-300 ? p : 1e-300;
          ipb[bf] = 1. / p;
        }
      }
      __syncthreads();
      const double ip = ipb[bf], nip = -ip;
      double v[CPTP], u[RPTP];
#pragma unroll
      for (int q = 0; q < CPTP; q += 2) {
        const doublex2 t = *(const doublex2*)(rb + tx * CPTP + q);
        v[q] = t[0] * nip;
        v[q + 1] = t[1] * nip;
      }
#pragma unroll
      for (int r = 0; r < RPTP; r += 2) {
        const doublex2 t = *(const doublex2*)(cb + ty * RPTP + r);
        u[r] = t[0];
        u[r + 1] = t[1];
      }
#pragma unroll
      for (int r = 0; r < RPT; ++r)
#pragma unroll
        for (int q = 0; q < CPT; ++q) a[r][q] = fma(u[r], v[q], a[r][q]);
      // row k <- A[k, :] / p, column k <- -A[:, k] / p, (k, k) <- 1 / p
      if (ty == ky) {
#pragma unroll
        for (int q = 0; q < CPT; ++q) a[kr][q] = -v[q];
      }
      if (tx == kx) {
#pragma unroll
        for (int r = 0; r < RPT; ++r) a[r][kq] = u[r] * nip;
        if (ty == ky) a[kr][kq] = ip;
      }
    }
  }
}

// One Newton step of every matrix of class WT that has not converged: H^{-1} by Gauss-Jordan,
// the scaling, W (into Y1, or straight into Z1 as K at step 0) and the new H.
//
// Element (i, j) lives in a[i / RG][j / CG] of thread (i % RG, j % CG); the owner of a pivot
// row / column follows from uniform arithmetic, so the only divergence is the owner's lane
// mask.  Pivots are taken one at a time: the first step inverts G / ||G||_F + shift, whose
// condition number reaches 1 / shift, and sequential elimination of an SPD matrix is
// backward stable where an explicit inverse of a pivot block is not.  Rows/columns >= s are
// identity padding.
template <int WT>
__global__ __launch_bounds__(GjCfg<WT>::RG * GjCfg<WT>::CG) __attribute__((amdgpu_waves_per_eu(GjCfg<WT>::WPE)))
void newton_gj_kernel(
    const FhNsGroup* __restrict__ groups, int n_groups, const int64_t* __restrict__ mat_begin,
    int* __restrict__ state, int step, double tol) {
  constexpr int RG = GjCfg<WT>::RG, CG = GjCfg<WT>::CG, SMAX = GjCfg<WT>::SMAX, NTH = RG * CG;
  constexpr int RPT = SMAX / RG, CPT = SMAX / CG;
  static_assert(RPT * RG == SMAX && CPT * CG == SMAX, "register grid must tile the matrix");
  __shared__ __attribute__((aligned(16))) double rowb[2 * GjBuf<RG, CG, SMAX>::ROW];
  __shared__ __attribute__((aligned(16))) double colb[2 * GjBuf<RG, CG, SMAX>::COL];
  __shared__ double ipb[2];
  __shared__ double red[2 * (NTH / 64)];
  const MatRef m = find_mat(groups, n_groups, mat_begin, blockIdx.x);
  const int s = (int)m.g.s, ld = (int)m.g.ld;
  if (ns_class(s) != WT) return;
  const int st = state[blockIdx.x];
  if (st & kNtDone) return;
  double* const Hk = m.g.Z0 + m.b * m.g.stride;
  double* const Wp = (step == 0 ? m.g.Z1 : m.g.Y1) + m.b * m.g.stride;
  const int tid = threadIdx.x, ty = tid / CG, tx = tid % CG;
  double a[RPT][CPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r)
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const int i = ty + RG * r, j = tx + CG * q;
      a[r][q] = (i < s && j < s) ? Hk[i * ld + j] : (i == j ? 1. : 0.);
    }
  gj_invert<RG, CG, SMAX>(a, s, rowb, colb, ipb);
  // scaling from the traces of H and H^{-1}
  double th = 0., ti = 0.;
#pragma unroll
  for (int r = 0; r < RPT; ++r)
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const int i = ty + RG * r, j = tx + CG * q;
      if (i == j && i < s) {
        ti += a[r][q];
        th += Hk[i * ld + i];
      }
    }
  block_sum2<NTH>(th, ti, red);
  const double ratio = ti / th;
  const bool scaling = !(st & kNtNoScale);
  const double z = (scaling && ratio > 0. && ratio < 1e300) ? sqrt(sqrt(ratio)) : 1.;
  const double hz = 0.5 * z, hzi = 0.5 / z, qz2 = 0.25 * z * z, qzi2 = 0.25 / (z * z);
  double dsq = 0., dummy = 0.;
#pragma unroll
  for (int r = 0; r < RPT; ++r)
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const int i = ty + RG * r, j = tx + CG * q;
      if (i < s && j < s) {
        const double hinv = a[r][q];
        const double dg = (i == j) ? 1. : 0.;
        const int o = i * ld + j;
        Wp[o] = hzi * hinv + hz * dg;
        const double hn = qz2 * Hk[o] + qzi2 * hinv + 0.5 * dg;
        Hk[o] = hn;
        dsq += (hn - dg) * (hn - dg);
      }
    }
  block_sum2<NTH>(dsq, dummy, red);
  if (tid == 0) {
    // K after this step: Z1 at step 0, then alternating T / Z1
    int ns = (st & ~kNtCount) | (step + 1);
    if (dsq < 1e-2) ns |= kNtNoScale;
    if (dsq < tol) ns |= kNtDone;
    state[blockIdx.x] = ns;
  }
}

// out = (G + eps I)^{-1} for every system of class WT (s <= 160): Gauss-Jordan in registers.
// Replaces torch.linalg.solve(A + 1e-10 I, B) of update_factor (parafac_integrative.py:12-15)
// by an explicit inverse applied as a GEMM.
template <int WT>
__global__ __launch_bounds__(GjCfg<WT>::RG * GjCfg<WT>::CG) void spd_gj_kernel(const FhSpdSys* __restrict__ sys) {
  constexpr int RG = GjCfg<WT>::RG, CG = GjCfg<WT>::CG, SMAX = GjCfg<WT>::SMAX;
  constexpr int RPT = SMAX / RG, CPT = SMAX / CG;
  __shared__ __attribute__((aligned(16))) double rowb[2 * GjBuf<RG, CG, SMAX>::ROW];
  __shared__ __attribute__((aligned(16))) double colb[2 * GjBuf<RG, CG, SMAX>::COL];
  __shared__ double ipb[2];
  const FhSpdSys q = sys[blockIdx.x];
  const int s = (int)q.s;
  if (ns_class(s) != WT) return;
  const int tid = threadIdx.x, ty = tid / CG, tx = tid % CG;
  double a[RPT][CPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r)
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const int i = ty + RG * r, j = tx + CG * c;
      a[r][c] = (i < s && j < s) ? q.G[i * q.ldg + j] + (i == j ? q.eps : 0.) : (i == j ? 1. : 0.);
    }
  gj_invert<RG, CG, SMAX>(a, s, rowb, colb, ipb);
#pragma unroll
  for (int r = 0; r < RPT; ++r)
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const int i = ty + RG * r, j = tx + CG * c;
      if (i < s && j < s) q.out[i * q.ldo + j] = a[r][c];
    }
}

// K <- K W for the matrices stepped in launch `step` (>= 1): K in Z1 (odd step) / T (even step)
template <int WT>
__global__ __launch_bounds__(64 * NsCfg<WT>::GM * NsCfg<WT>::GN) void newton_kw_kernel(
    const FhNsGroup* __restrict__ groups, int n_groups, const int64_t* __restrict__ mat_begin,
    int* __restrict__ state, int step) {
  using Cfg = NsCfg<WT>;
  __shared__ double smem[2 * 16 * Cfg::TM * Cfg::GM * 17];
  const MatRef m = find_mat(groups, n_groups, mat_begin, blockIdx.x);
  const int64_t s = m.g.s, ld = m.g.ld;
  if (ns_class(s) != WT) return;
  const int st = state[blockIdx.x];
  if ((st & kNtCount) != step + 1) return;     // converged before this step
  double* const K0 = m.g.Z1 + m.b * m.g.stride;
  double* const K1 = m.g.T + m.b * m.g.stride;
  const double* W = m.g.Y1 + m.b * m.g.stride;
  const bool from_z1 = (step & 1);             // K of step-1 in Z1 when step is odd
  full_gemm_f64_body<Cfg::GM, Cfg::GN, Cfg::TM, Cfg::TN>(from_z1 ? K0 : K1, W, from_z1 ? K1 : K0, s, ld, 1.0,
                                                         0.0, nullptr, smem);
  if (threadIdx.x == 0) state[blockIdx.x] = from_z1 ? (st | kNtKbuf) : (st & ~kNtKbuf);
}

// Correction on the true residual, K <- K (1.5 I - 0.5 K H0 K), then G^{-1/2} = K / sqrt(c)
// into Z0 and tr(G^{1/2}) = sum_ij G_ij (G^{-1/2})_ij.
template <int WT>
__global__ __launch_bounds__(64 * NsCfg<WT>::GM * NsCfg<WT>::GN) void newton_final_kernel(
    const FhNsGroup* __restrict__ groups, int n_groups, const int64_t* __restrict__ mat_begin,
    int* __restrict__ state, double* __restrict__ nuc) {
  using Cfg = NsCfg<WT>;
  constexpr int NTH = 64 * Cfg::GM * Cfg::GN;
  __shared__ double smem[2 * 16 * Cfg::TM * Cfg::GM * 17];
  __shared__ double red[2 * (NTH / 64)];
  const MatRef m = find_mat(groups, n_groups, mat_begin, blockIdx.x);
  const int64_t s = m.g.s, ld = m.g.ld;
  if (ns_class(s) != WT) return;
  const int st = state[blockIdx.x];
  double* const H0 = m.g.Y0 + m.b * m.g.stride;
  double* const Out = m.g.Z0 + m.b * m.g.stride;
  double* const Sc = m.g.Y1 + m.b * m.g.stride;
  double* const K = ((st & kNtKbuf) ? m.g.T : m.g.Z1) + m.b * m.g.stride;
  double* const M = ((st & kNtKbuf) ? m.g.Z1 : m.g.T) + m.b * m.g.stride;
  full_gemm_f64<Cfg::GM, Cfg::GN, Cfg::TM, Cfg::TN>(K, H0, Sc, s, ld, 1.0, 0.0, nullptr, smem);
  __threadfence_block();
  __syncthreads();
  full_gemm_f64<Cfg::GM, Cfg::GN, Cfg::TM, Cfg::TN>(Sc, K, M, s, ld, -0.5, 1.5, nullptr, smem);
  __threadfence_block();
  __syncthreads();
  full_gemm_f64<Cfg::GM, Cfg::GN, Cfg::TM, Cfg::TN>(K, M, Out, s, ld, 1.0, 0.0, nullptr, smem);
  __threadfence_block();
  __syncthreads();
  const double c = m.g.scale[m.b];
  const double irs = 1. / sqrt(c);
  double tr = 0., d2 = 0.;
  for (int64_t e = threadIdx.x; e < s * s; e += NTH) {
    const int64_t r = e / s, q = e % s;
    const double v = Out[r * ld + q] * irs;
    Out[r * ld + q] = v;
    tr += H0[r * ld + q] * v;
  }
  block_sum2<NTH>(tr, d2, red);
  if (threadIdx.x == 0) {
    nuc[blockIdx.x] = c * tr;
    state[blockIdx.x] = st & kNtCount;
  }
}

}  // namespace

// Scaled Newton solve (s <= 160; larger s use the Newton-Schulz solve): G (in Y0) ->
// G^{-1/2} in Z0, tr(G^{1/2}) in nuc, Newton steps in iters (required: it also carries the
// per-matrix state between the launches).  Y1, Z1, T are scratch.  max_iter launches of
// each step kernel; tol: a matrix stops when ||H - I||_F^2 < tol.
extern "C" int fh_polar_newton(const FhNsGroup* groups_dev, int n_groups, const int64_t* mat_begin_dev,
                               int64_t total_mats, int max_iter, double tol, double* nuc_dev, int* iters_dev,
                               hipStream_t stream) {
  FH_REQUIRE(max_iter >= 1 && max_iter <= 64, "polar_newton: max_iter out of range (1..64)");
  FH_REQUIRE(iters_dev != nullptr, "polar_newton: iters (the per-matrix state) is required");
  if (total_mats == 0) return FH_OK;
  FH_REQUIRE(total_mats < (1ll << 31), "polar_newton: too many matrices");
  const dim3 grid((unsigned)total_mats);
  hipLaunchKernelGGL(newton_init_kernel, grid, dim3(256), 0, stream, groups_dev, n_groups, mat_begin_dev,
                     iters_dev);
  for (int step = 0; step < max_iter; ++step) {
    hipLaunchKernelGGL(newton_gj_kernel<5>, grid, dim3(GjCfg<5>::RG * GjCfg<5>::CG), 0, stream, groups_dev, n_groups, mat_begin_dev,
                       iters_dev, step, tol);
    hipLaunchKernelGGL(newton_gj_kernel<4>, grid, dim3(GjCfg<4>::RG * GjCfg<4>::CG), 0, stream, groups_dev, n_groups, mat_begin_dev,
                       iters_dev, step, tol);
    hipLaunchKernelGGL(newton_gj_kernel<2>, grid, dim3(GjCfg<2>::RG * GjCfg<2>::CG), 0, stream, groups_dev, n_groups, mat_begin_dev,
                       iters_dev, step, tol);
    if (step > 0) {
      hipLaunchKernelGGL(newton_kw_kernel<5>, grid, dim3(640), 0, stream, groups_dev, n_groups, mat_begin_dev,
                         iters_dev, step);
      hipLaunchKernelGGL(newton_kw_kernel<4>, grid, dim3(512), 0, stream, groups_dev, n_groups, mat_begin_dev,
                         iters_dev, step);
      hipLaunchKernelGGL(newton_kw_kernel<2>, grid, dim3(256), 0, stream, groups_dev, n_groups, mat_begin_dev,
                         iters_dev, step);
    }
  }
  hipLaunchKernelGGL(newton_final_kernel<5>, grid, dim3(640), 0, stream, groups_dev, n_groups, mat_begin_dev,
                     iters_dev, nuc_dev);
  hipLaunchKernelGGL(newton_final_kernel<4>, grid, dim3(512), 0, stream, groups_dev, n_groups, mat_begin_dev,
                     iters_dev, nuc_dev);
  hipLaunchKernelGGL(newton_final_kernel<2>, grid, dim3(256), 0, stream, groups_dev, n_groups, mat_begin_dev,
                     iters_dev, nuc_dev);
  hipLaunchKernelGGL(ns_fused_kernel<0>, grid, dim3(256), 0, stream, groups_dev, n_groups, mat_begin_dev, 48,
                     1e-24, nuc_dev, iters_dev);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

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

// Fused solve: one launch, one workgroup per matrix; the result G^{-1/2} is left in Z0 and the
// iteration count of every matrix in iters (nullable).
extern "C" int fh_ns_fused(const FhNsGroup* groups_dev, int n_groups, const int64_t* mat_begin_dev,
                           int64_t total_mats, int max_iter, double tol, double* nuc_dev, int* iters_dev,
                           hipStream_t stream) {
  FH_REQUIRE(max_iter >= 1 && max_iter <= 200, "ns_fused: max_iter out of range");
  if (total_mats == 0) return FH_OK;
  FH_REQUIRE(total_mats < (1ll << 31), "ns_fused: too many matrices");
  const dim3 grid((unsigned)total_mats);
  hipLaunchKernelGGL(ns_fused_kernel<5>, grid, dim3(640), 0, stream, groups_dev, n_groups, mat_begin_dev, max_iter,
                     tol, nuc_dev, iters_dev);
  hipLaunchKernelGGL(ns_fused_kernel<4>, grid, dim3(512), 0, stream, groups_dev, n_groups, mat_begin_dev, max_iter,
                     tol, nuc_dev, iters_dev);
  hipLaunchKernelGGL(ns_fused_kernel<2>, grid, dim3(256), 0, stream, groups_dev, n_groups, mat_begin_dev, max_iter,
                     tol, nuc_dev, iters_dev);
  hipLaunchKernelGGL(ns_fused_kernel<0>, grid, dim3(256), 0, stream, groups_dev, n_groups, mat_begin_dev, max_iter,
                     tol, nuc_dev, iters_dev);
  FH_CHECK_LAUNCH();
  return FH_OK;
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
  // Scaled iteration (Chen & Chow, "A stable scaling of Newton-Schulz for improving the sign
  // function computation of a Hermitian matrix", 2014): in sign-function form the step is
  // x <- a x (3 - a^2 x^2) / 2 on the singular values x of Y_0^{1/2} (in (0, 1] after the
  // Frobenius scaling), with a = sqrt(3 / (1 + l + l^2)) for a lower bound l of the interval;
  // the map keeps (0, 1] in (0, 1], lifts small values by up to 1.5 sqrt(3) per step instead
  // of 1.5, and l follows the same map.  l_0 = 1e-4 (condition number 1e8 of G); smaller
  // singular values still converge, as in the unscaled iteration once a reaches 1.  The
  // coupled form: T = (a / 2) (3 I - a^2 Z Y).  The limit (G / c)^{+-1/2} does not change.
  double ell = 1e-4;
  for (int it = 0; it < n_iter; ++it) {
    double a = 1.;
    if (ell < 1. - 1e-3) {
      a = sqrt(3. / (1. + ell + ell * ell));
      ell = a * ell * (3. - a * a * ell * ell) / 2.;
    }
    hipLaunchKernelGGL(ns_t_kernel, dim3((unsigned)total_tiles), dim3(NT), 0, stream, groups_dev, n_groups, it, tol,
                       -0.5 * a * a * a, 1.5 * a);
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
  const dim3 grid(n);
  if (max_s > 128) hipLaunchKernelGGL(spd_gj_kernel<5>, grid, dim3(GjCfg<5>::RG * GjCfg<5>::CG), 0, stream, sys_dev);
  if (max_s > 64) hipLaunchKernelGGL(spd_gj_kernel<4>, grid, dim3(GjCfg<4>::RG * GjCfg<4>::CG), 0, stream, sys_dev);
  hipLaunchKernelGGL(spd_gj_kernel<2>, grid, dim3(GjCfg<2>::RG * GjCfg<2>::CG), 0, stream, sys_dev);
  if (max_s > 160) hipLaunchKernelGGL(spd_inverse_kernel, grid, dim3(NT), lds, stream, sys_dev, status_dev);
  FH_CHECK_LAUNCH();
  return FH_OK;
}
