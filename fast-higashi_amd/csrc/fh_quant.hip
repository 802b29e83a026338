// FP8 (OCP e4m3, largest finite 448) quantisation of the streamed GEMM operands.
//
//  * fh_amax_rows, fh_amax_col_chunks / fh_quant_col_chunks: the data X of one chromosome
//    (positions x cells) with a rank-one scale, X = row_div[w] * col_scale[n] * e4m3: each row
//    first divided by its amax (the contact maps' magnitude falls with the distance from the
//    diagonal, along the rows), then one scale per chunk of cells, amax(chunk) / 448 (1 for an
//    all-zero chunk).  The row factor is applied in the Y = X C epilogue and folded into P for
//    T_hat = P^T X; the cell factor is folded into C before its quantisation and applied in the
//    T_hat epilogue.
//  * fh_quant_rows: a table of row-major matrices, one scale per row (amax over the row of
//    x * colscale[c], / 448), for C^T (rows r, cells) and P^T (rows p, positions).
// Columns from `cols` to the destination's leading dimension are written as zero, so every
// k-chunk the GEMM reads past K is zero.  e4m3 conversion: v_cvt_pk_fp8_f32 (round to nearest
// even) of values already scaled into [-448, 448].
#include "fh_common.h"

namespace {
constexpr int NT = 256;
constexpr float kE4M3Max = 448.f;

__device__ __forceinline__ float load_f(const void* p, int dt, int64_t i) {
  return dt == FH_BF16 ? (float)((const __bf16*)p)[i] : ((const float*)p)[i];
}

__device__ __forceinline__ uint8_t to_e4m3(float v) {
  v = fminf(fmaxf(v, -kE4M3Max), kE4M3Max);
  return (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(v, 0.f, 0, false) & 0xff);
}

__device__ __forceinline__ float scale_of(float amax) { return amax > 0.f ? amax / kE4M3Max : 1.f; }

__global__ void zero_kernel(float* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) p[i] = 0.f;
}

// one wave per row: amax over the row
__global__ __launch_bounds__(NT) void amax_rows_kernel(const void* X, int dt, int64_t ld, int64_t rows, int64_t cols,
                                                       float* out) {
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  if (r >= rows) return;
  float m = 0.f;
  for (int64_t c = lane; c < cols; c += 64) m = fmaxf(m, fabsf(load_f(X, dt, r * ld + c)));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if (lane == 0) out[r] = m;
}

// 1 / the row divisor (rows that are all zero divide by 1)
__device__ __forceinline__ float inv_row(const float* rs, int64_t r) {
  if (!rs) return 1.f;
  const float v = rs[r];
  return v > 0.f ? 1.f / v : 1.f;
}

// block (rb, cb): rows [rb * RPB, +RPB) of column chunk cb; |x| >= 0, so the float max is the
// unsigned max of the bit patterns (order independent, deterministic)
__global__ void amax_kernel(const void* X, int dt, int64_t ld, int64_t rows, int64_t cols, int64_t chunk,
                            int64_t rpb, const float* row_div, float* amax) {
  const int64_t cb = blockIdx.y;
  const int64_t c0 = cb * chunk, c1 = min(c0 + chunk, cols);
  const int64_t r0 = (int64_t)blockIdx.x * rpb, r1 = min(r0 + rpb, rows);
  const int64_t w = c1 - c0;
  float m = 0.f;
  for (int64_t e = threadIdx.x; e < (r1 - r0) * w; e += NT) {
    const int64_t r = r0 + e / w, c = c0 + e % w;
    m = fmaxf(m, fabsf(load_f(X, dt, r * ld + c) * inv_row(row_div, r)));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax((unsigned int*)(amax + cb), __float_as_uint(m));
}

__global__ void quant_cols_kernel(const void* X, int dt, int64_t ld, int64_t rows, int64_t cols, int64_t chunk,
                                  const float* row_div, const float* amax, uint8_t* dst, int64_t ldd,
                                  float* col_scale) {
  const int64_t n = rows * ldd;
  for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < n; e += (int64_t)gridDim.x * NT) {
    const int64_t r = e / ldd, c = e - r * ldd;
    uint8_t q = 0;
    if (c < cols) {
      const float s = scale_of(amax[c / chunk]);
      q = to_e4m3(load_f(X, dt, r * ld + c) * inv_row(row_div, r) / s);
      if (r == 0 && col_scale) col_scale[c] = s;
    }
    dst[e] = q;
  }
}

// one wave per row: amax of x * colscale over the row, then the quantised row
__global__ __launch_bounds__(NT) void quant_rows_kernel(const FhQuantRows* __restrict__ tab, int n_tab,
                                                        int64_t total_rows) {
  const int lane = threadIdx.x & 63;
  const int64_t row_g = (int64_t)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  if (row_g >= total_rows) return;
  int lo = 0, hi = n_tab - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tab[mid].row_begin <= row_g) lo = mid; else hi = mid - 1;
  }
  const FhQuantRows q = tab[lo];
  const int64_t r = row_g - q.row_begin;
  const float* cs = q.colscale ? q.colscale + (r / q.cs_rows) * q.cs_stride : nullptr;
  float m = 0.f;
  for (int64_t c = lane; c < q.cols; c += 64) {
    float v = load_f(q.src, q.sdt, r * q.lds + c);
    if (cs) v *= cs[c];
    m = fmaxf(m, fabsf(v));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  const float s = scale_of(m);
  if (lane == 0) q.scale[r] = s;
  for (int64_t c = lane; c < q.ldd; c += 64) {
    uint8_t o = 0;
    if (c < q.cols) {
      float v = load_f(q.src, q.sdt, r * q.lds + c);
      if (cs) v *= cs[c];
      o = to_e4m3(v / s);
    }
    q.dst[r * q.ldd + c] = o;
  }
}

static inline unsigned grid_for(int64_t n) {
  int64_t g = (n + NT - 1) / NT;
  return (unsigned)(g < 1 ? 1 : (g > 65535 ? 65535 : g));
}

}  // namespace

extern "C" int fh_amax_rows(const void* X, int dt, int64_t ld, int64_t rows, int64_t cols, float* amax,
                            hipStream_t st) {
  FH_REQUIRE(dt == FH_F32 || dt == FH_BF16, "amax_rows: dtype %d (fp32 or bf16)", dt);
  FH_REQUIRE(cols > 0 && ld >= cols, "amax_rows: bad extents");
  if (rows <= 0) return FH_OK;
  const int64_t blocks = fh_cdiv(rows, NT / 64);
  FH_REQUIRE(blocks < ((int64_t)1 << 31), "amax_rows: too many rows");
  hipLaunchKernelGGL(amax_rows_kernel, dim3((unsigned)blocks), dim3(NT), 0, st, X, dt, ld, rows, cols, amax);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

extern "C" int fh_amax_col_chunks(const void* X, int dt, int64_t ld, int64_t rows, int64_t cols, int64_t chunk,
                                  const float* row_div, float* amax, hipStream_t st) {
  FH_REQUIRE(dt == FH_F32 || dt == FH_BF16, "amax: dtype %d (fp32 or bf16)", dt);
  FH_REQUIRE(chunk > 0 && cols > 0 && rows >= 0 && ld >= cols, "amax: bad extents");
  const int64_t nch = fh_cdiv(cols, chunk);
  FH_REQUIRE(nch <= 65535, "amax: %lld chunks (at most 65535)", (long long)nch);
  hipLaunchKernelGGL(zero_kernel, dim3(grid_for(nch)), dim3(NT), 0, st, amax, nch);
  if (rows > 0) {
    const int64_t rpb = 256;
    const int64_t rb = fh_cdiv(rows, rpb);
    FH_REQUIRE(rb < ((int64_t)1 << 31), "amax: too many row blocks");
    hipLaunchKernelGGL(amax_kernel, dim3((unsigned)rb, (unsigned)nch), dim3(NT), 0, st, X, dt, ld, rows, cols,
                       chunk, rpb, row_div, amax);
  }
  FH_CHECK_LAUNCH();
  return FH_OK;
}

extern "C" int fh_quant_col_chunks(const void* X, int dt, int64_t ld, int64_t rows, int64_t cols, int64_t chunk,
                                   const float* row_div, const float* amax, uint8_t* dst, int64_t ldd,
                                   float* col_scale, hipStream_t st) {
  FH_REQUIRE(dt == FH_F32 || dt == FH_BF16, "quant: dtype %d (fp32 or bf16)", dt);
  FH_REQUIRE(chunk > 0 && ld >= cols && ldd >= cols, "quant: bad extents");
  if (rows <= 0) return FH_OK;
  hipLaunchKernelGGL(quant_cols_kernel, dim3(grid_for(rows * ldd)), dim3(NT), 0, st, X, dt, ld, rows, cols, chunk,
                     row_div, amax, dst, ldd, col_scale);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

extern "C" int64_t fh_quant_rows_prepare(FhQuantRows* tab_host, int n) {
  int64_t total = 0;
  for (int i = 0; i < n; ++i) {
    if (tab_host[i].cs_rows <= 0) tab_host[i].cs_rows = tab_host[i].rows > 0 ? tab_host[i].rows : 1;
    tab_host[i].row_begin = total;
    total += tab_host[i].rows;
  }
  return total;
}

extern "C" int fh_quant_rows(const FhQuantRows* tab_dev, int n_tab, int64_t total_rows, hipStream_t st) {
  if (total_rows <= 0) return FH_OK;
  FH_REQUIRE(tab_dev != nullptr && n_tab > 0, "quant_rows: empty table");
  const int64_t blocks = fh_cdiv(total_rows, NT / 64);
  FH_REQUIRE(blocks < ((int64_t)1 << 31), "quant_rows: too many rows");
  hipLaunchKernelGGL(quant_rows_kernel, dim3((unsigned)blocks), dim3(NT), 0, st, tab_dev, n_tab, total_rows);
  FH_CHECK_LAUNCH();
  return FH_OK;
}
