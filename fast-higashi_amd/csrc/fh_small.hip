// Element-wise and reduction kernels around the sweep's GEMMs.
// All reductions are deterministic (fixed partition, fixed order): repeated runs on
// the same inputs give bit-identical results.
#include "fh_common.h"

namespace {
constexpr int NT = 256;

__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) t += red[w];
  return t;
}

template <typename T> __device__ __forceinline__ double ld_as_f64(const void* p, int64_t i) {
  return (double)((const T*)p)[i];
}
__device__ __forceinline__ double load_any(const void* p, int dt, int64_t i) {
  switch (dt) {
    case FH_F32: return ((const float*)p)[i];
    case FH_F64: return ((const double*)p)[i];
    case FH_BF16: return (double)(float)((const __bf16*)p)[i];
    default: return 0.;
  }
}
__device__ __forceinline__ void store_any(void* p, int dt, int64_t i, double v) {
  switch (dt) {
    case FH_F32: ((float*)p)[i] = (float)v; break;
    case FH_F64: ((double*)p)[i] = v; break;
    case FH_BF16: ((uint16_t*)p)[i] = fh_f32_to_bf16_bits((float)v); break;
    default: break;
  }
}

__global__ void fill_kernel(void* p, int esize, int64_t n, uint64_t pattern) {
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    if (esize == 4) ((uint32_t*)p)[i] = (uint32_t)pattern;
    else if (esize == 2) ((uint16_t*)p)[i] = (uint16_t)pattern;
    else if (esize == 8) ((uint64_t*)p)[i] = pattern;
    else ((uint8_t*)p)[i] = (uint8_t)pattern;
  }
}

// dst[r*dld + c] = src[r*srs + c*scs]   (dtype conversion included)
__global__ void copy2d_kernel(const void* src, int sdt, int64_t srs, int64_t scs, void* dst, int ddt,
                              int64_t dld, int64_t rows, int64_t cols, double alpha) {
  const int64_t n = rows * cols;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    int64_t r = i / cols, c = i % cols;
    store_any(dst, ddt, r * dld + c, alpha * load_any(src, sdt, r * srs + c * scs));
  }
}

// Row scaling of per-bin blocks:  Y[r][j] *= A[bin0 + r / w][j] * rowf[r] * colf[j]   (j < s; A,
// rowf, colf nullable), one tile per blockIdx.y.  One wave per row (the bin index once per row);
// with ldy, lda multiples of 4 and 16-byte aligned bases each lane scales four columns with
// 16-byte accesses, else one column at a time.
__global__ __launch_bounds__(NT) void scale_rows_kernel(const FhRowScale* __restrict__ tab) {
  const FhRowScale t = tab[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  const int64_t nw = (int64_t)gridDim.x * (NT / 64);
  const bool vec = ((t.ldy | t.lda) & 3) == 0 &&
                   ((((uintptr_t)t.Y) | ((uintptr_t)t.A) | ((uintptr_t)t.colf)) & 15) == 0;
  for (int64_t r = wave; r < t.nrows; r += nw) {
    float* y = t.Y + r * t.ldy;
    const float* a = t.A ? t.A + (r / t.w) * t.lda : nullptr;
    const float rf = t.rowf ? t.rowf[r] : 1.f;
    const float* cf = t.colf;
    if (vec) {
      for (int64_t j = 4 * lane; j < t.s; j += 256) {
        if (j + 4 <= t.s) {
          floatx4 v = *(floatx4*)(y + j) * rf;
          if (a) v *= *(const floatx4*)(a + j);
          if (cf) v *= *(const floatx4*)(cf + j);
          *(floatx4*)(y + j) = v;
        } else {
          for (int64_t e = j; e < t.s; ++e) y[e] *= (a ? a[e] : 1.f) * rf * (cf ? cf[e] : 1.f);
        }
      }
    } else {
      for (int64_t j = lane; j < t.s; j += 64) y[j] *= (a ? a[j] : 1.f) * rf * (cf ? cf[j] : 1.f);
    }
  }
}

// Khatri-Rao rows:  L[r][i*sp + j] = A[i][r] * B[j][r] (* colf[i*sp + j])  (j < s; zero for s <= j < sp)
__global__ void krp_kernel(const float* A, int64_t lda, const float* B, int64_t ldb, int64_t n, int64_t s,
                           int64_t sp, void* L, int ldt, const float* colf) {
  const int64_t cols = n * sp, tot = s * cols;
  for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < tot; e += (int64_t)gridDim.x * NT) {
    int64_t r = e / cols, q = e % cols, i = q / sp, j = q % sp;
    double v = (j < s) ? (double)A[i * lda + r] * (double)B[j * ldb + r] : 0.;
    if (colf) v *= (double)colf[q];
    store_any(L, ldt, e, (double)(float)v);
  }
}

// W1 rows (i*sp + j), columns r (ld = ldw).
// mode 0:  out[i][r] = sum_j W1[i*sp+j][r] * B[j][r]      (i < n)
// mode 1:  out[j][r] = sum_i W1[i*sp+j][r] * A[i][r]      (j < s)

// out = G1 o G2 (o G3)  s x s fp64

// balance_norm (parafac_integrative.py:19-26) on one chromosome, one block per column r:
// every factor column is unit-normalised, D's column also takes the product of the norms.

// out[0] = sum_{ij} G1 o G2 o G3  (= ||[[A,B,D]]||^2 from the three Grams)

// partial sums of squares over rows x cols (ld) blocks: part[blockIdx.x]
__global__ void sumsq_partial_kernel(const void* X, int dt, int64_t rows, int64_t cols, int64_t ld,
                                     int64_t rows_per_block, double* part) {
  __shared__ double red[NT / 64];
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < rows ? r0 + rows_per_block : rows;
  double p = 0.;
  for (int64_t r = r0; r < r1; ++r)
    for (int64_t c = threadIdx.x; c < cols; c += NT) {
      double v = load_any(X, dt, r * ld + c);
      p += v * v;
    }
  double t = block_sum(p, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// out[g] (+)= sum over segment g of vals (segments given by offsets[g]..offsets[g+1]), fixed order
__global__ void segment_sum_kernel(const double* vals, const int64_t* offsets, int n_seg, double* out,
                                   int accumulate) {
  __shared__ double red[NT / 64];
  const int g = blockIdx.x;
  if (g >= n_seg) return;
  double p = 0.;
  for (int64_t i = offsets[g] + threadIdx.x; i < offsets[g + 1]; i += NT) p += vals[i];
  double t = block_sum(p, red);
  if (threadIdx.x == 0) out[g] = accumulate ? out[g] + t : t;
}

// <X, Y> over rows x cols with independent strides (fp32/fp64 inputs), written to out[0]

// Sweep error bookkeeping (parafac2_intergrative.py:667-668,710):
//   err_U[c] = tnorm[c] + core[c] - 2 xU[c],  err_V = sum tnorm + sum core - 2 xV
//   hist: [rec_error, err_V, xV, then per chrom: rec_block[c], xU[c], core[c]]
__global__ void sweep_errors_kernel(const double* tnorm, const double* core, const double* xU, const double* xV,
                                    int K, double* hist) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double st = 0., sc = 0.;
  for (int c = 0; c < K; ++c) { st += tnorm[c]; sc += core[c]; }
  double errV = st + sc - 2. * xV[0];
  hist[0] = sqrt(errV) / sqrt(st);
  hist[1] = errV;
  hist[2] = xV[0];
  for (int c = 0; c < K; ++c) {
    double eu = tnorm[c] + core[c] - 2. * xU[c];
    hist[3 + 3 * c] = sqrt(eu) / sqrt(tnorm[c]);
    hist[4 + 3 * c] = xU[c];
    hist[5 + 3 * c] = core[c];
  }
}

// out[e] = sum_b in[b * stride + e]  (fixed order), fp32 or fp64
template <typename T>
__global__ void batch_sum_kernel(const T* in, int64_t nb, int64_t stride, int64_t count, T* out) {
  for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < count; e += (int64_t)gridDim.x * NT) {
    double acc = 0.;
    for (int64_t b = 0; b < nb; ++b) acc += (double)in[b * stride + e];
    out[e] = (T)acc;
  }
}

// X[b][i][i] = v (set) or += v, i < n
__global__ void diag_kernel(float* X, int64_t nb, int64_t stride, int64_t ld, int64_t n, float v, int add) {
  for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < nb * n; e += (int64_t)gridDim.x * NT) {
    int64_t b = e / n, i = e % n;
    float* p = X + b * stride + i * ld + i;
    *p = add ? *p + v : v;
  }
}

// CP-ALS inner-iteration test (parafac_integrative.py:80-108) from closed forms:
//   loss = ||T||^2 - 2 <M_D, D> + ||[[A,B,D]]||^2 ; wdiff = (prev - loss) / prev; stop if wdiff < 1e-5
// state: [0] = previous loss, [1] = have previous; flag: 1 once stopped.


// Grouped batched transpose: dst[b][c][r] = src[b][r][c] (r < rows, c < cols) with fp32/bf16
// conversion; one 64 x 64 tile per workgroup, staged through LDS so both sides are coalesced.
__global__ __launch_bounds__(NT) void transpose_grouped_kernel(const FhTransDesc* __restrict__ descs, int n) {
  __shared__ float tile[64][65];
  const int64_t t = blockIdx.x;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (descs[mid].tile_begin <= t) lo = mid; else hi = mid - 1;
  }
  const FhTransDesc d = descs[lo];
  int64_t local = t - d.tile_begin;
  const int64_t per = (int64_t)d.tiles_r * d.tiles_c;
  const int64_t b = local / per;
  local -= b * per;
  const int64_t r0 = (local / d.tiles_c) * 64, c0 = (local % d.tiles_c) * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int es = d.sdt == FH_BF16 ? 2 : (d.sdt == FH_FP8 ? 1 : 4), ed = d.ddt == FH_BF16 ? 2 : (d.ddt == FH_FP8 ? 1 : 4);
  const char* src = (const char*)d.src + b * d.sbs * es;
  char* dst = (char*)d.dst + b * d.sbd * ed;
  for (int i = ty; i < 64; i += NT / 64) {
    const int64_t r = r0 + i, c = c0 + tx;
    float v = 0.f;
    if (r < d.rows && c < d.cols) {
      const int64_t o = r * d.lds + c;
      v = es == 2 ? (float)((const __bf16*)src)[o] : (es == 1 ? (float)((const uint8_t*)src)[o] : ((const float*)src)[o]);
    }
    tile[i][tx] = v;
  }
  __syncthreads();
  for (int i = ty; i < 64; i += NT / 64) {
    const int64_t c = c0 + i, r = r0 + tx;
    if (r < d.rows && c < d.cols) {
      const int64_t o = c * d.ldd + r;
      const float v = tile[tx][i];
      if (ed == 2) ((uint16_t*)dst)[o] = fh_f32_to_bf16_bits(v);
      else if (ed == 1) ((uint8_t*)dst)[o] = (uint8_t)v;     // FP8 bytes (FP8 source only)
      else ((float*)dst)[o] = v;
    }
  }
}
}  // namespace

extern "C" {

static inline unsigned grid_for(int64_t n) {
  int64_t g = (n + NT - 1) / NT;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

int fh_fill(void* p, int esize, int64_t n, uint64_t pattern, hipStream_t st) {
  if (n <= 0) return FH_OK;
  FH_REQUIRE(esize == 1 || esize == 2 || esize == 4 || esize == 8, "fill: bad element size");
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(NT), 0, st, p, esize, n, pattern);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

int fh_copy2d(const void* src, int sdt, int64_t srs, int64_t scs, void* dst, int ddt, int64_t dld,
              int64_t rows, int64_t cols, double alpha, hipStream_t st) {
  if (rows <= 0 || cols <= 0) return FH_OK;
  hipLaunchKernelGGL(copy2d_kernel, dim3(grid_for(rows * cols)), dim3(NT), 0, st, src, sdt, srs, scs, dst, ddt,
                     dld, rows, cols, alpha);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

int fh_scale_rows(const FhRowScale* tab_dev, int n_tiles, int64_t max_elems, hipStream_t st) {
  if (n_tiles <= 0) return FH_OK;
  // max_elems bounds rows x s of a tile; one wave per row, NT / 64 rows per workgroup
  int64_t gx64 = (max_elems + 4 * NT - 1) / (4 * NT);
  unsigned gx = (unsigned)(gx64 < 1 ? 1 : gx64 > 1024 ? 1024 : gx64);
  hipLaunchKernelGGL(scale_rows_kernel, dim3(gx, n_tiles), dim3(NT), 0, st, tab_dev);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

int fh_krp(const float* A, int64_t lda, const float* B, int64_t ldb, int64_t n, int64_t s, int64_t sp,
           void* L, int l_dtype, const float* colf, hipStream_t st) {
  hipLaunchKernelGGL(krp_kernel, dim3(grid_for(s * n * sp)), dim3(NT), 0, st, A, lda, B, ldb, n, s, sp, L, l_dtype,
                     colf);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

// partial sums of squares; part must hold ceil(rows / rows_per_block) doubles
int fh_sumsq_partial(const void* X, int dt, int64_t rows, int64_t cols, int64_t ld, int64_t rows_per_block,
                     double* part, hipStream_t st) {
  if (rows <= 0) return FH_OK;
  int64_t nb = (rows + rows_per_block - 1) / rows_per_block;
  FH_REQUIRE(nb < (1ll << 31), "sumsq: too many blocks");
  hipLaunchKernelGGL(sumsq_partial_kernel, dim3((unsigned)nb), dim3(NT), 0, st, X, dt, rows, cols, ld,
                     rows_per_block, part);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

int fh_segment_sum(const double* vals, const int64_t* offsets_dev, int n_seg, double* out, int accumulate,
                   hipStream_t st) {
  if (n_seg <= 0) return FH_OK;
  hipLaunchKernelGGL(segment_sum_kernel, dim3(n_seg), dim3(NT), 0, st, vals, offsets_dev, n_seg, out, accumulate);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

int fh_sweep_errors(const double* tnorm, const double* core, const double* xU, const double* xV, int K,
                    double* hist, hipStream_t st) {
  hipLaunchKernelGGL(sweep_errors_kernel, dim3(1), dim3(64), 0, st, tnorm, core, xU, xV, K, hist);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

int fh_batch_sum(const void* in, int dt, int64_t nb, int64_t stride, int64_t count, void* out, hipStream_t st) {
  if (count <= 0) return FH_OK;
  if (dt == FH_F64)
    hipLaunchKernelGGL((batch_sum_kernel<double>), dim3(grid_for(count)), dim3(NT), 0, st, (const double*)in, nb,
                       stride, count, (double*)out);
  else
    hipLaunchKernelGGL((batch_sum_kernel<float>), dim3(grid_for(count)), dim3(NT), 0, st, (const float*)in, nb,
                       stride, count, (float*)out);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

int fh_diag(float* X, int64_t nb, int64_t stride, int64_t ld, int64_t n, float v, int add, hipStream_t st) {
  if (nb * n <= 0) return FH_OK;
  hipLaunchKernelGGL(diag_kernel, dim3(grid_for(nb * n)), dim3(NT), 0, st, X, nb, stride, ld, n, v, add);
  FH_CHECK_LAUNCH();
  return FH_OK;
}


int fh_transpose_grouped(const FhTransDesc* descs_dev, int n_desc, int64_t total_tiles, hipStream_t st) {
  if (total_tiles <= 0) return FH_OK;
  FH_REQUIRE(total_tiles < (1ll << 31), "transpose_grouped: too many tiles");
  hipLaunchKernelGGL(transpose_grouped_kernel, dim3((unsigned)total_tiles), dim3(NT), 0, st, descs_dev, n_desc);
  FH_CHECK_LAUNCH();
  return FH_OK;
}

}  // extern "C"
