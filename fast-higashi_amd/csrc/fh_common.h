// Shared definitions for the Fast-Higashi gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "fh_api.h"

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

#define FH_WAVE 64

// thread-local last error message (fh_api.cpp)
void fh_set_error(const char* fmt, ...);

#define FH_CHECK_HIP(expr)                                                  \
  do {                                                                      \
    hipError_t _e = (expr);                                                 \
    if (_e != hipSuccess) {                                                 \
      fh_set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,               \
                   hipGetErrorString(_e));                                  \
      return FH_ERR_HIP;                                                    \
    }                                                                       \
  } while (0)

#define FH_CHECK_LAUNCH()                                                   \
  do {                                                                      \
    hipError_t _e = hipGetLastError();                                      \
    if (_e != hipSuccess) {                                                 \
      fh_set_error("%s:%d kernel launch: %s", __FILE__, __LINE__,           \
                   hipGetErrorString(_e));                                  \
      return FH_ERR_HIP;                                                    \
    }                                                                       \
  } while (0)

#define FH_REQUIRE(cond, ...)                                               \
  do {                                                                      \
    if (!(cond)) {                                                          \
      fh_set_error(__VA_ARGS__);                                            \
      return FH_ERR_ARG;                                                    \
    }                                                                       \
  } while (0)

static inline int64_t fh_cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

__device__ __forceinline__ float fh_bf16_to_f32(__bf16 v) { return (float)v; }

// Round-to-nearest-even float -> bf16 (bit manipulation; no NaN inputs expected).
__device__ __forceinline__ uint16_t fh_f32_to_bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}
