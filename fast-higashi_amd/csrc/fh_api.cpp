// C-ABI core: error reporting and library identity.  Every compute entry point of
// libfh_hip.so returns an int status (0 = OK); fh_last_error() gives the message of
// the last failure on the calling thread.
#include <hip/hip_runtime.h>

#include "fh_api.h"
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

static thread_local char g_err[1024] = {0};

void fh_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" {

const char* fh_last_error(void) { return g_err; }

void fh_clear_error(void) { g_err[0] = 0; }

int fh_version(void) { return 100; }

// Name of the offload target the kernels were built for (compile-time constant).
const char* fh_target(void) { return "gfx950"; }

// Synchronise a stream and report any asynchronous kernel error.
int fh_stream_sync(hipStream_t stream) {
  hipError_t e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    fh_set_error("stream sync: %s", hipGetErrorString(e));
    return -2;
  }
  return 0;
}

}  // extern "C"
