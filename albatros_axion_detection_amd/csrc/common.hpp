// Shared helpers for the ALBATROS MI355X kernels: error reporting, complex
// arithmetic, launch checks and the per-device twiddle-table registry.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>
#include <string.h>

#define ALB_OK 0
#define ALB_EINVAL 1
#define ALB_EHIP 2
#define ALB_ENOTSUP 3

namespace alb {

// Last error message of the calling host thread (read through alb_last_error()).
void set_error(const char* fmt, ...);
const char* last_error();

inline int hip_fail(hipError_t e, const char* what) {
  set_error("%s: %s", what, hipGetErrorString(e));
  return ALB_EHIP;
}

#define ALB_REQUIRE(cond, ...)            \
  do {                                    \
    if (!(cond)) {                        \
      ::alb::set_error(__VA_ARGS__);      \
      return ALB_EINVAL;                  \
    }                                     \
  } while (0)

#define ALB_HIP(call)                                       \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return ::alb::hip_fail(e_, #call); \
  } while (0)

#define ALB_LAUNCH_CHECK(name)                                   \
  do {                                                           \
    hipError_t e_ = hipGetLastError();                           \
    if (e_ != hipSuccess) return ::alb::hip_fail(e_, name);      \
  } while (0)

static inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

static inline bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }
static inline int ilog2(int64_t n) {
  int l = 0;
  while ((int64_t(1) << l) < n) ++l;
  return l;
}

// ---------------------------------------------------------------- complex
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// -------------------------------------------------------- twiddle tables
// Two-level table of W_N^m = exp(-2 pi i m / N):  W_N^m = A[m >> lb] * B[m & (2^lb - 1)]
// with A[i] = W_N^(i << lb) (N >> lb entries) and B[i] = W_N^i (2^lb entries),
// both rounded from float64.  One table per (device, N), built on first use
// (or ahead of time by alb_prepare_transform / alb_prepare_repfb) in a fixed
// device array; a first use on a stream that is being captured into a graph is
// refused, since building synchronises.
struct Twiddle {
  const float2* A;
  const float2* B;
  int lb;     // bits of the low table
  int nA;     // entries of A
  int nB;     // entries of B
};
int get_twiddle(int64_t N, Twiddle* out, hipStream_t s);

// Re-PFB for osamp 128..65536 (repfb_large.hip): stage 1 into the Z[b][q][p]
// workspace, band-limited stage 2 from it (argument meaning of alb_repfb_stage1
// and alb_repfb_stage2_inputs, a_stride in complex elements).
int repfb_large_stage1(const float* ts, const float* win, int ntap, int64_t nblock, int64_t R1, float2* Z,
                       hipStream_t s);
int repfb_large_stage2(const float2* Z, int64_t a_stride, int64_t ninput, int64_t nblock, int64_t R1, int64_t c_lo,
                       int64_t c_hi, float2* out, int64_t sm, int64_t sb, int64_t sc, int64_t bpj, int64_t sj,
                       hipStream_t s);

// Tuning knob set through alb_tune_set (0 when unset).
int64_t tune_get(const char* key);

}  // namespace alb
