// Library runtime: per-thread error string, device twiddle-table registry and
// the small C-ABI housekeeping entry points (version, last error, sync).
#include <mutex>
#include <map>
#include <utility>
#include "common.hpp"

namespace alb {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char* last_error() { return g_err; }

// Build W_N^m = exp(-2 pi i m / N) for m = i * step in float64, rounded to float32.
__global__ void twiddle_kernel(float2* out, int n, int64_t N, int64_t step) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t m = (int64_t(i) * step) % N;
  // exact argument reduction: angle = -2 pi m / N = -pi * (2m / N)
  double x = -2.0 * double(m) / double(N);
  double s, c;
  sincospi(x, &s, &c);
  out[i] = make_float2(float(c), float(s));
}

struct TwKey {
  int dev;
  int64_t N;
  bool operator<(const TwKey& o) const { return dev < o.dev || (dev == o.dev && N < o.N); }
};
static std::mutex g_tw_mu;
static std::map<TwKey, Twiddle> g_tw;

int get_twiddle(int64_t N, Twiddle* out) {
  int dev = 0;
  ALB_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto it = g_tw.find(TwKey{dev, N});
  if (it != g_tw.end()) { *out = it->second; return ALB_OK; }
  int lg = ilog2(N);
  ALB_REQUIRE((int64_t(1) << lg) == N && lg >= 1 && lg <= 30, "twiddle size %lld is not a power of two",
              (long long)N);
  int lb = (lg + 1) / 2;
  int nB = 1 << lb;
  int nA = int(N >> lb);
  if (nA < 1) nA = 1;
  float2* buf = nullptr;
  ALB_HIP(hipMalloc(&buf, sizeof(float2) * size_t(nA + nB)));
  hipLaunchKernelGGL(twiddle_kernel, dim3((nA + 255) / 256), dim3(256), 0, 0, buf, nA, N, int64_t(nB));
  hipLaunchKernelGGL(twiddle_kernel, dim3((nB + 255) / 256), dim3(256), 0, 0, buf + nA, nB, N, int64_t(1));
  ALB_LAUNCH_CHECK("twiddle_kernel");
  ALB_HIP(hipDeviceSynchronize());
  Twiddle t{buf, buf + nA, lb, nA, nB};
  g_tw[TwKey{dev, N}] = t;
  *out = t;
  return ALB_OK;
}

}  // namespace alb

extern "C" {

const char* alb_last_error(void) { return alb::last_error(); }

int alb_version(void) { return 100; }

int alb_stream_synchronize(void* stream) {
  ALB_HIP(hipStreamSynchronize(alb::as_stream(stream)));
  return ALB_OK;
}

}  // extern "C"
