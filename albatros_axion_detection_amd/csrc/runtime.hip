// Library runtime: per-thread error string, device twiddle-table registry and
// the small C-ABI housekeeping entry points (version, last error, sync).
// Transform tables live in a fixed device array (kTablePool float2 per
// device): building one launches a kernel and synchronises but allocates
// nothing, so the allocation-free entry points stay so on their first call.
// Tables beyond the pool (only the largest transforms) fall back to hipMalloc.
#include <mutex>
#include <map>
#include <string>
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

constexpr int64_t kTablePool = int64_t(1) << 18;   // float2 per device (2 MiB)
constexpr int kPoolDevices = 64;
static __device__ __align__(256) float2 table_pool[kTablePool];
static void* g_pool_ptr[kPoolDevices];
static int64_t g_pool_used[kPoolDevices];

// n float2 of table storage on device dev (caller holds g_tw_mu): from the
// pool while it lasts (address of this device's copy from
// hipGetSymbolAddress, looked up once with the device current), else hipMalloc.
static int table_alloc(int dev, int64_t n, float2** out) {
  n = (n + 31) / 32 * 32;   // 256-byte aligned pieces
  if (dev >= 0 && dev < kPoolDevices) {
    if (g_pool_ptr[dev] == nullptr) {
      void* p = nullptr;
      ALB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(table_pool)));
      ALB_REQUIRE(p != nullptr, "no device address for the transform-table pool on device %d", dev);
      g_pool_ptr[dev] = p;
    }
    if (g_pool_used[dev] + n <= kTablePool) {
      *out = reinterpret_cast<float2*>(g_pool_ptr[dev]) + g_pool_used[dev];
      g_pool_used[dev] += n;
      return ALB_OK;
    }
  }
  ALB_HIP(hipMalloc(out, sizeof(float2) * size_t(n)));
  return ALB_OK;
}

int get_twiddle(int64_t N, Twiddle* out, hipStream_t s) {
  int dev = 0;
  ALB_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto it = g_tw.find(TwKey{dev, N});
  if (it != g_tw.end()) { *out = it->second; return ALB_OK; }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  ALB_HIP(hipStreamIsCapturing(s, &cap));
  ALB_REQUIRE(cap == hipStreamCaptureStatusNone,
              "the %lld-point twiddle table is not built yet and the stream is being captured (building it "
              "allocates and synchronises): call alb_prepare_transform(%lld) before the capture",
              (long long)N, (long long)N);
  int lg = ilog2(N);
  ALB_REQUIRE((int64_t(1) << lg) == N && lg >= 1 && lg <= 30, "twiddle size %lld is not a power of two",
              (long long)N);
  int lb = (lg + 1) / 2;
  int nB = 1 << lb;
  int nA = int(N >> lb);
  if (nA < 1) nA = 1;
  float2* buf = nullptr;
  const int rc = table_alloc(dev, nA + nB, &buf);
  if (rc != ALB_OK) return rc;
  hipLaunchKernelGGL(twiddle_kernel, dim3((nA + 255) / 256), dim3(256), 0, 0, buf, nA, N, int64_t(nB));
  hipLaunchKernelGGL(twiddle_kernel, dim3((nB + 255) / 256), dim3(256), 0, 0, buf + nA, nB, N, int64_t(1));
  ALB_LAUNCH_CHECK("twiddle_kernel");
  ALB_HIP(hipDeviceSynchronize());
  Twiddle t{buf, buf + nA, lb, nA, nB};
  g_tw[TwKey{dev, N}] = t;
  *out = t;
  return ALB_OK;
}


// Tuning knobs (0 = the built-in choice).
static std::mutex g_tune_mu;
static std::map<std::string, int64_t> g_tune;

int64_t tune_get(const char* key) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(key);
  return it == g_tune.end() ? 0 : it->second;
}

}  // namespace alb

extern "C" {

int alb_tune_set(const char* key, int64_t value) {
  ALB_REQUIRE(key, "alb_tune_set: null key");
  const std::string k(key);
  if (k == "ipfb_cols_width") {
    ALB_REQUIRE(value == 0 || value == 2 || value == 4 || value == 8,
                "alb_tune_set: ipfb_cols_width must be 0, 2, 4 or 8");
  } else if (k == "ipfb_cols_group") {
    ALB_REQUIRE(value == 0 || value == 2 || value == 4 || value == 8 || value == 16 || value == 32 || value == 64,
                "alb_tune_set: ipfb_cols_group must be 0, 2, 4, 8, 16, 32 or 64");
  } else if (k == "ipfb_rows_group") {
    ALB_REQUIRE(value == 0 || value == 2 || value == 4, "alb_tune_set: ipfb_rows_group must be 0, 2 or 4");
  } else if (k == "xcorr_int_method") {
    ALB_REQUIRE(value >= 0 && value <= 2, "alb_tune_set: xcorr_int_method must be 0, 1 or 2");
  } else if (k == "xcorr1_form") {
    ALB_REQUIRE(value >= 0 && value <= 2, "alb_tune_set: xcorr1_form must be 0, 1 or 2");
  } else if (k == "xcorr4_segment") {
    ALB_REQUIRE(value == 0 || value == 64 || value == 128, "alb_tune_set: xcorr4_segment must be 0, 64 or 128");
  } else if (k == "xcorr4_workgroups") {
    ALB_REQUIRE(value >= 0 && value <= (int64_t(1) << 20), "alb_tune_set: xcorr4_workgroups must be 0..2^20");
  } else {
    ALB_REQUIRE(false, "alb_tune_set: unknown key '%s'", key);
  }
  std::lock_guard<std::mutex> lk(alb::g_tune_mu);
  alb::g_tune[k] = value;
  return ALB_OK;
}

int64_t alb_tune_get(const char* key) { return key ? alb::tune_get(key) : 0; }

const char* alb_last_error(void) { return alb::last_error(); }

int alb_version(void) { return 101; }

// Build the twiddle table of an N-point transform on the current device now
// (first uses otherwise build it lazily, which a graph capture cannot do).
int alb_prepare_transform(int64_t N) {
  alb::Twiddle tw;
  return alb::get_twiddle(N, &tw, nullptr);
}

int alb_stream_synchronize(void* stream) {
  ALB_HIP(hipStreamSynchronize(alb::as_stream(stream)));
  return ALB_OK;
}

}  // extern "C"
