// Allocation-free form of the re-PFB (alb_repfb_noalloc): the stage-1
// workspace lives in a fixed device array that the library never allocates or
// frees, for callers that cannot call hipMalloc (hipGraph capture, serving
// loops, harnesses that forbid it).  Each device holds its own copy of the
// array until the process ends.  Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device: the array's name is a
// host-side stand-in, not a device address.
#include <mutex>
#include "common.hpp"

extern "C" {
int alb_repfb(const float* ts, const float* win, int ntap, int64_t nblock, int64_t osamp, void* A, int64_t c_lo,
              int64_t c_hi, void* out, int64_t sb, int64_t sc, int64_t bpj, int64_t sj, void* stream);
int64_t alb_repfb_workspace_bytes(int64_t nblock, int64_t osamp);
int alb_prepare_repfb(int64_t osamp);
}

namespace alb {

constexpr int64_t kStaticWorkspaceBytes = int64_t(64) << 20;
static __device__ __align__(256) float2 scratch[kStaticWorkspaceBytes / int64_t(sizeof(float2))];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

// Device address of the current device's copy of the array, looked up once
// per device (with that device current).
static int static_workspace(void** out) {
  int dev = 0;
  ALB_HIP(hipGetDevice(&dev));
  ALB_REQUIRE(dev >= 0 && dev < kMaxDevices, "alb_repfb_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    ALB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ALB_REQUIRE(p != nullptr, "alb_repfb_noalloc: no device address for the static workspace on device %d", dev);
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return ALB_OK;
}

}  // namespace alb

extern "C" {

// alb_repfb with its stage-1 workspace in the static array; nblock is limited
// to what kStaticWorkspaceBytes holds (alb_repfb_workspace_bytes(nblock, osamp)
// <= 64 MiB: 64 blocks at osamp 64).  Stage 1 writes every workspace slot stage
// 2 reads, so nothing needs clearing between calls; calls that overlap on one
// device share the array and race.
int alb_repfb_noalloc(const float* ts, const float* win, int ntap, int64_t nblock, int64_t osamp, int64_t c_lo,
                      int64_t c_hi, void* out, int64_t sb, int64_t sc, int64_t bpj, int64_t sj, void* stream) {
  ALB_REQUIRE(ts && win && out, "alb_repfb_noalloc: null pointer");
  ALB_REQUIRE(nblock >= 0 && osamp >= 1, "alb_repfb_noalloc: bad sizes");
  // osamp above 64 takes transform tables that are built (allocated) on first use
  ALB_REQUIRE(osamp <= 64, "alb_repfb_noalloc: osamp %lld above 64 is not allocation-free (use alb_repfb)",
              (long long)osamp);
  const int64_t need = alb_repfb_workspace_bytes(nblock, osamp);
  ALB_REQUIRE(need <= alb::kStaticWorkspaceBytes,
              "alb_repfb_noalloc: %lld blocks at osamp %lld need %lld bytes of workspace, more than the %lld-byte "
              "static workspace (use alb_repfb with a caller workspace)",
              (long long)nblock, (long long)osamp, (long long)need, (long long)alb::kStaticWorkspaceBytes);
  if (nblock == 0) return ALB_OK;
  // transform tables: built on a device's first call from the library's fixed
  // table array (no allocation; alb_prepare_repfb builds them ahead of time)
  const int rp = alb_prepare_repfb(osamp);
  if (rp != ALB_OK) return rp;
  void* A = nullptr;
  const int rc = alb::static_workspace(&A);
  if (rc != ALB_OK) return rc;
  return alb_repfb(ts, win, ntap, nblock, osamp, A, c_lo, c_hi, out, sb, sc, bpj, sj, stream);
}

// Size of the static workspace in bytes.
int64_t alb_repfb_noalloc_max_bytes(void) { return alb::kStaticWorkspaceBytes; }

}  // extern "C"
