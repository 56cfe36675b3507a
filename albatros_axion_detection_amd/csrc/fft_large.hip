// Batched 1-D complex FFTs of power-of-two length n = 2^15 .. 2^24 (pycufft.py:92-131
// _fft / ifft, pycufft.cu:122 cufft_c2c, for rows longer than the 16384 points of the
// single-workgroup row FFTs), and the FFT lag correlation of the ORBCOMM delay search
// (src/utils/orbcomm_utils_gpu.py:39-52 coarse_xcorr) built on them.
//
// Four-step form, n = P1 P2, element m = P2 m1 + m2 of a row:
//   pass 1 (columns): T[k1][m2] = W_n^(m2 k1) sum_m1 x[P2 m1 + m2] W_P1^(m1 k1)
//   pass 2 (rows):    X[k1 + P1 k2] = sum_m2 T[k1][m2] W_P2^(m2 k2)
// Pass 1 takes a strip of CP consecutive columns m2 (runs of 8 CP bytes per row),
// pass 2 CR consecutive rows k1 and writes runs of CR consecutive outputs; both hold
// their sequences in LDS (at most 16384 points per workgroup).  The inverse
// conjugates on load and on store and scales by 1/n (pycufft.py:107).
#include "fft.hpp"

namespace alb {

constexpr int kFT = 512;                  // threads per workgroup
constexpr int kFMax = 16384;              // points per workgroup

template <int P1, int CP>
__global__ __launch_bounds__(kFT) void fft_large_cols_kernel(const float2* __restrict__ x, float2* __restrict__ T,
                                                             int64_t n, int P2, int inverse, Twiddle twp, Twiddle twn) {
  constexpr int NP = Padded<P1>::value;
  extern __shared__ float2 smem[];
  float2* buf = smem;                    // [CP][NP]
  float2* TA = buf + CP * NP;
  float2* TB = TA + twp.nA;
  const int tid = threadIdx.x;
  load_twiddle(twp, TA, TB, tid, kFT);
  const int64_t b = blockIdx.y;
  const int m20 = blockIdx.x * CP;
  const float2* xb = x + b * n + m20;
  for (int e = tid; e < P1 * CP; e += kFT) {
    const int m1 = e / CP, c = e % CP;
    float2 v = xb[int64_t(m1) * P2 + c];
    if (inverse) v.y = -v.y;
    buf[c * NP + pad16(m1)] = v;
  }
  __syncthreads();
  fft_lds<P1, CP, kFT>(buf, TA, TB, twp.lb, 0, tid);
  float2* tb = T + b * n + m20;
  for (int e = tid; e < P1 * CP; e += kFT) {
    const int k1 = e / CP, c = e % CP;
    const float2 v = buf[c * NP + pad16(k1)];
    const int64_t m = (int64_t(m20 + c) * k1) & (n - 1);
    tb[int64_t(k1) * P2 + c] = cmul(v, tw_lookup(twn.A, twn.B, twn.lb, int(m)));
  }
}

template <int P2, int CR>
__global__ __launch_bounds__(kFT) void fft_large_rows_kernel(const float2* __restrict__ T, float2* __restrict__ out,
                                                             int64_t n, int P1, int inverse, float scale,
                                                             Twiddle twp) {
  constexpr int NP = Padded<P2>::value;
  extern __shared__ float2 smem[];
  float2* buf = smem;                    // [CR][NP]
  float2* TA = buf + CR * NP;
  float2* TB = TA + twp.nA;
  const int tid = threadIdx.x;
  load_twiddle(twp, TA, TB, tid, kFT);
  const int64_t b = blockIdx.y;
  const int k10 = blockIdx.x * CR;
  const float2* tb = T + b * n + int64_t(k10) * P2;
  for (int e = tid; e < CR * P2; e += kFT) {
    const int r = e / P2, m2 = e % P2;
    buf[r * NP + pad16(m2)] = tb[e];
  }
  __syncthreads();
  fft_lds<P2, CR, kFT>(buf, TA, TB, twp.lb, 0, tid);
  float2* ob = out + b * n + k10;
  for (int e = tid; e < CR * P2; e += kFT) {
    const int k2 = e / CR, r = e % CR;
    float2 v = buf[r * NP + pad16(k2)];
    if (inverse) v = make_float2(v.x * scale, -v.y * scale);
    ob[int64_t(k2) * P1 + r] = v;
  }
}

template <int P>
constexpr int per_wg() { return P >= 16 && kFMax / P >= 16 ? 16 : kFMax / P; }

template <int P1>
int launch_fft_large_cols(const float2* x, float2* T, int64_t n, int64_t nbatch, int inverse, hipStream_t s) {
  constexpr int CP = per_wg<P1>();
  const int P2 = int(n / P1);
  Twiddle twp, twn;
  int rc = get_twiddle(P1, &twp, s);
  if (rc) return rc;
  rc = get_twiddle(n, &twn, s);
  if (rc) return rc;
  const size_t lds = sizeof(float2) * (size_t(CP) * Padded<P1>::value + twp.nA + twp.nB);
  hipLaunchKernelGGL((fft_large_cols_kernel<P1, CP>), dim3(unsigned(P2 / CP), unsigned(nbatch)), dim3(kFT), lds, s, x,
                     T, n, P2, inverse, twp, twn);
  ALB_LAUNCH_CHECK("fft_large_cols_kernel");
  return ALB_OK;
}

template <int P2>
int launch_fft_large_rows(const float2* T, float2* out, int64_t n, int64_t nbatch, int inverse, hipStream_t s) {
  constexpr int CR = per_wg<P2>();
  const int P1 = int(n / P2);
  Twiddle twp;
  int rc = get_twiddle(P2, &twp, s);
  if (rc) return rc;
  const size_t lds = sizeof(float2) * (size_t(CR) * Padded<P2>::value + twp.nA + twp.nB);
  hipLaunchKernelGGL((fft_large_rows_kernel<P2, CR>), dim3(unsigned(P1 / CR), unsigned(nbatch)), dim3(kFT), lds, s, T,
                     out, n, P1, inverse, float(1.0 / double(n)), twp);
  ALB_LAUNCH_CHECK("fft_large_rows_kernel");
  return ALB_OK;
}

int fft_large(const float2* x, float2* work, float2* out, int64_t n, int64_t nbatch, int inverse, hipStream_t s) {
  const int lg = ilog2(n);
  const int P1 = 1 << (lg / 2);
  const int P2 = int(n / P1);
  int rc = ALB_ENOTSUP;
  switch (P1) {
#define CASE(p) case p: rc = launch_fft_large_cols<p>(x, work, n, nbatch, inverse, s); break;
    CASE(128) CASE(256) CASE(512) CASE(1024) CASE(2048) CASE(4096)
#undef CASE
    default: set_error("alb_fft_large: length %lld not supported", (long long)n);
  }
  if (rc) return rc;
  switch (P2) {
#define CASE(p) case p: return launch_fft_large_rows<p>(work, out, n, nbatch, inverse, s);
    CASE(256) CASE(512) CASE(1024) CASE(2048) CASE(4096)
#undef CASE
    default:
      set_error("alb_fft_large: length %lld not supported", (long long)n);
      return ALB_ENOTSUP;
  }
}

// ---------------------------------------------------------- lag correlation
// orbcomm_utils_gpu.py:39-52: per channel c, the spectra f[t][c] (t < N) of both
// inputs zero-padded to P >= N + dN points, FFT, F1 conj(F2), inverse FFT; lag
// l in [-dN, dN) is element l mod P divided by N - |l|.
__global__ void lag_pad_kernel(const float2* __restrict__ f, int64_t N, int64_t nchan, float2* __restrict__ out,
                               int64_t P) {
  const int64_t c = blockIdx.y;
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < P; t += int64_t(gridDim.x) * blockDim.x)
    out[c * P + t] = t < N ? f[t * nchan + c] : make_float2(0.f, 0.f);
}

__global__ void lag_cross_kernel(float2* __restrict__ a, const float2* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    a[i] = cmulc(a[i], b[i]);
}

__global__ void lag_extract_kernel(const float2* __restrict__ xc, int64_t P, int64_t N, int64_t dN,
                                   float2* __restrict__ out) {
  const int64_t c = blockIdx.y;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < 2 * dN; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t lag = i - dN;
    const float norm = float(N - (lag < 0 ? -lag : lag));
    const float2 v = xc[c * P + (lag < 0 ? lag + P : lag)];
    out[c * 2 * dN + i] = make_float2(v.x / norm, v.y / norm);
  }
}

// orbcomm_utils_gpu.py:13-37 apply_delay: out[t][c] = arr[t][c] exp(-2 pi i freqs[c] delay[t]),
// the phase y = (-2 pi freqs[c]) delay[t] and its cosine and sine in float64 (the
// reference's complex128 phase), the product rounded to complex64.
__global__ void apply_delay_kernel(const float2* __restrict__ arr, const double* __restrict__ delay,
                                   const double* __restrict__ freqs, int64_t nspec, int64_t nchan,
                                   float2* __restrict__ out) {
  const int64_t n = nspec * nchan;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t t = i / nchan, c = i - t * nchan;
    const double y = (-2.0 * 3.141592653589793 * freqs[c]) * delay[t];
    double sn, cs;
    sincos(y, &sn, &cs);
    const float2 a = arr[i];
    out[i] = make_float2(float(double(a.x) * cs - double(a.y) * sn), float(double(a.x) * sn + double(a.y) * cs));
  }
}

}  // namespace alb

using namespace alb;

extern "C" {

int alb_fft_rows(const void* in, int64_t ldi, void* out, int64_t ldo, int64_t nrows, int64_t n, int direction,
                 void* stream);

// pycufft.py:92-131 _fft along rows of n = 2^15..2^24 complex64 points: x and out
// [nbatch][n] (may be the same), work [nbatch][n] scratch (neither of them).
int alb_fft_large(const void* x, void* work, void* out, int64_t n, int64_t nbatch, int direction, void* stream) {
  ALB_REQUIRE(x && work && out, "alb_fft_large: null pointer");
  ALB_REQUIRE(direction == -1 || direction == 1, "alb_fft_large: direction must be -1 or +1");
  ALB_REQUIRE(is_pow2(n) && n >= (1 << 15) && n <= (1 << 24), "alb_fft_large: n must be a power of two 2^15..2^24");
  ALB_REQUIRE(nbatch >= 0 && nbatch <= 65535, "alb_fft_large: at most 65535 rows per call");
  ALB_REQUIRE(work != x && work != out, "alb_fft_large: work must not alias x or out");
  if (nbatch == 0) return ALB_OK;
  return fft_large((const float2*)x, (float2*)work, (float2*)out, n, nbatch, direction == 1, as_stream(stream));
}

// orbcomm_utils_gpu.py:39-52 coarse_xcorr: f1, f2 complex64 [N][nchan] (C order);
// work: 3 * nchan * P complex64 with P the FFT length (power of two, P >= N + dN, at
// most 2^24); out complex64 [nchan][2 dN], lag -dN + i at column i, divided by
// N - |lag|.
int alb_lag_xcorr(const void* f1, const void* f2, int64_t N, int64_t nchan, int64_t dN, int64_t P, void* work,
                  void* out, void* stream) {
  ALB_REQUIRE(f1 && f2 && work && out, "alb_lag_xcorr: null pointer");
  ALB_REQUIRE(N >= 1 && nchan >= 1 && nchan <= 65535 && dN >= 1 && dN <= N, "alb_lag_xcorr: bad sizes (need 1 <= dN <= N)");
  ALB_REQUIRE(is_pow2(P) && P >= N + dN && P <= (1 << 24) && P >= 64,
              "alb_lag_xcorr: P must be a power of two, N + dN <= P <= 2^24");
  hipStream_t s = as_stream(stream);
  float2* A = (float2*)work;
  float2* B = A + nchan * P;
  float2* W = B + nchan * P;
  const dim3 gp(unsigned((P + 255) / 256 < 4096 ? (P + 255) / 256 : 4096), unsigned(nchan));
  hipLaunchKernelGGL(lag_pad_kernel, gp, dim3(256), 0, s, (const float2*)f1, N, nchan, A, P);
  hipLaunchKernelGGL(lag_pad_kernel, gp, dim3(256), 0, s, (const float2*)f2, N, nchan, B, P);
  ALB_LAUNCH_CHECK("lag_pad_kernel");
  int rc;
  if (P <= 16384) {
    rc = alb_fft_rows(A, P, A, P, nchan, P, -1, stream);
    if (!rc) rc = alb_fft_rows(B, P, B, P, nchan, P, -1, stream);
  } else {
    rc = fft_large(A, W, A, P, nchan, 0, s);
    if (!rc) rc = fft_large(B, W, B, P, nchan, 0, s);
  }
  if (rc) return rc;
  const int64_t tot = nchan * P;
  hipLaunchKernelGGL(lag_cross_kernel, dim3(unsigned((tot + 255) / 256 < 65536 ? (tot + 255) / 256 : 65536)),
                     dim3(256), 0, s, A, (const float2*)B, tot);
  ALB_LAUNCH_CHECK("lag_cross_kernel");
  rc = P <= 16384 ? alb_fft_rows(A, P, A, P, nchan, P, 1, stream) : fft_large(A, W, A, P, nchan, 1, s);
  if (rc) return rc;
  const dim3 ge(unsigned((2 * dN + 255) / 256 < 4096 ? (2 * dN + 255) / 256 : 4096), unsigned(nchan));
  hipLaunchKernelGGL(lag_extract_kernel, ge, dim3(256), 0, s, A, P, N, dN, (float2*)out);
  ALB_LAUNCH_CHECK("lag_extract_kernel");
  return ALB_OK;
}

// orbcomm_utils_gpu.py:13-37 apply_delay on arr [nspec][nchan] complex64, delay float64 [nspec],
// freqs float64 [nchan]; out may be arr.
int alb_apply_delay(const void* arr, const double* delay, const double* freqs, int64_t nspec, int64_t nchan, void* out,
                    void* stream) {
  ALB_REQUIRE(nspec >= 0 && nchan >= 0, "alb_apply_delay: bad sizes");
  if (nspec * nchan == 0) return ALB_OK;
  ALB_REQUIRE(arr && delay && freqs && out, "alb_apply_delay: null pointer");
  const int64_t n = nspec * nchan;
  hipLaunchKernelGGL(apply_delay_kernel, dim3(unsigned((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536)), dim3(256),
                     0, as_stream(stream), (const float2*)arr, delay, freqs, nspec, nchan, (float2*)out);
  ALB_LAUNCH_CHECK("apply_delay_kernel");
  return ALB_OK;
}

}  // extern "C"
