// Integer correlators on 4-bit and 1-bit baseband.
//
//  * pair kernels  : one baseline, optional spec_num row matching (2-antenna
//                    form), exact int64 sums per channel
//                    (correlations_cpu.c:39-271, :273-639)
//  * all-pairs     : every baseline i <= j of up to 16 inputs in one pass,
//                    int64 block sums [block][pair][chan][re,im]
//                    (the direct correlators of configurations 0 and 4).
// Products are x_i * conj(x_j):  Re = r_i r_j + i_i i_j,  Im = r_j i_i - r_i i_j
// (correlations_cpu.c:172-173).  Every sum is exact; partial sums are int32 in
// registers (bounded by the row counts checked on the host) and int64 in memory.
#include "common.hpp"

namespace alb {

__device__ __forceinline__ int nib4(int v) { return v > 8 ? v - 16 : v; }

// Two's-complement int64 accumulate through the unsigned 64-bit atomic.
__device__ __forceinline__ void atomic_add_i64(long long* p, long long v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v));
}

__global__ void fill_i64(int64_t* p, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = 0;
}
__global__ void fill_i32(int* p, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = 0;
}

// ------------------------------------------------------------ row matching
// For row i of list 0 find j with specnum1[j]-idx1 == specnum0[i]-idx0 (both
// lists sorted ascending, as the merge of correlations_cpu.c:201-214 assumes).
__global__ void match_rows_kernel(const int64_t* __restrict__ s0, const int64_t* __restrict__ s1, int64_t n0, int64_t n1,
                                  int64_t idx0, int64_t idx1, int* __restrict__ match, int* __restrict__ count) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n0) return;
  const int64_t target = s0[i] - idx0 + idx1;
  int64_t lo = 0, hi = n1;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (s1[mid] < target) lo = mid + 1; else hi = mid;
  }
  const int m = (lo < n1 && s1[lo] == target) ? int(lo) : -1;
  match[i] = m;
  if (m >= 0) atomicAdd(count, 1);
}

// --------------------------------------------------------------- 4-bit pair
// mode 0: autocorrelation sum r^2+i^2 (re only); mode 1: cross x0 conj(x1).
// Rows of data0 0..nrows-1; row of data1 = match ? match[r] : r (skip if < 0).
__global__ void pair4_kernel(const uint8_t* __restrict__ d0, const uint8_t* __restrict__ d1, int64_t nrows, int64_t ncol,
                             const int* __restrict__ match, int mode, int64_t rows_per, long long* __restrict__ out) {
  const int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (c >= ncol) return;
  const int64_t r0 = blockIdx.y * rows_per;
  const int64_t r1 = r0 + rows_per < nrows ? r0 + rows_per : nrows;
  int sr = 0, si = 0;
  for (int64_t r = r0; r < r1; ++r) {
    int64_t q = r;
    if (match) { q = match[r]; if (q < 0) continue; }
    const int a = d0[r * ncol + c];
    const int ar = nib4(a >> 4), ai = nib4(a & 15);
    if (mode == 0) {
      sr += ar * ar + ai * ai;
    } else {
      const int b = d1[q * ncol + c];
      const int br = nib4(b >> 4), bi = nib4(b & 15);
      sr += ar * br + ai * bi;
      si += br * ai - ar * bi;
    }
  }
  if (sr) atomic_add_i64(&out[2 * c], (long long)sr);
  if (si) atomic_add_i64(&out[2 * c + 1], (long long)si);
}

// --------------------------------------------------------------- 1-bit pair
// Packed format: per pol, 4 channels per byte, (re, im) bits MSB first.
// mode 1: avg_xcorr_1bit sums (correlations_cpu.c:327-334), out [nchan][2];
// mode 2: van Vleck sums R0 R1 IM0 IM1 (correlations_cpu.c:488-491), out [4][nchan].
__global__ void pair1_kernel(const uint8_t* __restrict__ d0, const uint8_t* __restrict__ d1, int64_t nrows, int64_t ncol,
                             int64_t nchan, const int* __restrict__ match, int mode, int64_t rows_per,
                             long long* __restrict__ out) {
  const int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;  // byte column
  if (j >= ncol) return;
  const int64_t r0 = blockIdx.y * rows_per;
  const int64_t r1 = r0 + rows_per < nrows ? r0 + rows_per : nrows;
  int s[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) s[a][b] = 0;
  for (int64_t r = r0; r < r1; ++r) {
    int64_t q = r;
    if (match) { q = match[r]; if (q < 0) continue; }
    const int a = d0[r * ncol + j];
    const int b = d1[q * ncol + j];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int ra = (a >> (7 - 2 * c)) & 1, ia = (a >> (6 - 2 * c)) & 1;
      const int rb = (b >> (7 - 2 * c)) & 1, ib = (b >> (6 - 2 * c)) & 1;
      if (mode == 1) {
        s[c][0] += 2 - 2 * ((ra ^ rb) + (ia ^ ib));
        s[c][1] += -2 * ((rb ^ ia) - (ra ^ ib));
      } else {
        s[c][0] += 1 - 2 * (ra ^ rb);
        s[c][1] += 1 - 2 * (ia ^ ib);
        s[c][2] += 1 - 2 * (ra ^ ib);
        s[c][3] += 1 - 2 * (rb ^ ia);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int64_t ch = 4 * j + c;
    if (ch >= nchan) break;
    if (mode == 1) {
      if (s[c][0]) atomic_add_i64(&out[2 * ch], (long long)s[c][0]);
      if (s[c][1]) atomic_add_i64(&out[2 * ch + 1], (long long)s[c][1]);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (s[c][k]) atomic_add_i64(&out[k * nchan + ch], (long long)s[c][k]);
    }
  }
}

// ------------------------------------------------------- 4-bit all pairs
struct InputPtrs { const uint8_t* p[16]; };

// Thread: 4 consecutive channels (one dword per input per spectrum); WG rows of
// one block split over RL row lanes; partials reduced through LDS, then one
// int64 atomic per (pair, channel, re/im).
template <int NIN>
__global__ __launch_bounds__(256) void allpairs4_kernel(InputPtrs in, int64_t ncol, int64_t blocklen, int nsplit,
                                                        long long* __restrict__ out) {
  constexpr int NP = NIN * (NIN + 1) / 2;
  constexpr int CL = 32;          // channel lanes (x 4 channels)
  constexpr int RL = 256 / CL;    // row lanes
  __shared__ int red[RL][CL * 4];
  const int cl = threadIdx.x % CL, rl = threadIdx.x / CL;
  const int64_t c0 = (int64_t(blockIdx.x) * CL + cl) * 4;
  const int64_t blk = blockIdx.y / nsplit, sp = blockIdx.y % nsplit;
  const int64_t rows_per = (blocklen + nsplit - 1) / nsplit;
  const int64_t rb = blk * blocklen + sp * rows_per;
  int64_t re_ = rb + rows_per;
  if (re_ > blk * blocklen + blocklen) re_ = blk * blocklen + blocklen;
  const bool active = c0 < ncol;
  int acc[NP][4][2];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int c = 0; c < 4; ++c) { acc[p][c][0] = 0; acc[p][c][1] = 0; }
  if (active) {
    for (int64_t r = rb + rl; r < re_; r += RL) {
      int xr[NIN][4], xi[NIN][4];
#pragma unroll
      for (int u = 0; u < NIN; ++u) {
        const uint32_t w = *reinterpret_cast<const uint32_t*>(in.p[u] + r * ncol + c0);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int b = (w >> (8 * c)) & 0xFF;
          xr[u][c] = nib4(b >> 4);
          xi[u][c] = nib4(b & 15);
        }
      }
      int p = 0;
#pragma unroll
      for (int a = 0; a < NIN; ++a)
#pragma unroll
        for (int b = a; b < NIN; ++b) {
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            acc[p][c][0] += xr[a][c] * xr[b][c] + xi[a][c] * xi[b][c];
            if (a != b) acc[p][c][1] += xr[b][c] * xi[a][c] - xr[a][c] * xi[b][c];
          }
          ++p;
        }
    }
  }
  // reduce over row lanes, pair by pair
#pragma unroll
  for (int p = 0; p < NP; ++p) {
#pragma unroll
    for (int ri = 0; ri < 2; ++ri) {
#pragma unroll
      for (int c = 0; c < 4; ++c) red[rl][cl * 4 + c] = acc[p][c][ri];
      __syncthreads();
      if (rl == 0 && active) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          long long s = 0;
          for (int k = 0; k < RL; ++k) s += red[k][cl * 4 + c];
          if (s) atomic_add_i64(&out[((blk * NP + p) * ncol + c0 + c) * 2 + ri], s);
        }
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------- 1-bit all pairs
// Thread: one packed byte column (4 channels) for a range of rows and a group
// of pairs; XOR form of correlations_cpu.c:327-334 per pair.
__global__ __launch_bounds__(256) void allpairs1_kernel(InputPtrs in, int nin, int64_t ncol, int64_t nchan,
                                                        int64_t blocklen, int nsplit, int pairs_per, int npairs,
                                                        long long* __restrict__ out) {
  const int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (j >= ncol) return;
  const int pg = blockIdx.z;
  const int64_t blk = blockIdx.y / nsplit, sp = blockIdx.y % nsplit;
  const int64_t rows_per = (blocklen + nsplit - 1) / nsplit;
  const int64_t rb = blk * blocklen + sp * rows_per;
  int64_t rend = rb + rows_per;
  if (rend > blk * blocklen + blocklen) rend = blk * blocklen + blocklen;
  for (int pp = 0; pp < pairs_per; ++pp) {
    const int p = pg * pairs_per + pp;
    if (p >= npairs) break;
    // decode pair index -> (a, b), a <= b, row-major
    int a = 0, rem = p;
    while (rem >= nin - a) { rem -= nin - a; ++a; }
    const int b = a + rem;
    const uint8_t* pa = in.p[a];
    const uint8_t* pb = in.p[b];
    int sre[4] = {0, 0, 0, 0}, sim[4] = {0, 0, 0, 0};
    for (int64_t r = rb; r < rend; ++r) {
      const int x = pa[r * ncol + j], y = pb[r * ncol + j];
      const int d = x ^ y;
      const int ysw = ((y & 0xAA) >> 1) | ((y & 0x55) << 1);
      const int e = x ^ ysw;  // bit hi: r_x ^ i_y, bit lo: i_x ^ r_y
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        sre[c] += 2 - 2 * (((d >> (7 - 2 * c)) & 1) + ((d >> (6 - 2 * c)) & 1));
        sim[c] += 2 * (((e >> (7 - 2 * c)) & 1) - ((e >> (6 - 2 * c)) & 1));
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t ch = 4 * j + c;
      if (ch >= nchan) break;
      if (sre[c]) atomic_add_i64(&out[((blk * npairs + p) * nchan + ch) * 2], (long long)sre[c]);
      if (sim[c]) atomic_add_i64(&out[((blk * npairs + p) * nchan + ch) * 2 + 1], (long long)sim[c]);
    }
  }
}

static inline dim3 grid1(int64_t n, int bs = 256) { return dim3(unsigned((n + bs - 1) / bs)); }

}  // namespace alb

using namespace alb;

extern "C" {

// correlations_cpu.c:39 avg_autocorr_4bit / :128 avg_xcorr_4bit / :188 avg_xcorr_4bit_2ant.
// out: int64 [ncol][2] sums; match: optional device int[nrows] row map of data1 (-1 = skip).
int alb_xcorr_4bit_pair(const uint8_t* data0, const uint8_t* data1, int64_t nrows, int64_t ncol, const int* match,
                        int autocorr, int64_t* out, void* stream) {
  ALB_REQUIRE(data0 && out && (autocorr || data1), "alb_xcorr_4bit_pair: null pointer");
  ALB_REQUIRE(nrows >= 0 && ncol > 0, "alb_xcorr_4bit_pair: bad sizes");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(fill_i64, grid1(2 * ncol), dim3(256), 0, s, out, 2 * ncol);
  if (nrows > 0) {
    const int64_t rows_per = 1024;  // |r r' + i i'| <= 128: int32 partials are exact
    hipLaunchKernelGGL(pair4_kernel, dim3(unsigned((ncol + 63) / 64), unsigned((nrows + rows_per - 1) / rows_per)),
                       dim3(64), 0, s, data0, autocorr ? data0 : data1, nrows, ncol, match, autocorr ? 0 : 1, rows_per,
                       (long long*)out);
  }
  ALB_LAUNCH_CHECK("pair4_kernel");
  return ALB_OK;
}

// Row match for the 2-antenna forms: match[i] = j or -1, *count = matched rows.
int alb_match_rows(const int64_t* specnum0, const int64_t* specnum1, int64_t n0, int64_t n1, int64_t idxstart0,
                   int64_t idxstart1, int* match, int* count, void* stream) {
  ALB_REQUIRE(match && count && (n0 == 0 || specnum0) && (n1 == 0 || specnum1), "alb_match_rows: null pointer");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, s, count, int64_t(1));
  if (n0 > 0)
    hipLaunchKernelGGL(match_rows_kernel, grid1(n0), dim3(256), 0, s, specnum0, specnum1, n0, n1, idxstart0, idxstart1,
                       match, count);
  ALB_LAUNCH_CHECK("match_rows_kernel");
  return ALB_OK;
}

// correlations_cpu.c:273 avg_xcorr_1bit (vanvleck=0, out int64 [nchan][2]) and
// :435 / :533 avg_xcorr_1bit_vanvleck(_2ant) (vanvleck=1, out int64 [4][nchan]).
int alb_xcorr_1bit_pair(const uint8_t* data0, const uint8_t* data1, int64_t nrows, int64_t ncol, int64_t nchan,
                        const int* match, int vanvleck, int64_t* out, void* stream) {
  ALB_REQUIRE(data0 && data1 && out, "alb_xcorr_1bit_pair: null pointer");
  ALB_REQUIRE(nrows >= 0 && ncol > 0 && nchan > 0 && nchan <= 4 * ncol, "alb_xcorr_1bit_pair: bad sizes");
  hipStream_t s = as_stream(stream);
  const int64_t nout = vanvleck ? 4 * nchan : 2 * nchan;
  hipLaunchKernelGGL(fill_i64, grid1(nout), dim3(256), 0, s, out, nout);
  if (nrows > 0) {
    const int64_t rows_per = 4096;
    hipLaunchKernelGGL(pair1_kernel, dim3(unsigned((ncol + 63) / 64), unsigned((nrows + rows_per - 1) / rows_per)),
                       dim3(64), 0, s, data0, data1, nrows, ncol, nchan, match, vanvleck ? 2 : 1, rows_per,
                       (long long*)out);
  }
  ALB_LAUNCH_CHECK("pair1_kernel");
  return ALB_OK;
}

// All baselines i <= j of ninput 4-bit inputs (per-pol bytes, [T][ncol] each),
// sums over consecutive blocks of blocklen spectra -> int64 [nblock][npair][ncol][2].
int alb_xcorr_4bit_allpairs(const uint8_t* const* inputs, int ninput, int64_t nrows, int64_t ncol, int64_t blocklen,
                            int64_t* out, void* stream) {
  ALB_REQUIRE(inputs && out, "alb_xcorr_4bit_allpairs: null pointer");
  ALB_REQUIRE(ninput >= 1 && ninput <= 8, "alb_xcorr_4bit_allpairs: 1..8 inputs supported (got %d)", ninput);
  ALB_REQUIRE(ncol % 4 == 0 && ncol > 0, "alb_xcorr_4bit_allpairs: ncol must be a positive multiple of 4");
  ALB_REQUIRE(blocklen > 0 && nrows >= blocklen, "alb_xcorr_4bit_allpairs: need at least one full block");
  InputPtrs ip{};
  for (int u = 0; u < ninput; ++u) {
    ALB_REQUIRE(inputs[u] && (reinterpret_cast<uintptr_t>(inputs[u]) % 4 == 0),
                "alb_xcorr_4bit_allpairs: input %d null or not 4-byte aligned", u);
    ip.p[u] = inputs[u];
  }
  const int64_t nblock = nrows / blocklen;
  const int64_t npair = int64_t(ninput) * (ninput + 1) / 2;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(fill_i64, grid1(nblock * npair * ncol * 2), dim3(256), 0, s, out, nblock * npair * ncol * 2);
  const int nsplit = int(blocklen >= 1024 ? 8 : 1);
  ALB_REQUIRE(blocklen / nsplit <= (int64_t(1) << 24), "alb_xcorr_4bit_allpairs: block too long for int32 partials");
  dim3 grid(unsigned((ncol / 4 + 31) / 32), unsigned(nblock * nsplit));
  switch (ninput) {
#define CASE(n) case n: hipLaunchKernelGGL(allpairs4_kernel<n>, grid, dim3(256), 0, s, ip, ncol, blocklen, nsplit, (long long*)out); break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
  }
  ALB_LAUNCH_CHECK("allpairs4_kernel");
  return ALB_OK;
}

// All baselines of ninput 1-bit inputs (sortpols 1-bit bytes, [T][ncol] each,
// nchan channels) -> int64 [nblock][npair][nchan][2] block sums.
int alb_xcorr_1bit_allpairs(const uint8_t* const* inputs, int ninput, int64_t nrows, int64_t ncol, int64_t nchan,
                            int64_t blocklen, int64_t* out, void* stream) {
  ALB_REQUIRE(inputs && out, "alb_xcorr_1bit_allpairs: null pointer");
  ALB_REQUIRE(ninput >= 1 && ninput <= 16, "alb_xcorr_1bit_allpairs: 1..16 inputs supported (got %d)", ninput);
  ALB_REQUIRE(ncol > 0 && nchan > 0 && nchan <= 4 * ncol, "alb_xcorr_1bit_allpairs: bad sizes");
  ALB_REQUIRE(blocklen > 0 && nrows >= blocklen, "alb_xcorr_1bit_allpairs: need at least one full block");
  InputPtrs ip{};
  for (int u = 0; u < ninput; ++u) {
    ALB_REQUIRE(inputs[u], "alb_xcorr_1bit_allpairs: input %d is null", u);
    ip.p[u] = inputs[u];
  }
  const int64_t nblock = nrows / blocklen;
  const int npair = ninput * (ninput + 1) / 2;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(fill_i64, grid1(nblock * npair * nchan * 2), dim3(256), 0, s, out, nblock * npair * nchan * 2);
  const int nsplit = int(blocklen >= 2048 ? 4 : 1);
  const int pairs_per = 8;
  dim3 grid(unsigned((ncol + 255) / 256), unsigned(nblock * nsplit), unsigned((npair + pairs_per - 1) / pairs_per));
  hipLaunchKernelGGL(allpairs1_kernel, grid, dim3(256), 0, s, ip, ninput, ncol, nchan, blocklen, nsplit, pairs_per,
                     npair, (long long*)out);
  ALB_LAUNCH_CHECK("allpairs1_kernel");
  return ALB_OK;
}

}  // extern "C"
