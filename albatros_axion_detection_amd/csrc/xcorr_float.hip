// Per-channel cross-antenna correlation out[i,j,f] = sum_t x[i,t,f] conj(x[j,t,f]) * (1/K)
// (correlations_gpu.py:95-121, cgemm_batch.cu:8-54).  x and out are column-major
// (F-order): x[(m, t, f)] at m + M*(t + K*f), out[(i, j, f)] at i + M*(j + M*f).
//
// M <= 8: the realified Gram G = Y Y^T of Y = [Re x; Im x] (16 x K) is built on the
// FP32 matrix cores with v_mfma_f32_16x16x4f32.  Lane l feeds Y[l % 16][t0 + l / 16]
// to both operands (A = Y, B = Y^T share registers).  Row m < 8 of Y is input
// m % SLOT (real part; row m + 8 the imaginary part) of channel m / SLOT of the
// wavefront: SLOT = 8 for M > 4 (one channel per wavefront), 4 for M = 3, 4 and
// 2 for M <= 2, so 2 or 4 channels share one product and every lane loads.  A
// channel's block of G is its own Gram (each element sums its two rows only, in
// the same order whatever the other rows hold).  Then, with i, j offset by the
// channel's slot, Re = G[i][j] + G[i+8][j+8], Im = G[i+8][j] - G[i][j+8].
// Other M: a VALU kernel, workgroup per channel with its inputs staged through LDS
// and one thread per baseline (F-order and strided inputs alike).
//
// The correlator kernels also take a strided input: channel c = s F + f of
// S jobs of F channels, element (m, t) of it at x[s ss + f sf + t sk + m]
// (complex elements; the F-order layout is sk = M, sf = M K, ss = 0 with S = 1).
#include "common.hpp"

namespace alb {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Input addressing of the matrix-core correlators (complex-element strides).
struct XLayout {
  int64_t sk, sf, ss, F;   // time, channel and job strides; channels per job
  __device__ __forceinline__ int64_t channel(int64_t c) const {
    const int64_t s = c / F;
    return s * ss + (c - s * F) * sf;
  }
};

// Channels per wavefront and input slots per channel of the matrix-core forms.
template <int M>
struct XPack {
  static constexpr int SLOT = M <= 2 ? 2 : (M <= 4 ? 4 : 8);
  static constexpr int CPW = 8 / SLOT;
};

// Row of Y fed by lane r = l % 16: its channel (f0 + sub) and input (mi); the
// rows of absent inputs and channels are zero.
template <int M>
struct XRow {
  int64_t f0;
  int sub, mi, part;
  bool ok;
  __device__ __forceinline__ XRow(int64_t wave, int r, int64_t F) {
    const int m = r % 8;
    f0 = wave * XPack<M>::CPW;
    sub = m / XPack<M>::SLOT;
    mi = m % XPack<M>::SLOT;
    part = r / 8;
    ok = f0 + sub < F && mi < M;
  }
};

// Re/Im of the channels of one wavefront from its Gram G (16 x 16, row-major):
// lane e < CPW M M writes baseline (i, j) of channel f0 + e / (M M).
template <int M>
__device__ __forceinline__ void xgram_store(const float* g, int lane, int64_t f0, int64_t F,
                                            float2* __restrict__ out, float scale) {
  constexpr int SLOT = XPack<M>::SLOT;
  if (lane < XPack<M>::CPW * M * M) {
    const int sub = lane / (M * M), e = lane % (M * M);
    const int i = e % M, j = e / M;
    const int64_t f = f0 + sub;
    if (f < F) {
      const int io = sub * SLOT + i, jo = sub * SLOT + j;
      const float re = g[io * 16 + jo] + g[(io + 8) * 16 + (jo + 8)];
      const float im = g[(io + 8) * 16 + jo] - g[io * 16 + (jo + 8)];
      out[f * M * M + int64_t(j) * M + i] = make_float2(re * scale, im * scale);
    }
  }
}

template <int M>
__global__ __launch_bounds__(256) void xcorr_mfma16_kernel(const float* __restrict__ x, int64_t K, int64_t F,
                                                           XLayout L, float2* __restrict__ out, float scale) {
  __shared__ float G[4][16 * 16];
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int r = lane % 16, slot = lane / 16;
  const XRow<M> row(int64_t(blockIdx.x) * 4 + w, r, F);
  const bool row_ok = row.ok;
  const float* base = x + 2 * L.channel(row_ok ? row.f0 + row.sub : 0) + 2 * row.mi + row.part;
  const int64_t st = 2 * L.sk;
  floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = acc0, acc2 = acc0, acc3 = acc0;
  int64_t t0 = 0;
  for (; t0 + 16 <= K; t0 += 16) {
    const float a0 = row_ok ? base[st * (t0 + slot)] : 0.f;
    const float a1 = row_ok ? base[st * (t0 + 4 + slot)] : 0.f;
    const float a2 = row_ok ? base[st * (t0 + 8 + slot)] : 0.f;
    const float a3 = row_ok ? base[st * (t0 + 12 + slot)] : 0.f;
    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, a0, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, a1, acc1, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, a2, acc2, 0, 0, 0);
    acc3 = __builtin_amdgcn_mfma_f32_16x16x4f32(a3, a3, acc3, 0, 0, 0);
  }
  for (; t0 < K; t0 += 4) {
    const int64_t t = t0 + slot;
    const float a0 = (row_ok && t < K) ? base[st * t] : 0.f;
    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, a0, acc0, 0, 0, 0);
  }
  // D layout: lane l holds column l % 16, rows 4*(l/16) + v.
#pragma unroll
  for (int v = 0; v < 4; ++v) G[w][(4 * slot + v) * 16 + r] = (acc0[v] + acc1[v]) + (acc2[v] + acc3[v]);
  __syncthreads();
  xgram_store<M>(G[w], lane, row.f0, F, out, scale);
}

// Reduced-precision forms of the same Gram product on the 16x16x32 matrix cores
// (FP32 accumulation), one wavefront per channel:
//   PREC_BF16X3  x = hi + lo with hi = bf16(x), lo = bf16(x - hi);
//                G = hi hi^T + (hi lo^T + lo hi^T)   (lo lo^T dropped)
//   PREC_FP16    x -> fp16, one product
//   PREC_FP16X3  x = hi + 2^-11 lo with hi = fp16(x), lo = fp16(2^11 (x - hi)) (the
//                scaled residual stays out of the fp16 subnormal range);
//                G = hi hi^T + 2^-11 (hi lo^T + lo hi^T)
// Lane l feeds time samples t0 + 8 (l / 16) + i (i < 8) of realified row l % 16 to
// both operands; A and B share one k order, so any assignment of samples to k is a
// valid Gram sum.
enum { PREC_FP32 = 0, PREC_BF16X3 = 1, PREC_FP16 = 2, PREC_FP16X3 = 3 };
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <int M, int MODE>
__global__ __launch_bounds__(256) void xcorr_mfma16_lp_kernel(const float* __restrict__ x, int64_t K, int64_t F,
                                                              XLayout L, float2* __restrict__ out, float scale) {
  __shared__ float G[4][16 * 16];
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int r = lane % 16, grp = lane / 16;
  const XRow<M> row(int64_t(blockIdx.x) * 4 + w, r, F);
  const bool row_ok = row.ok;
  const float* base = x + 2 * L.channel(row_ok ? row.f0 + row.sub : 0) + 2 * row.mi + row.part;
  const int64_t st = 2 * L.sk;
  floatx4 accH = {0.f, 0.f, 0.f, 0.f}, accX = accH;
  for (int64_t t0 = 0; t0 < K; t0 += 32) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int64_t t = t0 + 8 * grp + i;
      v[i] = (row_ok && t < K) ? base[st * t] : 0.f;
    }
    if constexpr (MODE == PREC_BF16X3) {
      bf16x8 hi, lo;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        hi[i] = (__bf16)v[i];
        lo[i] = (__bf16)(v[i] - (float)hi[i]);
      }
      accH = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi, hi, accH, 0, 0, 0);
      accX = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi, lo, accX, 0, 0, 0);
      accX = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lo, hi, accX, 0, 0, 0);
    } else if constexpr (MODE == PREC_FP16) {
      halfx8 h;
#pragma unroll
      for (int i = 0; i < 8; ++i) h[i] = (_Float16)v[i];
      accH = __builtin_amdgcn_mfma_f32_16x16x32_f16(h, h, accH, 0, 0, 0);
    } else {
      halfx8 hi, lo;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        hi[i] = (_Float16)v[i];
        lo[i] = (_Float16)((v[i] - (float)hi[i]) * 2048.f);
      }
      accH = __builtin_amdgcn_mfma_f32_16x16x32_f16(hi, hi, accH, 0, 0, 0);
      accX = __builtin_amdgcn_mfma_f32_16x16x32_f16(hi, lo, accX, 0, 0, 0);
      accX = __builtin_amdgcn_mfma_f32_16x16x32_f16(lo, hi, accX, 0, 0, 0);
    }
  }
  constexpr float xs = MODE == PREC_FP16X3 ? 1.f / 2048.f : 1.f;
#pragma unroll
  for (int v = 0; v < 4; ++v) G[w][(4 * grp + v) * 16 + r] = accH[v] + accX[v] * xs;
  __syncthreads();
  xgram_store<M>(G[w], lane, row.f0, F, out, scale);
}

// Generic VALU form (any M): workgroup per channel c of the XLayout, thread per
// baseline i <= j (more than 256 baselines loop in batches).  The channel's M
// inputs are staged through LDS kVT time samples at a time (one contiguous run
// of M complex per sample in every layout used here), so the loads coalesce;
// FP32 products into FP32 partial sums, flushed into float64 every 256 samples.
constexpr int kVT = 32;

__global__ __launch_bounds__(256) void xcorr_valu_kernel(const float2* __restrict__ x, int M, int64_t K, XLayout L,
                                                         float2* __restrict__ out, float scale) {
  extern __shared__ float2 xs[];   // [kVT][M]
  const int64_t c = blockIdx.x;
  const float2* xc = x + L.channel(c);
  const int npair = M * (M + 1) / 2;
  for (int p0 = 0; p0 < npair; p0 += blockDim.x) {
    const int p = p0 + threadIdx.x;
    int i = 0, j = 0;
    if (p < npair) {
      int rem = p;
      while (rem >= M - i) { rem -= M - i; ++i; }
      j = i + rem;
    }
    double sr = 0.0, si = 0.0;
    float cr = 0.f, ci = 0.f;
    for (int64_t t0 = 0; t0 < K; t0 += kVT) {
      const int nt = K - t0 < kVT ? int(K - t0) : kVT;
      __syncthreads();
      for (int e = threadIdx.x; e < nt * M; e += blockDim.x) {
        const int tt = e / M, m = e - tt * M;
        xs[e] = xc[(t0 + tt) * L.sk + m];
      }
      __syncthreads();
      if (p < npair) {
        for (int tt = 0; tt < nt; ++tt) {
          const float2 a = xs[tt * M + i], b = xs[tt * M + j];
          cr = fmaf(a.x, b.x, fmaf(a.y, b.y, cr));
          ci = fmaf(a.y, b.x, fmaf(-a.x, b.y, ci));
        }
        if (((t0 / kVT) & 7) == 7) { sr += cr; si += ci; cr = 0.f; ci = 0.f; }
      }
    }
    if (p < npair) {
      sr += cr;
      si += ci;
      const float re = float(sr), im = float(si);
      out[c * M * M + int64_t(j) * M + i] = make_float2(re * scale, im * scale);
      if (i != j) out[c * M * M + int64_t(i) * M + j] = make_float2(re * scale, -im * scale);
    }
  }
}

// correlations.py:115-151 avg_xcorr_all_ant_cpu: every product x_i x_j^* rounded
// to FP32 (two FP32 multiplies and one FP32 add per component, no fused
// multiply-add, as the complex64 product of the reference's Numba loop), summed
// in float64, divided by K in float64 and rounded to complex64; the lower
// triangle is the conjugate.  Thread per channel f (consecutive f are
// neighbouring lanes), grid y = baseline.  Element (m, k, f) at
// x[m*sm + k*sk + f*sf], output (i, j, f) at out[i*oi + j*oj + f*of].
__global__ __launch_bounds__(256) void xcorr_f64acc_kernel(const float2* __restrict__ x, int M, int64_t K, int64_t F,
                                                           int64_t sm, int64_t sk, int64_t sf,
                                                           float2* __restrict__ out, int64_t oi, int64_t oj,
                                                           int64_t of) {
#pragma clang fp contract(off)
  const int64_t f = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (f >= F) return;
  int i = 0, rem = int(blockIdx.y);
  while (rem >= M - i) { rem -= M - i; ++i; }
  const int j = i + rem;
  const float2* xi = x + i * sm + f * sf;
  const float2* xj = x + j * sm + f * sf;
  double sr = 0.0, si = 0.0;
  for (int64_t k = 0; k < K; ++k) {
    const float2 a = xi[k * sk], b = xj[k * sk];
    const float pr = a.x * b.x + a.y * b.y;
    const float pi = a.y * b.x - a.x * b.y;
    sr += double(pr);
    si += double(pi);
  }
  const double dk = double(K);
  const double re = sr / dk, im = si / dk;
  out[i * oi + j * oj + f * of] = make_float2(float(re), float(im));
  if (i != j) out[j * oi + i * oj + f * of] = make_float2(float(re), float(-im));
}

// cgemm_batch.cu:8-54 cgemm_strided_batched: C = A B^H per batch, column-major,
// A (M, K), B (N, K), C (M, N), batch strides M K, N K, M N.  A 16 x 16 tile of
// C per workgroup (thread per element), A and B staged through LDS 16 columns
// of K at a time; FP32 products and sums (alpha 1, beta 0, as the reference).
__global__ __launch_bounds__(256) void cgemm_abh_kernel(const float2* __restrict__ A, const float2* __restrict__ B,
                                                        float2* __restrict__ C, int64_t M, int64_t N, int64_t K) {
  __shared__ float2 As[16][17];   // [k][m]
  __shared__ float2 Bs[16][17];   // [k][n]
  const int64_t b = blockIdx.x;
  const int64_t m0 = int64_t(blockIdx.y) * 16, n0 = int64_t(blockIdx.z) * 16;
  const int tm = threadIdx.x % 16, tn = threadIdx.x / 16;
  const float2* Ab = A + b * M * K;
  const float2* Bb = B + b * N * K;
  float cr = 0.f, ci = 0.f;
  for (int64_t k0 = 0; k0 < K; k0 += 16) {
    // lane (tm, tn) loads A[m0 + tm, k0 + tn] and B[n0 + tm, k0 + tn]
    const int64_t k = k0 + tn;
    As[tn][tm] = (m0 + tm < M && k < K) ? Ab[k * M + m0 + tm] : make_float2(0.f, 0.f);
    Bs[tn][tm] = (n0 + tm < N && k < K) ? Bb[k * N + n0 + tm] : make_float2(0.f, 0.f);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      const float2 a = As[kk][tm], bb = Bs[kk][tn];
      cr = fmaf(a.x, bb.x, fmaf(a.y, bb.y, cr));
      ci = fmaf(a.y, bb.x, fmaf(-a.x, bb.y, ci));
    }
    __syncthreads();
  }
  if (m0 + tm < M && n0 + tn < N) C[b * M * N + (n0 + tn) * M + m0 + tm] = make_float2(cr, ci);
}

}  // namespace alb

using namespace alb;

extern "C" {

int alb_xcorr_all_ant_strided(const void* x, int64_t M, int64_t K, int64_t F, int64_t S, int64_t sk, int64_t sf,
                              int64_t ss, void* out, float scale, int mode, void* stream);

// correlations_gpu.py:95-121 avg_xcorr_all_ant_gpu (split = 1): out = x x^H * scale
// per channel; scale is normally fl32(1/K).
int alb_xcorr_all_ant(const void* x, int64_t M, int64_t K, int64_t F, void* out, float scale, void* stream) {
  ALB_REQUIRE(x && out, "alb_xcorr_all_ant: null pointer");
  ALB_REQUIRE(M >= 1 && K >= 0 && F >= 0, "alb_xcorr_all_ant: bad sizes");
  if (F == 0) return ALB_OK;
  return alb_xcorr_all_ant_strided(x, M, K, F, 1, M, M * K, 0, out, scale, PREC_FP32, stream);
}

// The same product in a chosen arithmetic: 0 FP32 (as alb_xcorr_all_ant), 1 BF16 hi/lo
// (three products), 2 FP16, 3 FP16 hi/lo; all accumulate in FP32 on the matrix
// cores.  Modes 1-3 take M <= 8.
int alb_xcorr_all_ant_prec(const void* x, int64_t M, int64_t K, int64_t F, void* out, float scale, int mode,
                           void* stream) {
  ALB_REQUIRE(mode >= 0 && mode <= 3, "alb_xcorr_all_ant_prec: mode must be 0..3");
  if (mode == PREC_FP32) return alb_xcorr_all_ant(x, M, K, F, out, scale, stream);
  return alb_xcorr_all_ant_strided(x, M, K, F, 1, M, M * K, 0, out, scale, mode, stream);
}

// Strided input: job s, channel f, time k, input m at x[s*ss + f*sf + k*sk + m]; out as
// alb_xcorr_all_ant over F*S channels (job-major), M <= 8.
int alb_xcorr_all_ant_strided(const void* x, int64_t M, int64_t K, int64_t F, int64_t S, int64_t sk, int64_t sf,
                              int64_t ss, void* out, float scale, int mode, void* stream) {
  ALB_REQUIRE(mode >= 0 && mode <= 3, "alb_xcorr_all_ant_strided: mode must be 0..3");
  ALB_REQUIRE(x && out, "alb_xcorr_all_ant_strided: null pointer");
  ALB_REQUIRE(M >= 1 && K >= 0 && F >= 0 && S >= 1 && sk >= M && sf >= 0 && ss >= 0,
              "alb_xcorr_all_ant_strided: bad sizes");
  if (M > 8 && mode != PREC_FP32) {
    set_error("alb_xcorr_all_ant_strided: the reduced-precision forms take M <= 8 (got %lld)", (long long)M);
    return ALB_ENOTSUP;
  }
  if (F == 0) return ALB_OK;
  const int64_t nch = F * S;
  if (M > 8) {
    ALB_REQUIRE(M * (M + 1) / 2 <= 4096, "alb_xcorr_all_ant_strided: M too large (at most 90 inputs)");
    ALB_REQUIRE(nch <= 0x7fffffff, "alb_xcorr_all_ant_strided: too many channels");
    const XLayout L{sk, sf, ss, F};
    hipLaunchKernelGGL(xcorr_valu_kernel, dim3(unsigned(nch)), dim3(256), sizeof(float2) * size_t(kVT) * size_t(M),
                       as_stream(stream), (const float2*)x, int(M), K, L, (float2*)out, scale);
    ALB_LAUNCH_CHECK("xcorr_valu_kernel");
    return ALB_OK;
  }
  const int64_t cpw = M <= 2 ? 4 : (M <= 4 ? 2 : 1);       // channels per wavefront (XPack)
  const int64_t nwave = (nch + cpw - 1) / cpw;
  ALB_REQUIRE((nwave + 3) / 4 <= 0x7fffffff, "alb_xcorr_all_ant_strided: too many channels");
  hipStream_t s = as_stream(stream);
  const dim3 grid(unsigned((nwave + 3) / 4));
  const XLayout L{sk, sf, ss, F};
  if (mode == PREC_FP32) {
    switch (M) {
#define CASE(m) case m: hipLaunchKernelGGL(xcorr_mfma16_kernel<m>, grid, dim3(256), 0, s, (const float*)x, K, nch, L, (float2*)out, scale); break;
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
      default: break;
    }
  } else {
#define CASE(m, md) case m: hipLaunchKernelGGL((xcorr_mfma16_lp_kernel<m, md>), grid, dim3(256), 0, s, (const float*)x, K, nch, L, (float2*)out, scale); break;
#define MODES(md) switch (M) { CASE(1, md) CASE(2, md) CASE(3, md) CASE(4, md) CASE(5, md) CASE(6, md) CASE(7, md) CASE(8, md) default: break; }
    if (mode == PREC_BF16X3) { MODES(PREC_BF16X3) }
    else if (mode == PREC_FP16) { MODES(PREC_FP16) }
    else { MODES(PREC_FP16X3) }
#undef MODES
#undef CASE
  }
  ALB_LAUNCH_CHECK("xcorr_all_ant_strided");
  return ALB_OK;
}

// correlations.py:115-151 avg_xcorr_all_ant_cpu (FP32 products, float64 sums and
// division by K, complex64 out); strides in complex elements.
int alb_xcorr_all_ant_f64acc(const void* x, int64_t M, int64_t K, int64_t F, int64_t sm, int64_t sk, int64_t sf,
                             void* out, int64_t oi, int64_t oj, int64_t of, void* stream) {
  ALB_REQUIRE(M >= 1 && K >= 0 && F >= 0, "alb_xcorr_all_ant_f64acc: bad sizes");
  ALB_REQUIRE(M * (M + 1) / 2 <= 65535, "alb_xcorr_all_ant_f64acc: M too large (at most 361 inputs)");
  if (F == 0) return ALB_OK;
  ALB_REQUIRE(x && out, "alb_xcorr_all_ant_f64acc: null pointer");
  ALB_REQUIRE((F + 255) / 256 <= 0x7fffffff, "alb_xcorr_all_ant_f64acc: too many channels");
  const dim3 grid(unsigned((F + 255) / 256), unsigned(M * (M + 1) / 2));
  hipLaunchKernelGGL(xcorr_f64acc_kernel, grid, dim3(256), 0, as_stream(stream), (const float2*)x, int(M), K, F, sm,
                     sk, sf, (float2*)out, oi, oj, of);
  ALB_LAUNCH_CHECK("xcorr_f64acc_kernel");
  return ALB_OK;
}

// cgemm_batch.cu:8-54 cgemm_strided_batched: C = A B^H, column-major batches.
int alb_cgemm_strided_batched(const void* A, const void* B, void* C, int64_t M, int64_t N, int64_t K,
                              int64_t batch_count, void* stream) {
  ALB_REQUIRE(M >= 0 && N >= 0 && K >= 0 && batch_count >= 0, "alb_cgemm_strided_batched: bad sizes");
  if (M == 0 || N == 0 || batch_count == 0) return ALB_OK;
  ALB_REQUIRE(A && B && C, "alb_cgemm_strided_batched: null pointer");
  ALB_REQUIRE((M + 15) / 16 <= 65535 && (N + 15) / 16 <= 65535 && batch_count <= 0x7fffffff,
              "alb_cgemm_strided_batched: shape too large for one launch");
  const dim3 grid(unsigned(batch_count), unsigned((M + 15) / 16), unsigned((N + 15) / 16));
  hipLaunchKernelGGL(cgemm_abh_kernel, grid, dim3(256), 0, as_stream(stream), (const float2*)A, (const float2*)B,
                     (float2*)C, M, N, K);
  ALB_LAUNCH_CHECK("cgemm_abh_kernel");
  return ALB_OK;
}

}  // extern "C"
