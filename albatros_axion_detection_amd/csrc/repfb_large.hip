// Re-PFB (pfb_gpu_utils.py:105-121 with lblock L = 4096*R1) for R1 = osamp of 128
// up to 65536, the upchannelization of the reference's production runs
// (scripts/xcorr/config_axion_gpu.json:13, osamp 65536).
//
// A re-PFB block b is y[r][k] (r < R1 rows of 4096 columns), the polyphase sum
// y[r][k] = sum_j ts[(b + j) R1 + r][k] win[j R1 + r][k].  With L = 4096 R1,
// n = 4096 r + k and fine channel f = q + R1 f2 (q < R1):
//     Y[f] = sum_k X_k[q] W_L^(f k),   X_k[q] = sum_r y[r][k] W_R1^(q r).
// Stage 1 computes X for every column as R1-point FFTs down the columns, two
// real columns at a time (z = y[.][2p] + i y[.][2p+1], p < 2048), in two passes
// of R1 = P1 P2 (r = r1 + P1 r2, q = P2 q1 + q2):
//     pass A: B[r1 P2 + q2][p] = W_R1^(q2 r1) sum_r2 z[r1 + P1 r2][p] W_P2^(q2 r2)
//     pass B: Z[q1 P2 + q2][p] = sum_r1 B[r1 P2 + q2][p] W_P1^(q1 r1)
// Pass B reads and writes the same rows {x P2 + q2} of one workgroup, so it
// runs in place: the workspace holds Z[b][q][p] (q < R1, natural order),
// complex64, R1 * 2048 per block -- the size alb_repfb_workspace_bytes gives.
// Stage 2 is band-limited: the kept channels are few per q (one coarse channel
// gives one per q), so each is a direct 4096-term sum over the columns, with
// X_2p[q] = (Z[q][p] + conj Z[R1-q][p]) / 2 and
// X_2p+1[q] = (Z[q][p] - conj Z[R1-q][p]) / 2i; one workgroup takes q and its
// mirror R1 - q, which read the same two rows.  Twiddles W_L^(f k) come from
// float64 sincospi of the exactly reduced exponent (f k) mod L.
#include "fft.hpp"

namespace alb {

constexpr int kLT = 256;                 // threads per workgroup
constexpr int kLP = 16;                  // column pairs per workgroup of the two column passes
constexpr int kLStrips = 2048 / kLP;
constexpr int kLRB = 32;                 // outputs reduced per round in stage 2

template <int P2>
__global__ __launch_bounds__(kLT) void repfb_large_a_kernel(const float* __restrict__ ts,
                                                            const float* __restrict__ win, int ntap, int64_t R1,
                                                            int P1, float2* __restrict__ Z, Twiddle twp,
                                                            Twiddle twr) {
  constexpr int NP = Padded<P2>::value;
  __shared__ float2 buf[kLP * NP];
  __shared__ float2 TA[16], TB[16];
  const int tid = threadIdx.x;
  load_twiddle(twp, TA, TB, tid, kLT);
  const int strip = blockIdx.x, r1 = blockIdx.y;
  const int64_t b = blockIdx.z;
  const int64_t col0 = int64_t(2) * strip * kLP;
  // consecutive lanes take consecutive column pairs of one row (128-byte runs)
  for (int e = tid; e < P2 * kLP; e += kLT) {
    const int r2 = e / kLP, pp = e % kLP;
    const int64_t r = r1 + int64_t(P1) * r2;
    const int64_t col = col0 + 2 * pp;
    float re = 0.f, im = 0.f;
    for (int j = 0; j < ntap; ++j) {
      const float2 t = *reinterpret_cast<const float2*>(ts + ((b + j) * R1 + r) * 4096 + col);
      const float2 w = *reinterpret_cast<const float2*>(win + (int64_t(j) * R1 + r) * 4096 + col);
      re = fmaf(t.x, w.x, re);
      im = fmaf(t.y, w.y, im);
    }
    buf[pp * NP + pad16(r2)] = make_float2(re, im);
  }
  __syncthreads();
  fft_lds<P2, kLP, kLT>(buf, TA, TB, twp.lb, 0, tid);
  float2* zb = Z + b * R1 * 2048 + col0 / 2;
  for (int e = tid; e < P2 * kLP; e += kLT) {
    const int q2 = e / kLP, pp = e % kLP;
    float2 v = buf[pp * NP + pad16(q2)];
    if (P1 > 1) v = cmul(v, tw_lookup(twr.A, twr.B, twr.lb, int((int64_t(q2) * r1) & (R1 - 1))));
    zb[(int64_t(r1) * P2 + q2) * 2048 + pp] = v;
  }
}

template <int P1>
__global__ __launch_bounds__(kLT) void repfb_large_b_kernel(float2* __restrict__ Z, int64_t R1, int P2, Twiddle twp) {
  constexpr int NP = Padded<P1>::value;
  __shared__ float2 buf[kLP * NP];
  __shared__ float2 TA[16], TB[16];
  const int tid = threadIdx.x;
  load_twiddle(twp, TA, TB, tid, kLT);
  const int strip = blockIdx.x, q2 = blockIdx.y;
  const int64_t b = blockIdx.z;
  float2* zb = Z + b * R1 * 2048 + int64_t(strip) * kLP;
  for (int e = tid; e < P1 * kLP; e += kLT) {
    const int r1 = e / kLP, pp = e % kLP;
    buf[pp * NP + pad16(r1)] = zb[(int64_t(r1) * P2 + q2) * 2048 + pp];
  }
  __syncthreads();
  fft_lds<P1, kLP, kLT>(buf, TA, TB, twp.lb, 0, tid);
  for (int e = tid; e < P1 * kLP; e += kLT) {
    const int q1 = e / kLP, pp = e % kLP;
    zb[(int64_t(q1) * P2 + q2) * 2048 + pp] = buf[pp * NP + pad16(q1)];
  }
}

// W_L^e for an exponent e already reduced into [0, L), in float64.
__device__ __forceinline__ double2 wl64(int64_t e, int64_t L) {
  double s, c;
  sincospi(-2.0 * double(e) / double(L), &s, &c);
  return make_double2(c, s);
}

__device__ __forceinline__ float2 wave_sum(float2 v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    v.x += __shfl_xor(v.x, off);
    v.y += __shfl_xor(v.y, off);
  }
  return v;
}

// Workgroup (q <= R1/2, b, input): the kept channels f = q + R1 f2 and
// f = (R1 - q) + R1 f2 in [c_lo, c_hi), each a 4096-term sum, written to
// out[m sm + (b/bpj) sj + (b%bpj) sb + (f - c_lo) sc].
__global__ __launch_bounds__(kLT) void repfb_large_stage2_kernel(const float2* __restrict__ Z, int64_t a_stride,
                                                                 int64_t R1, int64_t c_lo, int64_t c_hi,
                                                                 float2* __restrict__ out, int64_t sm, int64_t sb,
                                                                 int64_t sc, int64_t bpj, int64_t sj) {
  __shared__ float2 X[2][4096];
  __shared__ float2 red[kLT / 64][kLRB];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t q = blockIdx.x;
  const int64_t b = blockIdx.y;
  const int64_t m_in = blockIdx.z;
  const int64_t qm = (R1 - q) & (R1 - 1);
  const int64_t L = 4096 * R1;
  const float2* zq = Z + m_in * a_stride + (b * R1 + q) * 2048;
  const float2* zm = Z + m_in * a_stride + (b * R1 + qm) * 2048;
  for (int p = tid; p < 2048; p += kLT) {
    const float2 a = zq[p], m = zm[p];
    X[0][2 * p] = make_float2(0.5f * (a.x + m.x), 0.5f * (a.y - m.y));
    X[0][2 * p + 1] = make_float2(0.5f * (a.y + m.y), -0.5f * (a.x - m.x));
    X[1][2 * p] = make_float2(0.5f * (m.x + a.x), 0.5f * (m.y - a.y));
    X[1][2 * p + 1] = make_float2(0.5f * (m.y + a.y), -0.5f * (m.x - a.x));
  }
  __syncthreads();
  float2* ob = out + m_in * sm + (b / bpj) * sj + (b % bpj) * sb;
  const int nsets = qm != q ? 2 : 1;
  for (int set = 0; set < nsets; ++set) {
    const int64_t qq = set ? qm : q;
    // f2 with c_lo <= qq + R1 f2 < c_hi
    const int64_t f2lo = c_lo > qq ? (c_lo - qq + R1 - 1) / R1 : 0;
    const int64_t f2hi = c_hi - 1 - qq >= 0 ? (c_hi - 1 - qq) / R1 : -1;
    const float2* xs = X[set];
    for (int64_t f2a = f2lo; f2a <= f2hi; f2a += kLRB) {
      const int nb = int(f2hi - f2a + 1 < kLRB ? f2hi - f2a + 1 : kLRB);
      for (int i = 0; i < nb; ++i) {
        const int64_t f = qq + R1 * (f2a + i);
        // this thread's columns k = tid + 256 j: W_L^(f k) = W_L^(f tid) (W_L^(256 f))^j
        double2 w = wl64((f * tid) & (L - 1), L);
        const double2 st = wl64((f * 256) & (L - 1), L);
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const float2 x = xs[tid + 256 * j];
          const float wr = float(w.x), wi = float(w.y);
          acc.x = fmaf(x.x, wr, fmaf(-x.y, wi, acc.x));
          acc.y = fmaf(x.x, wi, fmaf(x.y, wr, acc.y));
          w = make_double2(w.x * st.x - w.y * st.y, w.x * st.y + w.y * st.x);
        }
        acc = wave_sum(acc);
        if (lane == 0) red[wv][i] = acc;
      }
      __syncthreads();
      if (tid < nb) {
        float2 s = red[0][tid];
#pragma unroll
        for (int u = 1; u < kLT / 64; ++u) s = cadd(s, red[u][tid]);
        ob[(qq + R1 * (f2a + tid) - c_lo) * sc] = s;
      }
      __syncthreads();
    }
  }
}

int repfb_large_stage1(const float* ts, const float* win, int ntap, int64_t nblock, int64_t R1, float2* Z,
                       hipStream_t s) {
  ALB_REQUIRE(is_pow2(R1) && R1 >= 128 && R1 <= 65536, "alb_repfb: osamp %lld not supported (power of two up to 65536)",
              (long long)R1);
  ALB_REQUIRE(nblock >= 0 && nblock <= 65535, "alb_repfb: at most 65535 blocks per call at osamp %lld",
              (long long)R1);
  if (nblock == 0) return ALB_OK;
  const int P2 = R1 >= 256 ? 256 : int(R1);
  const int P1 = int(R1 / P2);
  Twiddle twp2{}, twr{}, twp1{};
  int rc = get_twiddle(P2, &twp2, s);
  if (rc) return rc;
  if (P1 > 1) {
    rc = get_twiddle(R1, &twr, s);
    if (rc) return rc;
    rc = get_twiddle(P1, &twp1, s);
    if (rc) return rc;
  }
  const dim3 ga(kLStrips, unsigned(P1), unsigned(nblock));
  if (P2 == 256)
    hipLaunchKernelGGL(repfb_large_a_kernel<256>, ga, dim3(kLT), 0, s, ts, win, ntap, R1, P1, Z, twp2, twr);
  else
    hipLaunchKernelGGL(repfb_large_a_kernel<128>, ga, dim3(kLT), 0, s, ts, win, ntap, R1, P1, Z, twp2, twr);
  ALB_LAUNCH_CHECK("repfb_large_a_kernel");
  if (P1 == 1) return ALB_OK;
  const dim3 gb(kLStrips, unsigned(P2), unsigned(nblock));
  switch (P1) {
#define CASE(p) case p: hipLaunchKernelGGL(repfb_large_b_kernel<p>, gb, dim3(kLT), 0, s, Z, R1, P2, twp1); break;
    CASE(2) CASE(4) CASE(8) CASE(16) CASE(32) CASE(64) CASE(128) CASE(256)
#undef CASE
    default:
      set_error("alb_repfb: osamp %lld not supported", (long long)R1);
      return ALB_ENOTSUP;
  }
  ALB_LAUNCH_CHECK("repfb_large_b_kernel");
  return ALB_OK;
}

int repfb_large_stage2(const float2* Z, int64_t a_stride, int64_t ninput, int64_t nblock, int64_t R1, int64_t c_lo,
                       int64_t c_hi, float2* out, int64_t sm, int64_t sb, int64_t sc, int64_t bpj, int64_t sj,
                       hipStream_t s) {
  ALB_REQUIRE(is_pow2(R1) && R1 >= 128 && R1 <= 65536, "alb_repfb: osamp %lld not supported (power of two up to 65536)",
              (long long)R1);
  ALB_REQUIRE(nblock <= 65535 && ninput <= 65535, "alb_repfb_stage2: at most 65535 blocks and inputs per call");
  if (nblock == 0) return ALB_OK;
  const dim3 g(unsigned(R1 / 2 + 1), unsigned(nblock), unsigned(ninput));
  hipLaunchKernelGGL(repfb_large_stage2_kernel, g, dim3(kLT), 0, s, Z, a_stride, R1, c_lo, c_hi, out, sm, sb, sc, bpj,
                     sj);
  ALB_LAUNCH_CHECK("repfb_large_stage2_kernel");
  return ALB_OK;
}

}  // namespace alb
