// Device FFT building blocks for gfx950.
//
//  * DFT<R>        : in-register forward DFT of R = 1..64 points (recursive DIT,
//                    compile-time twiddles, trivial rotations specialised).
//  * fft_lds<N,F,T>: forward complex FFT of F batched sequences of length N held
//                    in LDS (padded layout, one float2 of pad per 16), by T
//                    threads, as radix-16 Stockham passes (natural order in
//                    and out).  Inverse transforms conjugate on load and store.
//  * Two-level twiddles W_N^m = A[m >> lb] * B[m & mask] read from LDS.
//
// The forward sign is exp(-2 pi i n k / N) (numpy / pycufft convention,
// pycufft.py:92-161); no normalisation is applied here.
#pragma once
#include <type_traits>
#include "common.hpp"

namespace alb {

// ------------------------------------------------------- compile-time trig
struct SinCos { double s, c; };
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double cx_sin_t(double x) {
  double x2 = x * x, term = x, sum = x;
  for (int n = 1; n < 14; ++n) { term *= -x2 / double((2 * n) * (2 * n + 1)); sum += term; }
  return sum;
}
constexpr double cx_cos_t(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int n = 1; n < 14; ++n) { term *= -x2 / double((2 * n - 1) * (2 * n)); sum += term; }
  return sum;
}
// sin / cos of 2 pi m / N with exact quadrant reduction.
constexpr SinCos cx_sincos2pi(long m, long N) {
  m %= N;
  if (m < 0) m += N;
  long q = (4 * m) / N;
  long r4 = 4 * m - q * N;                       // angle in quadrant = (pi/2) r4 / N
  double a = (kPi / 2) * double(r4) / double(N);
  double sa = 0, ca = 0;
  if (2 * r4 <= N) { sa = cx_sin_t(a); ca = cx_cos_t(a); }
  else { double b = kPi / 2 - a; sa = cx_cos_t(b); ca = cx_sin_t(b); }
  switch (q) {
    case 0: return SinCos{sa, ca};
    case 1: return SinCos{ca, -sa};
    case 2: return SinCos{-sa, -ca};
    default: return SinCos{-ca, sa};
  }
}

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// v * W_N^M with W_N = exp(-2 pi i / N), trivial rotations specialised.
template <int N, int M>
__device__ __forceinline__ float2 twc(float2 v) {
  constexpr int m = ((M % N) + N) % N;
  constexpr float h = 0.70710678118654752440f;
  if constexpr (m == 0) return v;
  else if constexpr (4 * m == N) return make_float2(v.y, -v.x);
  else if constexpr (2 * m == N) return make_float2(-v.x, -v.y);
  else if constexpr (4 * m == 3 * N) return make_float2(-v.y, v.x);
  else if constexpr (8 * m == N) return make_float2(h * (v.x + v.y), h * (v.y - v.x));
  else if constexpr (8 * m == 3 * N) return make_float2(h * (v.y - v.x), -h * (v.x + v.y));
  else if constexpr (8 * m == 5 * N) return make_float2(-h * (v.x + v.y), h * (v.x - v.y));
  else if constexpr (8 * m == 7 * N) return make_float2(h * (v.x - v.y), h * (v.x + v.y));
  else {
    constexpr SinCos sc = cx_sincos2pi(m, N);
    constexpr float c = float(sc.c);
    constexpr float s = float(-sc.s);
    return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
  }
}

template <int R>
struct DFT {
  static __device__ __forceinline__ void run(float2* v) {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) { e[i] = v[2 * i]; o[i] = v[2 * i + 1]; }
    DFT<R / 2>::run(e);
    DFT<R / 2>::run(o);
    static_for<0, R / 2>([&](auto K) {
      constexpr int k = decltype(K)::value;
      float2 t = twc<R, k>(o[k]);
      v[k] = cadd(e[k], t);
      v[k + R / 2] = csub(e[k], t);
    });
  }
};
template <>
struct DFT<1> {
  static __device__ __forceinline__ void run(float2*) {}
};
template <>
struct DFT<2> {
  static __device__ __forceinline__ void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

// ----------------------------------------------------------- LDS layout
__device__ __forceinline__ int pad16(int i) { return i + (i >> 4); }
template <int N>
struct Padded { static constexpr int value = N + N / 16; };
// W_256^(m r) table of the radix-16 pass at sub-transform size 16 (fill_w16)
constexpr int kW16Row = 17;
constexpr int kW16Size = 16 * kW16Row;

// Twiddle W_N^m from the two-level LDS tables.
__device__ __forceinline__ float2 tw_lookup(const float2* A, const float2* B, int lb, int m) {
  return cmul(A[m >> lb], B[m & ((1 << lb) - 1)]);
}

// Copy a two-level twiddle table from global memory into LDS.
__device__ __forceinline__ void load_twiddle(const Twiddle& tw, float2* A, float2* B, int tid, int nthreads) {
  for (int i = tid; i < tw.nA; i += nthreads) A[i] = tw.A[i];
  for (int i = tid; i < tw.nB; i += nthreads) B[i] = tw.B[i];
}

// Padded offset of element i + k (k a compile-time step) from the padded
// offset of i, when the pad cannot change between them: K a multiple of 16
// (pad16(i + K) = pad16(i) + 17 K / 16), or i a multiple of 16 and K < 16.
template <int K>
__device__ __forceinline__ constexpr int pad_step() { return K % 16 == 0 ? K + K / 16 : K; }

// One radix-R Stockham pass of F FFTs of length N (sub-transform size NS).
// Butterfly (f, j) reads j + r NB and writes d + r NS, d = (j / NS) NS R + j % NS.
// With NB (and NS) multiples of 16, or NS = 1 (d a multiple of R, R <= 16),
// every padded address is one base plus compile-time offsets.
// TWC 1: also keep this pass's twiddles w[1..R-1] in twc; TWC 2: take them from
// twc (a later transform of the same size on the same thread mapping, one
// butterfly per thread).
template <int N, int R, int NS, int F, int T, bool TAB = false, int TWC = 0, int RS = Padded<N>::value>
__device__ __forceinline__ void fft_pass(float2* lds, const float2* A, const float2* B, int lb, int tsh, int tid,
                                         const float2* W16 = nullptr, float2* twc = nullptr) {
  constexpr int NB = N / R;
  constexpr int TOT = NB * F;
  constexpr int PER = (TOT + T - 1) / T;
  constexpr int NP = RS;   // row stride of the F sequences (>= the padded length)
  constexpr bool RD_FAST = NB % 16 == 0;
  constexpr bool WR_FAST = NS % 16 == 0 || (NS == 1 && 16 % R == 0);
  float2 v[PER][R];
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const unsigned b = unsigned(tid) + unsigned(t * T);
    if (TOT % T == 0 || b < unsigned(TOT)) {
      const unsigned f = b / NB, j = b % NB;
      const float2* base = lds + f * NP;
      if constexpr (RD_FAST) {
        const float2* bj = base + pad16(int(j));
#pragma unroll
        for (int r = 0; r < R; ++r) v[t][r] = bj[r * pad_step<NB>()];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) v[t][r] = base[pad16(int(j) + r * NB)];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const unsigned b = unsigned(tid) + unsigned(t * T);
    if (TOT % T == 0 || b < unsigned(TOT)) {
      const unsigned f = b / NB, j = b % NB;
      if constexpr (TAB && NS == 16 && R == 16) {
        // sub-transform size 16: W_N^(r (j%16) N/256) = W_256^(r m), m = j % 16,
        // from the table W16[17 m + r]
        const float2* wt = W16 + (j % 16) * kW16Row;
#pragma unroll
        for (int r = 1; r < R; ++r) v[t][r] = cmul(v[t][r], wt[r]);
      } else if constexpr (NS > 1 && TWC == 2) {
        static_assert(T % NB == 0, "cached twiddles need every butterfly of a thread at one j");
#pragma unroll
        for (int r = 1; r < R; ++r) v[t][r] = cmul(v[t][r], twc[r - 1]);
      } else if constexpr (NS > 1) {
        // W^(r*base) for r = 1..R-1 from the looked-up powers W^base, W^2base,
        // W^4base, W^8base (products of at most four factors)
        constexpr int STEP = N / (NS * R);
        const int base = int(((j % NS) * STEP) << tsh);
        float2 wp[4];
#pragma unroll
        for (int e = 0; (1 << e) < R; ++e) wp[e] = tw_lookup(A, B, lb, base << e);
        float2 w[R];
        w[0] = make_float2(1.f, 0.f);
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const int hi = 31 - __builtin_clz(r);  // highest set bit (compile time after unrolling)
          w[r] = (r == (1 << hi)) ? wp[hi] : cmul(w[r - (1 << hi)], wp[hi]);
        }
        if constexpr (TWC == 1) {
          static_assert(PER == 1, "storing cached twiddles needs one butterfly per thread");
#pragma unroll
          for (int r = 1; r < R; ++r) twc[r - 1] = w[r];
        }
#pragma unroll
        for (int r = 1; r < R; ++r) v[t][r] = cmul(v[t][r], w[r]);
      }
      DFT<R>::run(v[t]);
      const int d = int((j / NS) * NS * R + (j % NS));
      float2* base = lds + f * NP;
      if constexpr (WR_FAST) {
        float2* bd = base + pad16(d);
#pragma unroll
        for (int r = 0; r < R; ++r) bd[r * pad_step<NS>()] = v[t][r];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) base[pad16(d + r * NS)] = v[t][r];
      }
    }
  }
  __syncthreads();
}

template <int REM>
struct PickRadix {
  static constexpr int value = REM >= 16 ? 16 : REM;
};

// Forward FFT (unnormalised, natural order) of F sequences of length N in LDS.
// The twiddle tables may belong to a larger power-of-two size N << tsh.
// Passes run while the sub-transform size is below NSE (NSE < N stops early:
// the caller finishes the last pass itself, e.g. for only some output bins).
// The caller synchronises before the call; the call ends with a barrier.
// With TAB the radix-16 pass at sub-transform size 16 takes its twiddles from
// W16 (fill_w16).  TWC (see fft_pass) applies to the pass at sub-transform
// size 256.
template <int N, int F, int T, int NS = 1, int NSE = N, bool TAB = false, int TWC = 0, int RS = Padded<N>::value>
__device__ __forceinline__ void fft_lds(float2* lds, const float2* A, const float2* B, int lb, int tsh, int tid,
                                        const float2* W16 = nullptr, float2* twc = nullptr) {
  if constexpr (NS < NSE) {
    constexpr int R = PickRadix<N / NS>::value;
    fft_pass<N, R, NS, F, T, TAB, NS == 256 ? TWC : 0, RS>(lds, A, B, lb, tsh, tid, W16, twc);
    fft_lds<N, F, T, NS * R, NSE, TAB, TWC, RS>(lds, A, B, lb, tsh, tid, W16, twc);
  }
}

// Twiddles w[r - 1] = W^(r base), r = 1..R-1, of the radix-R pass at
// sub-transform size NS for butterfly index j = tid % (N / R) (the values that
// fft_pass computes), for use with TWC = 2.
template <int N, int R, int NS>
__device__ __forceinline__ void pass_twiddles(const float2* A, const float2* B, int lb, int tsh, int tid, float2* twc) {
  constexpr int NB = N / R;
  constexpr int STEP = N / (NS * R);
  const int base = int(((unsigned(tid) % NB % NS) * STEP) << tsh);
  float2 wp[4];
#pragma unroll
  for (int e = 0; (1 << e) < R; ++e) wp[e] = tw_lookup(A, B, lb, base << e);
  float2 w[R];
  w[0] = make_float2(1.f, 0.f);
#pragma unroll
  for (int r = 1; r < R; ++r) {
    const int hi = 31 - __builtin_clz(r);
    w[r] = (r == (1 << hi)) ? wp[hi] : cmul(w[r - (1 << hi)], wp[hi]);
  }
#pragma unroll
  for (int r = 1; r < R; ++r) twc[r - 1] = w[r];
}

// W16[17 m + r] = W_256^(m r) (m, r < 16; rows of 17 so that the 16 rows
// start in different LDS banks) from the two-level tables of a power-of-two
// size NT >= 256 (W_256^e = W_NT^(e NT / 256)).  kW16Size entries; the caller
// synchronises before the first use.
__device__ __forceinline__ void fill_w16(float2* W16, const float2* A, const float2* B, int lb, int NT, int tid,
                                         int nthreads) {
  for (int i = tid; i < 256; i += nthreads) {
    const int m = i >> 4, r = i & 15;
    W16[m * kW16Row + r] = tw_lookup(A, B, lb, ((m * r) & 255) * (NT / 256));
  }
}

}  // namespace alb
