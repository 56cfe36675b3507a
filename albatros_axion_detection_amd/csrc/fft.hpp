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

// Twiddle W_N^m from the two-level LDS tables.
__device__ __forceinline__ float2 tw_lookup(const float2* A, const float2* B, int lb, int m) {
  return cmul(A[m >> lb], B[m & ((1 << lb) - 1)]);
}

// Copy a two-level twiddle table from global memory into LDS.
__device__ __forceinline__ void load_twiddle(const Twiddle& tw, float2* A, float2* B, int tid, int nthreads) {
  for (int i = tid; i < tw.nA; i += nthreads) A[i] = tw.A[i];
  for (int i = tid; i < tw.nB; i += nthreads) B[i] = tw.B[i];
}

// One radix-R Stockham pass of F FFTs of length N (sub-transform size NS).
template <int N, int R, int NS, int F, int T>
__device__ __forceinline__ void fft_pass(float2* lds, const float2* A, const float2* B, int lb, int tsh, int tid) {
  constexpr int NB = N / R;
  constexpr int TOT = NB * F;
  constexpr int PER = (TOT + T - 1) / T;
  constexpr int NP = Padded<N>::value;
  float2 v[PER][R];
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int b = tid + t * T;
    if (TOT % T == 0 || b < TOT) {
      const int f = b / NB, j = b % NB;
      const float2* base = lds + f * NP;
#pragma unroll
      for (int r = 0; r < R; ++r) v[t][r] = base[pad16(j + r * NB)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int b = tid + t * T;
    if (TOT % T == 0 || b < TOT) {
      const int f = b / NB, j = b % NB;
      if constexpr (NS > 1) {
        // W^(r*base) for r = 1..R-1 from the looked-up powers W^base, W^2base,
        // W^4base, W^8base (products of at most four factors)
        constexpr int STEP = N / (NS * R);
        const int base = ((j % NS) * STEP) << tsh;
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
#pragma unroll
        for (int r = 1; r < R; ++r) v[t][r] = cmul(v[t][r], w[r]);
      }
      DFT<R>::run(v[t]);
      const int d = (j / NS) * NS * R + (j % NS);
      float2* base = lds + f * NP;
#pragma unroll
      for (int r = 0; r < R; ++r) base[pad16(d + r * NS)] = v[t][r];
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
template <int N, int F, int T, int NS = 1, int NSE = N>
__device__ __forceinline__ void fft_lds(float2* lds, const float2* A, const float2* B, int lb, int tsh, int tid) {
  if constexpr (NS < NSE) {
    constexpr int R = PickRadix<N / NS>::value;
    fft_pass<N, R, NS, F, T>(lds, A, B, lb, tsh, tid);
    fft_lds<N, F, T, NS * R, NSE>(lds, A, B, lb, tsh, tid);
  }
}

}  // namespace alb
