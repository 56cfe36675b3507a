// Polyphase filter bank kernels.
//
//  matft_filter   W[k,f] = sum_j w[j L + k] exp(-2 pi i f j / nslice) and the
//                 Wiener filter W (1+t^2)/(|W|^2+t^2) in float64
//                 (pfb_gpu_utils.py:29-36, :123-137).
//  irfft_rows     row inverse real FFT (N = 2(nchan-1)), 1/N scaling, DC and
//                 Nyquist imaginary parts ignored (pycufft.py:139-161).  Loaders:
//                 dense complex rows, or 4-bit packed bytes with the spec_num
//                 zero fill and channel scatter fused in
//                 (baseband_data_classes.py:48-73, unpacking.c:155-205).
//  rfft_rows      row forward real FFT (pycufft.py:120-137), optionally with a
//                 polyphase window fused in front (pfb_gpu_utils.py:105-121).
//  ipfb_cols      column pass of the inverse PFB: along time, forward real FFT,
//                 multiply by the filter, inverse real FFT, 1/nslice scaling,
//                 rows [cut, nslice-cut) written to the output stream
//                 (pfb_gpu_utils.py:84-91, helper_gpu_stream.py:199).
//  repfb_stage1   re-PFB, four-step first half: polyphase window + R1-point real
//                 DFT down the R1 rows of each block, per column.
//  repfb_stage2   re-PFB second half: twiddle, 4096-point FFT along columns,
//                 only the requested fine channels written.
#include <atomic>
#include "fft.hpp"
#include "packed.hpp"

namespace alb {

constexpr int kTPB = 256;
constexpr int kBlockListCap = 64;

// ----------------------------------------------------------- matft / filter
__global__ void matft_filter_kernel(const float* __restrict__ win, int ntap, int64_t lblock, int64_t nslice, int64_t nf,
                                    double thresh, float2* __restrict__ matft, float2* __restrict__ filt) {
  const int64_t n = lblock * nf;
  for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < n; idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t k = idx / nf, f = idx - k * nf;
    double wr = 0.0, wi = 0.0;
    for (int j = 0; j < ntap; ++j) {
      const double w = double(win[int64_t(j) * lblock + k]);
      const int64_t m = (f * j) % nslice;
      double s, c;
      sincospi(-2.0 * double(m) / double(nslice), &s, &c);
      wr += w * c;
      wi += w * s;
    }
    if (matft) matft[idx] = make_float2(float(wr), float(wi));
    if (filt) {
      const double a = wr * wr + wi * wi;
      double sc;
      if (thresh > 0.0) sc = (1.0 + thresh * thresh) / (a + thresh * thresh);
      else sc = 1.0 / a;
      filt[idx] = make_float2(float(wr * sc), float(wi * sc));
    }
  }
}

// ------------------------------------------------------------- row loaders
// Dense complex rows: X[r][c] for c <= M.
struct DenseRows {
  const float2* x;
  int64_t ld;
};


// -------------------------------------------------------- irfft along rows
// F rows per workgroup, N = 2M real outputs per row.  The packed half spectrum
// Z[m] = (X[m] + conj X[M-m]) + i conj(W_N^m) (X[m] - conj X[M-m]) is formed
// straight into LDS (stored conjugated, so a forward FFT gives the inverse),
// then y[2n] + i y[2n+1] = IFFT_M(Z)[n] / N.

// y[2n] + i y[2n+1] = conj(Z[f][n]) / N for rows row0 + f < nrows: row-major
// with row stride ldo, or (blocked) in the tile layout out[t][k / 8][i][k % 8]
// (tile t = row / nslice, i = row % nslice).
// ZERO: also clear every entry of Z it reads (rows past nrows included), so Z
// is all zeros again for the next group's scatter.
template <int M, int F, int TT, bool ZERO = false, int MP = Padded<M>::value>
__device__ __forceinline__ void irfft_store(float2* Z, int64_t row0, int64_t nrows, int64_t nslice,
                                            float* __restrict__ out, int64_t ldo, bool blocked, int tid) {
  constexpr int N = 2 * M;
  const float sc = 1.0f / float(N);
  if constexpr (M % TT == 0 && TT % (4 * F) == 0) {
    if (blocked) {
      // Lanes 4F apart cover one block (8 columns) of all F rows, which are
      // adjacent in the layout, so a store instruction writes runs of 32 F
      // contiguous bytes: thread (b0, f, c) = (tid / 4F, (tid / 4) % F, tid % 4)
      // stores column pair n = 4 (b0 + (TT / 4F) j) + c of row row0 + f.
      constexpr int LB = 4 * F;               // lanes per block
      constexpr int BPS = TT / LB;            // blocks per store instruction
      static_assert((M / 4) % BPS == 0, "blocks must split evenly over the store instructions");
      const int c = tid & 3, f = (tid >> 2) % F, b0 = tid / LB;
      const int64_t row = row0 + f;
      const bool ok = row < nrows;
      const int64_t t = row / nslice, i = row - t * nslice;
      float* base = out + t * int64_t(N) * nslice + i * 8 + int64_t(b0) * 8 * nslice + c * 2;
      float2* zf = Z + f * MP;
#pragma unroll
      for (int j = 0; j < (M / 4) / BPS; ++j) {
        float2* zp = zf + pad16(4 * (b0 + BPS * j) + c);
        const float2 v = *zp;
        if constexpr (ZERO) *zp = make_float2(0.f, 0.f);
        // Whole 128-byte lines (F >= 4) are stored non-temporally: the column
        // pass reads the tile only after the whole launch (2% faster column
        // pass on 16-tile launches, row pass unchanged).  64-byte runs (F = 2)
        // keep cached stores, which merge the two halves of a line in L2
        // (non-temporal there measured 26% slower).
        if (ok) {
          const float2 o = make_float2(v.x * sc, -v.y * sc);
          float* dst = base + int64_t(j) * BPS * 8 * nslice;
          if constexpr (F >= 4)
            __builtin_nontemporal_store(__builtin_bit_cast(unsigned long long, o),
                                        reinterpret_cast<unsigned long long*>(dst));
          else
            *reinterpret_cast<float2*>(dst) = o;
        }
      }
      return;
    }
  }
  for (int idx = tid; idx < F * M; idx += TT) {
    const int f = idx / M, n = idx - f * M;
    const int64_t row = row0 + f;
    float2* zp = Z + f * MP + pad16(n);
    const float2 v = *zp;
    if constexpr (ZERO) *zp = make_float2(0.f, 0.f);
    if (row >= nrows) continue;
    reinterpret_cast<float2*>(out + row * ldo)[n] = make_float2(v.x * sc, -v.y * sc);
  }
}

template <int M, int F, int TT = kTPB>
__global__ __launch_bounds__(TT) void irfft_rows_kernel(DenseRows dense, int64_t nrows, float* __restrict__ out,
                                                        int64_t ldo, Twiddle tw) {
  constexpr int MP = Padded<M>::value;
  extern __shared__ float2 smem[];
  float2* Z = smem;                       // F x MP
  float2* TA = Z + F * MP;
  float2* TB = TA + tw.nA;
  const int tid = threadIdx.x;
  const int64_t row0 = int64_t(blockIdx.x) * F;
  load_twiddle(tw, TA, TB, tid, TT);
  __syncthreads();
  for (int idx = tid; idx < F * (M / 2 + 1); idx += TT) {
    const int f = idx / (M / 2 + 1), m = idx - f * (M / 2 + 1);
    const int64_t row = row0 + f;
    float2* Zf = Z + f * MP;
    float2 a = make_float2(0.f, 0.f), b = a;
    if (row < nrows) {
      const float2* X = dense.x + row * dense.ld;
      a = X[m];
      b = X[M - m];
    }
    if (m == 0) { a.y = 0.f; b.y = 0.f; }
    const float2 w = tw_lookup(TA, TB, tw.lb, m);           // W_N^m
    const float2 s = make_float2(a.x + b.x, a.y - b.y);     // a + conj b
    const float2 d = make_float2(a.x - b.x, a.y + b.y);     // a - conj b
    const float2 cd = cmulc(d, w);                          // d * conj(W^m)
    Zf[pad16(m)] = make_float2(s.x - cd.y, -(s.y + cd.x));  // conj(s + i cd)
    if (m != 0 && 2 * m != M) {
      const float2 s2 = make_float2(b.x + a.x, b.y - a.y);  // b + conj a
      const float2 d2 = make_float2(b.x - a.x, b.y + a.y);  // b - conj a
      const float2 wd = cmul(d2, w);                        // W^m d2
      Zf[pad16(M - m)] = make_float2(s2.x + wd.y, -(s2.y - wd.x));  // conj(s2 - i W^m d2)
    }
  }
  __syncthreads();
  fft_lds<M, F, TT>(Z, TA, TB, tw.lb, 1, tid);
  irfft_store<M, F, TT>(Z, row0, nrows, 1, out, ldo, false, tid);
}

// Packed 4-bit rows (2049-bin spectra, N = 4096): channel ch of a row
// contributes x (1 + i conj W^ch) to Z[ch] and conj(x) (1 + i W^ch) to
// Z[M-ch]; columns with a negative channel number are skipped.
//
// Persistent form: workgroup w runs on XCD w % 8 and walks the row groups
// x per + it S + s of its XCD's range (S workgroups per XCD, slot s), so the
// workgroups of one XCD transform neighbouring rows together and the blocked
// layout's 32-byte row pieces meet in that XCD's L2.  Each thread's first
// kRowItems (row, column) items keep their channel weights in registers; the
// packed bytes of the next group and the row-map entries of the group after
// it are fetched while the current group is transformed.  KI = kRowItems
// covers F * ncols <= KI * TT (100-channel spectra: 2; 300 channels: 4).
// Row stride of the packed row pass's LDS spectra: the padded length, plus 8
// complex entries with 4 rows per group, so that the rows start 16 banks
// apart (the store reads all 4 rows in one instruction).
template <int F>
constexpr int packed4_row_stride() { return Padded<2048>::value + (F >= 4 ? 8 : 0); }

template <int F, int TT, int kRowItems>
__global__ __launch_bounds__(TT, 1024 / TT) void irfft_packed4_kernel(Packed4Rows pk, int64_t nrows, int64_t nslice,
                                                           float* __restrict__ out, int64_t ldo, Twiddle tw) {
  constexpr int M = 2048;
  constexpr int MP = packed4_row_stride<F>();
  extern __shared__ float2 smem[];
  float2* Z = smem;                       // F x MP
  float2* TA = Z + F * MP;
  float2* TB = TA + tw.nA;
  float2* W16 = TB + tw.nB;               // W_256^(m r), fill_w16
  float* Zr = reinterpret_cast<float*>(Z);
  const int tid = threadIdx.x;
  load_twiddle(tw, TA, TB, tid, TT);
  fill_w16(W16, tw.A, tw.B, tw.lb, tw.nA * tw.nB, tid, TT);
  const int64_t ngroups = (nrows + F - 1) / F;
  const int64_t per = (ngroups + 7) / 8;
  const int64_t S = gridDim.x / 8;
  const int64_t xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  auto group = [&](int64_t it) -> int64_t {   // row group of iteration it, -1 past the end
    const int64_t j = it * S + slot;
    const int64_t g = xcd * per + j;
    return (j < per && g < ngroups) ? g : -1;
  };
  const int ncols = int(pk.ncols), nit = F * ncols;
  int fi[kRowItems], ci[kRowItems], ch[kRowItems];
#pragma unroll
  for (int k = 0; k < kRowItems; ++k) {
    const int idx = tid + k * TT;
    fi[k] = 0;
    ci[k] = 0;
    ch[k] = -1;
    if (idx < nit) {
      fi[k] = idx / ncols;
      ci[k] = idx - fi[k] * ncols;
      const int cv = pk.chan[ci[k]];
      ch[k] = cv <= M ? cv : -1;   // channel numbers outside 0..2048 are skipped like negative ones
    }
  }
  // Row-map entry of item k in group g and the packed byte at packed row p.
  // Both loads are issued unconditionally (from a readable dummy address when
  // the item has no data) with the validity kept apart, so nothing waits for
  // them before their use one group later.  Rows < 2^31 (checked by the host).
  const int ns = int(nslice), nr = int(nrows), cut2 = int(pk.cut2), acc = int(pk.acclen), tile0 = int(pk.tile0);
  auto prow = [&](int k, int64_t g, bool* ok) -> int {
    const int row = int(g) * F + fi[k];
    bool valid = g >= 0 && ch[k] >= 0 && row < nr;
    const int rr = valid ? row : 0;
    const int tile = rr / ns, r = rr - tile * ns;
    const int chunk = tile0 + tile - (r < cut2 ? 1 : 0);
    const int crow = r < cut2 ? acc - cut2 + r : r - cut2;
    valid = valid && chunk >= 0;
    *ok = valid;
    return pk.rowmap[valid ? int64_t(chunk) * acc + crow : 0];
  };
  auto pbyte = [&](int k, int p, bool ok, bool* okb) -> int {
    const bool valid = ok && p >= 0;
    *okb = valid;
    const uint8_t* a = valid ? pk.bytes + int64_t(p) * ncols + ci[k] : reinterpret_cast<const uint8_t*>(pk.chan);
    return int(*a);
  };
  __syncthreads();  // twiddles
  // this thread's twiddles of the last (radix-8, sub-transform 256) pass: the
  // same in every group
  float2 tw8[7];
  pass_twiddles<M, 8, 256>(TA, TB, tw.lb, 1, tid, tw8);
  // Channel c contributes to Z[c] and Z[M - c]; a bin with one contributor
  // (no repeat of c, no channel M - c) is written with a plain store, the rest
  // are summed with LDS atomics, which cost about ten times a store.  The
  // channel counts live in Z until it is cleared for the first group.
  int* cnt = reinterpret_cast<int*>(Z);
  for (int idx = tid; idx <= M; idx += TT) cnt[idx] = 0;
  __syncthreads();
  for (int c = tid; c < ncols; c += TT) {
    const int h = pk.chan[c];
    if (h >= 0 && h <= M) atomicAdd(&cnt[h], 1);
  }
  __syncthreads();
  // item constants: weights 1 + i conj(w), 1 + i w and the float offsets of Z[ch], Z[M-ch]
  float2 ga[kRowItems], gb[kRowItems];
  int oa[kRowItems], ob[kRowItems];
  bool own[kRowItems];
#pragma unroll
  for (int k = 0; k < kRowItems; ++k) {
    const int c = ch[k] < 0 ? 0 : ch[k];
    const float2 w = tw_lookup(TA, TB, tw.lb, c);
    ga[k] = make_float2(1.f + w.y, w.x);
    gb[k] = make_float2(1.f - w.y, w.x);
    oa[k] = fi[k] * 2 * MP + 2 * pad16(c);
    ob[k] = fi[k] * 2 * MP + 2 * pad16(M - c);
    own[k] = ch[k] >= 0 && ch[k] <= M && cnt[c] == 1 && (2 * c == M || cnt[M - c] == 0);
  }
  __syncthreads();  // every thread has read the channel counts: Z starts all zeros
  for (int idx = tid; idx < F * MP; idx += TT) Z[idx] = make_float2(0.f, 0.f);
  // one row group: zero fill, channel scatter, inverse transform, store
  auto transform = [&](int64_t g, const int* bc, const bool* vc) {
    // an opaque copy of the thread index: the transform's LDS addresses are
    // recomputed per group instead of being held in registers across the loop
    int tl = tid;
    asm volatile("" : "+v"(tl));
    __syncthreads();  // Z is all zeros: the initial fill, or the previous group's store cleared it
#pragma unroll
    for (int k = 0; k < kRowItems; ++k) {
      if (vc[k]) {
        const int c = ch[k];
        float2 x = make_float2(float(nib4f(bc[k] >> 4)), float(nib4f(bc[k] & 15)));
        if (c == 0 || c == M) x.y = 0.f;
        const float2 va = cmul(x, ga[k]);
        const float2 vb = cmul(cconj(x), gb[k]);
        if (own[k]) {
          if (2 * c == M) {
            *reinterpret_cast<float2*>(&Zr[oa[k]]) = make_float2(va.x + vb.x, -va.y + -vb.y);
          } else {
            if (c < M) *reinterpret_cast<float2*>(&Zr[oa[k]]) = make_float2(va.x, -va.y);
            if (c > 0) *reinterpret_cast<float2*>(&Zr[ob[k]]) = make_float2(vb.x, -vb.y);
          }
        } else {
          if (c < M) {
            atomicAdd(&Zr[oa[k]], va.x);
            atomicAdd(&Zr[oa[k] + 1], -va.y);
          }
          if (c > 0) {
            atomicAdd(&Zr[ob[k]], vb.x);
            atomicAdd(&Zr[ob[k] + 1], -vb.y);
          }
        }
      }
    }
    // items beyond the prefetched ones (more than kRowItems * TT / F channels)
    for (int idx = tl + kRowItems * TT; idx < nit; idx += TT) {
      const int f = idx / ncols;
      const int c = idx - f * ncols;
      const int64_t row = g * F + f;
      const int cc = pk.chan[c];
      if (row >= nrows || cc < 0 || cc > M) continue;
      const int64_t p = packed_row(pk, row, nslice);
      if (p < 0) continue;
      const int byte = pk.bytes[p * ncols + c];
      float2 x = make_float2(float(nib4f(byte >> 4)), float(nib4f(byte & 15)));
      if (cc == 0 || cc == M) x.y = 0.f;
      const float2 w = tw_lookup(TA, TB, tw.lb, cc);
      float* Zf = Zr + f * 2 * MP;
      if (cc < M) {
        const float2 v = cmul(x, make_float2(1.f + w.y, w.x));
        atomicAdd(&Zf[2 * pad16(cc)], v.x);
        atomicAdd(&Zf[2 * pad16(cc) + 1], -v.y);
      }
      if (cc > 0) {
        const float2 v = cmul(cconj(x), make_float2(1.f - w.y, w.x));
        atomicAdd(&Zf[2 * pad16(M - cc)], v.x);
        atomicAdd(&Zf[2 * pad16(M - cc) + 1], -v.y);
      }
    }
    __syncthreads();
    fft_lds<M, F, TT, 1, M, true, 2, MP>(Z, TA, TB, tw.lb, 1, tl, W16, tw8);
    irfft_store<M, F, TT, true, MP>(Z, g * F, nrows, nslice, out, ldo, ldo == 8, tl);
  };
  // Two sets of prefetch registers used in turn (no copies of values still in
  // flight): while group it is transformed, the bytes of group it + 1 and the
  // row map of group it + 2 arrive.
  int pA[kRowItems], bA[kRowItems], pB[kRowItems], bB[kRowItems];
  bool qA[kRowItems], vA[kRowItems], qB[kRowItems], vB[kRowItems];
#pragma unroll
  for (int k = 0; k < kRowItems; ++k) {
    bool ok;
    const int p = prow(k, group(0), &ok);
    bA[k] = pbyte(k, p, ok, &vA[k]);
    pA[k] = prow(k, group(1), &qA[k]);
  }
  for (int64_t it = 0;; it += 2) {
    int64_t g = group(it);
    if (g < 0) break;
    const int64_t g2 = group(it + 2);
#pragma unroll
    for (int k = 0; k < kRowItems; ++k) {
      bB[k] = pbyte(k, pA[k], qA[k], &vB[k]);
      pB[k] = prow(k, g2, &qB[k]);
    }
    transform(g, bA, vA);
    g = group(it + 1);
    if (g < 0) break;
    const int64_t g3 = group(it + 3);
#pragma unroll
    for (int k = 0; k < kRowItems; ++k) {
      bA[k] = pbyte(k, pB[k], qB[k], &vA[k]);
      pA[k] = prow(k, g3, &qA[k]);
    }
    transform(g, bB, vB);
  }
}

// --------------------------------------------------------- rfft along rows
// Optional polyphase front end: y[k] = sum_j ts[(row+j) L + k] w[j L + k].
// Output columns [c0, c1) of the N/2+1 bins, row stride ldo (complex).
template <int M, int F>
__global__ __launch_bounds__(kTPB) void rfft_rows_kernel(const float* __restrict__ x, int64_t ldx, int64_t nrows,
                                                         const float* __restrict__ win, int ntap, int64_t c0, int64_t c1,
                                                         float2* __restrict__ out, int64_t ldo, Twiddle tw) {
  constexpr int N = 2 * M;
  constexpr int MP = Padded<M>::value;
  extern __shared__ float2 smem[];
  float2* Z = smem;
  float2* TA = Z + F * MP;
  float2* TB = TA + tw.nA;
  const int tid = threadIdx.x;
  load_twiddle(tw, TA, TB, tid, kTPB);
  const int64_t row0 = int64_t(blockIdx.x) * F;
  for (int idx = tid; idx < F * M; idx += kTPB) {
    const int f = idx / M, n = idx - f * M;
    const int64_t row = row0 + f;
    float2 v = make_float2(0.f, 0.f);
    if (row < nrows) {
      if (win) {
        float ax = 0.f, ay = 0.f;
        for (int j = 0; j < ntap; ++j) {
          const float2 t = reinterpret_cast<const float2*>(x + (row + j) * ldx)[n];
          const float2 w = reinterpret_cast<const float2*>(win + int64_t(j) * N)[n];
          ax = fmaf(t.x, w.x, ax);
          ay = fmaf(t.y, w.y, ay);
        }
        v = make_float2(ax, ay);
      } else {
        v = reinterpret_cast<const float2*>(x + row * ldx)[n];
      }
    }
    Z[f * MP + pad16(n)] = v;
  }
  __syncthreads();
  fft_lds<M, F, kTPB>(Z, TA, TB, tw.lb, 1, tid);
  const int64_t nc = c1 - c0;
  for (int64_t idx = tid; idx < F * nc; idx += kTPB) {
    const int f = int(idx / nc);
    const int64_t m = c0 + (idx - f * nc);
    const int64_t row = row0 + f;
    if (row >= nrows) continue;
    const float2 a = Z[f * MP + pad16(int(m % M))];
    const float2 b = Z[f * MP + pad16(int((M - m % M) % M))];
    const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));   // (a + conj b)/2
    const float2 d = make_float2(0.5f * (a.x - b.x), 0.5f * (a.y + b.y));   // (a - conj b)/2
    const float2 o = make_float2(d.y, -d.x);                                // d / i
    float2 w;
    if (m == M) w = make_float2(-1.f, 0.f);
    else w = tw_lookup(TA, TB, tw.lb, int(m));
    const float2 wo = cmul(w, o);
    out[row * ldo + (m - c0)] = make_float2(e.x + wo.x, e.y + wo.y);
  }
}

// ------------------------------------------------------ iPFB column pass
// One workgroup = CW adjacent columns of one tile (all nslice rows) held in LDS
// as CW/2 complex sequences u = y[:, k] + i y[:, k+1].
//
// Workgroup order (workgroup i runs on XCD i % 8): each XCD owns a contiguous
// range of strips and walks it in groups of G strips, all tiles of a group
// back to back.  Workgroups in flight together read neighbouring columns of
// the same rows (whole cache lines, long DRAM bursts) and a group's filter
// columns are re-read by the next tile while still cached.
__device__ __forceinline__ void cols_item(int64_t nstrips, int group, int64_t* sidx, int64_t* tile) {
  const int64_t ntiles = gridDim.x / nstrips;
  if (nstrips % 8 == 0) {
    const int64_t P = nstrips / 8;                        // strips per XCD
    const int64_t G = (P % group == 0) ? group : P;       // strips per group
    const int64_t slot = blockIdx.x / 8;
    const int64_t grp = slot / (G * ntiles), rem = slot - grp * (G * ntiles);
    *tile = rem / G;
    *sidx = (blockIdx.x % 8) * P + grp * G + (rem - *tile * G);
  } else {
    *tile = blockIdx.x / nstrips;
    *sidx = blockIdx.x - *tile * nstrips;
  }
}

// Load the strip, forward FFT along time, Wiener filter with 1/nslice, inverse
// FFT.  On return (after a barrier) time row i of column k0 + 2p is
// U[p][i].x and of column k0 + 2p + 1 is -U[p][i].y.
// Tile layout: element (i, k) at yt[(k / ldy) * ldy * NSL + i * ldy + k % ldy]:
// row-major with row stride ldy when ldy >= the column count, or blocks of
// ldy columns (each block's rows contiguous, ldy a multiple of CW) otherwise.
template <int NSL, int CW, int T, int KPN = 2>
__device__ __forceinline__ void cols_core(const float* __restrict__ yt, int64_t ldy, const float2* __restrict__ filt,
                                          int64_t ldf, int64_t k0, float2* U, const float2* TA, const float2* TB,
                                          int lb, int tid, const float2* W16) {
  constexpr int NP = CW / 2;
  constexpr int PADN = Padded<NSL>::value;
  const float* ys = yt + (k0 / ldy) * ldy * NSL + (k0 % ldy);
  // First radix-16 pass of the forward transform straight from memory:
  // butterfly (j, p) takes rows j + NB r (r < 16) of sequence p (threads run p
  // fastest, so a wave reads whole row pieces) and stores the Stockham pass
  // output U[p][16 j + r]; the other passes start at sub-transform size 16.
  {
    constexpr int NB = NSL / 16;
    constexpr int TOT = NB * NP;
    constexpr int PER = (TOT + T - 1) / T;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int b = tid + e * T;
      if (TOT % T == 0 || b < TOT) {
        const int p = b % NP, j = b / NP;
        float2 v[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = reinterpret_cast<const float2*>(ys + int64_t(j + NB * r) * ldy)[p];
        DFT<16>::run(v);
        float2* Up = U + p * PADN;
#pragma unroll
        for (int r = 0; r < 16; ++r) Up[pad16(16 * j + r)] = v[r];
      }
    }
  }
  // Filter taps: thread handles sequence fp = tid / TPS, bins
  // f = tid % TPS + m TPS (m < KF) up to NSL / 2.  The first KP bins' taps are
  // fetched now so that their loads are in flight during the forward passes
  // (the rest are issued at the start of the filter loop, behind them).
  // KPN: the fused column pass at nslice 4096 (KF = 9) takes 5, 3% faster
  // than 2 at 111 VGPRs; all 9 take it past 128 VGPRs, below the two
  // workgroups per CU that its LDS allows, and 50% slower.  The plain column
  // pass keeps 2 (5 would take it to 127 VGPRs).
  constexpr int TPS = T / NP;
  constexpr int NF = NSL / 2 + 1;
  constexpr int KF = (NF + TPS - 1) / TPS;
  constexpr int KP = KF < KPN ? KF : KPN;
  const int fp = tid / TPS, f0 = tid - fp * TPS;
  const float2* h0p = filt + (k0 + 2 * fp) * ldf;
  const float2* h1p = h0p + ldf;
  float2 h0[KP], h1[KP];
#pragma unroll
  for (int m = 0; m < KP; ++m) {
    const int f = f0 + m * TPS;
    if ((m + 1) * TPS <= NF || f < NF) {
      h0[m] = h0p[f];
      h1[m] = h1p[f];
    }
  }
  __syncthreads();
  // the sub-transform-256 twiddles of this thread are the same in the inverse
  // transform: kept in registers when a thread owns one butterfly per pass
  constexpr bool KEEP = (NSL / 16) * NP == T && NSL >= 4096;
  float2 tw256[15];
  fft_lds<NSL, NP, T, 16, NSL, true, KEEP ? 1 : 0>(U, TA, TB, lb, 0, tid, W16, tw256);
  // filter on the real-pair spectra, stored conjugated for the inverse
  const float sc = 1.0f / float(NSL);
  float2* Uf = U + fp * PADN;
#pragma unroll
  for (int m = 0; m < KF; ++m) {
    const int f = f0 + m * TPS;
    if ((m + 1) * TPS <= NF || f < NF) {
      const float2 a = Uf[pad16(f)];
      const float2 b = Uf[pad16((NSL - f) & (NSL - 1))];
      const float2 y0 = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));   // (a + conj b)/2
      const float2 dd = make_float2(0.5f * (a.x - b.x), 0.5f * (a.y + b.y));   // (a - conj b)/2
      const float2 y1 = make_float2(dd.y, -dd.x);                              // -i dd
      const float2 hv0 = m < KP ? h0[m] : h0p[f];
      const float2 hv1 = m < KP ? h1[m] : h1p[f];
      // explicit fused forms, so every build of this step rounds alike
      float2 g0 = make_float2(fmaf(hv0.x, y0.x, -(hv0.y * y0.y)) * sc, fmaf(hv0.x, y0.y, hv0.y * y0.x) * sc);
      float2 g1 = make_float2(fmaf(hv1.x, y1.x, -(hv1.y * y1.y)) * sc, fmaf(hv1.x, y1.y, hv1.y * y1.x) * sc);
      if (f == 0 || 2 * f == NSL) { g0.y = 0.f; g1.y = 0.f; }
      const float2 uf = make_float2(g0.x - g1.y, g0.y + g1.x);     // g0 + i g1
      Uf[pad16(f)] = cconj(uf);
      if (f != 0 && 2 * f != NSL) {
        const float2 ub = make_float2(g0.x + g1.y, -g0.y + g1.x);  // conj g0 + i conj g1
        Uf[pad16(NSL - f)] = cconj(ub);
      }
    }
  }
  __syncthreads();
  fft_lds<NSL, NP, T, 1, NSL, true, KEEP ? 2 : 0>(U, TA, TB, lb, 0, tid, W16, tw256);
}

// Spectral-order column core (see ipfb_rows_spectral_kernel): the strip's
// time spectra are already transformed, Ds[4 f + c] = ddft[k0 + c][f] for
// f <= NSL/2 without the 1/4096 of the channel transform.  Wiener filter with
// 1/(4096 nslice) (a power of two: exact), inverse FFT; on return U holds the
// time rows as cols_core leaves them.
template <int NSL, int CW, int T>
__device__ __forceinline__ void cols_core_spec(const float2* __restrict__ Ds, const float2* __restrict__ filt,
                                               int64_t ldf, int64_t k0, float2* U, const float2* TA, const float2* TB,
                                               int lb, int tid, const float2* W16) {
  static_assert(CW == 4, "the spectral strip layout holds 4 columns");
  constexpr int NP = CW / 2;
  constexpr int PADN = Padded<NSL>::value;
  constexpr int NF = NSL / 2 + 1;
  constexpr int TOT = NF * NP;
  constexpr int PER = (TOT + T - 1) / T;
  const float sc = 1.0f / (float(NSL) * 4096.0f);
  // lane (f, p), p fastest: a wave reads 32 consecutive 32-byte pieces
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = tid + e * T;
    if (idx < TOT) {
      const int p = idx % NP, f = idx / NP;
      const float4 d = *reinterpret_cast<const float4*>(Ds + int64_t(f) * 4 + 2 * p);
      const float2 y0 = make_float2(d.x, d.y), y1 = make_float2(d.z, d.w);
      const float2* h0p = filt + (k0 + 2 * p) * ldf;
      const float2 hv0 = h0p[f];
      const float2 hv1 = h0p[ldf + f];
      float2 g0 = make_float2(fmaf(hv0.x, y0.x, -(hv0.y * y0.y)) * sc, fmaf(hv0.x, y0.y, hv0.y * y0.x) * sc);
      float2 g1 = make_float2(fmaf(hv1.x, y1.x, -(hv1.y * y1.y)) * sc, fmaf(hv1.x, y1.y, hv1.y * y1.x) * sc);
      if (f == 0 || 2 * f == NSL) { g0.y = 0.f; g1.y = 0.f; }
      float2* Up = U + p * PADN;
      Up[pad16(f)] = make_float2(g0.x - g1.y, -(g0.y + g1.x));               // conj(g0 + i g1)
      if (f != 0 && 2 * f != NSL) Up[pad16(NSL - f)] = make_float2(g0.x + g1.y, g0.y - g1.x);   // conj(conj g0 + i conj g1)
    }
  }
  __syncthreads();
  fft_lds<NSL, NP, T, 1, NSL, true, 0>(U, TA, TB, lb, 0, tid, W16);
}

template <int NSL, int CW, int T>
__global__ __launch_bounds__(T) void ipfb_cols_kernel(const float* __restrict__ y, int64_t ldy, int64_t tile_stride,
                                                         const float2* __restrict__ filt, int64_t ldf, int64_t ncols,
                                                         int cut, float* __restrict__ out, int64_t ldo,
                                                         int64_t out_tile_stride, int group, Twiddle tw) {
  constexpr int NP = CW / 2;
  constexpr int PADN = Padded<NSL>::value;
  extern __shared__ float2 smem[];
  float2* U = smem;
  float2* TA = U + NP * PADN;
  float2* TB = TA + tw.nA;
  float2* W16 = TB + tw.nB;
  const int tid = threadIdx.x;
  load_twiddle(tw, TA, TB, tid, T);
  fill_w16(W16, tw.A, tw.B, tw.lb, tw.nA * tw.nB, tid, T);
  int64_t sidx, tile;
  cols_item(ncols / CW, group, &sidx, &tile);
  const int64_t k0 = sidx * CW;
  cols_core<NSL, CW, T>(y + tile * tile_stride, ldy, filt, ldf, k0, U, TA, TB, tw.lb, tid, W16);
  float* ot = out + tile * out_tile_stride;
  const int nkeep = NSL - 2 * cut;
  for (int idx = tid; idx < nkeep * NP; idx += T) {
    const int i = idx / NP, p = idx - i * NP;
    const float2 v = U[p * PADN + pad16(i + cut)];
    reinterpret_cast<float2*>(ot + int64_t(i) * ldo + k0)[p] = make_float2(v.x, -v.y);
  }
}

// ------------------------------------------- re-PFB stage-1 workspace layout
// A[b] holds the bins q = 0..H (H = R1/2) of the R1-point real DFT down each of
// the 4096 columns, grouped by 4 columns: bin slot s of columns 4g..4g+3 is
// A[((b * 1024 + g) * QP + s) * 4 + k % 4].  Bins 0 and H of a real sequence
// are real, so both share slot 0 (.x = bin 0, .y = bin H) and slots 1..H-1
// hold the other bins: QP = H slots, and the bins of 4 columns of one block
// are one contiguous run of 32 QP bytes (whole 128-byte lines for R1 >= 8).
template <int R1>
struct StageA {
  static constexpr int H = R1 / 2;
  static constexpr int QP = H > 0 ? H : 1;
  static constexpr int64_t kBlock = int64_t(1024) * QP * 4;   // float2 per block
  __host__ __device__ static int64_t at(int64_t b, int s, int64_t k) {
    return (b * 1024 + (k >> 2)) * (QP * 4) + s * 4 + (k & 3);
  }
};

// Row walk of the fused polyphase: sums of rows xrow + R1 (bl + j), j < ntap
// (taps j >= ntap have weight 0), for the nb <= SEG blocks bl of one lane,
// with the window weights of the lane's row at wl[j R1 CW] (.x and .y for the
// two columns of the pair; Up holds conj for the second).  Rows past the
// lane's last block feed only block sums that are never stored, so every
// read is unconditional (clamped into the sequence).  MASK: rows outside
// [lo, hi) (outside the tile's kept rows) count as zero: reads clamp at both
// ends, and the caller has zeroed the rows of Up outside [lo, hi) (the cut
// rows, lo >= 1), so a clamped read is a zero.  Ends with the block sums
// written to Yp[pad16((bl0 + b) R1 + rb)] after a workgroup barrier (Yp may
// overlap Up).
constexpr int kWalkSeg = 16;
template <int NSL, int R1, int CW, int NT, int SEG = kWalkSeg, bool MASK = false>
__device__ __forceinline__ void polyphase_walk(const float2* Up, float2* Yp, int xrow, const float* wl, int ntap,
                                               int nb, int bl0, int rb, int lo = 0, int hi = NSL) {
  float w0[NT], w1[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const float2 w = *reinterpret_cast<const float2*>(wl + j * R1 * CW);
    w0[j] = j < ntap ? w.x : 0.f;
    w1[j] = j < ntap ? w.y : 0.f;
  }
  float2 acc[SEG];
#pragma unroll
  for (int b = 0; b < SEG; ++b) acc[b] = make_float2(0.f, 0.f);
#pragma unroll
  for (int m = 0; m < SEG + NT - 1; ++m) {
    const int x = xrow + R1 * m;
    int xc = x < NSL - 1 ? x : NSL - 1;
    if constexpr (MASK) xc = xc > 0 ? xc : 0;
    const float2 v = Up[pad16(xc)];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      if (m - j >= 0 && m - j < SEG) {
        acc[m - j].x = fmaf(v.x, w0[j], acc[m - j].x);
        acc[m - j].y = fmaf(-v.y, w1[j], acc[m - j].y);
      }
    }
  }
  __syncthreads();  // every lane has read its rows: U becomes Y
#pragma unroll
  for (int b = 0; b < SEG; ++b)
    if (b < nb) Yp[pad16((bl0 + b) * R1 + rb)] = acc[b];
}

// Split form of the fused column pass (tile-boundary blocks in two halves):
// each tile computes re-PFB stage 1 of every block that touches its kept rows,
// over those rows only (the others count as zero).  The tile that holds a
// block's first row writes the result to the workspace A; the next tile
// writes its half to the boundary array P, slot
// ntap * (that tile's index in the stream window - 1) + (b - first block that
// straddles its first row); stage 2 adds the two halves.  Blocks span at most
// two tiles (R1 ntap <= kept rows).  Up to 4 taps: the walk segments hold
// kWalkSegSplit blocks and the block sums (Y) and the DFT exchange live in an
// area of SplitArea<NSL, R1, NP>::value complex entries.
constexpr int kWalkSegSplit = 17;
template <int NSL, int R1, int NP>
struct SplitArea {
  static constexpr int kBlocks = NSL / R1 + 4;                        // blocks touching one tile, at most
  static constexpr int kYStride = Padded<kBlocks * R1>::value;        // Y row of one column pair
  static constexpr int kSP = R1 / 4 >= 16 ? R1 / 4 + 1 : R1 / 4;      // exchange row (as the column pass)
  static constexpr int kExch = kBlocks * NP * 4 * kSP;
  static constexpr int kFFT = NP * Padded<NSL>::value;
  static constexpr int a = kFFT > NP * kYStride ? kFFT : NP * kYStride;
  static constexpr int value = a > kExch ? a : kExch;
};

// ------------------------------------ iPFB column pass fused with re-PFB stage 1
// Stream row g = row0 + tile * rows_per_tile + (i - cut) holds kept row i of a
// tile.  Re-PFB block b (R1 rows per block row, ntap taps) reads stream rows
// [R1 b, R1 (b + ntap)).  The blocks whose rows all lie in this tile,
// [bi0, bi1) below, get their stage-1 result A[b][q][k] (q <= R1/2, StageA layout) straight
// from LDS; only rows that other blocks read go to the stream buffer.
// A tile may start before stream row 0 (g0 < 0: a window of the stream whose
// row 0 is the first row of its first block); blocks start at 0.
__host__ __device__ inline int64_t floor_div(int64_t a, int64_t b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
__host__ __device__ inline void repfb_interior(int64_t g0, int64_t nkeep, int64_t R1, int64_t ntap, int64_t nblocks,
                                               int64_t* bi0, int64_t* bi1) {
  const int64_t lo = g0 > 0 ? (g0 + R1 - 1) / R1 : 0;
  int64_t hi = floor_div(g0 + nkeep, R1) - ntap + 1;
  if (hi > nblocks) hi = nblocks;
  *bi0 = lo;
  *bi1 = hi > lo ? hi : lo;
}

// Split form: the blocks [*lo, *hi) (of nblocks) that touch a tile's kept rows
// [g0, g0 + nkeep), and *e0, the first block that straddles its first row g0.
__host__ __device__ inline void split_blocks(int64_t g0, int64_t nkeep, int64_t R1, int64_t ntap, int64_t nblocks,
                                             int64_t* lo, int64_t* hi, int64_t* e0) {
  *e0 = floor_div(g0, R1) - ntap + 1;
  const int64_t a = *e0 > 0 ? *e0 : 0;
  int64_t b = floor_div(g0 + nkeep - 1, R1) + 1;
  if (b > nblocks) b = nblocks;
  *lo = a;
  *hi = b > a ? b : a;
}

// Stage 1 for both columns of a pair at once: z[r] = y0[r] + i y1[r] (y0, y1
// the polyphase sums of columns k0+2p and k0+2p+1), R1-point DFT split over 4
// lanes as in the split kernel below (lane t: rows 4s + t, S-point DFT, radix-4
// combine through LDS), then Y0[q] = (Z[q] + conj Z[-q]) / 2 and
// Y1[q] = (Z[q] - conj Z[-q]) / 2i.
// One tile of the fused column pass: column pass of strip k0 of tile `tile`,
// stream rows for the tile-boundary blocks, stage 1 of the interior blocks.
// SPLIT: the split form above (P, tile0 = this launch's first tile's index
// in the stream window; ts unused).
template <int NSL, int CW, int T, int R1, bool SPEC = false, bool SPLIT = false>
__device__ __forceinline__ void cols_repfb_tile(const float* __restrict__ y, int64_t ldy, int64_t tile_stride,
                                                const float2* __restrict__ filt, int64_t ldf, int cut,
                                                float* __restrict__ ts, int64_t row0, int64_t rows_per_tile,
                                                int ntap, int64_t nblocks, float2* __restrict__ A, int64_t tile,
                                                int64_t k0, float2* U, const float2* TA, const float2* TB,
                                                const float2* W16, const float2* WR, const float* WIN, int lb,
                                                int tid, float2* __restrict__ P = nullptr, int64_t tile0 = 0) {
  constexpr int NP = CW / 2;
  constexpr int PADN = Padded<NSL>::value;
  constexpr int S = R1 / 4;
  constexpr int H = R1 / 2;
  constexpr int SP = S >= 16 ? S + 1 : S;   // row of the exchange buffer (padded when it fits)
  constexpr int64_t C = 4096;
  if constexpr (SPEC)
    cols_core_spec<NSL, CW, T>(reinterpret_cast<const float2*>(y) + (tile * (4096 / 4) + k0 / 4) * int64_t(NSL / 2 + 1) * 4,
                               filt, ldf, k0, U, TA, TB, lb, tid, W16);
  else
    cols_core<NSL, CW, T, 5>(y + tile * tile_stride, ldy, filt, ldf, k0, U, TA, TB, lb, tid, W16);

  const int nkeep = NSL - 2 * cut;
  const int64_t g0 = row0 + tile * rows_per_tile;
  int64_t bi0, bi1, e0 = 0;   // e0: first block straddling this tile's first row (split form)
  if constexpr (SPLIT) {
    split_blocks(g0, nkeep, R1, ntap, nblocks, &bi0, &bi1, &e0);   // every block touching the kept rows
  } else {
    repfb_interior(g0, nkeep, R1, ntap, nblocks, &bi0, &bi1);
  }
  const int nint = int(bi1 - bi0);
  // Rows read by blocks outside [bi0, bi1) go to the stream buffer: local
  // rows [0, n1) and [n2, nkeep) (all rows when the two ranges meet).
  if constexpr (!SPLIT) {
    int n1 = nkeep, n2 = nkeep;
    if (nint > 0) {
      n1 = int(R1 * (bi0 + ntap - 1) - g0);
      n2 = int(R1 * bi1 - g0);
      if (n1 > nkeep) n1 = nkeep;
      if (n2 <= n1) n1 = n2 = nkeep;
    }
    const int nw = (n1 + nkeep - n2) * NP;
    for (int idx = tid; idx < nw; idx += T) {
      const int w = idx / NP, p = idx - w * NP;
      const int i = w < n1 ? w : w - n1 + n2;
      if (g0 + i < 0) continue;   // before the window's first block: read by no block of this call
      const float2 v = U[p * PADN + pad16(i + cut)];
      reinterpret_cast<float2*>(ts + (g0 + i) * C + k0)[p] = make_float2(v.x, -v.y);
    }
  }
  if (nint == 0) return;

  // Polyphase sums, row-walk form: lane (rb, p, seg) owns row rb of the
  // blocks bl in [seg SL, seg SL + SL) and walks the block rows m once
  // (row rb + R1 m of the interior run is tap m - bl of block bl), so each
  // stream row is read once and the lane's window weights stay in registers.
  // Block sums are kept until every lane has read U, then written back as
  // Y[p][bl R1 + rb] for the DFT step.  Taps are summed in order
  // j = 0..ntap-1, as in the stage-1 kernels.
  constexpr int NSEG = T / (R1 * NP);        // segments per (rb, p)
  static_assert(NSEG * R1 * NP == T, "one walk lane per thread");
  // Y (block sums) row stride of one column pair: the FFT rows' in the
  // reference form, the longer split-area row in the split form
  constexpr int YS = SPLIT ? SplitArea<NSL, R1, NP>::kYStride : PADN;
  if constexpr (SPLIT) {
    // the cut rows (outside the kept rows) read as zero in the walk
    if (tid < 2 * cut * NP) {
      const int p = tid / (2 * cut), c = tid % (2 * cut);
      U[p * PADN + pad16(c < cut ? c : NSL - 2 * cut + c)] = make_float2(0.f, 0.f);
    }
    __syncthreads();
  }
  {
    const int SL = (nint + NSEG - 1) / NSEG;  // <= the walk's segment (checked by the host)
    const int rb = tid % R1, pw = (tid / R1) % NP, seg = tid / (R1 * NP);
    const int bl0 = seg * SL;
    int nb = nint - bl0;
    nb = nb < 0 ? 0 : (nb > SL ? SL : nb);
    const int xrow = int(R1 * (bi0 + bl0) - g0) + cut + rb;   // local row of block row m = 0
    const float* wl = WIN + rb * CW + 2 * pw;
    if constexpr (SPLIT) {
      polyphase_walk<NSL, R1, CW, 4, kWalkSegSplit, true>(U + pw * PADN, U + pw * YS, xrow, wl, ntap, nb, bl0, rb, cut,
                                                          cut + nkeep);
    } else {
      if (ntap <= 4) polyphase_walk<NSL, R1, CW, 4>(U + pw * PADN, U + pw * PADN, xrow, wl, ntap, nb, bl0, rb);
      else polyphase_walk<NSL, R1, CW, 8>(U + pw * PADN, U + pw * PADN, xrow, wl, ntap, nb, bl0, rb);
    }
  }
  __syncthreads();

  // S-point DFTs: lane u = task * 4 + t, task = bl + nint * p, takes the
  // block sums of rows r = 4 s + t (s < S).
  const int nlanes = nint * NP * 4;
  constexpr int PER = 2;                     // lane slots per thread (nint * NP * 4 <= PER * T is checked by the host)
  float2 z[PER][S];
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int u = tid + e * T;
    if (u < nlanes) {
      const int t = u & 3, task = u >> 2;
      const int p = task / nint, bl = task - p * nint;
      const float2* Yp = U + p * YS + bl * (R1 + R1 / 16) + t;   // pad16(bl R1 + 4 s + t), R1 a multiple of 16
#pragma unroll
      for (int sI = 0; sI < S; ++sI) z[e][sI] = Yp[4 * sI + sI / 4];
      DFT<S>::run(z[e]);
    }
  }
  __syncthreads();  // every lane is done reading U: reuse it as the exchange buffer
  float2* X = U;
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int u = tid + e * T;
    if (u < nlanes) {
#pragma unroll
      for (int q0 = 0; q0 < S; ++q0) X[u * SP + q0] = z[e][q0];
    }
  }
  __syncthreads();
  // radix-4 combine: lane (task, t) owns q0 = t + 4 m; Y[q0 + S pp] goes to slot (pp, q0)
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int u = tid + e * T;
    if (u < nlanes) {
      const int t = u & 3, task = u >> 2;
      float2* Xt = X + task * 4 * SP;
#pragma unroll
      for (int m = 0; m < S / 4; ++m) {
        const int q0 = t + 4 * m;
        float2 v[4];
        v[0] = Xt[q0];
#pragma unroll
        for (int tt = 1; tt < 4; ++tt) v[tt] = cmul(Xt[tt * SP + q0], WR[(q0 * tt) & (R1 - 1)]);
        const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const float2 s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
        Xt[0 * SP + q0] = cadd(s02, s13);                                    // pp = 0
        Xt[1 * SP + q0] = make_float2(d02.x + d13.y, d02.y - d13.x);         // pp = 1: d02 - i d13
        Xt[2 * SP + q0] = csub(s02, s13);                                    // pp = 2
        Xt[3 * SP + q0] = make_float2(d02.x - d13.y, d02.y + d13.x);         // pp = 3: d02 + i d13
      }
    }
  }
  __syncthreads();
  // Separate the two real columns k0 + 2p, k0 + 2p + 1 and store their bins
  // in the StageA layout.  Thread: fixed p = tid % NP, slots r = bl H + q
  // stepping by T / NP; slot 0 carries bins 0 and H (both real:
  // Y0[0] = Z[0].x, Y1[0] = Z[0].y and likewise for Z[H]).
  {
    static_assert(H >= 1 && CW % 2 == 0, "fused stage 1 needs R1 >= 2 and column pairs");
    constexpr int RS = T / NP;
    constexpr int QA = RS % H, BA = RS / H;
    const int p = tid % NP;
    int r = tid / NP;
    int q = r % H, bl = r / H;
    const int nr = nint * H;
    const float2* Xp = X + nint * p * 4 * SP;
    for (; r < nr; r += RS) {
      const float2* Xt = Xp + bl * 4 * SP;
      float4 o;
      if (q == 0) {
        const float2 a = Xt[0];
        const float2 c = Xt[(H / S) * SP + (H & (S - 1))];
        o = make_float4(a.x, c.x, a.y, c.y);
      } else {
        const int qm = R1 - q;
        const float2 a = Xt[(q / S) * SP + (q & (S - 1))];
        const float2 c = Xt[(qm / S) * SP + (qm & (S - 1))];
        o = make_float4(0.5f * (a.x + c.x), 0.5f * (a.y - c.y),      // (a + conj c) / 2
                        0.5f * (a.y + c.y), -0.5f * (a.x - c.x));    // (a - conj c) / 2i
      }
      // non-temporal: stage 2 reads A only after the column passes of every
      // tile (and, in the batched pass, every input): 6% faster column pass
      // on the batched workload, stage 2 unchanged
      typedef float f4v __attribute__((ext_vector_type(4)));
      float2* dst = A + StageA<R1>::at(bi0 + bl, q, k0 + 2 * p);
      if constexpr (SPLIT) {
        // a block whose first row lies before this tile: its second half
        const int64_t b = bi0 + bl;
        if (R1 * b < g0)
          dst = P + (ntap * (tile0 + tile - 1) + (b - e0)) * StageA<R1>::kBlock + StageA<R1>::at(0, q, k0 + 2 * p);
      }
      __builtin_nontemporal_store(f4v{o.x, o.y, o.z, o.w}, reinterpret_cast<f4v*>(dst));
      q += QA;
      bl += BA;
      if (q >= H) { q -= H; ++bl; }
    }
  }
}

// SPEC: y is the spectral-order strip array of ipfb_rows_spectral_kernel
// (complex [tiles][4096/4][nslice/2+1][4]; ldy and tile_stride unused).
// SPLIT: the split form of cols_repfb_tile (P, tile0).
template <int NSL, int CW, int T, int R1, bool SPEC = false, bool SPLIT = false>
__global__ __launch_bounds__(T) void ipfb_cols_repfb_kernel(const float* __restrict__ y, int64_t ldy,
                                                               int64_t tile_stride, const float2* __restrict__ filt,
                                                               int64_t ldf, int cut, float* __restrict__ ts,
                                                               int64_t row0, int64_t rows_per_tile,
                                                               const float* __restrict__ win, int ntap,
                                                               int64_t nblocks, float2* __restrict__ A, int group,
                                                               Twiddle tw, float2* __restrict__ P, int64_t tile0) {
  constexpr int NP = CW / 2;
  constexpr int PADN = Padded<NSL>::value;
  constexpr int S = R1 / 4;
  constexpr int H = R1 / 2;
  constexpr int SP = S >= 16 ? S + 1 : S;   // row of the exchange buffer (padded when it fits)
  constexpr int64_t C = 4096;
  constexpr int UA = SPLIT ? SplitArea<NSL, R1, NP>::value : NP * PADN;   // transform / stage-1 area
  extern __shared__ float2 smem[];
  float2* U = smem;
  float2* TA = U + UA;
  float2* TB = TA + tw.nA;
  float2* W16 = TB + tw.nB;                                 // W_256^(m r), fill_w16
  float2* WR = W16 + kW16Size;                                   // W_R1^m, m < R1
  float* WIN = reinterpret_cast<float*>(WR + R1);           // window strip [ntap*R1][CW]
  const int tid = threadIdx.x;
  load_twiddle(tw, TA, TB, tid, T);
  fill_w16(W16, tw.A, tw.B, tw.lb, tw.nA * tw.nB, tid, T);
  int64_t sidx, tile;
  cols_item(C / CW, group, &sidx, &tile);
  const int64_t k0 = sidx * CW;
  if (tid < R1) {
    double sn, cs;
    sincospi(-2.0 * double(tid) / double(R1), &sn, &cs);
    WR[tid] = make_float2(float(cs), float(sn));
  }
  for (int idx = tid; idx < ntap * R1 * CW; idx += T) {
    const int r = idx / CW, c = idx - r * CW;
    WIN[idx] = win[int64_t(r) * C + k0 + c];
  }
  cols_repfb_tile<NSL, CW, T, R1, SPEC, SPLIT>(y, ldy, tile_stride, filt, ldf, cut, ts, row0, rows_per_tile, ntap,
                                               nblocks, A, tile, k0, U, TA, TB, W16, WR, WIN, tw.lb, tid, P, tile0);
}

// ------------------------------------------- spectral-order inverse PFB
// cupy_ipfb (pfb_gpu_utils.py:56-91) takes dd = irfft(dat, axis=1) and then
// ddft = rfft(dd.T, axis=1), a transform along time of every one of the 4096
// samples of a spectrum.  The two transforms are linear and act on different
// axes, so the one along time can run first, on the ncols non-zero channel
// columns of dat only.  With E[f][c] = FFT_nslice(dat[:, c])[f] (the
// imaginary parts of channels 0 and 2048 dropped first, as irfft drops them):
//   ddft[j][f] = (1/4096) sum_k X_f[k] exp(+2 pi i j k / 4096),
//   X_f[chan c] = E[f][c],  X_f[4096 - chan c] = conj(E[-f][c]) (0 < chan c < 2048)
// for f = 0..nslice/2.  That is nslice/2+1 inverse 4096-point transforms plus
// ncols transforms along time in place of nslice inverse transforms of the
// spectra plus 4096 real transforms along time: the column pass then has only
// its inverse transform left.

// Channel transform along time: E[t][f][c] = FFT_nslice over the tile's rows of
// the decoded 4-bit values of column c (zero for lost spectra, out-of-range
// channels and the rows before the first chunk).  One workgroup = CA columns
// of one tile; workgroups of one XCD take neighbouring columns.
template <int NSL, int CA, int T>
__global__ __launch_bounds__(T) void ipfb_chan_fft_kernel(Packed4Rows pk, int64_t ntiles, float2* __restrict__ E,
                                                          Twiddle tw) {
  constexpr int PADN = Padded<NSL>::value;
  extern __shared__ float2 smem[];
  float2* U = smem;                        // CA x PADN
  float2* TA = U + CA * PADN;
  float2* TB = TA + tw.nA;
  float2* W16 = TB + tw.nB;
  const int tid = threadIdx.x;
  load_twiddle(tw, TA, TB, tid, T);
  fill_w16(W16, tw.A, tw.B, tw.lb, tw.nA * tw.nB, tid, T);
  const int ncols = int(pk.ncols);
  const int64_t ngrp = (ncols + CA - 1) / CA;
  const int64_t nitem = ngrp * ntiles;
  // XCD-contiguous item order (workgroup i runs on XCD i % 8)
  const int64_t per = (nitem + 7) / 8;
  const int64_t item = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (item >= nitem) return;               // whole workgroup: no barrier is skipped by part of it
  const int64_t tile = item / ngrp;
  const int c0 = int(item - tile * ngrp) * CA;
  // first radix-16 pass straight from the packed bytes: butterfly (j, p) takes
  // rows j + NB r of column c0 + p (p fastest: neighbouring lanes read
  // neighbouring bytes of one row)
  {
    constexpr int NB = NSL / 16;
    constexpr int TOT = NB * CA;
    constexpr int PER = (TOT + T - 1) / T;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int b = tid + e * T;
      if (TOT % T == 0 || b < TOT) {
        const int p = b % CA, j = b / CA;
        const int c = c0 + p;
        const int ch = c < ncols ? pk.chan[c] : -1;
        const bool use = ch >= 0 && ch <= 2048;
        float2 v[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int64_t prow = use ? packed_row(pk, tile * NSL + j + NB * r, NSL) : -1;
          const int byte = prow >= 0 ? int(pk.bytes[prow * ncols + c]) : 0;
          v[r] = make_float2(float(nib4f(byte >> 4)), (ch == 0 || ch == 2048) ? 0.f : float(nib4f(byte & 15)));
        }
        DFT<16>::run(v);
        float2* Up = U + p * PADN;
#pragma unroll
        for (int r = 0; r < 16; ++r) Up[pad16(16 * j + r)] = v[r];
      }
    }
  }
  __syncthreads();
  fft_lds<NSL, CA, T, 16, NSL, true>(U, TA, TB, tw.lb, 0, tid, W16);
  float2* Et = E + tile * NSL * int64_t(ncols);
  for (int idx = tid; idx < NSL * CA; idx += T) {
    const int f = idx / CA, p = idx - f * CA;
    if (c0 + p < ncols) Et[int64_t(f) * ncols + c0 + p] = U[p * PADN + pad16(f)];
  }
}

// Spectra along the sample axis: for each tile and f = 0..nslice/2,
// D[t][s][f][c] = 4096 ddft[4 s + c][f] (s < 1024 strips of 4 samples) from
// E[t][f][.] and E[t][-f][.], FB rows f per group in LDS.  A strip's rows f,
// f+1 are adjacent, so each store instruction writes 64-byte runs.
//
// Persistent: workgroup w runs on XCD w % 8 and walks that XCD's contiguous
// range of (tile, row group) items, so neighbouring groups (whose stores share
// lines) run on one XCD.  Thread item k is column (tid + k T) % ncols of row
// (tid + k T) / ncols of the group: its channel number is fixed, and the E
// values of the next group are fetched while the current one is transformed.
// The store reads Z and clears it, so the next scatter starts from zeros.
template <int FB, int T, int KI>
__global__ __launch_bounds__(T, 2) void ipfb_rows_spectral_kernel(const float2* __restrict__ E,
                                                                  const int* __restrict__ chan, int ncols,
                                                                  int64_t nslice, int64_t ntiles,
                                                                  float2* __restrict__ D, Twiddle tw) {
  constexpr int N = 4096;
  constexpr int NP = Padded<N>::value;
  extern __shared__ float2 smem[];
  float2* Z = smem;                        // FB x NP
  float2* TA = Z + FB * NP;
  float2* TB = TA + tw.nA;
  float2* W16 = TB + tw.nB;
  const int tid = threadIdx.x;
  load_twiddle(tw, TA, TB, tid, T);
  fill_w16(W16, tw.A, tw.B, tw.lb, tw.nA * tw.nB, tid, T);
  const int64_t nf = nslice / 2 + 1;
  const int64_t ngrp = (nf + FB - 1) / FB;
  const int64_t nitem = ngrp * ntiles;
  const int64_t per = (nitem + 7) / 8;
  const int64_t S = gridDim.x / 8;
  const int64_t xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  auto item_of = [&](int64_t it) -> int64_t {   // item of iteration it, -1 past the end
    const int64_t j = it * S + slot;
    const int64_t g = xcd * per + j;
    return (j < per && g < nitem) ? g : -1;
  };
  // the thread's items: row fb[k] of the group, column cc[k], channel ch[k] (-1: none)
  int fbk[KI], cc[KI], ch[KI];
#pragma unroll
  for (int k = 0; k < KI; ++k) {
    const int idx = tid + k * T;
    fbk[k] = idx / ncols;
    cc[k] = idx - fbk[k] * ncols;
    const int cv = idx < FB * ncols ? chan[cc[k]] : -1;
    ch[k] = (cv >= 0 && cv <= 2048) ? cv : -1;
  }
  // E values of item k in group g: a = E[f][c], b = E[-f][c]; loads from a
  // readable address when the item has none (kept apart by ch / f checks)
  auto fetch = [&](int64_t g, float2* av, float2* bv) {
#pragma unroll
    for (int k = 0; k < KI; ++k) {
      int64_t f = 0, t = 0;
      if (g >= 0) {
        t = g / ngrp;
        f = (g - t * ngrp) * FB + fbk[k];
      }
      const bool ok = g >= 0 && ch[k] >= 0 && f < nf;
      const int64_t fm = f == 0 ? 0 : nslice - f;
      const float2* Et = E + t * nslice * int64_t(ncols) + cc[k];
      av[k] = Et[(ok ? f : 0) * ncols];
      bv[k] = Et[(ok ? fm : 0) * ncols];
    }
  };
  for (int idx = tid; idx < FB * NP; idx += T) Z[idx] = make_float2(0.f, 0.f);
  float2 aA[KI], bA[KI], aB[KI], bB[KI];
  fetch(item_of(0), aA, bA);
  // one group: conj(X_f) into LDS (a forward transform of it is the
  // conjugated inverse; channel numbers are distinct, so no two items write
  // one bin), transform, store (clearing Z)
  auto transform = [&](int64_t g, const float2* av, const float2* bv) {
    const int64_t t = g / ngrp;
    const int64_t f0 = (g - t * ngrp) * FB;
    __syncthreads();   // Z is all zeros: the initial fill, or the previous group's store cleared it
#pragma unroll
    for (int k = 0; k < KI; ++k) {
      const int c = ch[k];
      if (c >= 0 && f0 + fbk[k] < nf) {
        float2* Zf = Z + fbk[k] * NP;
        Zf[pad16(c)] = cconj(av[k]);
        if (c > 0 && c < 2048) Zf[pad16(N - c)] = bv[k];
      }
    }
    __syncthreads();
    fft_lds<N, FB, T, 1, N, true>(Z, TA, TB, tw.lb, 0, tid, W16);
    // lane q = (s, fb, h): samples 4 s + 2 h, + 1 of row f0 + fb, one float4
    constexpr int NQ = (N / 2) * FB;
    float2* Dt = D + t * (N / 4) * nf * 4;
    for (int q = tid; q < NQ; q += T) {
      const int h = q & 1, fb = (q >> 1) % FB, sI = q / (2 * FB);
      const int64_t f = f0 + fb;
      float2* za = Z + fb * NP + pad16(4 * sI + 2 * h);
      const float2 va = za[0], vb = za[1];
      za[0] = make_float2(0.f, 0.f);
      za[1] = make_float2(0.f, 0.f);
      if (f < nf)
        *reinterpret_cast<float4*>(Dt + (int64_t(sI) * nf + f) * 4 + 2 * h) = make_float4(va.x, -va.y, vb.x, -vb.y);
    }
  };
  for (int64_t it = 0;; it += 2) {
    int64_t g = item_of(it);
    if (g < 0) break;
    fetch(item_of(it + 1), aB, bB);
    transform(g, aA, bA);
    g = item_of(it + 1);
    if (g < 0) break;
    fetch(item_of(it + 2), aA, bA);
    transform(g, aB, bB);
  }
}

// ---------------------------------------- iPFB column pass, any even nslice
// Chirp-z (Bluestein) form for lengths that are not powers of two: with
// c_m = exp(-i pi m^2 / n),  DFT_n(x)[k] = c_k sum_m (x_m c_m) conj(c_{k-m}),
// a circular convolution done with P-point FFTs (P >= 2n-1, power of two).
// chirp[m] = c_m (m < n) and bhat = FFT_P(b), b[m] = conj(c_m) for m < n,
// b[P-m] = conj(c_m) for 0 < m < n, are float64-derived tables from the caller.
// One workgroup = 2 adjacent columns (one complex sequence) of one tile.
template <int P>
__device__ __forceinline__ void bluestein_dft(float2* U, int n, const float2* __restrict__ chirp,
                                              const float2* __restrict__ bhat, const float2* TA, const float2* TB,
                                              int lb, int tid) {
  // U[0..n) in natural (padded) order -> DFT_n(U) in U[0..n)
  for (int m = tid; m < P; m += kTPB) {
    float2 v = make_float2(0.f, 0.f);
    if (m < n) v = cmul(U[pad16(m)], chirp[m]);
    U[pad16(m)] = v;
  }
  __syncthreads();
  fft_lds<P, 1, kTPB>(U, TA, TB, lb, 0, tid);
  const float sp = 1.0f / float(P);
  for (int m = tid; m < P; m += kTPB) U[pad16(m)] = cconj(cscale(cmul(U[pad16(m)], bhat[m]), sp));
  __syncthreads();
  fft_lds<P, 1, kTPB>(U, TA, TB, lb, 0, tid);  // conj(IFFT(.)) * P
  for (int k = tid; k < n; k += kTPB) U[pad16(k)] = cmul(cconj(U[pad16(k)]), chirp[k]);
  __syncthreads();
}

template <int P>
__global__ __launch_bounds__(kTPB) void ipfb_cols_bluestein_kernel(const float* __restrict__ y, int64_t ldy,
                                                                   int64_t tile_stride, const float2* __restrict__ filt,
                                                                   int64_t ldf, int n, int cut, float* __restrict__ out,
                                                                   int64_t ldo, int64_t out_tile_stride,
                                                                   const float2* __restrict__ chirp,
                                                                   const float2* __restrict__ bhat, Twiddle tw) {
  extern __shared__ float2 smem[];
  float2* U = smem;
  float2* TA = U + Padded<P>::value;
  float2* TB = TA + tw.nA;
  const int tid = threadIdx.x;
  load_twiddle(tw, TA, TB, tid, kTPB);
  const int64_t k0 = int64_t(blockIdx.x) * 2;
  const int64_t tile = blockIdx.y;
  const float* yt = y + tile * tile_stride;
  for (int i = tid; i < n; i += kTPB) U[pad16(i)] = reinterpret_cast<const float2*>(yt + int64_t(i) * ldy + k0)[0];
  __syncthreads();
  bluestein_dft<P>(U, n, chirp, bhat, TA, TB, tw.lb, tid);
  // filter the real pair, then conj(U') so that a forward DFT gives the inverse
  const float sc = 1.0f / float(n);
  for (int f = tid; f <= n / 2; f += kTPB) {
    const float2 a = U[pad16(f)];
    const float2 b = U[pad16((n - f) % n)];
    const float2 y0 = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
    const float2 dd = make_float2(0.5f * (a.x - b.x), 0.5f * (a.y + b.y));
    const float2 y1 = make_float2(dd.y, -dd.x);
    float2 g0 = cscale(cmul(filt[k0 * ldf + f], y0), sc);
    float2 g1 = cscale(cmul(filt[(k0 + 1) * ldf + f], y1), sc);
    if (f == 0 || 2 * f == n) { g0.y = 0.f; g1.y = 0.f; }
    U[pad16(f)] = cconj(make_float2(g0.x - g1.y, g0.y + g1.x));
    if (f != 0 && 2 * f != n) U[pad16(n - f)] = cconj(make_float2(g0.x + g1.y, -g0.y + g1.x));
  }
  __syncthreads();
  bluestein_dft<P>(U, n, chirp, bhat, TA, TB, tw.lb, tid);
  float* ot = out + tile * out_tile_stride;
  for (int i = tid; i < n - 2 * cut; i += kTPB) {
    const float2 v = U[pad16(i + cut)];
    reinterpret_cast<float2*>(ot + int64_t(i) * ldo + k0)[0] = make_float2(v.x, -v.y);
  }
}

// ------------------------------------------------------- re-PFB stage 1
// For block b and column k (row width 4096): y[r] = sum_j ts[(b+j)R1 + r][k] w[(j R1 + r) 4096 + k],
// then the R1-point real DFT down r -> A[b][q][k], q = 0..R1/2 (StageA layout).
template <int R1>
__global__ __launch_bounds__(kTPB) void repfb_stage1_kernel(const float* __restrict__ ts, const float* __restrict__ win,
                                                            int ntap, int64_t nblock, int64_t blocks_per_wg,
                                                            float2* __restrict__ A) {
  constexpr int KW = 64;
  constexpr int G = kTPB / KW;
  constexpr int H = R1 / 2;
  constexpr int64_t C = 4096;
  extern __shared__ float wl[];  // [ntap*R1][KW]
  const int tid = threadIdx.x;
  const int kk = tid % KW, g = tid / KW;
  const int64_t k = int64_t(blockIdx.x) * KW + kk;
  for (int idx = tid; idx < ntap * R1 * KW; idx += kTPB) {
    const int jr = idx / KW, c = idx - jr * KW;
    wl[idx] = win[int64_t(jr) * C + int64_t(blockIdx.x) * KW + c];
  }
  __syncthreads();
  const int64_t b0 = int64_t(blockIdx.y) * blocks_per_wg;
  int64_t b1 = b0 + blocks_per_wg;
  if (b1 > nblock) b1 = nblock;
  for (int64_t b = b0 + g; b < b1; b += G) {
    float y[R1];
#pragma unroll
    for (int r = 0; r < R1; ++r) y[r] = 0.f;
    for (int j = 0; j < ntap; ++j) {
      const float* src = ts + ((b + j) * R1) * C + k;
      const float* wj = wl + (j * R1) * KW + kk;
#pragma unroll
      for (int r = 0; r < R1; ++r) y[r] = fmaf(src[int64_t(r) * C], wj[r * KW], y[r]);
    }
    float2* dst = A + StageA<R1>::at(b, 0, k);
    if constexpr (R1 == 1) {
      dst[0] = make_float2(y[0], 0.f);
    } else {
      float2 z[H];
#pragma unroll
      for (int n = 0; n < H; ++n) z[n] = make_float2(y[2 * n], y[2 * n + 1]);
      DFT<H>::run(z);
      float bin0 = 0.f;
      static_for<0, H + 1>([&](auto Q) {
        constexpr int q = decltype(Q)::value;
        const float2 a = z[q % H];
        const float2 bb = z[(H - q) % H];
        const float2 e = make_float2(0.5f * (a.x + bb.x), 0.5f * (a.y - bb.y));
        const float2 d = make_float2(0.5f * (a.x - bb.x), 0.5f * (a.y + bb.y));
        const float2 o = make_float2(d.y, -d.x);
        const float2 wo = twc<R1, q>(o);
        // bins 0 and H are real and share slot 0
        if constexpr (q == 0) bin0 = e.x + wo.x;
        else if constexpr (q == H) dst[0] = make_float2(bin0, e.x + wo.x);
        else dst[q * 4] = make_float2(e.x + wo.x, e.y + wo.y);
      });
    }
  }
}

// Same result for R1 = 4S >= 16, with each column's R1-point DFT split over 4
// lanes: lane t owns rows r = 4s + t (s < S), does the polyphase sum and an
// S-point real DFT Z_t, then, after an exchange through LDS, the radix-4
// combine  Y[q0 + S p] = sum_t (W_R1^(q0 t) Z_t[q0]) W_4^(p t).
// A workgroup = 64 columns of one block; wave t holds lane t of all 64
// columns, so every global load and store of a wave is 64 consecutive
// columns; no spill-prone 64-register arrays.
// Block b of workgroup row y: y itself, or list.b[y] when a list is given.
struct BlockList {
  int n;
  int b[kBlockListCap];
};

// S-point real DFT of the polyphase sums y[s] (rows 4 s + t of column k of
// block b), radix-4 combine through the exchange buffer zx and the StageA store
// of the block's bins (when store is set); one lane per (t, column), four
// lanes per column.  Every lane of the workgroup calls it (one barrier).
template <int R1>
__device__ __forceinline__ void stage1_split_finish(const float* y, int t, int cl, int64_t b, int64_t k,
                                                    float2 (*zx)[R1 / 8 + 1][64], const float2* wr,
                                                    float2* __restrict__ A, bool store = true) {
  constexpr int S = R1 / 4;
  // S-point real DFT of y -> Z_t[q0], q0 = 0..S/2
  {
    float2 z[S / 2];
#pragma unroll
    for (int n = 0; n < S / 2; ++n) z[n] = make_float2(y[2 * n], y[2 * n + 1]);
    DFT<S / 2>::run(z);
    static_for<0, S / 2 + 1>([&](auto Q) {
      constexpr int q = decltype(Q)::value;
      const float2 a = z[q % (S / 2)];
      const float2 bb = z[(S / 2 - q) % (S / 2)];
      const float2 e = make_float2(0.5f * (a.x + bb.x), 0.5f * (a.y - bb.y));
      const float2 d = make_float2(0.5f * (a.x - bb.x), 0.5f * (a.y + bb.y));
      const float2 o = make_float2(d.y, -d.x);
      const float2 wo = twc<S, q>(o);
      zx[t][q][cl] = make_float2(e.x + wo.x, e.y + wo.y);
    });
  }
  __syncthreads();
  // lane u combines q0 = u, u+4, ...: outputs q = q0 (p=0), q0+S (p=1), and q = 2S = H for q0 = 0
  // (bins 0 and H are real and share slot 0 of the StageA layout)
  float2* dst = A + StageA<R1>::at(b, 0, k);
  for (int q0 = t; store && q0 < S; q0 += 4) {
    float2 v[4];
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) {
      float2 zz = (2 * q0 <= S) ? zx[tt][q0][cl] : cconj(zx[tt][S - q0][cl]);
      v[tt] = cmul(zz, wr[(q0 * tt) % R1]);
    }
    // 4-point DFT over tt: p = 0, 1, 2
    const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    const float2 s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
    if (q0 == 0) dst[0] = make_float2(s02.x + s13.x, s02.x - s13.x);        // p = 0 and p = 2 (bins 0, H)
    else dst[q0 * 4] = cadd(s02, s13);                                        // p = 0
    dst[(q0 + S) * 4] = make_float2(d02.x + d13.y, d02.y - d13.x);            // p = 1: d02 - i d13
  }
}


template <int R1>
__global__ __launch_bounds__(kTPB) void repfb_stage1_split_kernel(const float* __restrict__ ts,
                                                                  const float* __restrict__ win, int ntap,
                                                                  float2* __restrict__ A, BlockList list) {
  constexpr int S = R1 / 4;
  constexpr int ZH = S / 2 + 1;      // stored half spectrum of Z_t
  constexpr int64_t C = 4096;
  __shared__ float2 zx[4][ZH][64];
  __shared__ float2 wr[R1];
  const int tid = threadIdx.x;
  const int t = tid >> 6, cl = tid & 63;  // lane (= wave) within the column quad, column 0..63
  const int64_t k = int64_t(blockIdx.x) * 64 + cl;
  const int64_t b = list.n ? list.b[blockIdx.y] : int64_t(blockIdx.y);
  if (tid < R1) {
    double sn, cs;
    sincospi(-2.0 * double(tid) / double(R1), &sn, &cs);
    wr[tid] = make_float2(float(cs), float(sn));
  }
  // polyphase sums of this lane's rows
  float y[S];
#pragma unroll
  for (int sI = 0; sI < S; ++sI) y[sI] = 0.f;
  for (int j = 0; j < ntap; ++j) {
    const float* src = ts + ((b + j) * R1 + t) * C + k;
    const float* wj = win + (int64_t(j) * R1 + t) * C + k;
#pragma unroll
    for (int sI = 0; sI < S; ++sI) y[sI] = fmaf(src[int64_t(4 * sI) * C], wj[int64_t(4 * sI) * C], y[sI]);
  }
  stage1_split_finish<R1>(y, t, cl, b, k, zx, wr, A);
}

// Edge blocks in runs of up to kRunMax consecutive blocks b0 .. b0 + nb - 1
// (list entry b0 * 8 + nb).  Block b0 + bl reads rows R1 (b0 + bl + j) + r as
// tap j, so the run's nb + ntap - 1 block rows are walked once: 1024 lanes
// per 64 columns, lane (u0, column) owns rows r = u0 + 16 u (u < R1 / 16) of
// every block row and adds row m to block m - j as tap j, its window weights
// in registers.  The taps of each block are summed in the order
// j = 0 .. ntap - 1 with the same fmaf as the split kernel, so both give the
// same bits.  The block sums go through LDS to the split kernel's DFT lanes,
// 256 per block, all blocks of the run at once.
constexpr int kRunMax = 4;
struct RunList {
  int n;
  int e[kBlockListCap];
};

template <int R1>
constexpr size_t run_kernel_lds() {
  constexpr size_t y = size_t(kRunMax) * R1 * 64 * sizeof(float);
  constexpr size_t z = size_t(kRunMax) * 4 * (R1 / 8 + 1) * 64 * sizeof(float2);
  return y > z ? y : z;
}

template <int R1, int NT>
__global__ __launch_bounds__(1024) void repfb_stage1_run_kernel(const float* __restrict__ ts,
                                                                const float* __restrict__ win, int ntap,
                                                                float2* __restrict__ A, RunList list) {
  constexpr int S = R1 / 4;
  constexpr int ZH = S / 2 + 1;
  constexpr int RPT = R1 / 16;              // rows per lane of one block row
  constexpr int64_t C = 4096;
  static_assert(R1 % 16 == 0, "16 row lanes per column");
  extern __shared__ float run_smem[];
  float* Y = run_smem;                      // [kRunMax][R1][64] polyphase sums
  auto zx = reinterpret_cast<float2 (*)[ZH][64]>(run_smem);   // [kRunMax * 4][ZH][64] once Y is read
  __shared__ float2 wr[R1];
  const int tid = threadIdx.x;
  const int cl = tid & 63;
  const int64_t k = int64_t(blockIdx.x) * 64 + cl;
  const int e = list.e[blockIdx.y];
  const int64_t b0 = e >> 3;
  const int nb = e & 7;
  if (tid < R1) {
    double sn, cs;
    sincospi(-2.0 * double(tid) / double(R1), &sn, &cs);
    wr[tid] = make_float2(float(cs), float(sn));
  }
  {
    const int u0 = tid >> 6;
    float y[kRunMax][RPT], w[NT][RPT];
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
      const int64_t r = u0 + 16 * u;
#pragma unroll
      for (int bl = 0; bl < kRunMax; ++bl) y[bl][u] = 0.f;
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const float v = win[(int64_t(j < ntap ? j : 0) * R1 + r) * C + k];   // in bounds for every j
        w[j][u] = j < ntap ? v : 0.f;
      }
    }
    const int nrow = nb + ntap - 1;
#pragma unroll
    for (int m = 0; m < kRunMax + NT - 1; ++m) {
      if (m < nrow) {
#pragma unroll
        for (int u = 0; u < RPT; ++u) {
          const float x = ts[((b0 + m) * R1 + u0 + 16 * u) * C + k];
#pragma unroll
          for (int bl = 0; bl < kRunMax; ++bl) {
            const int j = m - bl;
            if (j >= 0 && j < NT && bl < nb && j < ntap) y[bl][u] = fmaf(x, w[j][u], y[bl][u]);
          }
        }
      }
    }
#pragma unroll
    for (int bl = 0; bl < kRunMax; ++bl)
#pragma unroll
      for (int u = 0; u < RPT; ++u) Y[(bl * R1 + u0 + 16 * u) * 64 + cl] = y[bl][u];
  }
  __syncthreads();
  const int bl = tid >> 8, t = (tid >> 6) & 3;
  float ys[S];
#pragma unroll
  for (int sI = 0; sI < S; ++sI) ys[sI] = Y[(bl * R1 + 4 * sI + t) * 64 + cl];
  __syncthreads();   // Y is read: its space becomes the blocks' exchange buffers
  stage1_split_finish<R1>(ys, t, cl, b0 + bl, k, zx + 4 * bl, wr, A, bl < nb);
}

// ------------------------------------------------------- re-PFB stage 2
// Workgroup (q1 <= R1/2, b): u[k] = A[b][q1][k] (StageA layout) W_L^(q1 k), U = FFT_4096(u).
// Fine channel q = q1 + R1 q2 is U[q2].  Because the block is real, channel
// q' = (R1 - q1) + R1 q2' is conj(X[L - q']) = conj(U[4095 - q2']), so one
// transform serves both q1 and its mirror R1 - q1 and A is read once.
// Kept channels q in [c_lo, c_hi) go to
// out[(b/bpj)*sj + (b%bpj)*sb + (q - c_lo)*sc] (complex elements; bpj = blocks per job).
__device__ __forceinline__ void q2_range(int64_t q1, int64_t R1, int64_t c_lo, int64_t c_hi, int64_t* lo, int64_t* hi) {
  int64_t a = (c_lo - q1 + R1 - 1) / R1;
  if (c_lo - q1 < 0) a = 0;
  const int64_t z = (c_hi - 1 - q1) >= 0 ? (c_hi - 1 - q1) / R1 : -1;
  *lo = a;
  *hi = z < 4095 ? z : 4095;
}

// Boundary halves of the split column pass, read by stage 2: block b of the
// stream window whose rows straddle the boundary between the tile holding its
// first row (window tile tr) and the next tile of the window also takes slot
// ntap tr + (b - first block straddling that boundary) of P (input m at
// P + m p_stride), added to its A before the transform.
struct SplitRead {
  const float2* P;     // nullptr: A alone
  int64_t p_stride;    // complex elements per input
  int64_t g_first;     // window row of the window's first tile's first kept row (<= 0)
  int64_t rows;        // rows per tile
  int64_t ntiles;      // tiles of the window
  int ntap;
};
// 32-bit unsigned arithmetic: block rows and window rows are >= 0 and below
// 2^31 (checked by the host), and the next tile's first row bt is > 0.
__host__ __device__ inline int64_t split_slot(const SplitRead& sp, int64_t b, int64_t R1) {
  const uint32_t r0 = uint32_t(R1 * b);
  const uint32_t off = uint32_t(r0 - sp.g_first);             // row of the block's first row from tile 0's first row
  const uint32_t tr = off / uint32_t(sp.rows);                 // tile holding the block's first row
  if (int64_t(tr) + 1 >= sp.ntiles) return -1;                // no next tile: rows past the window's tiles are zero
  const uint32_t bt = uint32_t(sp.g_first + (int64_t(tr) + 1) * sp.rows);   // the next tile's first row
  if (r0 + uint32_t(R1 * sp.ntap) <= bt) return -1;            // the block lies in its first tile
  return int64_t(sp.ntap) * tr + (b - (int64_t(bt / uint32_t(R1)) - sp.ntap + 1));
}

// SPLIT: add the tile-boundary halves of sp (a separate instantiation, so
// the plain form keeps its register budget).
template <int R1, bool SPLIT = false>
__global__ __launch_bounds__(kTPB) void repfb_stage2_kernel(const float2* __restrict__ A, int64_t a_stride, int ninput,
                                                            int64_t nblock, int64_t c_lo, int64_t c_hi,
                                                            float2* __restrict__ out, int64_t sm, int64_t sb, int64_t sc,
                                                            int64_t bpj, int64_t sj, Twiddle tw4, SplitRead sp) {
  constexpr int N = 4096;
  constexpr int H = R1 / 2;
  constexpr int NPAD = Padded<N>::value;
  extern __shared__ float2 smem[];
  float2* Z = smem;
  float2* TA = Z + NPAD;
  float2* TB = TA + tw4.nA;
  float2* W16 = TB + tw4.nB;              // W_256^(m r), fill_w16
  const int tid = threadIdx.x;
  load_twiddle(tw4, TA, TB, tid, kTPB);
  fill_w16(W16, tw4.A, tw4.B, tw4.lb, tw4.nA * tw4.nB, tid, kTPB);
  // Workgroup i runs on XCD i % 8: the (H+1) * ninput workgroups of one block
  // go to one XCD back to back, inputs fastest, so the neighbouring output
  // elements they write (input m of block b at out + m sm) merge in its L2.
  const int64_t total = int64_t(H + 1) * nblock * ninput;
  const int64_t per = (total + 7) / 8;
  const int64_t item = int64_t(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (item >= total) return;   // the whole workgroup: padding of the 8-way split
  const int m_in = int(item % ninput);
  const int64_t qb = item / ninput;
  const int q1 = int(qb % (H + 1));
  const int64_t b = qb / (H + 1);
  // column k = tid + 256 r of bin q1 (StageA layout; bins 0 and H from slot 0)
  const int slot = (q1 == 0 || q1 == H) ? 0 : q1;
  const float2* src = A + m_in * a_stride + StageA<R1>::at(b, slot, tid);
  // First radix-16 pass straight from memory: thread j takes k = j + 256 r
  // (r < 16), twiddles, transforms and stores the Stockham pass output
  // Z[16 j + r]; the remaining passes start at sub-transform size 16.  The
  // loads are issued before the barrier that publishes the LDS tables.
  static_assert(N / 16 == kTPB, "one first-pass butterfly per thread");
  float2 v[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = src[r * (64 * StageA<R1>::QP * 4)];
  if constexpr (SPLIT) {   // split column pass: add the block's second half (uniform per workgroup)
    const int64_t e = split_slot(sp, b, R1);
    if (e >= 0) {
      const float2* ps = sp.P + m_in * sp.p_stride + e * StageA<R1>::kBlock + StageA<R1>::at(0, slot, tid);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const float2 w = ps[r * (64 * StageA<R1>::QP * 4)];
        v[r] = make_float2(v[r].x + w.x, v[r].y + w.y);
      }
    }
  }
  // W_L^(q1 k) = W_L^(q1 c) W_4096^(q1 a) for k = R1 a + c; c = tid mod R1 is
  // the same for all of this thread's k (kTPB is a multiple of R1).  The
  // float64 sincospi here measured 3-5% faster than reading the same value
  // from an L2-resident table (the extra load waits behind the A loads).
  static_assert(kTPB % R1 == 0, "thread stride must keep k mod R1 fixed");
  float2 wc = make_float2(1.f, 0.f);
  if (q1 != 0) {
    double sn, cs;
    sincospi(-2.0 * double(q1 * (tid % R1)) / double(int64_t(N) * R1), &sn, &cs);
    wc = make_float2(float(cs), float(sn));
  }
  __syncthreads();
  {
    if (q1 == 0) {
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r].y = 0.f;
    } else if (q1 == H) {
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = make_float2(v[r].y, 0.f);
    }
    if (q1 != 0) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int k = tid + 256 * r;
        v[r] = cmul(v[r], cmul(wc, tw_lookup(TA, TB, tw4.lb, (q1 * (k / R1)) & (N - 1))));
      }
    }
    DFT<16>::run(v);
#pragma unroll
    for (int r = 0; r < 16; ++r) Z[pad16(16 * tid + r)] = v[r];
  }
  __syncthreads();
  int64_t alo, ahi, mlo = 0, mhi = -1;
  q2_range(q1, R1, c_lo, c_hi, &alo, &ahi);
  const bool mirror = q1 > 0 && q1 < H;
  if (mirror) q2_range(R1 - q1, R1, c_lo, c_hi, &mlo, &mhi);
  const int64_t na = ahi >= alo ? ahi - alo + 1 : 0;
  const int64_t nm = mhi >= mlo ? mhi - mlo + 1 : 0;
  float2* ob = out + m_in * sm + (b / bpj) * sj + (b % bpj) * sb;
  if (na + nm <= 512) {
    // passes NS = 1, 16; the last radix-16 pass only for the wanted bins:
    // U[j] = sum_m Z[(j % 256) + 256 m] W_4096^(m j)
    fft_lds<N, 1, kTPB, 16, 256, true>(Z, TA, TB, tw4.lb, 0, tid, W16);
    for (int64_t t = tid; t < na + nm; t += kTPB) {
      const bool dir = t < na;
      const int64_t q2 = dir ? alo + t : mlo + (t - na);
      const int j = dir ? int(q2) : int(N - 1 - q2);
      const int jl = j & 255;
      float ax = 0.f, ay = 0.f;
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const float2 v = Z[pad16(jl + 256 * m)];
        const float2 w = tw_lookup(TA, TB, tw4.lb, (m * j) & (N - 1));
        ax = fmaf(v.x, w.x, fmaf(-v.y, w.y, ax));
        ay = fmaf(v.x, w.y, fmaf(v.y, w.x, ay));
      }
      const int64_t q = dir ? q1 + int64_t(R1) * q2 : (R1 - q1) + int64_t(R1) * q2;
      ob[(q - c_lo) * sc] = dir ? make_float2(ax, ay) : make_float2(ax, -ay);
    }
  } else {
    fft_lds<N, 1, kTPB, 16, N, true>(Z, TA, TB, tw4.lb, 0, tid, W16);
    for (int64_t t = tid; t < na + nm; t += kTPB) {
      const bool dir = t < na;
      const int64_t q2 = dir ? alo + t : mlo + (t - na);
      const float2 v = Z[pad16(dir ? int(q2) : int(N - 1 - q2))];
      const int64_t q = dir ? q1 + int64_t(R1) * q2 : (R1 - q1) + int64_t(R1) * q2;
      ob[(q - c_lo) * sc] = dir ? v : cconj(v);
    }
  }
}

// ------------------------------------------------------------ c2c rows
// direction -1: forward; +1: inverse with 1/N (pycufft.py:92-112).
template <int N, int F>
__global__ __launch_bounds__(kTPB) void fft_rows_kernel(const float2* __restrict__ x, int64_t ldx, int64_t nrows,
                                                        float2* __restrict__ out, int64_t ldo, int inverse, Twiddle tw) {
  constexpr int NPD = Padded<N>::value;
  extern __shared__ float2 smem[];
  float2* Z = smem;
  float2* TA = Z + F * NPD;
  float2* TB = TA + tw.nA;
  const int tid = threadIdx.x;
  load_twiddle(tw, TA, TB, tid, kTPB);
  const int64_t row0 = int64_t(blockIdx.x) * F;
  for (int idx = tid; idx < F * N; idx += kTPB) {
    const int f = idx / N, n = idx - f * N;
    const int64_t row = row0 + f;
    float2 v = row < nrows ? x[row * ldx + n] : make_float2(0.f, 0.f);
    if (inverse) v.y = -v.y;
    Z[f * NPD + pad16(n)] = v;
  }
  __syncthreads();
  fft_lds<N, F, kTPB>(Z, TA, TB, tw.lb, 0, tid);
  const float sc = inverse ? 1.0f / float(N) : 1.0f;
  for (int idx = tid; idx < F * N; idx += kTPB) {
    const int f = idx / N, n = idx - f * N;
    const int64_t row = row0 + f;
    if (row >= nrows) continue;
    float2 v = Z[f * NPD + pad16(n)];
    if (inverse) v = make_float2(v.x * sc, -v.y * sc);
    out[row * ldo + n] = v;
  }
}

// filt = 1/conj(W) [* |W|^2/(t^2+|W|^2)(1+t^2)] from a given W, in float64.
__global__ void filter_from_matft_kernel(const float2* __restrict__ W, int64_t n, double thresh, float2* __restrict__ filt) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const double wr = W[i].x, wi = W[i].y;
    const double a = wr * wr + wi * wi;
    const double sc = thresh > 0.0 ? (1.0 + thresh * thresh) / (a + thresh * thresh) : 1.0 / a;
    filt[i] = make_float2(float(wr * sc), float(wi * sc));
  }
}

__global__ void rowmap_fill_kernel(int* rowmap, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) rowmap[i] = -1;
}
__global__ void rowmap_scatter_kernel(const int64_t* __restrict__ specnum, int64_t npresent, int64_t start, int64_t acclen,
                                      int64_t row_base, int* __restrict__ rowmap) {
  const int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (p >= npresent) return;
  int64_t r = (specnum[p] - start) % acclen;
  if (r < 0) r += acclen;
  // spectra that land on one row (a repeated spec_num, or spec_nums acclen
  // apart): the last one in packet order is kept, as numpy's fancy assignment
  // of make_continuous keeps it
  atomicMax(&rowmap[r], int(row_base + p));
}

}  // namespace alb

using namespace alb;

// ---------------------------------------------------------------- host side
namespace {

// columns per workgroup of the nslice = 4096 column passes when the knob is unset
constexpr int64_t kColsWidth = 4;

// strips per group of the column-pass workgroup order (tuning knob).  Unset:
// 2 for launches of up to 4 tiles (the tiles of a strip pair run back to back
// on one XCD, so its filter taps are fetched from HBM once per launch; 5-8%
// faster than 32 or 64 at 4 tiles) and 64 above (the workgroups resident on
// an XCD then cover one tile's strips; 10% faster than 2 at 16 tiles).
int cols_group(int64_t ntiles) {
  const int64_t g = tune_get("ipfb_cols_group");
  return g > 0 ? int(g) : (ntiles <= 4 ? 2 : 64);
}

template <int M, int F, int TT = kTPB>
int launch_irfft(DenseRows d, int64_t nrows, float* out, int64_t ldo, hipStream_t s) {
  Twiddle tw;
  int rc = get_twiddle(2 * M, &tw, s);
  if (rc) return rc;
  const size_t lds = sizeof(float2) * (size_t(F) * Padded<M>::value + tw.nA + tw.nB);
  const int64_t ngroups = (nrows + F - 1) / F;
  ALB_REQUIRE(ngroups < (int64_t(1) << 31), "irfft: too many rows");
  hipLaunchKernelGGL((irfft_rows_kernel<M, F, TT>), dim3(unsigned(ngroups)), dim3(TT), lds, s, d, nrows, out, ldo, tw);
  ALB_LAUNCH_CHECK("irfft_rows_kernel");
  return ALB_OK;
}

// Persistent packed-row launch: as many workgroups as fit on the device at
// once (a multiple of 8, so every XCD gets the same count), at most one per
// row group.
template <int F, int TT, int KI>
int launch_irfft_packed4(const Packed4Rows& pk, int64_t nrows, int64_t nslice, float* out, int64_t ldo,
                         hipStream_t s) {
  Twiddle tw;
  int rc = get_twiddle(4096, &tw, s);
  if (rc) return rc;
  const size_t lds = sizeof(float2) * (size_t(F) * packed4_row_stride<F>() + tw.nA + tw.nB + kW16Size);
  const int64_t ngroups = (nrows + F - 1) / F;
  ALB_REQUIRE(nrows < (int64_t(1) << 31) - F && pk.tile0 >= 0 && pk.tile0 < (int64_t(1) << 30),
              "irfft: too many rows");
  static std::atomic<int> resident[64];   // one cache per instantiation
  int dev = 0;
  ALB_HIP(hipGetDevice(&dev));
  int nres = dev < 64 ? resident[dev].load() : 0;
  if (nres == 0) {
    int ncu = 0, occ = 0;
    ALB_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    ALB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, irfft_packed4_kernel<F, TT, KI>, TT, lds));
    nres = ncu * (occ > 0 ? occ : 1);
    if (dev < 64) resident[dev].store(nres);
  }
  const int64_t per = (ngroups + 7) / 8;
  int64_t S = nres / 8 > 0 ? nres / 8 : 1;
  if (S > per) S = per;
  hipLaunchKernelGGL((irfft_packed4_kernel<F, TT, KI>), dim3(unsigned(8 * S)), dim3(TT), lds, s, pk, nrows, nslice,
                     out, ldo, tw);
  ALB_LAUNCH_CHECK("irfft_packed4_kernel");
  return ALB_OK;
}

// packed 4-bit rows (2049-channel spectra, n = 4096): two rows per 256-thread
// workgroup, or four per 512 threads (each store instruction writing whole
// 128-byte lines of the blocked layout).  Unset knob: two up to 4 tiles per
// launch, whose output stays in the Infinity Cache (within 3% either way),
// four above (10% faster per tile at 16 tiles, 100 and 300 channels).
int irfft_packed4_dispatch(const Packed4Rows& pk, int64_t nrows, int64_t nslice, float* out, int64_t ldo,
                           hipStream_t s) {
  const int64_t knob = tune_get("ipfb_rows_group");
  const int group = knob > 0 ? int(knob) : (nrows > 4 * nslice ? 4 : 2);
  // prefetched (row, column) items per thread: enough for every item when it fits
  if (group == 4) {
    if (4 * pk.ncols <= 2 * 512) return launch_irfft_packed4<4, 512, 2>(pk, nrows, nslice, out, ldo, s);
    return launch_irfft_packed4<4, 512, 4>(pk, nrows, nslice, out, ldo, s);
  }
  if (2 * pk.ncols <= 2 * kTPB) return launch_irfft_packed4<2, kTPB, 2>(pk, nrows, nslice, out, ldo, s);
  return launch_irfft_packed4<2, kTPB, 4>(pk, nrows, nslice, out, ldo, s);
}

int irfft_dispatch(int64_t n, DenseRows d, int64_t nrows, float* out, int64_t ldo, hipStream_t s) {
  switch (n) {
    case 64: return launch_irfft<32, 64>(d, nrows, out, ldo, s);
    case 128: return launch_irfft<64, 32>(d, nrows, out, ldo, s);
    case 256: return launch_irfft<128, 16>(d, nrows, out, ldo, s);
    case 512: return launch_irfft<256, 8>(d, nrows, out, ldo, s);
    case 1024: return launch_irfft<512, 4>(d, nrows, out, ldo, s);
    case 2048: return launch_irfft<1024, 2>(d, nrows, out, ldo, s);
    case 4096: return launch_irfft<2048, 2>(d, nrows, out, ldo, s);
    case 8192: return launch_irfft<4096, 1>(d, nrows, out, ldo, s);
    case 16384: return launch_irfft<8192, 1>(d, nrows, out, ldo, s);
    default:
      set_error("irfft: length %lld not supported (power of two 64..16384)", (long long)n);
      return ALB_ENOTSUP;
  }
}

template <int M, int F>
int launch_rfft(const float* x, int64_t ldx, int64_t nrows, const float* win, int ntap, int64_t c0, int64_t c1,
                float2* out, int64_t ldo, hipStream_t s) {
  Twiddle tw;
  int rc = get_twiddle(2 * M, &tw, s);
  if (rc) return rc;
  const size_t lds = sizeof(float2) * (size_t(F) * Padded<M>::value + tw.nA + tw.nB);
  const unsigned grid = unsigned((nrows + F - 1) / F);
  hipLaunchKernelGGL((rfft_rows_kernel<M, F>), dim3(grid), dim3(kTPB), lds, s, x, ldx, nrows, win, ntap, c0, c1, out,
                     ldo, tw);
  ALB_LAUNCH_CHECK("rfft_rows_kernel");
  return ALB_OK;
}

int rfft_dispatch(int64_t n, const float* x, int64_t ldx, int64_t nrows, const float* win, int ntap, int64_t c0,
                  int64_t c1, float2* out, int64_t ldo, hipStream_t s) {
  switch (n) {
    case 64: return launch_rfft<32, 64>(x, ldx, nrows, win, ntap, c0, c1, out, ldo, s);
    case 128: return launch_rfft<64, 32>(x, ldx, nrows, win, ntap, c0, c1, out, ldo, s);
    case 256: return launch_rfft<128, 16>(x, ldx, nrows, win, ntap, c0, c1, out, ldo, s);
    case 512: return launch_rfft<256, 8>(x, ldx, nrows, win, ntap, c0, c1, out, ldo, s);
    case 1024: return launch_rfft<512, 4>(x, ldx, nrows, win, ntap, c0, c1, out, ldo, s);
    case 2048: return launch_rfft<1024, 2>(x, ldx, nrows, win, ntap, c0, c1, out, ldo, s);
    case 4096: return launch_rfft<2048, 2>(x, ldx, nrows, win, ntap, c0, c1, out, ldo, s);
    case 8192: return launch_rfft<4096, 1>(x, ldx, nrows, win, ntap, c0, c1, out, ldo, s);
    case 16384: return launch_rfft<8192, 1>(x, ldx, nrows, win, ntap, c0, c1, out, ldo, s);
    default:
      set_error("rfft: length %lld not supported (power of two 64..16384)", (long long)n);
      return ALB_ENOTSUP;
  }
}

template <int NSL, int CW, int T>
int launch_cols(const float* y, int64_t ldy, int64_t tile_stride, int64_t ntiles, const float2* filt, int64_t ldf,
                int64_t ncols, int cut, float* out, int64_t ldo, int64_t out_tile_stride, hipStream_t s) {
  Twiddle tw;
  int rc = get_twiddle(NSL, &tw, s);
  if (rc) return rc;
  const size_t lds = sizeof(float2) * (size_t(CW / 2) * Padded<NSL>::value + tw.nA + tw.nB + kW16Size);
  ALB_REQUIRE(ncols % CW == 0, "ipfb column pass: width %lld not a multiple of %d", (long long)ncols, CW);
  ALB_REQUIRE((ncols / CW) * ntiles < (int64_t(1) << 31), "ipfb column pass: too many tiles");
  hipLaunchKernelGGL((ipfb_cols_kernel<NSL, CW, T>), dim3(unsigned((ncols / CW) * ntiles)), dim3(T), lds, s, y,
                     ldy, tile_stride, filt, ldf, ncols, cut, out, ldo, out_tile_stride, cols_group(ntiles), tw);
  ALB_LAUNCH_CHECK("ipfb_cols_kernel");
  return ALB_OK;
}

int cols_dispatch(int64_t nslice, const float* y, int64_t ldy, int64_t tile_stride, int64_t ntiles, const float2* filt,
                  int64_t ldf, int64_t ncols, int cut, float* out, int64_t ldo, int64_t out_tile_stride,
                  hipStream_t s) {
  if (nslice == 4096) {
    // strip width (columns per workgroup): 8 -> one 1024-thread workgroup per
    // CU, 4 -> two of 512, 2 -> four of 256
    const int64_t kw = tune_get("ipfb_cols_width");
    switch (kw > 0 ? kw : kColsWidth) {
      case 2: return launch_cols<4096, 2, 256>(y, ldy, tile_stride, ntiles, filt, ldf, ncols, cut, out, ldo, out_tile_stride, s);
      case 4: return launch_cols<4096, 4, 512>(y, ldy, tile_stride, ntiles, filt, ldf, ncols, cut, out, ldo, out_tile_stride, s);
      default: break;
    }
  }
  switch (nslice) {
#define CASE(n, w, t) case n: return launch_cols<n, w, t>(y, ldy, tile_stride, ntiles, filt, ldf, ncols, cut, out, ldo, out_tile_stride, s);
    CASE(16, 8, 256) CASE(32, 8, 256) CASE(64, 8, 256) CASE(128, 8, 256) CASE(256, 8, 256) CASE(512, 8, 256)
    CASE(1024, 8, 256) CASE(2048, 8, 512) CASE(4096, 8, 1024) CASE(8192, 4, 1024) CASE(16384, 2, 1024)
#undef CASE
    default:
      set_error("ipfb: nslice %lld not supported (power of two 16..16384)", (long long)nslice);
      return ALB_ENOTSUP;
  }
}

template <int R1>
int launch_repfb_stage1(const float* ts, const float* win, int ntap, int64_t nblock, float2* A, hipStream_t s) {
  if constexpr (R1 >= 16) {
    ALB_REQUIRE(nblock <= 65535, "alb_repfb: at most 65535 blocks per call");
    BlockList none{};
    hipLaunchKernelGGL((repfb_stage1_split_kernel<R1>), dim3(4096 / 64, unsigned(nblock)), dim3(kTPB), 0, s, ts, win,
                       ntap, A, none);
    ALB_LAUNCH_CHECK("repfb_stage1_split_kernel");
    return ALB_OK;
  }
  const int64_t bpw = 16;
  const size_t lds1 = sizeof(float) * size_t(ntap) * R1 * 64;
  ALB_REQUIRE(lds1 <= 160 * 1024, "re-PFB: window strip does not fit in LDS");
  hipLaunchKernelGGL((repfb_stage1_kernel<R1>), dim3(4096 / 64, unsigned((nblock + bpw - 1) / bpw)), dim3(kTPB), lds1, s,
                     ts, win, ntap, nblock, bpw, A);
  ALB_LAUNCH_CHECK("repfb_stage1_kernel");
  return ALB_OK;
}

template <int R1>
int launch_repfb_stage1_list(const float* ts, const float* win, int ntap, const int* blocks, int64_t n, float2* A,
                             hipStream_t s) {
  if constexpr (R1 >= 16) {
    if (ntap <= 8) {
      // consecutive entries b, b + 1, ... form runs of up to kRunMax blocks
      RunList rl{};
      auto flush = [&]() -> int {
        if (rl.n == 0) return ALB_OK;
        constexpr size_t lds = run_kernel_lds<R1>();
        if (ntap <= 4)
          hipLaunchKernelGGL((repfb_stage1_run_kernel<R1, 4>), dim3(4096 / 64, unsigned(rl.n)), dim3(1024), lds, s,
                             ts, win, ntap, A, rl);
        else
          hipLaunchKernelGGL((repfb_stage1_run_kernel<R1, 8>), dim3(4096 / 64, unsigned(rl.n)), dim3(1024), lds, s,
                             ts, win, ntap, A, rl);
        ALB_LAUNCH_CHECK("repfb_stage1_run_kernel");
        rl.n = 0;
        return ALB_OK;
      };
      for (int64_t i = 0; i < n;) {
        int len = 1;
        while (len < kRunMax && i + len < n && blocks[i + len] == blocks[i] + len) ++len;
        if (rl.n == kBlockListCap) {
          const int rc = flush();
          if (rc != ALB_OK) return rc;
        }
        rl.e[rl.n++] = blocks[i] * 8 + len;
        i += len;
      }
      return flush();
    }
    for (int64_t i0 = 0; i0 < n; i0 += kBlockListCap) {
      BlockList bl{};
      bl.n = int(n - i0 < kBlockListCap ? n - i0 : kBlockListCap);
      for (int i = 0; i < bl.n; ++i) bl.b[i] = blocks[i0 + i];
      hipLaunchKernelGGL((repfb_stage1_split_kernel<R1>), dim3(4096 / 64, unsigned(bl.n)), dim3(kTPB), 0, s, ts, win,
                         ntap, A, bl);
      ALB_LAUNCH_CHECK("repfb_stage1_split_kernel");
    }
    return ALB_OK;
  } else {
    set_error("alb_repfb_stage1_list: osamp %d not supported (16..64)", R1);
    return ALB_ENOTSUP;
  }
}

// Shape limits of the fused kernel: lane slots (2 per thread), the exchange
// buffer (reuses the strip's LDS) and the window strip.
template <int NSL, int CW, int T, int R1, bool SPLIT = false>
int check_cols_repfb(int cut, int ntap) {
  constexpr int NP = CW / 2, S = R1 / 4;
  constexpr int SP = S >= 16 ? S + 1 : S;
  const int64_t nkeep = NSL - 2 * int64_t(cut);
  const int64_t nint_max = nkeep / R1 - ntap + 1;
  if (ntap < 1 || ntap > 8) {
    set_error("fused iPFB/re-PFB: ntap must be 1..8");
    return ALB_ENOTSUP;
  }
  if constexpr (SPLIT) {
    // blocks touching one tile: at most ceil((nkeep - 1) / R1) + ntap (the split area holds NSL / R1 + 4)
    const int64_t nblk_max = (nkeep + R1 - 2) / R1 + ntap;
    if (ntap > 4 || cut < 1 || 2 * cut * NP > T || nkeep < int64_t(R1) * ntap ||
        nblk_max > int64_t(T / (R1 * NP)) * kWalkSegSplit ||
        nblk_max * NP * 4 > 2 * T || nblk_max > SplitArea<NSL, R1, NP>::kBlocks) {
      set_error("fused iPFB/re-PFB split form: needs ntap <= 4, 1 <= cut <= %d and at most %d blocks per tile at osamp %d",
                T / (2 * NP),
                int(T / (R1 * NP)) * kWalkSegSplit, R1);
      return ALB_ENOTSUP;
    }
    return ALB_OK;
  }
  // walk segments (T / (R1 NP) per row, at most 16 blocks each) and DFT lanes
  if (nint_max > int64_t(T / (R1 * NP)) * kWalkSeg || nint_max * NP * 4 > 2 * T ||
      nint_max * NP * 4 * SP > NP * Padded<NSL>::value) {
    set_error("fused iPFB/re-PFB: tile too long for osamp %d", R1);
    return ALB_ENOTSUP;
  }
  return ALB_OK;
}

// SPLIT: the split form (P, tile0; ts unused).
template <int NSL, int CW, int T, int R1, bool SPEC = false, bool SPLIT = false>
int launch_cols_repfb(const float* y, int64_t ldy, int64_t tile_stride, int64_t ntiles, const float2* filt, int64_t ldf,
                      int cut, float* ts, int64_t row0, int64_t rows_per_tile, const float* win, int ntap,
                      int64_t nblocks, float2* A, hipStream_t s, float2* P = nullptr, int64_t tile0 = 0) {
  constexpr int NP = CW / 2;
  constexpr int UA = SPLIT ? SplitArea<NSL, R1, NP>::value : NP * Padded<NSL>::value;
  int rc = check_cols_repfb<NSL, CW, T, R1, SPLIT>(cut, ntap);
  if (rc || ntiles == 0) return rc;
  ALB_REQUIRE(y && filt && win && A && (SPLIT ? P != nullptr : ts != nullptr), "alb_ipfb_columns_repfb: null pointer");
  ALB_REQUIRE((4096 / CW) * ntiles < (int64_t(1) << 31), "fused iPFB/re-PFB: too many tiles");
  ALB_REQUIRE(!SPLIT || (tile0 >= 0 && rows_per_tile == NSL - 2 * int64_t(cut)),
              "fused iPFB/re-PFB split form: consecutive tiles must be nslice - 2 cut rows apart and tile0 >= 0");
  Twiddle tw;
  rc = get_twiddle(NSL, &tw, s);
  if (rc) return rc;
  const size_t lds = sizeof(float2) * (size_t(UA) + tw.nA + tw.nB + kW16Size + R1) +
                     sizeof(float) * size_t(ntap) * R1 * CW;
  ALB_REQUIRE(lds <= 160 * 1024, "fused iPFB/re-PFB: LDS budget exceeded");
  hipLaunchKernelGGL((ipfb_cols_repfb_kernel<NSL, CW, T, R1, SPEC, SPLIT>), dim3(unsigned((4096 / CW) * ntiles)), dim3(T),
                     lds, s, y, ldy, tile_stride, filt, ldf, cut, ts, row0, rows_per_tile, win, ntap, nblocks, A,
                     cols_group(ntiles), tw, P, tile0);
  ALB_LAUNCH_CHECK("ipfb_cols_repfb_kernel");
  return ALB_OK;
}

// Item-per-workgroup launches of the spectral-order iPFB: a multiple of 8
// workgroups (the kernels map them to XCD-contiguous items).
inline unsigned xcd_grid(int64_t nitem) { return unsigned((nitem + 7) / 8 * 8); }

template <int NSL, int CA, int T>
int launch_chan_fft(const Packed4Rows& pk, int64_t ntiles, float2* E, hipStream_t s) {
  Twiddle tw;
  int rc = get_twiddle(NSL, &tw, s);
  if (rc) return rc;
  const size_t lds = sizeof(float2) * (size_t(CA) * Padded<NSL>::value + tw.nA + tw.nB + kW16Size);
  const int64_t nitem = (pk.ncols + CA - 1) / CA * ntiles;
  ALB_REQUIRE(nitem < (int64_t(1) << 31) - 8, "alb_ipfb_chan_fft_packed4: too many tiles");
  hipLaunchKernelGGL((ipfb_chan_fft_kernel<NSL, CA, T>), dim3(xcd_grid(nitem)), dim3(T), lds, s, pk, ntiles, E, tw);
  ALB_LAUNCH_CHECK("ipfb_chan_fft_kernel");
  return ALB_OK;
}

// Persistent launch: as many workgroups as fit on the device at once (a
// multiple of 8, so every XCD gets the same count), at most one per item.
template <int FB, int T, int KI>
int launch_rows_spectral(const float2* E, const int* chan, int64_t ncols, int64_t nslice, int64_t ntiles, float2* D,
                         hipStream_t s) {
  Twiddle tw;
  int rc = get_twiddle(4096, &tw, s);
  if (rc) return rc;
  const size_t lds = sizeof(float2) * (size_t(FB) * Padded<4096>::value + tw.nA + tw.nB + kW16Size);
  const int64_t nitem = (nslice / 2 + 1 + FB - 1) / FB * ntiles;
  ALB_REQUIRE(nitem < (int64_t(1) << 31) - 8, "alb_ipfb_rows_spectral: too many tiles");
  static std::atomic<int> resident[64];
  int dev = 0;
  ALB_HIP(hipGetDevice(&dev));
  int nres = dev < 64 ? resident[dev].load() : 0;
  if (nres == 0) {
    int ncu = 0, occ = 0;
    ALB_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    ALB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, ipfb_rows_spectral_kernel<FB, T, KI>, T, lds));
    nres = ncu * (occ > 0 ? occ : 1);
    if (dev < 64) resident[dev].store(nres);
  }
  const int64_t per = (nitem + 7) / 8;
  int64_t S = nres / 8 > 0 ? nres / 8 : 1;
  if (S > per) S = per;
  hipLaunchKernelGGL((ipfb_rows_spectral_kernel<FB, T, KI>), dim3(unsigned(8 * S)), dim3(T), lds, s, E, chan,
                     int(ncols), nslice, ntiles, D, tw);
  ALB_LAUNCH_CHECK("ipfb_rows_spectral_kernel");
  return ALB_OK;
}

template <int R1>
int launch_repfb_stage2(const float2* A, int64_t a_stride, int64_t ninput, int64_t nblock, int64_t c_lo, int64_t c_hi,
                        float2* out, int64_t sm, int64_t sb, int64_t sc, int64_t bpj, int64_t sj, hipStream_t s,
                        SplitRead sp = SplitRead{nullptr, 0, 0, 1, 0, 1}) {
  Twiddle tw4;
  int rc = get_twiddle(4096, &tw4, s);
  if (rc) return rc;
  const size_t lds2 = sizeof(float2) * (size_t(Padded<4096>::value) + tw4.nA + tw4.nB + kW16Size);
  const int64_t total = int64_t(R1 / 2 + 1) * nblock * ninput;
  ALB_REQUIRE(total < (int64_t(1) << 31) - 8, "alb_repfb: too many blocks");
  if (sp.P != nullptr)
    hipLaunchKernelGGL((repfb_stage2_kernel<R1, true>), dim3(unsigned((total + 7) / 8 * 8)), dim3(kTPB), lds2, s, A,
                       a_stride, int(ninput), nblock, c_lo, c_hi, out, sm, sb, sc, bpj, sj, tw4, sp);
  else
    hipLaunchKernelGGL((repfb_stage2_kernel<R1, false>), dim3(unsigned((total + 7) / 8 * 8)), dim3(kTPB), lds2, s, A,
                       a_stride, int(ninput), nblock, c_lo, c_hi, out, sm, sb, sc, bpj, sj, tw4, sp);
  ALB_LAUNCH_CHECK("repfb_stage2_kernel");
  return ALB_OK;
}

template <int N, int F>
int launch_fft(const float2* x, int64_t ldx, int64_t nrows, float2* out, int64_t ldo, int inverse, hipStream_t s) {
  Twiddle tw;
  int rc = get_twiddle(N, &tw, s);
  if (rc) return rc;
  const size_t lds = sizeof(float2) * (size_t(F) * Padded<N>::value + tw.nA + tw.nB);
  hipLaunchKernelGGL((fft_rows_kernel<N, F>), dim3(unsigned((nrows + F - 1) / F)), dim3(kTPB), lds, s, x, ldx, nrows,
                     out, ldo, inverse, tw);
  ALB_LAUNCH_CHECK("fft_rows_kernel");
  return ALB_OK;
}

int fft_dispatch(int64_t n, const float2* x, int64_t ldx, int64_t nrows, float2* out, int64_t ldo, int inverse,
                 hipStream_t s) {
  switch (n) {
    case 16: return launch_fft<16, 64>(x, ldx, nrows, out, ldo, inverse, s);
    case 32: return launch_fft<32, 64>(x, ldx, nrows, out, ldo, inverse, s);
    case 64: return launch_fft<64, 64>(x, ldx, nrows, out, ldo, inverse, s);
    case 128: return launch_fft<128, 32>(x, ldx, nrows, out, ldo, inverse, s);
    case 256: return launch_fft<256, 16>(x, ldx, nrows, out, ldo, inverse, s);
    case 512: return launch_fft<512, 8>(x, ldx, nrows, out, ldo, inverse, s);
    case 1024: return launch_fft<1024, 4>(x, ldx, nrows, out, ldo, inverse, s);
    case 2048: return launch_fft<2048, 2>(x, ldx, nrows, out, ldo, inverse, s);
    case 4096: return launch_fft<4096, 1>(x, ldx, nrows, out, ldo, inverse, s);
    case 8192: return launch_fft<8192, 1>(x, ldx, nrows, out, ldo, inverse, s);
    default:
      set_error("fft: length %lld not supported (power of two 16..8192)", (long long)n);
      return ALB_ENOTSUP;
  }
}

template <int P>
int launch_cols_bluestein(const float* y, int64_t ldy, int64_t tile_stride, int64_t ntiles, int n, const float2* filt,
                          int64_t ldf, int64_t ncols, int cut, float* out, int64_t ldo, int64_t out_tile_stride,
                          const float2* chirp, const float2* bhat, hipStream_t s) {
  Twiddle tw;
  int rc = get_twiddle(P, &tw, s);
  if (rc) return rc;
  const size_t lds = sizeof(float2) * (size_t(Padded<P>::value) + tw.nA + tw.nB);
  hipLaunchKernelGGL((ipfb_cols_bluestein_kernel<P>), dim3(unsigned(ncols / 2), unsigned(ntiles)), dim3(kTPB), lds, s, y,
                     ldy, tile_stride, filt, ldf, n, cut, out, ldo, out_tile_stride, chirp, bhat, tw);
  ALB_LAUNCH_CHECK("ipfb_cols_bluestein_kernel");
  return ALB_OK;
}

}  // namespace

extern "C" {

// pfb_gpu_utils.py:123-137 get_matft + :29-36 calculate_filter, in float64.
// win: float32 [ntap*lblock]; outputs [lblock][nslice/2+1] complex64 (either may be null).
int alb_matft_filter(const float* win, int ntap, int64_t lblock, int64_t nslice, double thresh, void* matft, void* filt,
                     void* stream) {
  ALB_REQUIRE(win && (matft || filt), "alb_matft_filter: null pointer");
  ALB_REQUIRE(ntap > 0 && lblock > 0 && nslice > 0, "alb_matft_filter: bad sizes");
  const int64_t nf = nslice / 2 + 1;
  const int64_t n = lblock * nf;
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(matft_filter_kernel, dim3(unsigned(g)), dim3(256), 0, as_stream(stream), win, ntap, lblock, nslice,
                     nf, thresh, (float2*)matft, (float2*)filt);
  ALB_LAUNCH_CHECK("matft_filter_kernel");
  return ALB_OK;
}

// pycufft.py:139-161 irfft along rows: in complex [nrows][ldi] (n/2+1 used), out float [nrows][ldo], 1/n scaling.
int alb_irfft_rows(const void* in, int64_t ldi, float* out, int64_t ldo, int64_t nrows, int64_t n, void* stream) {
  ALB_REQUIRE(in && out, "alb_irfft_rows: null pointer");
  ALB_REQUIRE(ldi >= n / 2 + 1 && ldo >= n && nrows >= 0, "alb_irfft_rows: bad strides");
  if (nrows == 0) return ALB_OK;
  DenseRows d{(const float2*)in, ldi};
  return irfft_dispatch(n, d, nrows, out, ldo, as_stream(stream));
}

// pycufft.py:120-137 rfft along rows (unnormalised), bins [c0, c1) written.
int alb_rfft_rows(const float* in, int64_t ldi, void* out, int64_t ldo, int64_t nrows, int64_t n, int64_t c0,
                  int64_t c1, void* stream) {
  ALB_REQUIRE(in && out, "alb_rfft_rows: null pointer");
  ALB_REQUIRE(ldi >= n && c0 >= 0 && c1 <= n / 2 + 1 && c1 > c0 && ldo >= c1 - c0, "alb_rfft_rows: bad sizes");
  if (nrows == 0) return ALB_OK;
  return rfft_dispatch(n, in, ldi, nrows, nullptr, 1, c0, c1, (float2*)out, ldo, as_stream(stream));
}

// pfb_gpu_utils.py:105-121 cupy_pfb for lblock <= 16384: polyphase + rfft, channels [c0, c1).
int alb_pfb_small(const float* ts, int64_t nblock, int64_t lblock, const float* win, int ntap, int64_t c0, int64_t c1,
                  void* out, int64_t ldo, void* stream) {
  ALB_REQUIRE(ts && win && out, "alb_pfb_small: null pointer");
  ALB_REQUIRE(nblock >= 0 && ntap >= 1 && c0 >= 0 && c1 <= lblock / 2 + 1 && c1 > c0, "alb_pfb_small: bad sizes");
  if (nblock == 0) return ALB_OK;
  return rfft_dispatch(lblock, ts, lblock, nblock, win, ntap, c0, c1, (float2*)out, ldo, as_stream(stream));
}

// Column pass of the inverse PFB over ntiles tiles of [nslice][ncols] float rows
// (row stride ldy, tile stride tile_stride); filt [ncols][ldf] complex64; kept rows
// [cut, nslice-cut) of tile t go to out + t*out_tile_stride (row stride ldo).
// In place (out == y, cut == 0) is allowed.
int alb_ipfb_columns(const float* y, int64_t ldy, int64_t tile_stride, int64_t ntiles, int64_t nslice, int64_t ncols,
                     const void* filt, int64_t ldf, int cut, float* out, int64_t ldo, int64_t out_tile_stride,
                     void* stream) {
  ALB_REQUIRE(y && filt && out, "alb_ipfb_columns: null pointer");
  ALB_REQUIRE((ldy >= ncols || (ldy == 8 && ncols % 8 == 0)) && ldo >= ncols && ldf >= nslice / 2 + 1 && cut >= 0 &&
                  2 * cut < nslice && ntiles >= 0,
              "alb_ipfb_columns: bad sizes");
  if (ntiles == 0) return ALB_OK;
  return cols_dispatch(nslice, y, ldy, tile_stride, ntiles, (const float2*)filt, ldf, ncols, cut, out, ldo,
                       out_tile_stride, as_stream(stream));
}

// Row pass of the fused 4-bit iPFB: unpack + spec_num zero fill + channel scatter
// + irfft (n = 4096) for ntiles tiles of nslice rows, tiles tile0.. of the chunk
// sequence.  bytes [rows][ncols] packed per-pol 4-bit spectra, rowmap
// [nchunk][acclen] = row of bytes or -1 (missing), chan[ncols] channel numbers.
int alb_ipfb_rows_packed4(const uint8_t* bytes, const int* rowmap, const int* chan, int64_t ncols, int64_t acclen,
                          int cut, int64_t tile0, int64_t ntiles, float* out, void* stream) {
  ALB_REQUIRE(bytes && rowmap && chan && out, "alb_ipfb_rows_packed4: null pointer");
  const int64_t nslice = acclen + 2 * cut;
  ALB_REQUIRE(acclen >= 2 * cut && ncols > 0 && ncols <= 2049, "alb_ipfb_rows_packed4: bad sizes");
  Packed4Rows pk{bytes, rowmap, chan, ncols, acclen, 2 * int64_t(cut), tile0};
  if (ntiles <= 0) return ALB_OK;
  return irfft_packed4_dispatch(pk, nslice * ntiles, nslice, out, 4096, as_stream(stream));
}

// The same row pass written in the blocked tile layout out[ntiles][4096/8][nslice][8]
// (the column passes read it with ldy = 8: one contiguous strip per workgroup).
int alb_ipfb_rows_packed4_blocked(const uint8_t* bytes, const int* rowmap, const int* chan, int64_t ncols,
                                  int64_t acclen, int cut, int64_t tile0, int64_t ntiles, float* out, void* stream) {
  ALB_REQUIRE(bytes && rowmap && chan && out, "alb_ipfb_rows_packed4_blocked: null pointer");
  const int64_t nslice = acclen + 2 * cut;
  ALB_REQUIRE(acclen >= 2 * cut && ncols > 0 && ncols <= 2049, "alb_ipfb_rows_packed4_blocked: bad sizes");
  Packed4Rows pk{bytes, rowmap, chan, ncols, acclen, 2 * int64_t(cut), tile0};
  if (ntiles <= 0) return ALB_OK;
  return irfft_packed4_dispatch(pk, nslice * ntiles, nslice, out, 8, as_stream(stream));
}

// Re-PFB (pfb_gpu_utils.py:105-121 with lblock = 4096*osamp) of nblock output
// spectra from the row-major stream ts [(nblock+ntap-1)*osamp][4096]; fine
// channels [c_lo, c_hi) of spectrum b to out[(b/bpj)*sj + (b%bpj)*sb + (c-c_lo)*sc];
// A is workspace of nblock*(osamp/2+1)*4096 complex64.
int alb_repfb_stage1(const float* ts, const float* win, int ntap, int64_t nblock, int64_t osamp, void* A,
                     void* stream) {
  ALB_REQUIRE(ts && win && A, "alb_repfb_stage1: null pointer");
  ALB_REQUIRE(nblock >= 0 && ntap >= 1, "alb_repfb_stage1: bad sizes");
  if (nblock == 0) return ALB_OK;
  ALB_REQUIRE(nblock <= 65535 * 16, "alb_repfb_stage1: too many blocks per call");
  hipStream_t s = as_stream(stream);
  switch (osamp) {
#define CASE(r) case r: return launch_repfb_stage1<r>(ts, win, ntap, nblock, (float2*)A, s);
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32) CASE(64)
#undef CASE
    default:
      if (is_pow2(osamp) && osamp <= 65536) return repfb_large_stage1(ts, win, ntap, nblock, osamp, (float2*)A, s);
      set_error("alb_repfb: osamp %lld not supported (power of two 1..65536)", (long long)osamp);
      return ALB_ENOTSUP;
  }
}

int alb_repfb_stage2_inputs(const void* A, int64_t a_stride, int64_t ninput, int64_t nblock, int64_t osamp,
                            int64_t c_lo, int64_t c_hi, void* out, int64_t sm, int64_t sb, int64_t sc, int64_t bpj,
                            int64_t sj, void* stream);
int64_t alb_repfb_workspace_bytes(int64_t nblock, int64_t osamp);

int alb_repfb_stage2(const void* A, int64_t nblock, int64_t osamp, int64_t c_lo, int64_t c_hi, void* out, int64_t sb,
                     int64_t sc, int64_t bpj, int64_t sj, void* stream) {
  return alb_repfb_stage2_inputs(A, 0, 1, nblock, osamp, c_lo, c_hi, out, 0, sb, sc, bpj, sj, stream);
}

// Stage 2 of ninput inputs at once: input m's workspace at A + m * a_stride
// bytes, its channels at out + m * sm (complex elements) + the alb_repfb_stage2
// addressing.
int alb_repfb_stage2_inputs(const void* A, int64_t a_stride, int64_t ninput, int64_t nblock, int64_t osamp,
                            int64_t c_lo, int64_t c_hi, void* out, int64_t sm, int64_t sb, int64_t sc, int64_t bpj,
                            int64_t sj, void* stream) {
  ALB_REQUIRE(A && out, "alb_repfb_stage2: null pointer");
  ALB_REQUIRE(nblock >= 0 && ninput >= 1 && ninput <= 1024 && c_lo >= 0 && c_hi > c_lo &&
                  c_hi <= 2048 * osamp + 1 && bpj >= 1,
              "alb_repfb_stage2: bad sizes");
  ALB_REQUIRE(a_stride % int64_t(sizeof(float2)) == 0 && (ninput == 1 || a_stride >= alb_repfb_workspace_bytes(nblock, osamp)),
              "alb_repfb_stage2: workspace stride must be a multiple of 8 bytes and hold nblock blocks");
  if (nblock == 0) return ALB_OK;
  hipStream_t s = as_stream(stream);
  const int64_t as = a_stride / int64_t(sizeof(float2));
  switch (osamp) {
#define CASE(r) case r: return launch_repfb_stage2<r>((const float2*)A, as, ninput, nblock, c_lo, c_hi, (float2*)out, sm, sb, sc, bpj, sj, s);
    CASE(1) CASE(2) CASE(4) CASE(8) CASE(16) CASE(32) CASE(64)
#undef CASE
    default:
      if (is_pow2(osamp) && osamp <= 65536)
        return repfb_large_stage2((const float2*)A, as, ninput, nblock, osamp, c_lo, c_hi, (float2*)out, sm, sb, sc,
                                  bpj, sj, s);
      set_error("alb_repfb: osamp %lld not supported (power of two 1..65536)", (long long)osamp);
      return ALB_ENOTSUP;
  }
}

// Re-PFB (pfb_gpu_utils.py:105-121 with lblock = 4096*osamp) of nblock output
// spectra from the row-major stream ts [(nblock+ntap-1)*osamp][4096]; fine
// channels [c_lo, c_hi) of spectrum b to out[(b/bpj)*sj + (b%bpj)*sb + (c-c_lo)*sc];
// A is workspace of nblock*(osamp/2+1)*4096 complex64.
int alb_repfb(const float* ts, const float* win, int ntap, int64_t nblock, int64_t osamp, void* A, int64_t c_lo,
              int64_t c_hi, void* out, int64_t sb, int64_t sc, int64_t bpj, int64_t sj, void* stream) {
  ALB_REQUIRE(ts && win && A && out, "alb_repfb: null pointer");
  int rc = alb_repfb_stage1(ts, win, ntap, nblock, osamp, A, stream);
  if (rc) return rc;
  return alb_repfb_stage2(A, nblock, osamp, c_lo, c_hi, out, sb, sc, bpj, sj, stream);
}

// Stage 1 of the re-PFB for a list of blocks (host array of n block indices):
// block b reads ts rows [osamp*b, osamp*(b+ntap)) and writes A[b].
int alb_repfb_stage1_list(const float* ts, const float* win, int ntap, const int* blocks, int64_t n, int64_t osamp,
                          void* A, void* stream) {
  ALB_REQUIRE(n >= 0, "alb_repfb_stage1_list: bad sizes");
  if (n == 0) return ALB_OK;
  ALB_REQUIRE(ts && win && A && blocks, "alb_repfb_stage1_list: null pointer");
  ALB_REQUIRE(ntap >= 1, "alb_repfb_stage1_list: bad ntap");
  for (int64_t i = 0; i < n; ++i)
    ALB_REQUIRE(blocks[i] >= 0 && blocks[i] < (1 << 27), "alb_repfb_stage1_list: block index out of range");
  hipStream_t s = as_stream(stream);
  switch (osamp) {
#define CASE(r) case r: return launch_repfb_stage1_list<r>(ts, win, ntap, blocks, n, (float2*)A, s);
    CASE(16) CASE(32) CASE(64)
#undef CASE
    default:
      set_error("alb_repfb_stage1_list: osamp %lld not supported (16, 32, 64)", (long long)osamp);
      return ALB_ENOTSUP;
  }
}

// Interior re-PFB blocks of a tile whose kept rows are stream rows
// [g0, g0 + nkeep): [*lo, *hi) (empty when *hi == *lo).
int alb_repfb_interior(int64_t g0, int64_t nkeep, int64_t osamp, int ntap, int64_t nblocks, int64_t* lo, int64_t* hi) {
  ALB_REQUIRE(lo && hi && nkeep >= 0 && osamp >= 1 && ntap >= 1 && nblocks >= 0, "alb_repfb_interior: bad arguments");
  repfb_interior(g0, nkeep, osamp, ntap, nblocks, lo, hi);
  return ALB_OK;
}

// Inverse-PFB column pass fused with re-PFB stage 1 (see ipfb_cols_repfb_kernel).
int alb_ipfb_columns_repfb(const float* y, int64_t ldy, int64_t tile_stride, int64_t ntiles, int64_t nslice,
                           const void* filt, int64_t ldf, int cut, float* ts, int64_t row0, int64_t rows_per_tile,
                           const float* win, int ntap, int64_t osamp, int64_t nblocks, void* A, void* stream) {
  ALB_REQUIRE(ntiles >= 0 && nblocks >= 0 && rows_per_tile >= 0 && row0 + ntiles * rows_per_tile >= 0,
              "alb_ipfb_columns_repfb: bad sizes");
  ALB_REQUIRE(cut >= 0 && 2 * cut < nslice, "alb_ipfb_columns_repfb: bad cut");
  ALB_REQUIRE(ntiles == 0 || ((ldy >= 4096 || ldy == 8) && ldf >= nslice / 2 + 1),
              "alb_ipfb_columns_repfb: bad strides (ldy >= 4096 row-major, or 8 for the blocked tile layout)");
  hipStream_t s = as_stream(stream);
  if (nslice == 4096 && (osamp == 32 || osamp == 64)) {
    // strip width: the knob, else 4 columns (two 512-thread workgroups per CU)
    const int64_t kw = tune_get("ipfb_cols_width");
    const int64_t w = kw > 0 ? kw : kColsWidth;
#define CASE(W, T, r) if (w == W && osamp == r) return launch_cols_repfb<4096, W, T, r>(y, ldy, tile_stride, ntiles, (const float2*)filt, ldf, cut, ts, row0, rows_per_tile, win, ntap, nblocks, (float2*)A, s);
    CASE(2, 256, 32) CASE(2, 256, 64) CASE(4, 512, 32) CASE(4, 512, 64)
#undef CASE
    if (osamp == 32)
      return launch_cols_repfb<4096, 8, 1024, 32>(y, ldy, tile_stride, ntiles, (const float2*)filt, ldf, cut, ts, row0,
                                                  rows_per_tile, win, ntap, nblocks, (float2*)A, s);
    return launch_cols_repfb<4096, 8, 1024, 64>(y, ldy, tile_stride, ntiles, (const float2*)filt, ldf, cut, ts, row0,
                                                rows_per_tile, win, ntap, nblocks, (float2*)A, s);
  }
  set_error("alb_ipfb_columns_repfb: nslice %lld with osamp %lld not supported (nslice 4096, osamp 32 or 64)",
            (long long)nslice, (long long)osamp);
  return ALB_ENOTSUP;
}

// Spectral-order inverse PFB (cupy_ipfb, pfb_gpu_utils.py:56-91, with its
// transform along time taken first), step 1: E[t][f][c] (complex64,
// [ntiles][nslice][ncols]) = FFT along the nslice rows of tile t of the
// decoded 4-bit column c (rows as in alb_ipfb_rows_packed4; channels 0 and
// 2048 keep their real part only; out-of-range channels give zeros).
int alb_ipfb_chan_fft_packed4(const uint8_t* bytes, const int* rowmap, const int* chan, int64_t ncols, int64_t acclen,
                              int cut, int64_t tile0, int64_t ntiles, void* E, void* stream) {
  ALB_REQUIRE(bytes && rowmap && chan && E, "alb_ipfb_chan_fft_packed4: null pointer");
  const int64_t nslice = acclen + 2 * int64_t(cut);
  ALB_REQUIRE(acclen >= 2 * cut && cut >= 0 && ncols > 0 && ncols <= 2049 && ntiles >= 0 && tile0 >= 0,
              "alb_ipfb_chan_fft_packed4: bad sizes");
  ALB_REQUIRE(nslice == 4096, "alb_ipfb_chan_fft_packed4: nslice %lld not supported (4096)", (long long)nslice);
  if (ntiles == 0) return ALB_OK;
  Packed4Rows pk{bytes, rowmap, chan, ncols, acclen, 2 * int64_t(cut), tile0};
  return launch_chan_fft<4096, 2, 512>(pk, ntiles, (float2*)E, as_stream(stream));
}

// Step 2: D[t][s][f][c] (complex64, [ntiles][1024][nslice/2+1][4]) =
// 4096 * ddft[4 s + c][f], ddft = rfft(irfft(dat, axis=1), axis=0) of tile t,
// from step 1's E.  chan: channel number of each column (repeats marked -1).
int alb_ipfb_rows_spectral(const void* E, const int* chan, int64_t ncols, int64_t nslice, int64_t ntiles, void* D,
                           void* stream) {
  ALB_REQUIRE(E && chan && D, "alb_ipfb_rows_spectral: null pointer");
  ALB_REQUIRE(ncols > 0 && ncols <= 2049 && ntiles >= 0, "alb_ipfb_rows_spectral: bad sizes");
  ALB_REQUIRE(nslice == 4096, "alb_ipfb_rows_spectral: nslice %lld not supported (4096)", (long long)nslice);
  if (ntiles == 0) return ALB_OK;
  // items per thread: every (row, column) of a group of two rows
  hipStream_t s = as_stream(stream);
  if (2 * ncols <= 512) return launch_rows_spectral<2, 512, 1>((const float2*)E, chan, ncols, nslice, ntiles, (float2*)D, s);
  if (2 * ncols <= 1024) return launch_rows_spectral<2, 512, 2>((const float2*)E, chan, ncols, nslice, ntiles, (float2*)D, s);
  if (2 * ncols <= 2048) return launch_rows_spectral<2, 512, 4>((const float2*)E, chan, ncols, nslice, ntiles, (float2*)D, s);
  return launch_rows_spectral<2, 512, 9>((const float2*)E, chan, ncols, nslice, ntiles, (float2*)D, s);
}

// Step 3: alb_ipfb_columns_repfb on step 2's D (Wiener filter, inverse
// transform along time, re-PFB stage 1 of the interior blocks, edge rows).
int alb_ipfb_columns_repfb_spectral(const void* D, int64_t ntiles, int64_t nslice, const void* filt, int64_t ldf,
                                    int cut, float* ts, int64_t row0, int64_t rows_per_tile, const float* win,
                                    int ntap, int64_t osamp, int64_t nblocks, void* A, void* stream) {
  ALB_REQUIRE(ntiles >= 0 && nblocks >= 0 && rows_per_tile >= 0 && row0 + ntiles * rows_per_tile >= 0,
              "alb_ipfb_columns_repfb_spectral: bad sizes");
  ALB_REQUIRE(cut >= 0 && 2 * cut < nslice, "alb_ipfb_columns_repfb_spectral: bad cut");
  ALB_REQUIRE(ntiles == 0 || ldf >= nslice / 2 + 1, "alb_ipfb_columns_repfb_spectral: bad filter stride");
  hipStream_t s = as_stream(stream);
  if (nslice == 4096 && osamp == 64)
    return launch_cols_repfb<4096, 4, 512, 64, true>((const float*)D, 0, 0, ntiles, (const float2*)filt, ldf, cut, ts,
                                                     row0, rows_per_tile, win, ntap, nblocks, (float2*)A, s);
  if (nslice == 4096 && osamp == 32)
    return launch_cols_repfb<4096, 4, 512, 32, true>((const float*)D, 0, 0, ntiles, (const float2*)filt, ldf, cut, ts,
                                                     row0, rows_per_tile, win, ntap, nblocks, (float2*)A, s);
  set_error("alb_ipfb_columns_repfb_spectral: nslice %lld with osamp %lld not supported (nslice 4096, osamp 32 or 64)",
            (long long)nslice, (long long)osamp);
  return ALB_ENOTSUP;
}

// Split form of alb_ipfb_columns_repfb (nslice 4096, osamp 64, ntap <= 4,
// strip width 4): stage 1 of every block touching a tile's kept rows, over
// those rows; a block's first tile writes A, the next tile writes its half to
// P slot ntap * (tile0 + t - 1) + (b - first block straddling tile t's first
// row), where t is the tile within this launch and tile0 this launch's first
// tile's index in the stream window (tiles are rows_per_tile = nslice - 2 cut
// rows apart).  P holds p_slots slots of alb_repfb_workspace_bytes(1, osamp)
// bytes.  No stream rows are written.
// The window's tile 0 must start at or before window row 0 (row0 - tile0
// rows_per_tile <= 0): its column pass then has no block starting before its
// first kept row, which would have no P slot.
static int check_split_slots(int64_t ntiles, int64_t tile0, int ntap, int64_t p_slots, int64_t row0,
                             int64_t rows_per_tile) {
  ALB_REQUIRE(ntiles == 0 || (tile0 >= 0 && ntap >= 1 && int64_t(ntap) * (tile0 + ntiles - 1) <= p_slots),
              "split column pass: P holds %lld slots, tiles %lld..%lld need %lld", (long long)p_slots,
              (long long)tile0, (long long)(tile0 + ntiles - 1), (long long)(int64_t(ntap) * (tile0 + ntiles - 1)));
  ALB_REQUIRE(ntiles == 0 || row0 - tile0 * rows_per_tile <= 0,
              "split column pass: the window's tile 0 starts at row %lld (row0 - tile0 rows_per_tile), after row 0",
              (long long)(row0 - tile0 * rows_per_tile));
  return ALB_OK;
}

int alb_ipfb_columns_repfb_split(const float* y, int64_t ldy, int64_t tile_stride, int64_t ntiles, int64_t nslice,
                                 const void* filt, int64_t ldf, int cut, int64_t row0, int64_t rows_per_tile,
                                 const float* win, int ntap, int64_t osamp, int64_t nblocks, void* A, void* P,
                                 int64_t tile0, int64_t p_slots, void* stream) {
  ALB_REQUIRE(ntiles >= 0 && nblocks >= 0 && rows_per_tile >= 0 && row0 + ntiles * rows_per_tile >= 0,
              "alb_ipfb_columns_repfb_split: bad sizes");
  ALB_REQUIRE(cut >= 0 && 2 * cut < nslice, "alb_ipfb_columns_repfb_split: bad cut");
  ALB_REQUIRE(ntiles == 0 || ((ldy >= 4096 || ldy == 8) && ldf >= nslice / 2 + 1),
              "alb_ipfb_columns_repfb_split: bad strides (ldy >= 4096 row-major, or 8 for the blocked tile layout)");
  if (nslice != 4096 || osamp != 64) {
    set_error("alb_ipfb_columns_repfb_split: nslice %lld with osamp %lld not supported (nslice 4096, osamp 64)",
              (long long)nslice, (long long)osamp);
    return ALB_ENOTSUP;
  }
  int rc = check_split_slots(ntiles, tile0, ntap, p_slots, row0, rows_per_tile);
  if (rc) return rc;
  return launch_cols_repfb<4096, 4, 512, 64, false, true>(y, ldy, tile_stride, ntiles, (const float2*)filt, ldf, cut,
                                                          nullptr, row0, rows_per_tile, win, ntap, nblocks, (float2*)A,
                                                          as_stream(stream), (float2*)P, tile0);
}

// The same on the spectral-order strips of alb_ipfb_rows_spectral.
int alb_ipfb_columns_repfb_spectral_split(const void* D, int64_t ntiles, int64_t nslice, const void* filt, int64_t ldf,
                                          int cut, int64_t row0, int64_t rows_per_tile, const float* win, int ntap,
                                          int64_t osamp, int64_t nblocks, void* A, void* P, int64_t tile0,
                                          int64_t p_slots, void* stream) {
  ALB_REQUIRE(ntiles >= 0 && nblocks >= 0 && rows_per_tile >= 0 && row0 + ntiles * rows_per_tile >= 0,
              "alb_ipfb_columns_repfb_spectral_split: bad sizes");
  ALB_REQUIRE(cut >= 0 && 2 * cut < nslice, "alb_ipfb_columns_repfb_spectral_split: bad cut");
  ALB_REQUIRE(ntiles == 0 || ldf >= nslice / 2 + 1, "alb_ipfb_columns_repfb_spectral_split: bad filter stride");
  if (nslice != 4096 || osamp != 64) {
    set_error("alb_ipfb_columns_repfb_spectral_split: nslice %lld with osamp %lld not supported (nslice 4096, osamp 64)",
              (long long)nslice, (long long)osamp);
    return ALB_ENOTSUP;
  }
  int rc = check_split_slots(ntiles, tile0, ntap, p_slots, row0, rows_per_tile);
  if (rc) return rc;
  return launch_cols_repfb<4096, 4, 512, 64, true, true>((const float*)D, 0, 0, ntiles, (const float2*)filt, ldf, cut,
                                                         nullptr, row0, rows_per_tile, win, ntap, nblocks, (float2*)A,
                                                         as_stream(stream), (float2*)P, tile0);
}

// Stage 2 after the split column pass: alb_repfb_stage2_inputs (osamp 64) with
// each tile-boundary block's second half added from P (input m at P + m
// p_stride bytes): the stream window's tiles start at window row g_first (<= 0;
// tile c at g_first + c rows_per_tile), ntiles of them were processed.
int alb_repfb_stage2_split(const void* A, int64_t a_stride, const void* P, int64_t p_stride, int64_t ninput,
                           int64_t nblock, int64_t osamp, int64_t c_lo, int64_t c_hi, void* out, int64_t sm, int64_t sb,
                           int64_t sc, int64_t bpj, int64_t sj, int64_t g_first, int64_t rows_per_tile, int64_t ntiles,
                           int ntap, void* stream) {
  ALB_REQUIRE(A && out && (P || ntiles <= 1), "alb_repfb_stage2_split: null pointer");
  ALB_REQUIRE(nblock >= 0 && ninput >= 1 && ninput <= 1024 && c_lo >= 0 && c_hi > c_lo &&
                  c_hi <= 2048 * osamp + 1 && bpj >= 1,
              "alb_repfb_stage2_split: bad sizes");
  ALB_REQUIRE(osamp == 64, "alb_repfb_stage2_split: osamp must be 64");
  if (nblock == 0) return ALB_OK;   // e.g. a window of jobs past the tiles with data
  ALB_REQUIRE(ntap >= 1 && ntap <= 4 && g_first <= 0 && rows_per_tile >= int64_t(osamp) * ntap && ntiles >= 0,
              "alb_repfb_stage2_split: bad tile geometry");
  const int64_t blk = alb_repfb_workspace_bytes(1, osamp);
  ALB_REQUIRE(a_stride % int64_t(sizeof(float2)) == 0 && (ninput == 1 || a_stride >= alb_repfb_workspace_bytes(nblock, osamp)),
              "alb_repfb_stage2_split: workspace stride must be a multiple of 8 bytes and hold nblock blocks");
  ALB_REQUIRE(p_stride % int64_t(sizeof(float2)) == 0 && (ninput == 1 || p_stride >= blk * ntap * (ntiles > 1 ? ntiles - 1 : 0)),
              "alb_repfb_stage2_split: boundary stride must be a multiple of 8 bytes and hold ntap (ntiles - 1) blocks");
  // every block's reads lie inside the tiles: its first row in tile 0..ntiles-1
  ALB_REQUIRE(nblock == 0 || ntiles == 0 || int64_t(osamp) * (nblock - 1) < g_first + ntiles * rows_per_tile,
              "alb_repfb_stage2_split: blocks past the window's tiles");
  ALB_REQUIRE(int64_t(osamp) * (nblock + ntap) - g_first < (int64_t(1) << 31) && g_first > -rows_per_tile,
              "alb_repfb_stage2_split: window too long (rows below 2^31) or g_first <= -rows_per_tile");
  const SplitRead sp{ntiles > 1 ? (const float2*)P : nullptr, p_stride / int64_t(sizeof(float2)), g_first, rows_per_tile,
                     ntiles, ntap};
  return launch_repfb_stage2<64>((const float2*)A, a_stride / int64_t(sizeof(float2)), ninput, nblock, c_lo, c_hi,
                                 (float2*)out, sm, sb, sc, bpj, sj, as_stream(stream), sp);
}

int64_t alb_repfb_split_slot_bytes(int64_t osamp) { return alb_repfb_workspace_bytes(1, osamp); }

// The split form's index functions, as the kernels use them (host-only: for
// callers that plan their own launches, and for tests).
int alb_repfb_split_blocks(int64_t g0, int64_t nkeep, int64_t osamp, int ntap, int64_t nblocks, int64_t* lo,
                           int64_t* hi, int64_t* e0) {
  ALB_REQUIRE(lo && hi && e0, "alb_repfb_split_blocks: null pointer");
  ALB_REQUIRE(osamp >= 1 && ntap >= 1 && nkeep >= 1 && nblocks >= 0, "alb_repfb_split_blocks: bad sizes");
  split_blocks(g0, nkeep, osamp, ntap, nblocks, lo, hi, e0);
  return ALB_OK;
}

int64_t alb_repfb_split_slot(int64_t b, int64_t g_first, int64_t rows_per_tile, int64_t ntiles, int ntap,
                             int64_t osamp) {
  if (b < 0 || g_first > 0 || g_first <= -rows_per_tile || rows_per_tile < osamp * ntap || ntap < 1 ||
      osamp * (b + ntap) - g_first >= (int64_t(1) << 31))
    return -2;
  const SplitRead sp{nullptr, 0, g_first, rows_per_tile, ntiles, ntap};
  return split_slot(sp, b, osamp);
}

// Build the transform tables of alb_repfb / alb_repfb_stage2 at this osamp
// (1..64) on the current device now (stage 1 takes none; stage 2 the
// 4096-point twiddles); later calls, and hipGraph captures of them, then find
// them built.  Builds allocate nothing (a fixed device array holds them) but
// synchronise the device.
int alb_prepare_repfb(int64_t osamp) {
  ALB_REQUIRE(osamp >= 1 && osamp <= 64 && is_pow2(osamp), "alb_prepare_repfb: osamp must be a power of two 1..64");
  Twiddle tw;
  return get_twiddle(4096, &tw, nullptr);
}

int64_t alb_repfb_workspace_bytes(int64_t nblock, int64_t osamp) {
  // StageA layout: max(osamp / 2, 1) bin slots per column
  return nblock * (osamp / 2 > 0 ? osamp / 2 : 1) * 4096 * int64_t(sizeof(float2));
}

int alb_fft_rows(const void* in, int64_t ldi, void* out, int64_t ldo, int64_t nrows, int64_t n, int direction,
                 void* stream) {
  ALB_REQUIRE(in && out, "alb_fft_rows: null pointer");
  ALB_REQUIRE(direction == -1 || direction == 1, "alb_fft_rows: direction must be -1 or +1");
  ALB_REQUIRE(ldi >= n && ldo >= n && nrows >= 0, "alb_fft_rows: bad strides");
  if (nrows == 0) return ALB_OK;
  return fft_dispatch(n, (const float2*)in, ldi, nrows, (float2*)out, ldo, direction == 1, as_stream(stream));
}

int alb_filter_from_matft(const void* matft, int64_t n, double thresh, void* filt, void* stream) {
  ALB_REQUIRE(matft && filt, "alb_filter_from_matft: null pointer");
  if (n <= 0) return ALB_OK;
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(filter_from_matft_kernel, dim3(unsigned(g)), dim3(256), 0, as_stream(stream), (const float2*)matft,
                     n, thresh, (float2*)filt);
  ALB_LAUNCH_CHECK("filter_from_matft_kernel");
  return ALB_OK;
}

int alb_build_rowmap(const int64_t* specnum, int64_t npresent, int64_t start, int64_t acclen, int64_t row_base,
                     int* rowmap, void* stream) {
  ALB_REQUIRE(rowmap && (npresent == 0 || specnum), "alb_build_rowmap: null pointer");
  ALB_REQUIRE(acclen > 0 && npresent >= 0 && npresent <= acclen && row_base >= 0 && row_base + npresent <= 0x7fffffff,
              "alb_build_rowmap: bad sizes");
  hipStream_t s = as_stream(stream);
  int64_t g = (acclen + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(rowmap_fill_kernel, dim3(unsigned(g)), dim3(256), 0, s, rowmap, acclen);
  if (npresent > 0)
    hipLaunchKernelGGL(rowmap_scatter_kernel, dim3(unsigned((npresent + 255) / 256)), dim3(256), 0, s, specnum, npresent,
                       start, acclen, row_base, rowmap);
  ALB_LAUNCH_CHECK("rowmap kernels");
  return ALB_OK;
}

// Column pass of the inverse PFB for any even nslice (chirp-z form; the
// power-of-two alb_ipfb_columns is the fast path).  chirp: complex64 [nslice]
// exp(-i pi m^2/nslice); bhat: complex64 [P] = FFT_P of the conjugate chirp
// filter, P = the power of two >= 2*nslice-1 (P <= 16384).
int alb_ipfb_columns_any(const float* y, int64_t ldy, int64_t tile_stride, int64_t ntiles, int64_t nslice,
                         int64_t ncols, const void* filt, int64_t ldf, int cut, float* out, int64_t ldo,
                         int64_t out_tile_stride, const void* chirp, const void* bhat, int64_t P, void* stream) {
  ALB_REQUIRE(y && filt && out && chirp && bhat, "alb_ipfb_columns_any: null pointer");
  ALB_REQUIRE(nslice >= 2 && nslice % 2 == 0, "alb_ipfb_columns_any: nslice must be even (got %lld)",
              (long long)nslice);
  ALB_REQUIRE(P >= 2 * nslice - 1 && is_pow2(P), "alb_ipfb_columns_any: P must be a power of two >= 2*nslice-1");
  ALB_REQUIRE(ncols % 2 == 0 && ldy >= ncols && ldo >= ncols && ldf >= nslice / 2 + 1 && cut >= 0 && 2 * cut < nslice,
              "alb_ipfb_columns_any: bad sizes");
  if (ntiles == 0) return ALB_OK;
  hipStream_t s = as_stream(stream);
  const int n = int(nslice);
#define ARGS y, ldy, tile_stride, ntiles, n, (const float2*)filt, ldf, ncols, cut, out, ldo, out_tile_stride, \
             (const float2*)chirp, (const float2*)bhat, s
  switch (P) {
    case 16: return launch_cols_bluestein<16>(ARGS);
    case 32: return launch_cols_bluestein<32>(ARGS);
    case 64: return launch_cols_bluestein<64>(ARGS);
    case 128: return launch_cols_bluestein<128>(ARGS);
    case 256: return launch_cols_bluestein<256>(ARGS);
    case 512: return launch_cols_bluestein<512>(ARGS);
    case 1024: return launch_cols_bluestein<1024>(ARGS);
    case 2048: return launch_cols_bluestein<2048>(ARGS);
    case 4096: return launch_cols_bluestein<4096>(ARGS);
    case 8192: return launch_cols_bluestein<8192>(ARGS);
    case 16384: return launch_cols_bluestein<16384>(ARGS);
    default:
      set_error("alb_ipfb_columns_any: P %lld not supported (<= 16384)", (long long)P);
      return ALB_ENOTSUP;
  }
#undef ARGS
}

}  // extern "C"
