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

// ------------------------------------------------ all pairs on the int8 matrix cores
// One wavefront = one dword column w (4 bytes of every spectrum) of one block;
// lane l takes realified row l % 16 and the 16 spectra t0 + 16 (l / 16) + i of
// each 64-spectrum step, so every v_mfma_i32_16x16x64_i8 adds 64 spectra to a
// 16 x 16 Gram matrix G = Y Y^T (A and B share registers and k order).  Sums are
// exact in int32 (|G| <= 64 * blocklen, checked on the host).
typedef int intx4 __attribute__((ext_vector_type(4)));

// byte j of x0..x3 -> bytes 0..3 (v_perm_b32; selector 0x0c gives 0x00)
__device__ __forceinline__ uint32_t gather_byte(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3, int j) {
  const uint32_t s01 = uint32_t(j) | (uint32_t(4 + j) << 8) | 0x0c0c0000u;
  const uint32_t s23 = 0x00000c0cu | (uint32_t(j) << 16) | (uint32_t(4 + j) << 24);
  return __builtin_amdgcn_perm(x1, x0, s01) | __builtin_amdgcn_perm(x3, x2, s23);
}

// Work item of a workgroup: XCD-aware so the column groups of one block run
// together on one XCD (workgroup i is dispatched to XCD i % 8) and share the
// cache lines of the block's rows.
__device__ __forceinline__ void mfma_item(int64_t ngx, int64_t nblock, int64_t* blk, int64_t* gx) {
  const int64_t total = ngx * nblock;
  int64_t item = blockIdx.x;
  if (total % 8 == 0) item = (blockIdx.x % 8) * (total / 8) + blockIdx.x / 8;
  *blk = item / ngx;
  *gx = item - *blk * ngx;
}

// 4-bit: rows 0..7 = real parts of inputs 0..7, rows 8..15 = imaginary parts;
// one channel (byte) per product.
// Re_ab = G[a][b] + G[8+a][8+b],  Im_ab = G[8+a][b] - G[a][8+b]  (x_a conj x_b).
// A pair (a, b) reads only rows a, b, 8+a, 8+b, so rows of absent inputs
// (u >= NIN) may hold anything: they read input 0's bytes or the zeroed slots.
//
// Workgroup: 4 waves, 4 dword columns (16 channels) of one block.  Each
// 64-spectrum step the 16-byte row pieces of the inputs are loaded once into
// LDS (double-buffered, [spectrum][dword][input slot], 33-dword spectrum
// stride so the operand reads of 4 spectrum groups land in different banks;
// the real and imaginary rows of an input read the same word).  Wave w takes
// dword w.  LS input slots are loaded (4 for up to 4 inputs, else 8), one
// 16-byte load per (slot, spectrum); kPF4 steps of loads are in flight per
// thread.  The output staging reuses the step buffers.
constexpr int kDW4 = 4;          // dword columns per workgroup
constexpr int kTS4 = 33;         // LDS dwords per spectrum
constexpr int kPF4 = 4;          // 64-spectrum steps in flight per thread

// With nsplit > 1 the spectra of a block are shared by nsplit workgroups whose
// exact partial sums meet in the zero-filled output through int64 atomics.
template <int NIN>
__global__ __launch_bounds__(256) void allpairs4_mfma_kernel(InputPtrs in, int64_t ncol, int64_t blocklen,
                                                             int64_t nblock, int nsplit, long long* __restrict__ out) {
  constexpr int NP = NIN * (NIN + 1) / 2;
  constexpr int LS = NIN <= 4 ? 4 : 8;      // loaded input slots
  constexpr int SP = 256 / LS;              // spectra per load pass
  constexpr int LP = 64 / SP;               // 16-byte loads per thread per step
  __shared__ uint32_t buf[2][64 * kTS4];
  static_assert(4 * 4 * 256 <= 2 * 64 * kTS4, "output staging must fit the step buffers");
  const int tid = threadIdx.x;
  const int wv = tid / 64, lane = tid % 64;
  int64_t bs, gx;
  mfma_item((ncol / 4) / kDW4, nblock * nsplit, &bs, &gx);
  const int64_t blk = bs / nsplit, sp = bs - blk * nsplit;
  const int r = lane % 16, g = lane / 16;
  const int u = r % 8, part = r / 8;
  const int lu = tid % LS, lt = tid / LS;
  const uint8_t* lsrc = in.p[lu < NIN ? lu : 0] + 16 * gx;
  const int64_t nsteps = blocklen / 64 / nsplit;
  const int64_t t_begin = blk * blocklen + sp * nsteps * 64;
  if (LS < 8) {
    // slots LS..7 are read as operand rows of absent inputs: give them defined values
    for (int i = tid; i < 2 * 64 * 4; i += 256) {
      uint32_t* b = buf[i / 256] + ((i % 256) / 4) * kTS4 + (i % 4) * 8;
#pragma unroll
      for (int k = LS; k < 8; ++k) b[k] = 0;
    }
  }
  // ring slot q holds the steps st with st % kPF4 == q; unconditional loads
  // (past the range they reload its last step) keep the wait counts exact
  uint4 ld[kPF4][LP];
  auto fetch = [&](int q, int64_t st) {
    const uint8_t* p = lsrc + (t_begin + (st < nsteps ? st : nsteps - 1) * 64 + lt) * ncol;
#pragma unroll
    for (int e = 0; e < LP; ++e) ld[q][e] = *reinterpret_cast<const uint4*>(p + int64_t(SP) * e * ncol);
  };
  auto stash = [&](uint32_t* b, int q) {
#pragma unroll
    for (int e = 0; e < LP; ++e) {
      uint32_t* d = b + (lt + SP * e) * kTS4 + lu;
      d[0 * 8] = ld[q][e].x;
      d[1 * 8] = ld[q][e].y;
      d[2 * 8] = ld[q][e].z;
      d[3 * 8] = ld[q][e].w;
    }
  };
#pragma unroll
  for (int q = 0; q < kPF4; ++q) fetch(q, q);
  stash(buf[0], 0);
  fetch(0, kPF4);
  __syncthreads();
  intx4 acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = intx4{0, 0, 0, 0};
  for (int64_t s0 = 0; s0 < nsteps; s0 += kPF4) {
#pragma unroll
    for (int q = 0; q < kPF4; ++q) {
      const int64_t st = s0 + q;
      if (st < nsteps) {   // uniform
        const uint32_t* b = buf[q & 1];
        uint32_t x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = b[(16 * g + i) * kTS4 + wv * 8 + u];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          intx4 v;
#pragma unroll
          for (int d = 0; d < 4; ++d) {
            const uint32_t pd = gather_byte(x[4 * d], x[4 * d + 1], x[4 * d + 2], x[4 * d + 3], j);
            const uint32_t nib = part ? (pd & 0x0F0F0F0Fu) : ((pd >> 4) & 0x0F0F0F0Fu);
            // v > 8 -> v - 16 per byte: set the high nibble of those bytes
            uint32_t sgn = ((nib + 0x07070707u) >> 4) & 0x01010101u;
            sgn <<= 4;
            sgn |= sgn << 1;
            sgn |= sgn << 2;
            v[d] = int(nib | sgn);
          }
          acc[j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(v, v, acc[j], 0, 0, 0);
        }
      }
      // buffer (q + 1) & 1 was last read one step ago; the ring slot is free after the stash
      if (st + 1 < nsteps) stash(buf[(q + 1) & 1], (q + 1) % kPF4);
      fetch((q + 1) % kPF4, st + 1 + kPF4);
      __syncthreads();
    }
  }
  // D layout: lane l holds G[4 (l / 16) + v][l % 16]; staging [wave][j][256]
  int* S = reinterpret_cast<int*>(&buf[0][0]);
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) S[(wv * 4 + j) * 256 + (4 * g + v) * 16 + r] = acc[j][v];
  __syncthreads();
  const int64_t w = gx * kDW4 + wv;
  for (int it = lane; it < NP * 4; it += 64) {
    const int j = it % 4, p = it / 4;
    int a = 0, rem = p;
    while (rem >= NIN - a) { rem -= NIN - a; ++a; }
    const int bb = a + rem;
    const int* G = S + (wv * 4 + j) * 256;
    const long long re = (long long)G[a * 16 + bb] + G[(8 + a) * 16 + 8 + bb];
    const long long im = (long long)G[(8 + a) * 16 + bb] - G[a * 16 + 8 + bb];
    const int64_t ch = 4 * w + j;
    long long* o = out + ((blk * NP + p) * ncol + ch) * 2;
    if (nsplit > 1) {
      if (re) atomic_add_i64(o, re);
      if (im) atomic_add_i64(o + 1, im);
    } else {
      o[0] = re;
      o[1] = im;
    }
  }
}

// 4-bit, 4 SEGD bytes of every row (64 or 128 channels) per workgroup, so
// each line is fetched once and used whole; a wave's load instruction covers
// 4 SEGD contiguous bytes of each of its rows (128-byte segments reach about
// 5.7 TB/s of reads, 64-byte segments about 3.4).  Up to 4 inputs (LS = 4 loaded slots)
// a product carries two channels: row r = 8 h + 4 part + u (channel half h,
// part 0 = real, 1 = imaginary, input u); the two diagonal 8 x 8 blocks of G
// are the two channels' Grams (the off-diagonal blocks mix channels and are
// not read).  With 5..8 inputs (LS = 8) a product carries one channel: row
// r = 8 part + u.  Rows of absent inputs read input 0's bytes (or duplicate
// slots) and never reach an output.  Per channel half:
//   Re_ab = G[a][b] + G[P + a][P + b],  Im_ab = G[P + a][b] - G[a][P + b]
// with P = 4 (two channels per product) or 8, indices within the half's block.
//
// Workgroup: SEGD / 4 waves; wave w takes dwords 4w .. 4w + 3 (16 channels).
// Each 64-spectrum step the 16-byte row pieces (SEGD / 4 lanes per segment) go
// through registers into LDS [spectrum][SEGD dwords][LS slots] (spectrum stride
// SEGD LS + 1: the operand reads of 4 spectrum groups land in different banks;
// the lanes of one word broadcast), double-buffered, with kPF4w steps of loads
// in flight per thread.
constexpr int kPF4w = 2;

template <int NIN, int SEGD>
__global__ __launch_bounds__(16 * SEGD) void allpairs4_wide_kernel(InputPtrs in, int64_t ncol, int64_t blocklen,
                                                                   int64_t nblock, int nsplit,
                                                                   long long* __restrict__ out) {
  constexpr int NP = NIN * (NIN + 1) / 2;
  constexpr int NWAVE = SEGD / 4;            // waves (4 dwords each)
  constexpr int LPR = SEGD / 4;              // loader lanes per row segment
  constexpr int RPW = 64 / LPR;              // rows per wave load instruction
  constexpr int LS = NIN <= 4 ? 4 : 8;       // loaded input slots
  constexpr int PR = LS;                     // rows per part within a channel block
  constexpr int SS = SEGD * LS + 1;          // LDS dwords per spectrum
  constexpr int E = LS;                      // 16-byte pieces per thread per step
  constexpr int CH = LS == 4 ? 2 : 1;        // channels per product
  __shared__ uint32_t buf[2][64 * SS];
  static_assert(NWAVE * 4 * 256 <= 2 * 64 * SS, "output staging must fit the step buffers");
  static_assert(NWAVE * RPW == 64, "the waves' rows cover a step");
  const int tid = threadIdx.x;
  const int wv = __builtin_amdgcn_readfirstlane(tid / 64), lane = tid % 64;
  int64_t bs, gx;
  mfma_item(ncol / (4 * SEGD), nblock * nsplit, &bs, &gx);
  const int64_t blk = bs / nsplit, sp = bs - blk * nsplit;
  const int64_t nsteps = blocklen / 64 / nsplit;
  const int64_t t_begin = blk * blocklen + sp * nsteps * 64;
  // loader: lane = (16-byte piece q, row r of RPW); piece e: input slot e; wave wv takes
  // spectra RPW wv .. RPW wv + RPW - 1 of each step
  const int q = lane % LPR, r = lane / LPR;
  uint4 ld[kPF4w][E];
  auto fetch = [&](int slot, int64_t st) {
    const int64_t row = t_begin + (st < nsteps ? st : nsteps - 1) * 64 + RPW * wv + r;
#pragma unroll
    for (int e = 0; e < E; ++e)
      ld[slot][e] = *reinterpret_cast<const uint4*>(in.p[e < NIN ? e : 0] + row * ncol + 4 * SEGD * gx + 16 * q);
  };
  auto stash = [&](uint32_t* b, int slot) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      uint32_t* d = b + (RPW * wv + r) * SS + (4 * q) * LS + e;
      d[0 * LS] = ld[slot][e].x;
      d[1 * LS] = ld[slot][e].y;
      d[2 * LS] = ld[slot][e].z;
      d[3 * LS] = ld[slot][e].w;
    }
  };
#pragma unroll
  for (int x = 0; x < kPF4w; ++x) fetch(x, x);
  stash(buf[0], 0);
  fetch(0, kPF4w);
  __syncthreads();
  const int rr = lane % 16, g = lane / 16;
  // operand row: LS = 4: h = rr / 8, part = (rr / 4) % 2, u = rr % 4; LS = 8: h = 0, part = rr / 8, u = rr % 8
  const int h = LS == 4 ? rr / 8 : 0;
  const int part = LS == 4 ? (rr / 4) % 2 : rr / 8;
  const int u = LS == 4 ? rr % 4 : rr % 8;
  constexpr int NPROD = 16 / CH;             // products per wave per step (16 channels)
  intx4 acc[NPROD];
#pragma unroll
  for (int k = 0; k < NPROD; ++k) acc[k] = intx4{0, 0, 0, 0};
  for (int64_t s0 = 0; s0 < nsteps; s0 += 2) {   // s0 even and a multiple of kPF4w: static buffer and slot
#pragma unroll
    for (int hb = 0; hb < 2; ++hb) {
      const int64_t st = s0 + hb;
      if (st < nsteps) {   // uniform
#pragma unroll
        for (int dk = 0; dk < 4 / CH; ++dk) {
          // the lane's dword: 4 wv + dk (+ 2 for channel half 1 when two channels share a product)
          const int dw = 4 * wv + dk + (CH == 2 ? 2 * h : 0);
          if (dk > 0) asm volatile("" ::: "memory");
          const uint32_t* b = buf[hb] + (16 * g) * SS + dw * LS + u;
          uint32_t x[16];
#pragma unroll
          for (int i = 0; i < 16; ++i) x[i] = b[i * SS];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            intx4 v;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
              const uint32_t pd = gather_byte(x[4 * d], x[4 * d + 1], x[4 * d + 2], x[4 * d + 3], j);
              const uint32_t nib = part ? (pd & 0x0F0F0F0Fu) : ((pd >> 4) & 0x0F0F0F0Fu);
              // v > 8 -> v - 16 per byte: set the high nibble of those bytes
              uint32_t sgn = ((nib + 0x07070707u) >> 4) & 0x01010101u;
              sgn <<= 4;
              sgn |= sgn << 1;
              sgn |= sgn << 2;
              v[d] = int(nib | sgn);
            }
            acc[4 * dk + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(v, v, acc[4 * dk + j], 0, 0, 0);
          }
        }
      }
      // buffer hb ^ 1 was last read one step ago; its ring slot then takes step st + 1 + kPF4w
      if (st + 1 < nsteps) stash(buf[hb ^ 1], int((st + 1) % kPF4w));
      fetch(int((st + 1) % kPF4w), st + 1 + kPF4w);
      __syncthreads();
    }
  }
  // D layout: lane l holds G[4 (l / 16) + v][l % 16]; staging [wave][4 products][256]
  int* S = reinterpret_cast<int*>(&buf[0][0]) + wv * 1024;
#pragma unroll
  for (int k0 = 0; k0 < NPROD; k0 += 4) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int vv = 0; vv < 4; ++vv) S[k * 256 + (4 * g + vv) * 16 + rr] = acc[k0 + k][vv];
    __syncthreads();
    // 4 products x CH channels, each NP pairs
    for (int it = lane; it < 4 * CH * NP; it += 64) {
      const int p = it % NP, kc = it / NP, k = kc / CH, hc = kc % CH;
      int a = 0, rem = p;
      while (rem >= NIN - a) { rem -= NIN - a; ++a; }
      const int bb = a + rem;
      const int* G = S + k * 256;
      const int o0 = 8 * hc;   // channel block
      const long long re = (long long)G[(o0 + a) * 16 + o0 + bb] + G[(o0 + PR + a) * 16 + o0 + PR + bb];
      const long long im = (long long)G[(o0 + PR + a) * 16 + o0 + bb] - G[(o0 + a) * 16 + o0 + PR + bb];
      // product k0 + k: dword 4 wv + dk (+ 2 hc), byte j
      const int kk = k0 + k, dk = kk / 4, j = kk % 4;
      const int64_t ch = 4 * SEGD * gx + 4 * (4 * wv + dk + (CH == 2 ? 2 * hc : 0)) + j;
      long long* o = out + ((blk * NP + p) * ncol + ch) * 2;
      if (nsplit > 1) {
        if (re) atomic_add_i64(o, re);
        if (im) atomic_add_i64(o + 1, im);
      } else {
        o[0] = re;
        o[1] = im;
      }
    }
    __syncthreads();
  }
}

// 1-bit: bit b -> x = 2b - 1.  The products run on the bits themselves
// (0/1 int8): with G' = B B^T,  sum x_a x_b = 4 G'_ab - 2 G'_aa - 2 G'_bb + K
// (b^2 = b, so the diagonal holds the bit counts).  Row u = input u; three
// products per channel: re re^T, im im^T, re im^T.
//
// Workgroup: 8 waves, 4 dword columns (16 bytes = 64 channels of every row)
// of one block.  Each 64-spectrum step the workgroup loads the 16-byte row
// pieces of all inputs (one 16-byte load per (input, spectrum)) into LDS,
// double-buffered, laid out [spectrum][dword][input] with a 65-dword spectrum
// stride so the matrix-operand reads (16 inputs x 4 spectrum groups per
// instruction) hit 64 different banks.  Wave w takes bytes 2 (w % 2), +1 of
// dword w / 2: 8 channels, 24 accumulator tiles.  The row pieces of kPF1
// steps are in flight per thread (a register ring), which covers the HBM
// latency at one workgroup per CU.
//
// MASKED: the operand bytes are the raw bytes ANDed with the bit's mask (one
// operation per 4 spectra), so a set bit weighs its place value s (-128 for
// bit 7, signed int8) and each product sum is s_a s_b times the bit count:
// an exact power-of-two multiple, removed by a shift.  The largest weight is
// 128^2 = 2^14, so the int32 sums stay exact while blocklen <= 2^16.
// Otherwise each bit is shifted to place 0 first.
constexpr int kDW1 = 4;          // dword columns per workgroup
constexpr int kTS1 = 65;         // LDS dwords per spectrum
constexpr int kT1 = 512;         // threads per workgroup
constexpr int kPF1 = 4;          // 64-spectrum steps in flight per thread
constexpr int64_t kMasked1Max = int64_t(1) << 16;

template <int NIN, bool MASKED>
__global__ __launch_bounds__(kT1) void allpairs1_mfma_kernel(InputPtrs in, int64_t ncol, int64_t nchan,
                                                             int64_t blocklen, int64_t nblock,
                                                             long long* __restrict__ out) {
  constexpr int NP = NIN * (NIN + 1) / 2;
  __shared__ uint32_t buf[2][64 * kTS1];
  const int tid = threadIdx.x;
  const int wv = tid / 64, lane = tid % 64;
  int64_t blk, gx;
  mfma_item((ncol / 4) / kDW1, nblock, &blk, &gx);
  const int dw = wv / 2, j0 = (wv % 2) * 2;
  const int u = lane % 16, g = lane / 16;
  // loader: chunks c = tid + 512 e (e < 2): input lu = c % 16, spectrum lt = c / 16
  // Rows of absent inputs (lu >= NIN) load input 0's bytes: a pair (a, b) reads
  // only rows a and b of the products, so those rows never reach an output.
  const int lu = tid % 16, lt = tid / 16;           // lt in [0, 32); second chunk lt + 32
  const uint8_t* lsrc = in.p[lu < NIN ? lu : 0] + 16 * gx;
  const int64_t t_begin = blk * blocklen, nsteps = blocklen / 64;
  // Ring slot q holds the row pieces of the steps st with st % kPF1 == q.  The
  // loads are unconditional (steps past the block reload its last step, an L2
  // hit), so the compiler counts the loads in flight exactly and each stash
  // waits only for its own step.
  uint4 ld[kPF1][2];
  auto fetch = [&](int q, int64_t st) {
    const uint8_t* p = lsrc + (t_begin + (st < nsteps ? st : nsteps - 1) * 64 + lt) * ncol;
    ld[q][0] = *reinterpret_cast<const uint4*>(p);
    ld[q][1] = *reinterpret_cast<const uint4*>(p + 32 * ncol);
  };
  auto stash = [&](uint32_t* b, int q) {
    b[lt * kTS1 + 0 * 16 + lu] = ld[q][0].x;
    b[lt * kTS1 + 1 * 16 + lu] = ld[q][0].y;
    b[lt * kTS1 + 2 * 16 + lu] = ld[q][0].z;
    b[lt * kTS1 + 3 * 16 + lu] = ld[q][0].w;
    b[(lt + 32) * kTS1 + 0 * 16 + lu] = ld[q][1].x;
    b[(lt + 32) * kTS1 + 1 * 16 + lu] = ld[q][1].y;
    b[(lt + 32) * kTS1 + 2 * 16 + lu] = ld[q][1].z;
    b[(lt + 32) * kTS1 + 3 * 16 + lu] = ld[q][1].w;
  };
#pragma unroll
  for (int q = 0; q < kPF1; ++q) fetch(q, q);
  stash(buf[0], 0);
  fetch(0, kPF1);
  __syncthreads();
  intx4 acc[2][4][3];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int m = 0; m < 3; ++m) acc[j][c][m] = intx4{0, 0, 0, 0};
  // s0 is a multiple of kPF1 (even): step s0 + q reads buffer q & 1 and ring slot q
  for (int64_t s0 = 0; s0 < nsteps; s0 += kPF1) {
#pragma unroll
    for (int q = 0; q < kPF1; ++q) {
      const int64_t st = s0 + q;
      if (st < nsteps) {   // uniform
        const uint32_t* b = buf[q & 1];
        uint32_t x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = b[(16 * g + i) * kTS1 + dw * 16 + u];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          uint32_t pd[4];
#pragma unroll
          for (int d = 0; d < 4; ++d)
            pd[d] = gather_byte(x[4 * d], x[4 * d + 1], x[4 * d + 2], x[4 * d + 3], j0 + j);
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            intx4 re, im;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
              if (MASKED) {
                re[d] = int(pd[d] & (0x01010101u << (7 - 2 * c)));
                im[d] = int(pd[d] & (0x01010101u << (6 - 2 * c)));
              } else {
                re[d] = int((pd[d] >> (7 - 2 * c)) & 0x01010101u);
                im[d] = int((pd[d] >> (6 - 2 * c)) & 0x01010101u);
              }
            }
            acc[j][c][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(re, re, acc[j][c][0], 0, 0, 0);
            acc[j][c][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(im, im, acc[j][c][1], 0, 0, 0);
            acc[j][c][2] = __builtin_amdgcn_mfma_i32_16x16x64_i8(re, im, acc[j][c][2], 0, 0, 0);
          }
        }
      }
      // buffer (q + 1) & 1 was last read one step ago; the ring slot is free after the stash
      if (st + 1 < nsteps) stash(buf[(q + 1) & 1], (q + 1) % kPF1);
      fetch((q + 1) % kPF1, st + 1 + kPF1);
      __syncthreads();
    }
  }
  // outputs: per channel, the wave's three 16 x 16 products go through its LDS
  // slot (768 dwords of the staging buffers) and lanes form the pairs
  int* S = reinterpret_cast<int*>(&buf[0][0]) + wv * 768;
  const long long K = blocklen;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // bit counts: MASKED sums carry the weights 2^(14-4c) (re re), 2^(12-4c)
        // (im im) and 2^(13-4c) (re im, negative for c = 0: bit 7 weighs -128)
        int rr = acc[j][c][0][v], ii = acc[j][c][1][v], ri = acc[j][c][2][v];
        if (MASKED) {
          rr >>= 14 - 4 * c;
          ii >>= 12 - 4 * c;
          ri = c == 0 ? -(ri >> 13) : ri >> (13 - 4 * c);
        }
        S[0 * 256 + (4 * g + v) * 16 + u] = rr;
        S[1 * 256 + (4 * g + v) * 16 + u] = ii;
        S[2 * 256 + (4 * g + v) * 16 + u] = ri;
      }
      __syncthreads();
      const int64_t ch = 4 * (4 * (gx * kDW1 + dw) + j0 + j) + c;
      if (ch < nchan) {
        for (int p = lane; p < NP; p += 64) {
          int a = 0, rem = p;
          while (rem >= NIN - a) { rem -= NIN - a; ++a; }
          const int bb = a + rem;
          const int* RR = S;
          const int* II = S + 256;
          const int* RI = S + 512;
          const long long ra = RR[a * 17], rb = RR[bb * 17], ia = II[a * 17], ib = II[bb * 17];
          const long long gre = 4LL * RR[a * 16 + bb] - 2 * ra - 2 * rb + K;     // sum re_a re_b
          const long long gim = 4LL * II[a * 16 + bb] - 2 * ia - 2 * ib + K;     // sum im_a im_b
          const long long xab = 4LL * RI[a * 16 + bb] - 2 * ra - 2 * ib + K;     // sum re_a im_b
          const long long xba = 4LL * RI[bb * 16 + a] - 2 * rb - 2 * ia + K;     // sum re_b im_a
          long long* o = out + ((blk * NP + p) * nchan + ch) * 2;
          o[0] = gre + gim;
          o[1] = xba - xab;
        }
      }
      __syncthreads();
    }
  }
}

// 1-bit, one 16 x 16 tile per channel on the FP4 matrix cores.  With
// x = +-1 coded as the e2m1 nibble 0x2 | (b << 3) (+1 or -1: the sign of x
// flipped, which leaves every product below unchanged), P = x_re, Q = x_im,
// and k concatenating two copies of a step's 64 spectra,
//   MFMA(A = [Q | P], B = [Q | P])  = Q Q^T + P P^T
//   MFMA(A = [Q | P], B = [P | -Q]) = Q P^T - P Q^T
// sum to T = Re + Im with Re_ab = sum x_re,a x_re,b + x_im,a x_im,b
// (symmetric) and Im_ab = sum x_im,a x_re,b - x_re,a x_im,b (antisymmetric),
// so Re = (T + T^T) / 2 and Im = (T - T^T) / 2.  Every product is +-1 and
// every sum an integer with |T| <= 4 blocklen <= 2^24: the FP32 accumulation
// is exact while blocklen <= 2^22.  One tile per channel (a third of the
// accumulators of the three-product form) lets one workgroup own 64 bytes of
// every row (256 channels): each 64-byte line is fetched once and used whole.
//
// Workgroup: 8 waves, 64 bytes of columns of one block.  Each 64-spectrum step
// the 16-byte row pieces (4 lanes per 64-byte line) go through registers into
// LDS [spectrum][input][16 dwords] (strides kSS1t, kUS1t), double-buffered,
// with the next step's loads in flight.  Wave w takes dwords 2w and 2w + 1:
// 8 bytes, 32 channels, 32 tiles; lane (u, g) holds spectra 16 g .. 16 g + 15
// of input u, 8 per operand dword.  Rows of absent inputs (u >= NIN) read
// input 0's bytes and never reach an output.
constexpr int kW1t = 64;                 // row bytes per workgroup
constexpr int kUS1t = 18;                // LDS dwords per (spectrum, input)
constexpr int kSS1t = 16 * kUS1t + 6;    // LDS dwords per spectrum
constexpr int64_t kTile1Max = int64_t(1) << 22;

typedef int intx8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// D += A B^T over 128 e2m1 elements per row (unit block scales; the e2m1
// operands are the low 4 dwords of a and b)
__device__ __forceinline__ floatx4 mfma_fp4(intx8 a, intx8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 4, 4, 0, 0, 0, 0);
}

// shift left by s (right for s < 0)
__device__ __forceinline__ uint32_t shl(uint32_t v, int s) { return s >= 0 ? v << s : v >> -s; }

// e2m1 sign nibbles of 8 spectra.  pa, pb: one byte column of 4 spectra each.
// hi holds the high nibbles of pa (high) and pb (low) of every byte, lo the
// low nibbles; the channel bit at place p of both lands on bits 7 and 3 of
// (p >= 4 ? hi : lo) << (3 - p % 4), which ORed with 0x2 per nibble gives
// 0x2 | (bit << 3).  v_bitop3 truth tables over (s0, s1, s2) = (0xF0, 0xCC,
// 0xAA): 0xE4 = s2 ? s0 : s1, 0xEA = (s0 & s1) | s2.
__device__ __forceinline__ void nibble_halves(uint32_t pa, uint32_t pb, uint32_t* hi, uint32_t* lo) {
  *hi = __builtin_amdgcn_bitop3_b32(pa, pb >> 4, 0xF0F0F0F0u, 0xE4);
  *lo = __builtin_amdgcn_bitop3_b32(pa << 4, pb, 0xF0F0F0F0u, 0xE4);
}
__device__ __forceinline__ uint32_t fp4_signs(uint32_t src, int s) {
  return __builtin_amdgcn_bitop3_b32(src << s, 0x88888888u, 0x22222222u, 0xEA);
}

template <int NIN>
__global__ __launch_bounds__(512) void allpairs1_tile_kernel(InputPtrs in, int64_t ncol, int64_t nchan,
                                                             int64_t blocklen, int64_t nblock,
                                                             long long* __restrict__ out) {
  constexpr int NP = NIN * (NIN + 1) / 2;
  constexpr int NW = 8;                // waves
  constexpr int BW = kW1t / NW;        // byte columns per wave
  constexpr int IW = 16 / NW;          // inputs loaded per wave
  constexpr int E = 4 * IW;            // 16-byte pieces per thread per step
  constexpr int R = 1;                 // steps of pieces in registers (a deeper ring gains nothing:
                                       // the loads are bound by the requests in flight per CU)
  __shared__ uint32_t buf[2][64 * kSS1t];
  static_assert(NW * 4 * 256 <= 2 * 64 * kSS1t, "output staging must fit the step buffers");
  const int tid = threadIdx.x;
  const int wv = __builtin_amdgcn_readfirstlane(tid / 64), lane = tid % 64;   // wave index: uniform
  int64_t blk, gx;
  mfma_item(ncol / kW1t, nblock, &blk, &gx);
  // loader: wave wv loads inputs IW wv .. IW wv + IW - 1; lane = (16-byte piece q, spectrum r of 16)
  const int q = lane % 4, r = lane / 4;
  const uint8_t* src[IW];
#pragma unroll
  for (int h = 0; h < IW; ++h) src[h] = in.p[IW * wv + h < NIN ? IW * wv + h : 0] + kW1t * gx + 16 * q;
  const int64_t t_begin = blk * blocklen, nsteps = blocklen / 64;
  // ring slot x % R holds step x from its loads to its stash, R steps later
  uint4 ld[R][E];
  auto fetch = [&](int slot, int64_t st) {   // piece e: input IW wv + e % IW, spectrum r + 16 (e / IW)
    const int64_t row = t_begin + (st < nsteps ? st : nsteps - 1) * 64 + r;
#pragma unroll
    for (int e = 0; e < E; ++e)
      ld[slot][e] = *reinterpret_cast<const uint4*>(src[e % IW] + (row + 16 * (e / IW)) * ncol);
  };
  auto stash = [&](uint32_t* b, int slot) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      uint32_t* d = b + (r + 16 * (e / IW)) * kSS1t + (IW * wv + e % IW) * kUS1t + 4 * q;
      *reinterpret_cast<uint2*>(d) = make_uint2(ld[slot][e].x, ld[slot][e].y);
      *reinterpret_cast<uint2*>(d + 2) = make_uint2(ld[slot][e].z, ld[slot][e].w);
    }
  };
#pragma unroll
  for (int x = 0; x < R; ++x) fetch(x, x);
  stash(buf[0], 0);
  fetch(0, R);
  __syncthreads();
  const int u = lane % 16, g = lane / 16;
  floatx4 acc[BW][4];
#pragma unroll
  for (int jj = 0; jj < BW; ++jj)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[jj][c] = floatx4{0, 0, 0, 0};
  for (int64_t s0 = 0; s0 < nsteps; s0 += 2) {   // s0 even and a multiple of R: static buffer and slot
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t st = s0 + h;
      if (st < nsteps) {   // uniform
#pragma unroll
        for (int half = 0; half < BW / 4; ++half) {
          // dword BW / 4 wv + half of the 16 spectra 16 g .. 16 g + 15 of input u;
          // read after the previous half's products (16 live registers, not 32)
          if (half > 0) asm volatile("" ::: "memory");
          const uint32_t* b = buf[h] + (16 * g) * kSS1t + u * kUS1t + (BW / 4) * wv + half;
          uint32_t x[16];
#pragma unroll
          for (int i = 0; i < 16; ++i) x[i] = b[i * kSS1t];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int jj = 4 * half + j;
            uint32_t pd[4];
#pragma unroll
            for (int d = 0; d < 4; ++d) pd[d] = gather_byte(x[4 * d], x[4 * d + 1], x[4 * d + 2], x[4 * d + 3], j);
            // operands of two channels first, w = [P0 P1 Q0 Q1 -P0 -P1]:
            //   MFMA(A = w[0:4], B = w[0:4]) = P P^T + Q Q^T
            //   MFMA(A = w[2:6], B = w[0:4]) = Q P^T - P Q^T
            // (every operand a window of w), then the products of the two
            // interleaved; channel c's re bit is at place 7 - 2c, its im bit at 6 - 2c
            uint32_t hi[2], lo[2];
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) nibble_halves(pd[2 * hh], pd[2 * hh + 1], &hi[hh], &lo[hh]);
#pragma unroll
            for (int cp = 0; cp < 2; ++cp) {   // channels 2 cp, 2 cp + 1: from hi (cp = 0) or lo
              intx8 w[2];
#pragma unroll
              for (int k = 0; k < 2; ++k) {
#pragma unroll
                for (int hh = 0; hh < 2; ++hh) {
                  const uint32_t src = cp == 0 ? hi[hh] : lo[hh];
                  w[k][hh] = int(fp4_signs(src, 2 * k));             // x_re
                  w[k][2 + hh] = int(fp4_signs(src, 2 * k + 1));     // x_im
                  w[k][4 + hh] = w[k][hh] ^ int(0x88888888u);        // -x_re
                }
                w[k][6] = w[k][7] = 0;
              }
#pragma unroll
              for (int k = 0; k < 2; ++k) acc[jj][2 * cp + k] = mfma_fp4(w[k], w[k], acc[jj][2 * cp + k]);
#pragma unroll
              for (int k = 0; k < 2; ++k)
                acc[jj][2 * cp + k] = mfma_fp4(__builtin_shufflevector(w[k], w[k], 2, 3, 4, 5, 6, 7, 6, 7), w[k],
                                               acc[jj][2 * cp + k]);
            }
          }
        }
      }
      // buffer h ^ 1 was last read one step ago; its ring slot then takes step st + 1 + R
      if (st + 1 < nsteps) stash(buf[h ^ 1], int((st + 1) % R));
      fetch(int((st + 1) % R), st + 1 + R);
      __syncthreads();
    }
  }
  // outputs, 4 channels (one byte column) at a time: the wave's 4 tiles go
  // through its 4 KB of the staging buffers; lanes take (pair, channel) with
  // the channel fastest, so each pair's 4 channels are 64 contiguous bytes
  int* S = reinterpret_cast<int*>(&buf[0][0]) + wv * 1024;
#pragma unroll
  for (int jj = 0; jj < BW; ++jj) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) S[c * 256 + (4 * g + v) * 16 + u] = int(acc[jj][c][v]);   // exact integers
    __syncthreads();
    const int64_t ch0 = 4 * (kW1t * gx + BW * wv + jj);
    for (int it = lane; it < NP * 4; it += 64) {
      const int c = it % 4, p = it / 4;
      int a = 0, rem = p;
      while (rem >= NIN - a) { rem -= NIN - a; ++a; }
      const int bb = a + rem;
      const long long tab = S[c * 256 + a * 16 + bb], tba = S[c * 256 + bb * 16 + a];
      if (ch0 + c < nchan) {
        long long* o = out + ((blk * NP + p) * nchan + ch0 + c) * 2;
        o[0] = (tab + tba) >> 1;   // even: T + T^T = 2 Re
        o[1] = (tab - tba) >> 1;
      }
    }
    __syncthreads();
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
  const int64_t method = tune_get("xcorr_int_method");
  bool aligned16 = true;
  for (int u = 0; u < ninput; ++u) aligned16 = aligned16 && (reinterpret_cast<uintptr_t>(inputs[u]) % 16 == 0);
  const bool mfma_ok = blocklen % 64 == 0 && ncol % 16 == 0 && aligned16 && blocklen <= (int64_t(1) << 24);
  ALB_REQUIRE(method != 2 || mfma_ok,
              "alb_xcorr_4bit_allpairs: the matrix-core form needs blocklen %% 64 == 0, ncol %% 16 == 0 and "
              "16-byte aligned inputs");
  if (method != 1 && mfma_ok && ncol % 64 == 0) {
    // 128- or 64-byte column groups; enough workgroups to fill the chip: split
    // blocks (exact partial sums, combined by atomics into the zero-filled output)
    int nsplit = 1;
    const int segd = ncol % 128 == 0 && tune_get("xcorr4_segment") == 128 ? 32 : 16;
    const int64_t ngx = ncol / (4 * segd);
    const int64_t target = tune_get("xcorr4_workgroups") > 0 ? tune_get("xcorr4_workgroups") : 256;
    while (ngx * nblock * nsplit < target && (blocklen / 64) % (2 * nsplit) == 0) nsplit *= 2;
    const int64_t total = ngx * nblock * nsplit;
    ALB_REQUIRE(total < (int64_t(1) << 31), "alb_xcorr_4bit_allpairs: too many blocks");
    if (nsplit > 1)
      hipLaunchKernelGGL(fill_i64, grid1(nblock * npair * ncol * 2), dim3(256), 0, s, out, nblock * npair * ncol * 2);
    switch (ninput) {
#define CASE(n) case n: \
      if (segd == 32) hipLaunchKernelGGL((allpairs4_wide_kernel<n, 32>), dim3(unsigned(total)), dim3(512), 0, s, ip, ncol, blocklen, nblock, nsplit, (long long*)out); \
      else hipLaunchKernelGGL((allpairs4_wide_kernel<n, 16>), dim3(unsigned(total)), dim3(256), 0, s, ip, ncol, blocklen, nblock, nsplit, (long long*)out); \
      break;
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    }
    ALB_LAUNCH_CHECK("allpairs4_wide_kernel");
    return ALB_OK;
  }
  if (method != 1 && mfma_ok) {
    // enough workgroups to fill the chip: split blocks (exact partial sums,
    // combined by atomics into the zero-filled output)
    int nsplit = 1;
    const int64_t ngx = (ncol / 4) / kDW4;
    const int64_t target = tune_get("xcorr4_workgroups") > 0 ? tune_get("xcorr4_workgroups") : 2048;
    while (ngx * nblock * nsplit < target && (blocklen / 64) % (2 * nsplit) == 0) nsplit *= 2;
    const int64_t total = ngx * nblock * nsplit;
    ALB_REQUIRE(total < (int64_t(1) << 31), "alb_xcorr_4bit_allpairs: too many blocks");
    if (nsplit > 1)
      hipLaunchKernelGGL(fill_i64, grid1(nblock * npair * ncol * 2), dim3(256), 0, s, out, nblock * npair * ncol * 2);
    switch (ninput) {
#define CASE(n) case n: hipLaunchKernelGGL(allpairs4_mfma_kernel<n>, dim3(unsigned(total)), dim3(256), 0, s, ip, ncol, blocklen, nblock, nsplit, (long long*)out); break;
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    }
    ALB_LAUNCH_CHECK("allpairs4_mfma_kernel");
    return ALB_OK;
  }
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
  const int64_t method = tune_get("xcorr_int_method");
  bool aligned16 = true;
  for (int u = 0; u < ninput; ++u) aligned16 = aligned16 && (reinterpret_cast<uintptr_t>(inputs[u]) % 16 == 0);
  const bool mfma_ok = blocklen % 64 == 0 && ncol % 16 == 0 && aligned16 && blocklen <= (int64_t(1) << 28);
  ALB_REQUIRE(method != 2 || mfma_ok,
              "alb_xcorr_1bit_allpairs: the matrix-core form needs blocklen %% 64 == 0, ncol %% 16 == 0 and "
              "16-byte aligned inputs");
  // matrix-core forms: one tile per channel (64-byte column groups, blocks of
  // at most 2^16 spectra) or three products per channel (any other shape)
  const int64_t form = tune_get("xcorr1_form");
  const bool tile_ok = ncol % kW1t == 0 && blocklen <= kTile1Max;
  ALB_REQUIRE(form != 1 || (mfma_ok && tile_ok),
              "alb_xcorr_1bit_allpairs: the one-tile form needs ncol %% 64 == 0 and blocklen <= 2^22");
  if (method != 1 && mfma_ok && tile_ok && form != 2) {
    const int64_t total = (ncol / kW1t) * nblock;
    ALB_REQUIRE(total < (int64_t(1) << 31), "alb_xcorr_1bit_allpairs: too many blocks");
    switch (ninput) {
#define CASE(n) case n: hipLaunchKernelGGL(allpairs1_tile_kernel<n>, dim3(unsigned(total)), dim3(512), 0, s, ip, ncol, nchan, blocklen, nblock, (long long*)out); break;
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12) CASE(13)
      CASE(14) CASE(15) CASE(16)
#undef CASE
    }
    ALB_LAUNCH_CHECK("allpairs1_tile_kernel");
    return ALB_OK;
  }
  if (method != 1 && mfma_ok) {
    const int64_t total = ((ncol / 4) / kDW1) * nblock;
    ALB_REQUIRE(total < (int64_t(1) << 31), "alb_xcorr_1bit_allpairs: too many blocks");
    const bool masked = blocklen <= kMasked1Max;
    switch (ninput) {
#define CASE(n) case n: \
      if (masked) hipLaunchKernelGGL((allpairs1_mfma_kernel<n, true>), dim3(unsigned(total)), dim3(kT1), 0, s, ip, ncol, nchan, blocklen, nblock, (long long*)out); \
      else hipLaunchKernelGGL((allpairs1_mfma_kernel<n, false>), dim3(unsigned(total)), dim3(kT1), 0, s, ip, ncol, nchan, blocklen, nblock, (long long*)out); \
      break;
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12) CASE(13)
      CASE(14) CASE(15) CASE(16)
#undef CASE
    }
    ALB_LAUNCH_CHECK("allpairs1_mfma_kernel");
    return ALB_OK;
  }
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
