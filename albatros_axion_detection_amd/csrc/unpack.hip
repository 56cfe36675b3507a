// Baseband unpack / repack kernels: 4-bit and 1-bit decode to complex64,
// per-pol repack (sortpols), raw-value histograms, spec_num zero-fill scatter
// and the dropped-packet zero fill.  All memory-bound.  The unpack and sortpols
// kernels are row-tiled (a lane owns a column or column group and walks the
// rows of its workgroup); the others are grid-stride loops with 64-bit
// indexing, every thread owning whole output elements.
#include "common.hpp"

namespace alb {

// Signed 4-bit rule of the reference: v > 8 ? v - 16 : v   (8 -> +8).
__device__ __forceinline__ int nib(int v) { return v > 8 ? v - 16 : v; }

__global__ void fill_u8(uint8_t* p, int64_t n, uint8_t v) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = v;
}
__global__ void fill_u64(uint64_t* p, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = 0;
}
__global__ void fill_c64(float2* p, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    p[i] = make_float2(0.f, 0.f);
}

static inline dim3 grid_for(int64_t n, int bs = 256, int64_t cap = 65536) {
  int64_t g = (n + bs - 1) / bs;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return dim3(unsigned(g));
}

// unpacking.c:155-205: (rowend-rowstart, ncols) complex64 per pol, column k from
// byte 2*channels[k] (+1 for pol1) of spectra of 2*nchan bytes.
__global__ void unpack4_kernel(const uint8_t* __restrict__ data, float2* __restrict__ pol0, float2* __restrict__ pol1,
                               int64_t rowstart, int64_t nrows, const int64_t* __restrict__ channels, int64_t ncols,
                               int64_t nchan) {
  // lane = column k of every row r = blockIdx.y + gridDim.y * m
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k >= ncols) return;
  const int64_t off = 2 * channels[k];
  for (int64_t r = blockIdx.y; r < nrows; r += gridDim.y) {
    const int64_t i = r * ncols + k;
    const uint8_t* p = data + (rowstart + r) * 2 * nchan + off;
    const int b0 = p[0], b1 = p[1];
    pol0[i] = make_float2(float(nib(b0 >> 4)), float(nib(b0 & 15)));
    pol1[i] = make_float2(float(nib(b1 >> 4)), float(nib(b1 & 15)));
  }
}

// unpacking.c:207-246: byte j of a spectrum = channels 2j, 2j+1 of both pols,
// bits r0c0 i0c0 r1c0 i1c0 r0c1 i0c1 r1c1 i1c1; 1 -> +1, 0 -> -1.
__global__ void unpack1_kernel(const uint8_t* __restrict__ data, float2* __restrict__ pol0, float2* __restrict__ pol1,
                               int64_t rowstart, int64_t nrows, int64_t chanstart, int64_t nbytes_out, int64_t nchan) {
  const int64_t n = nrows * nbytes_out;
  const int64_t bps = nchan / 2;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = i / nbytes_out, j = i - r * nbytes_out;
    const int b = data[(r + rowstart) * bps + chanstart / 2 + j];
    auto pm = [&](int s) { return float(2 * ((b >> s) & 1) - 1); };
    const int64_t o = r * 2 * nbytes_out + 2 * j;
    pol0[o] = make_float2(pm(7), pm(6));
    pol0[o + 1] = make_float2(pm(3), pm(2));
    pol1[o] = make_float2(pm(5), pm(4));
    pol1[o + 1] = make_float2(pm(1), pm(0));
  }
}

// Row-tiled forms of the three kernels above and below for aligned shapes:
// x indexes a group of columns, rows run over blockIdx.y + gridDim.y * j (the
// sortpols forms) or over one contiguous block of rows per y (the unpack
// forms: their workgroups then write long contiguous runs, 9% faster for
// 1-bit unpack, 1% for the 4-bit pair form, while sortpols measured 6%
// slower that way), so no 64-bit division per element, and each lane writes
// 16 bytes per pol (4-bit and 1-bit unpack, two complex64 each) or 4 bytes
// per pol (sortpols, four channels).
// streaming 16-byte store: the unpacked pols are written once and not read back here
typedef float nt_float4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_nt(float4* p, float a, float b, float c, float d) {
  const nt_float4 v = {a, b, c, d};
  __builtin_nontemporal_store(v, reinterpret_cast<nt_float4*>(p));
}

// Rows [*rb, *re) of workgroup row block blockIdx.y: nrows split into
// gridDim.y contiguous blocks.
__device__ __forceinline__ void row_block(int64_t nrows, int64_t* rb, int64_t* re) {
  const int64_t per = (nrows + gridDim.y - 1) / gridDim.y;
  *rb = int64_t(blockIdx.y) * per;
  *re = *rb + per < nrows ? *rb + per : nrows;
}

// 4-bit unpack through LDS: kRows4 consecutive raw spectra (2 nchan bytes
// each, contiguous in memory) are read with one 16-byte load per lane into
// LDS, and each lane gathers the two bytes of each of its two channels from
// there; the next batch's load is issued before this batch's stores.  Against
// the pair form's per-lane 32-bit loads: 1.076 -> 0.970 ms at 2^18 spectra x
// 1024 channels, bit-identical.  One workgroup covers every column pair of a
// row (npair <= 512), kRows4 spectra fit the 512 lanes' loads
// (2 nchan kRows4 <= 8192 bytes).
constexpr int kRows4 = 4;
__global__ __launch_bounds__(512) void unpack4_lds_kernel(const uint8_t* __restrict__ data, float4* __restrict__ pol0,
                                                         float4* __restrict__ pol1, int64_t rowstart, int64_t nrows,
                                                         const int64_t* __restrict__ channels, int64_t npair,
                                                         int64_t nchan) {
  extern __shared__ uint4 rowbuf[];   // kRows4 spectra of raw bytes
  const int tid = threadIdx.x;
  const int64_t kp = blockIdx.x * int64_t(blockDim.x) + tid;
  const bool valid = kp < npair;
  const int o0 = valid ? int(2 * channels[2 * kp]) : 0, o1 = valid ? int(2 * channels[2 * kp + 1]) : 0;
  const int64_t rowbytes = 2 * nchan;
  int64_t rb, re;
  row_block(nrows, &rb, &re);
  if (rb >= re) return;   // the whole workgroup
  const int n16 = int(kRows4 * rowbytes / 16);   // 16-byte pieces of a full batch (<= blockDim.x)
  // lane tid's piece of the batch starting at spectrum r0 (zero past the block's rows)
  auto fetch = [&](int64_t r0) -> uint4 {
    const int64_t nr = re - r0 < kRows4 ? re - r0 : kRows4;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (tid < nr * rowbytes / 16) v = reinterpret_cast<const uint4*>(data + (rowstart + r0) * rowbytes)[tid];
    return v;
  };
  uint4 nxt = fetch(rb);
  for (int64_t r0 = rb; r0 < re; r0 += kRows4) {
    if (tid < n16) rowbuf[tid] = nxt;
    __syncthreads();
    if (r0 + kRows4 < re) nxt = fetch(r0 + kRows4);
    const int nr = int(re - r0 < kRows4 ? re - r0 : kRows4);
    if (valid) {
      const uint8_t* base = reinterpret_cast<const uint8_t*>(rowbuf);
      for (int rr = 0; rr < nr; ++rr) {
        const int a = *reinterpret_cast<const uint16_t*>(base + rr * rowbytes + o0);   // byte 0 pol0, byte 1 pol1
        const int b = *reinterpret_cast<const uint16_t*>(base + rr * rowbytes + o1);
        const int64_t r = r0 + rr;
        store_nt(&pol0[r * npair + kp], float(nib((a >> 4) & 15)), float(nib(a & 15)), float(nib((b >> 4) & 15)),
                 float(nib(b & 15)));
        store_nt(&pol1[r * npair + kp], float(nib((a >> 12) & 15)), float(nib((a >> 8) & 15)),
                 float(nib((b >> 12) & 15)), float(nib((b >> 8) & 15)));
      }
    }
    __syncthreads();   // every lane has read the batch before the next one is stored
  }
}

__global__ void unpack4_pair_kernel(const uint8_t* __restrict__ data, float4* __restrict__ pol0,
                                    float4* __restrict__ pol1, int64_t rowstart, int64_t nrows,
                                    const int64_t* __restrict__ channels, int64_t npair, int64_t nchan, bool word_rows) {
  const int64_t kp = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (kp >= npair) return;
  const int64_t o0 = 2 * channels[2 * kp], o1 = 2 * channels[2 * kp + 1];
  // consecutive channels starting on a 4-byte boundary: one 32-bit load
  const bool word = word_rows && o1 == o0 + 2 && (o0 & 3) == 0;
  int64_t rb, re;
  row_block(nrows, &rb, &re);
  for (int64_t r = rb; r < re; ++r) {
    const uint8_t* row = data + (rowstart + r) * 2 * nchan;
    int a, b;   // byte 0 pol0, byte 1 pol1 of each channel
    if (word) {
      const uint32_t w = *reinterpret_cast<const uint32_t*>(row + o0);
      a = int(w & 0xffffu);
      b = int(w >> 16);
    } else {
      a = *reinterpret_cast<const uint16_t*>(row + o0);
      b = *reinterpret_cast<const uint16_t*>(row + o1);
    }
    store_nt(&pol0[r * npair + kp], float(nib((a >> 4) & 15)), float(nib(a & 15)), float(nib((b >> 4) & 15)),
             float(nib(b & 15)));
    store_nt(&pol1[r * npair + kp], float(nib((a >> 12) & 15)), float(nib((a >> 8) & 15)),
             float(nib((b >> 12) & 15)), float(nib((b >> 8) & 15)));
  }
}

__global__ void unpack1_rows_kernel(const uint8_t* __restrict__ data, float4* __restrict__ pol0,
                                    float4* __restrict__ pol1, int64_t rowstart, int64_t nrows, int64_t chanstart,
                                    int64_t nb, int64_t nchan) {
  const int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (j >= nb) return;
  const int64_t bps = nchan / 2;
  int64_t rb, re;
  row_block(nrows, &rb, &re);
  for (int64_t r = rb; r < re; ++r) {
    const int b = data[(r + rowstart) * bps + chanstart / 2 + j];
    auto pm = [&](int s) { return float(2 * ((b >> s) & 1) - 1); };
    store_nt(&pol0[r * nb + j], pm(7), pm(6), pm(3), pm(2));
    store_nt(&pol1[r * nb + j], pm(5), pm(4), pm(1), pm(0));
  }
}

// four channels per lane: bytes c0p0 c0p1 c1p0 c1p1 c2p0 c2p1 c3p0 c3p1 -> one
// 32-bit word per pol
__global__ void sortpols4_quad_kernel(const uint8_t* __restrict__ data, uint32_t* __restrict__ pol0,
                                      uint32_t* __restrict__ pol1, int64_t rowstart, int64_t nrows, int64_t nquad,
                                      int64_t nchan, int64_t chanstart) {
  const int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (q >= nquad) return;
  for (int64_t r = blockIdx.y; r < nrows; r += gridDim.y) {
    const uint2 v = *reinterpret_cast<const uint2*>(data + (r + rowstart) * 2 * nchan + 2 * chanstart + 8 * q);
    pol0[r * nquad + q] = (v.x & 0xffu) | ((v.x >> 8) & 0xff00u) | ((v.y & 0xffu) << 16) | ((v.y << 8) & 0xff000000u);
    pol1[r * nquad + q] = ((v.x >> 8) & 0xffu) | ((v.x >> 16) & 0xff00u) | ((v.y << 8) & 0xff0000u) |
                          (v.y & 0xff000000u);
  }
}

static inline dim3 grid_rows(int64_t nx, int64_t nrows, int bs = 256) {
  const int64_t gx = (nx + bs - 1) / bs;
  int64_t gy = nrows < 4096 ? nrows : 4096;
  if (gy < 1) gy = 1;
  return dim3(unsigned(gx), unsigned(gy));
}

static inline bool aligned(const void* p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0; }

// unpacking.c:263-275: 4-bit per-pol byte split.
__global__ void sortpols4_kernel(const uint8_t* __restrict__ data, uint8_t* __restrict__ pol0, uint8_t* __restrict__ pol1,
                                 int64_t rowstart, int64_t nrows, int64_t ncols, int64_t nchan, int64_t chanstart) {
  const int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (k >= ncols) return;
  for (int64_t r = blockIdx.y; r < nrows; r += gridDim.y) {
    const int64_t i = r * ncols + k;
    const uint8_t* p = data + (r + rowstart) * 2 * nchan + 2 * (k + chanstart);
    pol0[i] = p[0];
    pol1[i] = p[1];
  }
}

// unpacking.c:281-366: 1-bit repack, 4 channels per output byte, ragged last byte.
__global__ void sortpols1_kernel(const uint8_t* __restrict__ data, uint8_t* __restrict__ pol0, uint8_t* __restrict__ pol1,
                                 int64_t rowstart, int64_t nrows, int64_t ncols, int64_t nchan, int64_t chanstart,
                                 int cndn) {
  // lane = output byte j of every row r = blockIdx.y + gridDim.y * k
  const int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (j >= ncols) return;
  const int64_t bps = nchan / 2;
  const int use = j == ncols - 1 ? cndn : 0;   // 0 = two full raw bytes
  for (int64_t r = blockIdx.y; r < nrows; r += gridDim.y) {
    const int64_t i = r * ncols + j;
    const int64_t idx = bps * (r + rowstart) + 2 * j + chanstart / 2;
    const int b1 = data[idx];
    const int b2 = (use == 0 || use == 3) ? int(data[idx + 1]) : 0;
    const int p0c0 = b1 & 192, p1c0 = b1 & 48, p0c1 = b1 & 12, p1c1 = b1 & 3;
    const int p0c2 = b2 & 192, p1c2 = b2 & 48, p0c3 = b2 & 12, p1c3 = b2 & 3;
    int o0, o1;
    switch (use) {
      case 0: o0 = p0c0 + (p0c1 << 2) + (p0c2 >> 4) + (p0c3 >> 2); o1 = (p1c0 << 2) + (p1c1 << 4) + (p1c2 >> 2) + p1c3; break;
      case 3: o0 = p0c0 + (p0c1 << 2) + (p0c2 >> 4); o1 = (p1c0 << 2) + (p1c1 << 4) + (p1c2 >> 2); break;
      case 2: o0 = p0c0 + (p0c1 << 2); o1 = (p1c0 << 2) + (p1c1 << 4); break;
      default: o0 = p0c0; o1 = p1c0 << 2; break;
    }
    pol0[i] = uint8_t(o0);
    pol1[i] = uint8_t(o1);
  }
}

// sortpols1_kernel for full byte pairs (channel span a multiple of 4), four
// output bytes per lane: raw bytes 2k (even) and 2k + 1 (odd) of the lane's 8
// make output byte k, so the even and odd bytes are gathered into one dword
// each (v_perm_b32) and every bit move runs on all four bytes at once.  Block
// (64, 4): x = quad of the row, y = row within the block's 4 rows.
__global__ __launch_bounds__(256) void sortpols1_quad_kernel(const uint8_t* __restrict__ data,
                                                             uint32_t* __restrict__ pol0, uint32_t* __restrict__ pol1,
                                                             int64_t rowstart, int64_t nrows, int64_t nquad,
                                                             int64_t nchan, int64_t chanstart) {
  const int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (q >= nquad) return;
  const int64_t bps = nchan / 2;
  for (int64_t r = int64_t(blockIdx.y) * blockDim.y + threadIdx.y; r < nrows; r += int64_t(gridDim.y) * blockDim.y) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(data + bps * (r + rowstart) + chanstart / 2) + 2 * q;
    const uint32_t w0 = src[0], w1 = src[1];
    const uint32_t e = __builtin_amdgcn_perm(w1, w0, 0x06040200u);   // raw bytes 0, 2, 4, 6
    const uint32_t o = __builtin_amdgcn_perm(w1, w0, 0x07050301u);   // raw bytes 1, 3, 5, 7
    const uint32_t o0 = (e & 0xC0C0C0C0u) | ((e & 0x0C0C0C0Cu) << 2) | ((o & 0xC0C0C0C0u) >> 4) |
                        ((o & 0x0C0C0C0Cu) >> 2);
    const uint32_t o1 = ((e & 0x30303030u) << 2) | ((e & 0x03030303u) << 4) | ((o & 0x30303030u) >> 2) |
                        (o & 0x03030303u);
    pol0[r * nquad + q] = o0;
    pol1[r * nquad + q] = o1;
  }
}

// unpacking.c:7-74: raw nibble histogram hist[v * nchan + j] over rows; mode 0/1/-1.
// One thread per (channel, row chunk), register counters, 64-bit atomics.
__global__ void hist4_kernel(const uint8_t* __restrict__ data, unsigned long long* hist, int64_t rowstart, int64_t nrows,
                             int64_t nchan, int mode, int64_t rows_per) {
  const int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (j >= nchan) return;
  const int64_t r0 = blockIdx.y * rows_per;
  const int64_t r1 = r0 + rows_per < nrows ? r0 + rows_per : nrows;
  unsigned cnt[16];
#pragma unroll
  for (int v = 0; v < 16; ++v) cnt[v] = 0;
  for (int64_t r = r0; r < r1; ++r) {
    const uint8_t* p = data + (r + rowstart) * 2 * nchan + 2 * j;
    if (mode == 0 || mode == -1) { const int b = p[0]; cnt[b >> 4]++; cnt[b & 15]++; }
    if (mode == 1 || mode == -1) { const int b = p[1]; cnt[b >> 4]++; cnt[b & 15]++; }
  }
#pragma unroll
  for (int v = 0; v < 16; ++v)
    if (cnt[v]) atomicAdd(&hist[v * nchan + j], (unsigned long long)cnt[v]);
}

// unpacking.c:76-153: 1-bit histogram, two bins per channel.
__global__ void hist1_kernel(const uint8_t* __restrict__ data, unsigned long long* hist, int64_t rowstart, int64_t nrows,
                             int64_t nchan, int mode, int64_t rows_per) {
  const int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;  // byte column
  const int64_t ncols = nchan / 2;
  if (j >= ncols) return;
  const int64_t r0 = blockIdx.y * rows_per;
  const int64_t r1 = r0 + rows_per < nrows ? r0 + rows_per : nrows;
  unsigned c[2][2] = {{0, 0}, {0, 0}};  // [channel 2j / 2j+1][bin]
  for (int64_t r = r0; r < r1; ++r) {
    const int b = data[(r + rowstart) * ncols + j];
    if (mode == 0 || mode == -1) {
      c[0][(b >> 7) & 1]++; c[0][(b >> 6) & 1]++; c[1][(b >> 3) & 1]++; c[1][(b >> 2) & 1]++;
    }
    if (mode == 1 || mode == -1) {
      c[0][(b >> 5) & 1]++; c[0][(b >> 4) & 1]++; c[1][(b >> 1) & 1]++; c[1][b & 1]++;
    }
  }
  for (int ch = 0; ch < 2; ++ch)
    for (int v = 0; v < 2; ++v)
      if (c[ch][v]) atomicAdd(&hist[v * nchan + 2 * j + ch], (unsigned long long)c[ch][v]);
}

// baseband_data_classes.py:48-73: out[(specnum[p]-start) % nspec, channels[k]] = spec[p, k].
__global__ void scatter_kernel(const float2* __restrict__ spec, int64_t spec_ld, const int64_t* __restrict__ specnum,
                               int64_t npresent, int64_t start, int64_t nspec, const int64_t* __restrict__ channels,
                               int64_t ncols, float2* __restrict__ out, int64_t nchans) {
  const int64_t n = npresent * ncols;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t p = i / ncols, k = i - p * ncols;
    const int64_t c = channels[k];
    if (c < 0 || c >= nchans) continue;   // dropped by the caller (an earlier copy of a repeated channel) or out of range
    int64_t row = (specnum[p] - start) % nspec;
    if (row < 0) row += nspec;
    out[row * nchans + c] = spec[p * spec_ld + k];
  }
}

// unpacking_gpu.cu:6-21: zero packet i's bytes when spec_num[i+1]-spec_num[i] != spectra_per_packet.
__global__ void dropped_kernel(uint8_t* data, const int64_t* spec_num, int64_t npackets, int64_t spp, int64_t bpp,
                               unsigned long long* ndropped) {
  // packets i = blockIdx.y + gridDim.y * m (any count), bytes of one packet
  // over the x dimension
  for (int64_t i = blockIdx.y; i < npackets - 1; i += gridDim.y) {
    if (spec_num[i + 1] - spec_num[i] == spp) continue;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(ndropped, 1ull);
    for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < bpp; j += int64_t(gridDim.x) * blockDim.x)
      data[i * bpp + j] = 0;
  }
}

}  // namespace alb

using namespace alb;

extern "C" {

int alb_unpack_4bit_float(const uint8_t* data, void* pol0, void* pol1, int64_t rowstart, int64_t rowend,
                          const int64_t* channels, int64_t ncols, int64_t nchan, void* stream) {
  ALB_REQUIRE(rowend >= rowstart && rowstart >= 0 && ncols >= 0 && nchan > 0, "alb_unpack_4bit_float: bad sizes");
  const int64_t nrows = rowend - rowstart;
  if (nrows * ncols == 0) return ALB_OK;
  ALB_REQUIRE(data && pol0 && pol1 && channels, "alb_unpack_4bit_float: null pointer");
  if (ncols % 2 == 0 && ncols / 2 <= 512 && aligned(data, 16) && (2 * nchan) % 16 == 0 && 2 * nchan * kRows4 <= 512 * 16 &&
      aligned(pol0, 16) && aligned(pol1, 16)) {
    hipLaunchKernelGGL(unpack4_lds_kernel, grid_rows(ncols / 2, nrows, 512), dim3(512), size_t(2 * nchan * kRows4),
                       as_stream(stream), data, (float4*)pol0, (float4*)pol1, rowstart, nrows, channels, ncols / 2, nchan);
    ALB_LAUNCH_CHECK("unpack4_lds_kernel");
    return ALB_OK;
  }
  if (ncols % 2 == 0 && aligned(data, 2) && aligned(pol0, 16) && aligned(pol1, 16) && ncols / 2 < (int64_t(1) << 31)) {
    hipLaunchKernelGGL(unpack4_pair_kernel, grid_rows(ncols / 2, nrows), dim3(256), 0, as_stream(stream), data,
                       (float4*)pol0, (float4*)pol1, rowstart, nrows, channels, ncols / 2, nchan,
                       aligned(data, 4) && nchan % 2 == 0);
    ALB_LAUNCH_CHECK("unpack4_pair_kernel");
    return ALB_OK;
  }
  ALB_REQUIRE(ncols < (int64_t(1) << 31), "alb_unpack_4bit_float: too many columns");
  hipLaunchKernelGGL(unpack4_kernel, grid_rows(ncols, nrows), dim3(256), 0, as_stream(stream), data, (float2*)pol0,
                     (float2*)pol1, rowstart, nrows, channels, ncols, nchan);
  ALB_LAUNCH_CHECK("unpack4_kernel");
  return ALB_OK;
}

int alb_unpack_1bit_float(const uint8_t* data, void* pol0, void* pol1, int64_t rowstart, int64_t rowend,
                          int64_t chanstart, int64_t chanend, int64_t nchan, void* stream) {
  ALB_REQUIRE(rowend >= rowstart && rowstart >= 0 && chanend >= chanstart && chanstart >= 0 && chanend <= nchan &&
                  nchan > 0,
              "alb_unpack_1bit_float: bad sizes (need 0 <= chanstart <= chanend <= nchan)");
  if ((rowend - rowstart) * (chanend - chanstart) == 0) return ALB_OK;
  ALB_REQUIRE(data && pol0 && pol1, "alb_unpack_1bit_float: null pointer");
  ALB_REQUIRE(chanstart % 2 == 0 && (chanend - chanstart) % 2 == 0,
              "alb_unpack_1bit_float: channel range must start even and hold an even count");
  const int64_t nrows = rowend - rowstart, nb = (chanend - chanstart) / 2;
  if (nrows * nb == 0) return ALB_OK;
  if (aligned(pol0, 16) && aligned(pol1, 16) && nb < (int64_t(1) << 31)) {
    hipLaunchKernelGGL(unpack1_rows_kernel, grid_rows(nb, nrows), dim3(256), 0, as_stream(stream), data, (float4*)pol0,
                       (float4*)pol1, rowstart, nrows, chanstart, nb, nchan);
    ALB_LAUNCH_CHECK("unpack1_rows_kernel");
    return ALB_OK;
  }
  hipLaunchKernelGGL(unpack1_kernel, grid_for(nrows * nb), dim3(256), 0, as_stream(stream), data, (float2*)pol0,
                     (float2*)pol1, rowstart, nrows, chanstart, nb, nchan);
  ALB_LAUNCH_CHECK("unpack1_kernel");
  return ALB_OK;
}

int alb_sortpols(const uint8_t* data, uint8_t* pol0, uint8_t* pol1, int64_t rowstart, int64_t rowend, int64_t ncols,
                 int64_t nchan, int bit_depth, int64_t chanstart, int64_t chanend, void* stream) {
  ALB_REQUIRE(chanend <= nchan, "ERROR: Trying to unpack more channels than present. Check chanend.");
  ALB_REQUIRE(rowend >= rowstart && rowstart >= 0 && chanend > chanstart && chanstart >= 0, "alb_sortpols: bad sizes");
  const int64_t nrows = rowend - rowstart;
  if (nrows == 0) return ALB_OK;
  ALB_REQUIRE(data && pol0 && pol1, "alb_sortpols: null pointer");
  if (bit_depth == 4) {
    ALB_REQUIRE(ncols == chanend - chanstart, "alb_sortpols: 4-bit ncols must equal chanend-chanstart");
    if (ncols % 4 == 0 && chanstart % 4 == 0 && nchan % 4 == 0 && aligned(data, 8) && aligned(pol0, 4) &&
        aligned(pol1, 4) && ncols / 4 < (int64_t(1) << 31)) {
      hipLaunchKernelGGL(sortpols4_quad_kernel, grid_rows(ncols / 4, nrows), dim3(256), 0, as_stream(stream), data,
                         (uint32_t*)pol0, (uint32_t*)pol1, rowstart, nrows, ncols / 4, nchan, chanstart);
      ALB_LAUNCH_CHECK("sortpols4_quad_kernel");
      return ALB_OK;
    }
    ALB_REQUIRE(ncols < (int64_t(1) << 31), "alb_sortpols: too many columns");
    hipLaunchKernelGGL(sortpols4_kernel, grid_rows(ncols, nrows), dim3(256), 0, as_stream(stream), data, pol0, pol1,
                       rowstart, nrows, ncols, nchan, chanstart);
  } else if (bit_depth == 1) {
    ALB_REQUIRE(chanstart % 2 == 0, "ERROR: Start channel index must be even.");
    ALB_REQUIRE(ncols == (chanend - chanstart + 3) / 4, "alb_sortpols: 1-bit ncols must be ceil((chanend-chanstart)/4)");
    ALB_REQUIRE(ncols < (int64_t(1) << 31), "alb_sortpols: too many columns");
    const int64_t bps = nchan / 2;
    if ((chanend - chanstart) % 4 == 0 && ncols % 4 == 0 && bps % 4 == 0 && (chanstart / 2) % 4 == 0 &&
        aligned(data, 4) && aligned(pol0, 4) && aligned(pol1, 4)) {
      const int64_t nquad = ncols / 4;
      int64_t gy = (nrows + 3) / 4;
      if (gy > 4096) gy = 4096;
      hipLaunchKernelGGL(sortpols1_quad_kernel, dim3(unsigned((nquad + 63) / 64), unsigned(gy)), dim3(64, 4), 0,
                         as_stream(stream), data, (uint32_t*)pol0, (uint32_t*)pol1, rowstart, nrows, nquad, nchan,
                         chanstart);
      ALB_LAUNCH_CHECK("sortpols1_quad_kernel");
      return ALB_OK;
    }
    hipLaunchKernelGGL(sortpols1_kernel, grid_rows(ncols, nrows), dim3(256), 0, as_stream(stream), data, pol0, pol1,
                       rowstart, nrows, ncols, nchan, chanstart, int((chanend - chanstart) % 4));
  } else {
    ALB_REQUIRE(false, "alb_sortpols: bit_depth must be 1 or 4 (got %d)", bit_depth);
  }
  ALB_LAUNCH_CHECK("sortpols_kernel");
  return ALB_OK;
}

int alb_hist(const uint8_t* data, uint64_t* hist, int64_t rowstart, int64_t rowend, int64_t nchan, int bit_depth,
             int mode, void* stream) {
  ALB_REQUIRE(data && hist, "alb_hist: null pointer");
  ALB_REQUIRE(bit_depth == 1 || bit_depth == 4, "bit_depth set to %d, only takes values 1 or 4", bit_depth);
  ALB_REQUIRE(mode == 0 || mode == 1 || mode == -1, "mode set to %d, only takes values 0,1,-1", mode);
  ALB_REQUIRE(rowend >= rowstart && rowstart >= 0 && nchan > 0, "alb_hist: bad sizes");
  const int64_t nbins = bit_depth == 4 ? 16 : 2;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(fill_u64, grid_for(nbins * nchan), dim3(256), 0, s, hist, nbins * nchan);
  const int64_t nrows = rowend - rowstart;
  if (nrows > 0) {
    const int64_t rows_per = 256;
    const unsigned gy = unsigned((nrows + rows_per - 1) / rows_per);
    if (bit_depth == 4) {
      hipLaunchKernelGGL(hist4_kernel, dim3(unsigned((nchan + 127) / 128), gy), dim3(128), 0, s, data,
                         (unsigned long long*)hist, rowstart, nrows, nchan, mode, rows_per);
    } else {
      ALB_REQUIRE(nchan % 2 == 0, "alb_hist: 1-bit nchan must be even");
      hipLaunchKernelGGL(hist1_kernel, dim3(unsigned((nchan / 2 + 127) / 128), gy), dim3(128), 0, s, data,
                         (unsigned long long*)hist, rowstart, nrows, nchan, mode, rows_per);
    }
  }
  ALB_LAUNCH_CHECK("hist_kernel");
  return ALB_OK;
}

int alb_make_continuous(const void* spec, int64_t spec_ld, const int64_t* specnum, int64_t npresent, int64_t start,
                        int64_t nspec, const int64_t* channels, int64_t ncols, void* out, int64_t nchans, void* stream) {
  ALB_REQUIRE(out && (npresent == 0 || (spec && specnum && channels)), "alb_make_continuous: null pointer");
  ALB_REQUIRE(nspec > 0 && nchans > 0 && npresent >= 0 && ncols >= 0 && spec_ld >= ncols,
              "alb_make_continuous: bad sizes");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(fill_c64, grid_for(nspec * nchans), dim3(256), 0, s, (float2*)out, nspec * nchans);
  if (npresent * ncols > 0)
    hipLaunchKernelGGL(scatter_kernel, grid_for(npresent * ncols), dim3(256), 0, s, (const float2*)spec, spec_ld,
                       specnum, npresent, start, nspec, channels, ncols, (float2*)out, nchans);
  ALB_LAUNCH_CHECK("scatter_kernel");
  return ALB_OK;
}

int alb_dropped_packets(uint8_t* data, const int64_t* spec_num, int64_t num_packets, int64_t spectra_per_packet,
                        int64_t nchan, int bit_depth, uint64_t* num_dropped, void* stream) {
  ALB_REQUIRE(data && spec_num && num_dropped, "alb_dropped_packets: null pointer");
  int64_t bpp;
  if (bit_depth == 4) bpp = 2 * nchan * spectra_per_packet;
  else if (bit_depth == 2) bpp = nchan * spectra_per_packet;
  else if (bit_depth == 1) bpp = nchan * spectra_per_packet / 2;
  else { set_error("alb_dropped_packets: bit_depth must be 1, 2 or 4"); return ALB_EINVAL; }
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, s, num_dropped, int64_t(1));
  if (num_packets >= 2) {
    ALB_REQUIRE(bpp > 0, "alb_dropped_packets: empty packets");
    const int64_t gy = num_packets - 1 < 65535 ? num_packets - 1 : 65535;
    hipLaunchKernelGGL(dropped_kernel, dim3(unsigned((bpp + 255) / 256 < 64 ? (bpp + 255) / 256 : 64), unsigned(gy)),
                       dim3(256), 0, s, data, spec_num, num_packets, spectra_per_packet, bpp,
                       (unsigned long long*)num_dropped);
  }
  ALB_LAUNCH_CHECK("dropped_kernel");
  return ALB_OK;
}

int alb_fill_zero(void* p, int64_t nbytes, void* stream) {
  ALB_REQUIRE(p || nbytes == 0, "alb_fill_zero: null pointer");
  if (nbytes > 0)
    hipLaunchKernelGGL(fill_u8, grid_for(nbytes), dim3(256), 0, as_stream(stream), (uint8_t*)p, nbytes, uint8_t(0));
  ALB_LAUNCH_CHECK("fill_u8");
  return ALB_OK;
}

}  // extern "C"
