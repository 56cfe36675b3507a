// Row pass of the streaming inverse PFB on the matrix cores.
//
// A tile row of the inverse PFB is irfft_4096 of a 2049-channel spectrum of
// which only the ncols packed channels are non-zero (pfb_gpu_utils.py:80 on the
// zero-filled spectrum of baseband_data_classes.py:48-73), so
//   y[k] = (1/4096) sum_c ( re_c a_c(k) + im_c b_c(k) ),
//   a_c(k) = 2 cos(2 pi ch k / 4096), b_c(k) = -2 sin(2 pi ch k / 4096)
// (ch = chan[c]; ch = 0 gives a = 1, b = 0 and ch = 2048 gives a = (-1)^k,
// b = 0: the imaginary parts of DC and Nyquist are ignored, pycufft.py:139-161).
// For a band of ncols channels that is a GEMM  Y[rows][4096] = X[rows][2 ncols]
// T[2 ncols][4096]: X holds the decoded 4-bit values (exact in FP16), T the
// coefficients split T = T_hi + T_lo in FP16 (about 22 significant bits), and
// v_mfma_f32_16x16x32_f16 accumulates X T_hi + X T_lo in FP32.
//
// Output: the iPFB tile in the blocked layout the column pass reads as one
// contiguous strip: element (row i, column k) of tile t at
//   out[t][k / 8][i][k % 8].
#include "packed.hpp"

namespace alb {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

constexpr int kRowsBlock = 8;  // columns per block of the tile layout

// Coefficient table in B-operand order: [ks][cb][hl][lane][8] halfs, where
// column = 16 cb + lane % 16 and k index kk = 32 ks + 8 (lane / 16) + i stands
// for channel c = kk / 2, real (kk even) or imaginary (odd) part; hl = 0 high
// part, 1 low part.  Built in float64 from the channel numbers.
__global__ void rows_coeff_kernel(const int* __restrict__ chan, int64_t ncols, int64_t nks, _Float16* __restrict__ T) {
  const int64_t n = nks * 256 * 2 * 64 * 8;
  for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < n; idx += int64_t(gridDim.x) * blockDim.x) {
    const int i = int(idx % 8);
    const int l = int((idx / 8) % 64);
    const int hl = int((idx / 512) % 2);
    const int64_t cb = (idx / 1024) % 256;
    const int64_t ks = idx / (1024 * 256);
    const int64_t col = 16 * cb + l % 16;
    const int64_t kk = 32 * ks + 8 * (l / 16) + i;
    const int64_t c = kk / 2;
    const int part = int(kk % 2);
    double v = 0.0;
    // a repeated channel number keeps only its last column (numpy's fancy
    // assignment in baseband_data_classes.py:72)
    bool last = c < ncols && chan[c] >= 0 && chan[c] <= 2048;
    for (int64_t c2 = c + 1; last && c2 < ncols; ++c2) last = chan[c2] != chan[c];
    if (last) {
      const int64_t ch = chan[c];
      if (ch == 0) {
        v = part ? 0.0 : 1.0;
      } else if (ch == 2048) {
        v = part ? 0.0 : ((col & 1) ? -1.0 : 1.0);
      } else {
        const int64_t m = (ch * col) % 4096;
        double sn, cs;
        sincospi(2.0 * double(m) / 4096.0, &sn, &cs);
        v = part ? -2.0 * sn : 2.0 * cs;
      }
    }
    const _Float16 hi = (_Float16)v;
    T[idx] = hl ? (_Float16)(v - (double)hi) : hi;
  }
}

// Workgroup: 4 waves, 256 rows x 64 columns of the tile sequence.  The
// workgroup first stages the packed bytes of its 256 rows in LDS (zero rows
// for lost or carried-in-from-nothing spectra, zero columns past ncols); wave
// w then takes rows 64 w .. 64 w + 63 (four 16-row tiles) against all four
// 16-column tiles, so each B operand (read once per wave from the table, the
// four waves sharing it through L1) feeds four row tiles.  Workgroup i runs on
// XCD i % 8; each XCD owns 8 of the 64 column tiles (its eighth of T stays in
// its L2) and walks the row tiles.
constexpr int kRowsPerWG = 256;

__global__ __launch_bounds__(256) void rows_mfma_kernel(Packed4Rows pk, int64_t nrows, int64_t nslice, int64_t nks,
                                                       const halfx8* __restrict__ T, float* __restrict__ out) {
  // LDS: [256][ldw] words of packed bytes, [256] packed rows, [256] decode table
  extern __shared__ uint32_t a_lds[];
  const int ldw = int(4 * nks);                // words per row: 16 nks bytes
  int64_t* p_lds = reinterpret_cast<int64_t*>(a_lds + kRowsPerWG * ldw);
  uint32_t* lut = reinterpret_cast<uint32_t*>(p_lds + kRowsPerWG);
  const int tid = threadIdx.x;
  const int wv = tid / 64, lane = tid % 64;
  const int xcd = blockIdx.x % 8;
  const int64_t slot = blockIdx.x / 8;
  const int64_t ct = xcd * 8 + slot % 8;       // column tile (64 columns)
  const int64_t rt = slot / 8;                 // row tile (256 rows)
  const int64_t row0 = rt * kRowsPerWG;
  // stage the bytes of this row tile (word w of row r: bytes 4w .. 4w+3);
  // consecutive threads read consecutive words of a row
  {
    const int64_t row = row0 + tid;
    p_lds[tid] = row < nrows ? packed_row(pk, row, nslice) : -1;
    // byte -> (real, imaginary) as two packed FP16 values
    const _Float16 re = (_Float16)nib4f(tid >> 4), im = (_Float16)nib4f(tid & 15);
    lut[tid] = uint32_t(__builtin_bit_cast(uint16_t, re)) | (uint32_t(__builtin_bit_cast(uint16_t, im)) << 16);
  }
  __syncthreads();
  const bool aligned = (pk.ncols % 4 == 0) && ((reinterpret_cast<uintptr_t>(pk.bytes) & 3) == 0);
  for (int idx = tid; idx < kRowsPerWG * ldw; idx += 256) {
    const int rr = idx / ldw, w = idx - rr * ldw;
    const int64_t p = p_lds[rr];
    uint32_t v = 0;
    if (p >= 0 && 4 * w < pk.ncols) {
      const uint8_t* src = pk.bytes + p * pk.ncols + 4 * w;
      if (aligned) {
        v = *reinterpret_cast<const uint32_t*>(src);
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (4 * w + q < pk.ncols) v |= uint32_t(src[q]) << (8 * q);
      }
    }
    a_lds[idx] = v;
  }
  __syncthreads();
  const int r = lane % 16, g = lane / 16;
  floatx4 acc[4][4];
#pragma unroll
  for (int rs = 0; rs < 4; ++rs)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[rs][nt] = floatx4{0.f, 0.f, 0.f, 0.f};
  const uint32_t* arow = a_lds + (64 * wv + r) * ldw + g;
  const halfx8* tcol = T + (ct * 4) * 2 * 64 + lane;
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  halfx8 bn[2][4];                              // B operands of the next step (prefetched)
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    bn[0][nt] = tcol[nt * 2 * 64];
    bn[1][nt] = tcol[nt * 2 * 64 + 64];
  }
  for (int64_t ks = 0; ks < nks; ++ks) {
    halfx8 b[2][4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      b[0][nt] = bn[0][nt];
      b[1][nt] = bn[1][nt];
    }
    if (ks + 1 < nks) {
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const halfx8* tb = tcol + ((ks + 1) * 256 + nt) * 2 * 64;
        bn[0][nt] = tb[0];
        bn[1][nt] = tb[64];
      }
    }
    // A: channels 16 ks + 4 g + q (q < 4) of rows 64 wv + 16 rs + r, decoded
    // through the byte table
    halfx8 a[4];
#pragma unroll
    for (int rs = 0; rs < 4; ++rs) {
      const uint32_t w4 = arow[16 * rs * ldw + 4 * ks];
      const u32x4 d = {lut[w4 & 0xFFu], lut[(w4 >> 8) & 0xFFu], lut[(w4 >> 16) & 0xFFu], lut[w4 >> 24]};
      a[rs] = __builtin_bit_cast(halfx8, d);
    }
    // the 16 high-part products, then the 16 low-part ones: no back-to-back
    // dependence on one accumulator
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int rs = 0; rs < 4; ++rs)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
          acc[rs][nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[rs], b[h][nt], acc[rs][nt], 0, 0, 0);
  }
  // D layout: lane l holds D[4 (l / 16) + v][l % 16]
  const float sc = 1.0f / 4096.0f;
  // rows of a lane are consecutive, so the tile / row split is stepped, not
  // divided per store
  const int64_t kb = ct * 64 + r;                     // column of tile nt = 0
#pragma unroll
  for (int rs = 0; rs < 4; ++rs) {
    const int64_t orow0 = row0 + 64 * wv + 16 * rs + 4 * g;
    int64_t t = orow0 / nslice, i = orow0 - t * nslice;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (orow0 + v < nrows) {
        float* o = out + t * 4096 * nslice + i * kRowsBlock;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
          const int64_t k = kb + 16 * nt;
          o[(k / kRowsBlock) * (kRowsBlock * nslice) + (k % kRowsBlock)] = acc[rs][nt][v] * sc;
        }
      }
      if (++i == nslice) { i = 0; ++t; }
    }
  }
}

}  // namespace alb

using namespace alb;

extern "C" {

// Bytes of the coefficient table of alb_ipfb_rows_packed4_mfma for ncols channels.
int64_t alb_ipfb_rows_coeff_bytes(int64_t ncols) {
  const int64_t nks = (ncols + 15) / 16;
  return nks * 256 * 2 * 64 * 8 * int64_t(sizeof(_Float16));
}

// Build the coefficient table from the channel numbers chan[ncols] (< 2049).
int alb_ipfb_rows_coeff(const int* chan, int64_t ncols, void* coeff, void* stream) {
  ALB_REQUIRE(chan && coeff, "alb_ipfb_rows_coeff: null pointer");
  ALB_REQUIRE(ncols > 0 && ncols <= 2049, "alb_ipfb_rows_coeff: bad ncols");
  const int64_t nks = (ncols + 15) / 16;
  const int64_t n = nks * 256 * 2 * 64 * 8;
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(rows_coeff_kernel, dim3(unsigned(g)), dim3(256), 0, as_stream(stream), chan, ncols, nks,
                     (_Float16*)coeff);
  ALB_LAUNCH_CHECK("rows_coeff_kernel");
  return ALB_OK;
}

// Row pass of ntiles iPFB tiles from 4-bit packed bytes (the arguments of
// alb_ipfb_rows_packed4) on the matrix cores, into the blocked tile layout
// out[ntiles][4096/8][nslice][8] float32.
int alb_ipfb_rows_packed4_mfma(const uint8_t* bytes, const int* rowmap, const int* chan, int64_t ncols, int64_t acclen,
                               int cut, int64_t tile0, int64_t ntiles, const void* coeff, float* out, void* stream) {
  ALB_REQUIRE(bytes && rowmap && chan && coeff && out, "alb_ipfb_rows_packed4_mfma: null pointer");
  ALB_REQUIRE(acclen >= 2 * cut && cut >= 0 && ncols > 0 && ncols <= 2049 && ntiles >= 0 && tile0 >= 0,
              "alb_ipfb_rows_packed4_mfma: bad sizes");
  if (ntiles == 0) return ALB_OK;
  const int64_t nslice = acclen + 2 * int64_t(cut);
  const int64_t nrows = nslice * ntiles;
  const int64_t nks = (ncols + 15) / 16;
  const int64_t nrt = (nrows + kRowsPerWG - 1) / kRowsPerWG;
  ALB_REQUIRE(nrt * 64 < (int64_t(1) << 31), "alb_ipfb_rows_packed4_mfma: too many rows");
  size_t lds = sizeof(uint32_t) * size_t(kRowsPerWG) * size_t(4 * nks) + sizeof(int64_t) * kRowsPerWG +
               sizeof(uint32_t) * 256;
  ALB_REQUIRE(lds <= 160 * 1024, "alb_ipfb_rows_packed4_mfma: too many channels for the LDS stage");
  Packed4Rows pk{bytes, rowmap, chan, ncols, acclen, 2 * int64_t(cut), tile0};
  hipLaunchKernelGGL(rows_mfma_kernel, dim3(unsigned(nrt * 64)), dim3(256), lds, as_stream(stream), pk, nrows, nslice,
                     nks, (const halfx8*)coeff, out);
  ALB_LAUNCH_CHECK("rows_mfma_kernel");
  return ALB_OK;
}

}  // extern "C"
