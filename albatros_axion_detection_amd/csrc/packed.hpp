// 4-bit packed rows of the streaming inverse PFB (shared by the row passes).
#pragma once
#include "common.hpp"

namespace alb {

// 4-bit packed rows of an iPFB tile.  Tile t, row r < 2*cut comes from chunk
// t-1 row acclen-2*cut+r (zeros for t == 0), row r >= 2*cut from chunk t row
// r-2*cut.  rowmap[chunk][row] = row p of the packed byte array [rows][ncols]
// holding that spectrum, or -1 (zero-filled spectrum); byte (p, col) goes to
// channel number chan[col] (baseband_data_classes.py:48-73 zero fill,
// helper_gpu_stream.py:195-197 carried rows).
struct Packed4Rows {
  const uint8_t* bytes;
  const int* rowmap;
  const int* chan;
  int64_t ncols, acclen, cut2, tile0;
};

// 4-bit two's complement nibble with the reference's 8 -> +8 rule (unpacking.c:186-202).
__device__ __forceinline__ int nib4f(int v) { return v > 8 ? v - 16 : v; }

// Packed row holding row `row` of the tile sequence (rows of nslice per tile), or -1.
__device__ __forceinline__ int64_t packed_row(const Packed4Rows& pk, int64_t row, int64_t nslice) {
  const int64_t tile = row / nslice, r = row - tile * nslice;
  int64_t chunk, crow;
  if (r < pk.cut2) {
    chunk = pk.tile0 + tile - 1;
    crow = pk.acclen - pk.cut2 + r;
  } else {
    chunk = pk.tile0 + tile;
    crow = r - pk.cut2;
  }
  if (chunk < 0) return -1;
  return pk.rowmap[chunk * pk.acclen + crow];
}

}  // namespace alb
