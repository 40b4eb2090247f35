// Device helpers shared by the MFMA GEMM kernels (FP64/FP32 and BF16 hi/lo).
#pragma once
#include <hip/hip_runtime.h>

namespace ecg {

// Offset of index x split as x = x0 + d0*x1 with strides s0, s1.
__device__ __forceinline__ long split_off(int x, int d0, long s0, long s1) {
  int x1 = x / d0;
  int x0 = x - x1 * d0;
  return (long)x0 * s0 + (long)x1 * s1;
}

// Workgroup id -> (tile_m, tile_n) for a grid of tiles_m * tiles_n workgroups.
// Consecutive hardware ids are dealt round-robin to the 8 XCDs; renumber so each
// XCD receives a contiguous range of logical tiles, then walk tiles in groups
// of 8 M-tiles so neighbours on one XCD share A and B panels in that XCD's L2.
__device__ __forceinline__ void tile_coords(int wg, int nwg, int tiles_m, int& tm, int& tn) {
  const int nx = 8;
  int q = nwg / nx, r = nwg % nx;
  int xcd = wg % nx, idx = wg / nx;
  int logical = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + idx;
  const int G = 8;
  int tn_count = nwg / tiles_m;
  int gid = logical / (G * tn_count);
  int first_m = gid * G;
  int gsize = min(tiles_m - first_m, G);
  int in = logical - gid * G * tn_count;
  tm = first_m + in % gsize;
  tn = in / gsize;
}

}  // namespace ecg
