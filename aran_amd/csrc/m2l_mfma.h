// m2l_mfma.h -- the far field (M2L) as dense BF16 matrix products on MFMA.
//
// For every offset code the rotate / z-translate / rotate-back M2L on
// scaled expansions is one real matrix O (ro x ki), the same at every level:
//   inputs  (ki = 2 mc - p):     Re M^_l^m for all (l, m), then Im M^_l^m for
//                                m >= 1 (l = 0..p-1, tri order);
//   outputs (ro = 2 lc - p - 1): Re L^_l^m for all (l, m), then Im L^_l^m for
//                                m >= 1 (l = 0..p).
// A tile of 64 same-octant targets walks its 189 candidate offsets; for
// each, the 64 source multipoles form the B operand (ki x 64) and O the A
// operand, accumulated in FP32 by v_mfma_f32_32x32x16_bf16.  Operands are
// split into BF16 parts x = h + m + l (h = bf16(x), m = bf16(x - h),
// l = bf16(x - h - m)); the products kept are
//   3 products: h h + h m + m h            (BF16x3)
//   6 products: + h l + l h + m m          (BF16x6, FP32-level accuracy).
#pragma once

#include <hip/hip_runtime.h>

#include "engine.h"

namespace aran {

struct MfmaDims {
  int mc, lc, ki, ro, ks, rb;  // ks: 16-wide k steps, rb: 32-row blocks
};
constexpr __host__ __device__ MfmaDims mfma_dims(int p) {
  const int mc = p * (p + 1) / 2, lc = (p + 1) * (p + 2) / 2;
  const int ki = 2 * mc - p, ro = 2 * lc - p - 1;
  return MfmaDims{mc, lc, ki, ro, (ki + 15) / 16, (ro + 31) / 32};
}
// Column blocks (32 targets each) per wave: 2 while the accumulators of a
// whole 64-target tile fit (rb <= 4, p <= 10), else 1 (two waves per tile).
constexpr __host__ __device__ int mfma_col_blocks(int p) {
  return mfma_dims(p).rb <= 4 ? 2 : 1;
}

// 16-byte units of the operator table of one code and of the packed
// multipole of one node (three BF16 parts of 8 values each per lane slot).
constexpr __host__ __device__ long mfma_ops_units(int p) {
  return (long)mfma_dims(p).rb * mfma_dims(p).ks * 64 * 3;
}
constexpr __host__ __device__ long mfma_mult_units(int p) {
  return (long)mfma_dims(p).ks * 2 * 3;
}

// Tables: dense FP64 operators O[343][ro][ki] (from launch_m2l_dense) to the
// A-operand layout [code][ks][rb][part][lane][8 x bf16].
hipError_t launch_mfma_pack_ops(int p, const double *O, void *ops,
                                hipStream_t s);
// Multipoles of every node, scaled by w^-l, to the B-operand layout
// [node][ks][2][part][8 x bf16] with the nparts parts the form reads (2 for
// three products, 3 for six), followed by one all-zero node (index nnodes)
// that stands in for a missing source.
hipError_t launch_mfma_pack_mult(const NodeArrays &nd, int nnodes, int p,
                                 int nparts, const double2 *M, void *mb,
                                 hipStream_t s);
// The tiled far field: nprod 3 or 6; wg_tiles[nwg][8 / (2 / col blocks)]
// lists the tiles of each workgroup (one (level, octant) group, Morton
// order, -1 padding); Lsm receives [split][tile][ro][64] FP32 partial
// scaled locals; sched picks the workgroup order (schedule 1 expects an
// octant-major wg_tiles list).
hipError_t launch_m2l_mfma(int nprod, int p, int nwg, const int *wg_tiles,
                           int ntiles, const int *tile_src,
                           const int *tile_oct, const short *oct_codes,
                           const void *ops, const void *mb, int zero_node,
                           int nsplit, int sched, float *Lsm, hipStream_t s);
// Tiles per workgroup of the MFMA far field at order p.
constexpr __host__ __device__ int mfma_wg_tiles(int p) {
  return 8 / (2 / mfma_col_blocks(p));
}
// L of the tiled targets = (sum of the nsplit slices, in slice order) /
// w^(l+1); rows of nodes [0, first) are zeroed.
hipError_t launch_unscale_mfma(const NodeArrays &nd, int p, int first,
                               int ntiles, const int *tile_tgt,
                               const float *Lsm, int nsplit, double2 *L,
                               hipStream_t s);
// The same from FP64 slices (the plain form's dense far field).
hipError_t launch_unscale_mfma(const NodeArrays &nd, int p, int first,
                               int ntiles, const int *tile_tgt,
                               const double *Lsm, int nsplit, double2 *L,
                               hipStream_t s);
// Offset ranges of the MFMA far field (fixed per order: a target's sums do
// not depend on the rest of the solve).  A workgroup takes one range of 8
// tiles and holds its CU, so the count sets how evenly the work fills the
// 256 CUs; 5 measured fastest at N = 2^20 (Plummer, p = 10: 690 workgroups,
// 2.7 rounds, against 3.2 for 6) for both BF16 forms.
constexpr int kFarRangesMfma = 5;
int m2l_mfma_split(int p);
// Workgroup schedules of the MFMA far field (same sums, different order of
// work over the XCDs): 0 space-major, 1 offset-range-major.
constexpr int kMfmaSchedules = 2;
// Schedule used without the autotuner: `fallback` unless ARAN_MFMA_SCHED
// overrides it (BF16 forms: 0, measured best; FP64: 1, its operators are
// four times the BF16 bytes and gain most from sharing an XCD's L2).
int mfma_default_schedule(int fallback);

}  // namespace aran
