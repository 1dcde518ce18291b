// m2l_f64.h -- the plain form's far field (M2L) as dense FP64 matrix
// products on MFMA (v_mfma_f64_16x16x4_f64).
//
// The operator of an offset code is the real matrix O (ro x ki) of the
// rotate / z-translate / rotate-back M2L on scaled expansions (the same
// matrix the BF16 forms split into parts; built in FP64 by k_m2l_dense):
//   inputs  (ki = 2 mc - p):     Re M^_l^m for all (l, m), then Im M^_l^m for
//                                m >= 1 (l = 0..p-1, tri order);
//   outputs (ro = 2 lc - p - 1): Re L^_l^m for all (l, m), then Im L^_l^m for
//                                m >= 1 (l = 0..p).
// Products and sums are FP64 throughout; a target's sum over one fixed range
// of offset slots is accumulated in a fixed order, and the ranges are added
// in range order, so a result does not depend on the rest of the batch.
//
// k order.  An MFMA step contracts 4 k values, one per quarter kq of the
// lanes; the kernel walks steps of 8 k (two MFMAs) and gives quarter kq the
// k values kq * kp + j (j < kp = 2 ks2), so that a lane's source entries of
// consecutive steps are consecutive in memory (one 16-byte load per step).
#pragma once

#include <hip/hip_runtime.h>

#include "engine.h"

namespace aran {

struct F64Dims {
  int mc, lc, ki, ro;
  int rb;   // 16-row blocks of the result
  int rs;   // row splits (waves that share a tile's columns)
  int rbw;  // row blocks per wave
  int cbw;  // 16-column blocks per wave
  int wpt;  // waves per 64-target tile
  int tpw;  // tiles per workgroup of 8 waves
  int ks2;  // steps of 8 k
  int kp;   // k positions per quarter (2 ks2)
};
constexpr __host__ __device__ F64Dims f64_dims(int p) {
  const int mc = p * (p + 1) / 2, lc = (p + 1) * (p + 2) / 2;
  const int ki = 2 * mc - p, ro = 2 * lc - p - 1;
  const int rb = (ro + 15) / 16;
  const int rs = rb > 8 ? 2 : 1;
  const int rbw = (rb + rs - 1) / rs;
  const int cbw = rbw <= 4 ? 4 : 2;
  const int wpt = rs * (4 / cbw);
  const int ks2 = (ki + 7) / 8;
  return F64Dims{mc, lc, ki, ro, rb, rs, rbw, cbw, wpt, 8 / wpt, ks2, 2 * ks2};
}
// doubles of the operator table of one code ([ks2][2][rs rbw][64]) and of
// the packed multipole of one node ([4][kp])
constexpr __host__ __device__ long f64_ops_doubles(int p) {
  return (long)f64_dims(p).ks2 * 2 * f64_dims(p).rs * f64_dims(p).rbw * 64;
}
constexpr __host__ __device__ long f64_mult_doubles(int p) {
  return 4L * f64_dims(p).kp;
}
// Tiles per workgroup of the FP64 far field at order p.
constexpr __host__ __device__ int f64_wg_tiles(int p) {
  return f64_dims(p).tpw;
}

// Dense FP64 operators O[343][ro][ki] (launch_m2l_dense) to the A-operand
// layout [code][step][half][row block][lane] (lane: row lane % 16 of the
// block, quarter lane / 16).
hipError_t launch_f64_pack_ops(int p, const double *O, double *ops,
                               hipStream_t s);
// Multipoles of every node, scaled by w^-l, to [node][quarter][kp], followed
// by one all-zero node (index nnodes) that stands in for a missing source.
hipError_t launch_f64_pack_mult(const NodeArrays &nd, int nnodes, int p,
                                const double2 *M, double *mb, hipStream_t s);
// The tiled FP64 far field: wg_tiles[nwg][f64_wg_tiles(p)] lists the tiles
// of each workgroup (one (level, octant) group, -1 padding); Lsm receives
// [split][tile][ro][64] FP64 partial scaled locals.
hipError_t launch_m2l_f64(int p, int nwg, const int *wg_tiles, int ntiles,
                          const int *tile_src, const int *tile_oct,
                          const short *oct_codes, const double *ops,
                          const double *mb, int zero_node, int nsplit,
                          int sched, double *Lsm, hipStream_t s);
// Offset ranges per tile (fixed per order; ARAN_M2L_SPLIT overrides it).
constexpr int kFarRangesF64 = 7;
int m2l_f64_split(int p);
// Whether the plain form takes this far field at order p (the per-pair
// rotate kernel otherwise); ARAN_PLAIN_FAR=rotate turns it off.
bool m2l_f64_enabled(int p);

}  // namespace aran
