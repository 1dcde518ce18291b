// tree_rules.h -- the octree rules shared by every tree kernel.
//
// LibVsg's VsgPRTree3d (the reference's tree, not in its source tree) is
// replaced by these rules; the FP64 oracle implements the same ones:
//   * integer coordinate i = floor((x - lo) * 2^21 / width), clamped to
//     [0, 2^21 - 1]; (x - lo) * scale is a subtraction then a product, so no
//     contraction can change it between host and device;
//   * 63-bit Morton key, octant digit = bx | by << 1 | bz << 2, most
//     significant digit first;
//   * a node splits while it holds more than leafcap particles and its depth
//     is below 21; empty octants create no node;
//   * the root is a cube centred on the caller's bounds [lb, ub], of half
//     width that of their widest side (1 when that is 0), doubled until it
//     holds every particle (root_cube).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace aran {

// ints per node in NodeArrays::anc (depths 0..21)
constexpr int kAncRow = 22;

struct NodeArrays {
  int *depth, *ix, *iy, *iz, *pbegin, *pend, *parent, *child, *nchild, *leaf;
  int *sys;           // system (independent particle set) of the node
  int *anc;           // [node][kAncRow]: ancestor at depth 0..depth (itself)
  int *cmap;          // [node][8]: child in each octant, -1: none
  uint64_t *key;
  double *cx, *cy, *cz;
  const double *box;  // per system: lo x, lo y, lo z, width of the root cube
};

// Width of a node's box.
__host__ __device__ inline double node_width(const NodeArrays &nd, int node) {
  return ldexp(nd.box[4 * nd.sys[node] + 3], -nd.depth[node]);
}

__host__ __device__ inline uint64_t spread3(uint64_t v) {
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x1f00000000ffffull;
  v = (v | (v << 16)) & 0x1f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

__host__ __device__ inline uint64_t morton_key(uint64_t ix, uint64_t iy,
                                               uint64_t iz) {
  return spread3(ix) | (spread3(iy) << 1) | (spread3(iz) << 2);
}

__host__ __device__ inline uint64_t icoord(double x, double lo, double scale) {
  double f = floor((x - lo) * scale);
  const double top = 2097151.0;  // 2^21 - 1
  f = f < 0. ? 0. : f;
  f = f > top ? top : f;
  return (uint64_t)f;
}

// Root cube (lo corner, width) of a system with bounds [lb, ub] whose
// particles span [pmin, pmax] (finite).
__host__ __device__ inline void root_cube(const double *lb, const double *ub,
                                          const double *pmin,
                                          const double *pmax, double *lo,
                                          double *width) {
  double c[3], h = 0.;
  for (int k = 0; k < 3; k++) {
    c[k] = 0.5 * (lb[k] + ub[k]);
    if (0.5 * (ub[k] - lb[k]) > h) h = 0.5 * (ub[k] - lb[k]);
  }
  if (!(h > 0.)) h = 1.;
  for (;;) {
    bool inside = true;
    for (int k = 0; k < 3; k++)
      if (pmin[k] < c[k] - h || pmax[k] > c[k] + h) inside = false;
    if (inside) break;
    h *= 2.;
  }
  for (int k = 0; k < 3; k++) lo[k] = c[k] - h;
  *width = 2. * h;
}

}  // namespace aran
