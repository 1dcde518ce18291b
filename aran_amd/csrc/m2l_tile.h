// m2l_tile.h -- launch interface of the tiled far-field kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "engine.h"

namespace aran {

// Offsets (target minus source, box units) a target can see in its far
// list, for one octant of its parent: 6^3 - 3^3 = 189.
constexpr int kOctOffsets = 189;

struct M2LTileArgs {
  const int *tile_tgt;    // [ntiles * 64] target node or -1
  const int *tile_src;    // [ntiles * 189 * 64] source node or -1
  const int *tile_oct;    // [ntiles] octant shared by the tile
  const short *oct_codes; // [8 * 189] offset code, grouped by theta
  const DirEntry *dir;    // [343]
  const void *ephi;       // [343][2(p+1)]   e^{i m phi} as T
  const void *rot;        // [ntheta][4][rot_block(p)] as T
  const void *vfac;       // [343][(p+1)(p+1)p] scaled vertical factors
  int nsplit;             // waves per tile (offset ranges), >= 1
  long split_stride;      // elements between partial-local slices
};

// Octant offset lists (host): codes[8][189] sorted by theta table index,
// code_to_k[8][343] position of a code in its octant list or -1.
void octant_offsets(const std::vector<DirEntry> &dirs, std::vector<short> &codes,
                    std::vector<short> &code_to_k);

hipError_t launch_m2l_tile(int precision, int p, int ntiles,
                           const M2LTileArgs &a, const void *Ms, void *Ls,
                           hipStream_t s);
hipError_t launch_tile_count(const NodeArrays &nd, int first, int count,
                             int *cnt, hipStream_t s);
hipError_t launch_tile_fill(const NodeArrays &nd, int first, int count,
                            const int *group_base, int *fill, int *tile_tgt,
                            int *node_slot, hipStream_t s);
hipError_t launch_tile_src(const NodeArrays &nd, int first, int count,
                           const int *vofs, const int *vsrc, const short *vcode,
                           const short *code_to_k, const int *node_slot,
                           int *tile_src, hipStream_t s);
hipError_t launch_fill_int(int *p, long n, int v, hipStream_t s);
hipError_t launch_scale_mult(int precision, const NodeArrays &nd, int nnodes,
                             int p, const double2 *M, void *Ms, hipStream_t s);
hipError_t launch_unscale_local(int precision, const NodeArrays &nd, int nnodes,
                                int p, const void *Ls, int nsplit,
                                long split_stride, double2 *L, hipStream_t s);
hipError_t launch_m2l_tables(int precision, int p, const DirEntry *dir,
                             const double *tm2l, const double *rot, long nrot,
                             const double2 *eiphi, void *vfac, void *rotT,
                             void *ephi, hipStream_t s);

}  // namespace aran
