// m2l_tile.h -- launch interface of the tiled far-field kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "engine.h"

namespace aran {

// Offsets (target minus source, box units) a target can see in its far
// list, for one octant of its parent: 6^3 - 3^3 = 189.
constexpr int kOctOffsets = 189;

// Tiles (one wave each) per workgroup of the tiled far-field kernel.
constexpr int kTileWaves = 8;

// Waves per SIMD of the tiled far-field kernel of order p: two while the
// expansions fit in half the register file, else one (the kernel then also
// keeps values in the accumulation registers instead of scratch memory).
constexpr __host__ __device__ int m2l_tile_waves(int p) { return p <= 10 ? 2 : 1; }

// Packed vertical factors of one offset: for l = 0..p, m = 0..min(l, p-1),
// n = m..p-1; vfac_ofs(p, l, m) is the position of (l, m, n = m).
constexpr __host__ __device__ int vfac_ofs(int p, int l, int m) {
  int o = 0;
  for (int a = 0; a < l; a++)
    for (int b = 0; b <= (a < p - 1 ? a : p - 1); b++) o += p - b;
  for (int b = 0; b < m; b++) o += p - b;
  return o;
}
constexpr __host__ __device__ int vfac_size(int p) {
  return vfac_ofs(p, p + 1, 0);
}

struct M2LTileArgs {
  const int *tile_src;    // [ntiles * 189 * 64] source node or -1
  const int *tile_oct;    // [ntiles] octant shared by the tile
  const short *oct_codes; // [8 * 189] offset code, grouped by theta
  const DirEntry *dir;    // [343]
  const void *ephi;       // [343][2(p+1)]   e^{i m phi} as T
  const void *rot;        // [ntheta][4][rot_block(p)] as T
  const void *vfac;       // [343][vfac_size(p)] scaled vertical factors
  int nsplit;             // offset ranges per tile (partial slices), >= 1
  int tile_major;         // XCD order: all ranges of a tile group together
};

// Octant offset lists (host): codes[8][189] sorted by theta table index,
// code_to_k[8][343] position of a code in its octant list or -1.
void octant_offsets(const std::vector<DirEntry> &dirs, std::vector<short> &codes,
                    std::vector<short> &code_to_k);

hipError_t launch_m2l_tile(int precision, int p, int ntiles,
                           const M2LTileArgs &a, const void *Ms, void *Ls,
                           hipStream_t s);
// Node counts of the (level, octant) groups from the octant-flag scan
// (lstart: level starts on the device).
hipError_t launch_group_counts(int first, int count, int depth,
                               const int *lstart, const int *rank_scan,
                               int *cnt, hipStream_t s);
hipError_t launch_octant_flags(const NodeArrays &nd, int first, int count,
                               int *flags, hipStream_t s);
hipError_t launch_tile_fill(const NodeArrays &nd, int first, int count,
                            const int *rank_scan, const int *lstart,
                            const int *gofs, const int *tile_of, int *tile_tgt,
                            int *node_slot, hipStream_t s);
// childmap[8 node + o]: child of every node in octant o, -1 when absent
hipError_t launch_tile_src(const NodeArrays &nd, int ntiles,
                           const int *tile_tgt, const int *tile_oct,
                           const int *col, const int *childmap,
                           const short *oct_codes, int *tile_src, int *nfar,
                           hipStream_t s);
hipError_t launch_fill_int(int *p, long n, int v, hipStream_t s);
hipError_t launch_scale_mult(int precision, const NodeArrays &nd, int nnodes,
                             int p, const double2 *M, void *Ms, hipStream_t s);
// L of the tiled targets = (sum of the nsplit slices [split][tile][coef][lane]
// of Ls, in slice order) / w^(l+1); rows of nodes [0, first) are zeroed.
hipError_t launch_unscale_local(int precision, const NodeArrays &nd, int p,
                                int first, int ntiles, const int *tile_tgt,
                                const void *Ls, int nsplit, double2 *L,
                                hipStream_t s);
// Offset ranges per tile of the FP32 far field at order p (fixed per order,
// so a target's sums do not depend on the rest of the solve).
constexpr int kFarRangesTile = 8;
int m2l_choose_split(int p);
// Workgroup order of that launch (M2LTileArgs::tile_major).
int m2l_tile_major();
hipError_t launch_m2l_tables(int precision, int p, const DirEntry *dir,
                             const double *tm2l, const double *rot, long nrot,
                             const double2 *eiphi, void *vfac, void *rotT,
                             void *ephi, hipStream_t s);

}  // namespace aran
