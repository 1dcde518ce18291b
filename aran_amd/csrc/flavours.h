// flavours.h -- the O(p^4) "normal" and Kylin translation flavours
// (flavours.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "engine.h"
#include "translate.h"

namespace aran {

enum { kFlavRotate = 0, kFlavNormal = 1, kFlavKylin = 2 };
enum { kFlavM2M = 0, kFlavM2L = 1, kFlavL2L = 2 };

// Translation vector xdst - xsrc in spherical coordinates.
struct FlavVec {
  double r, cost, sint, cosp, sinp;
};

// One translation added into dst (device arrays, m >= 0 packed): M2M with
// sdeg / ddeg multipole degree counts, L2L with local degrees, M2L from a
// multipole of sdeg degrees into a local of degree ddeg.  Kylin's M2M needs
// sdeg == ddeg (pad the source).
hipError_t launch_flavour_one(const OpTables &tb, int kind, int flavour,
                              int sdeg, const double2 *src, int ddeg,
                              double2 *dst, const FlavVec &v, hipStream_t s);
// The far field of a solve over the CSR far lists (writes L of every node).
hipError_t launch_flavour_m2l(const OpTables &tb, int flavour,
                              const NodeArrays &nd, int nnodes,
                              const int *vofs, const int *vsrc,
                              const double2 *M, double2 *L, hipStream_t s);
// M2M (writes the parents) or L2L (adds into the children) of one level.
hipError_t launch_flavour_level(const OpTables &tb, int kind, int flavour,
                                const NodeArrays &nd, const TransRow *rows,
                                int nrows, double2 *X, hipStream_t s);

}  // namespace aran
