// translate.h -- M2M and L2L of one tree level (translate.hip).
#pragma once

#include "engine.h"

namespace aran {

// One parent of a level: node, child in each octant (bit 0: x, 1: y, 2: z;
// -1 when absent), box width.
struct TransRow {
  int node;
  int child[8];
  int pad;
  double width;
};

// rows[j] of internal[j] (every internal node).
hipError_t launch_trans_rows(const NodeArrays &nd, const int *internal,
                             int ninternal, TransRow *rows, hipStream_t s);
// rows: the parents of one level.  M2M writes the parents' multipoles from
// their children's; L2L adds the parents' locals, translated, into their
// children's.
hipError_t launch_m2m_level(const OpTables &tb, const TransRow *rows,
                            int nrows, double2 *M, hipStream_t s);
hipError_t launch_l2l_level(const OpTables &tb, const TransRow *rows,
                            int nrows, double2 *L, hipStream_t s);

}  // namespace aran
