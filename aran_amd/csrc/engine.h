// engine.h -- internal interface of the MI355X FMM engine (not installed).
//
// Device data layout (HBM):
//   particles   : SoA, sorted by 63-bit Morton key (x, y, z, q as FP64)
//   nodes       : SoA, breadth-first, Morton-ordered inside each level
//   multipoles  : [node][p(p+1)/2]     complex FP64 (re, im interleaved)
//   locals      : [node][(p+1)(p+2)/2] complex FP64
//   lists       : CSR tables; far (M2L) list = (source node, offset code)
//                 per node, near (P2P) list = particle ranges per leaf
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "tree_rules.h"

namespace aran {

constexpr int kMaxDepth = 21;   // 3 x 21 key bits
static_assert(kAncRow == kMaxDepth + 1, "ancestor rows cover every depth");
// Orders: the plain FP64 form and the single-operator entry points take
// local degrees 0..kMaxOrder; the mixed and MFMA forms, and the kernels
// templated on the order (generic kernels take the orders above), go to
// kMaxOrderFast.
constexpr int kMaxOrder = 30;
constexpr int kMaxOrderFast = 15;
constexpr int kNumCodes = 343;  // translation offsets in [-3,3]^3

inline __host__ __device__ int code_of(int dx, int dy, int dz) {
  return (dx + 3) * 49 + (dy + 3) * 7 + (dz + 3);
}
inline __host__ __device__ int tri(int n) { return n * (n + 1) / 2; }
// Per-theta rotation tables: four (l+1)x(l+1) real matrices per degree l.
inline __host__ __device__ int rot_block(int L) {
  return (L + 1) * (L + 2) * (2 * L + 3) / 6;
}
inline __host__ __device__ int rot_ofs(int l) {
  return l * (l + 1) * (2 * l + 1) / 6;
}

// Translation vector v (source centre -> target centre, integer units)
// after the z flip of the reference (rotate.c:190-194).
struct DirEntry {
  int theta;    // index of the theta rotation table
  double c;     // +1, or -1 when v.z < 0 (v is then negated)
  double norm;  // |v|
};

struct OpTables {
  const double *rot;     // [ntheta][4][rot_block(p)]   E+ E- F+ F-
  const double2 *eiphi;  // [343][p+1]   e^{i m phi}
  const DirEntry *dir;   // [343]
  const double *tm2l;    // [p+1][p+1][p]   T(l,m,n)
  const double *tm2m;    // [p][p][p]
  const double *tl2l;    // [p+1][p+1][p+1]
  const double *ynorm;   // [tri(p+2)]  N_l^m
  const double *ogain;   // [tri(p+1)]  N_l^m (l+m)!
  const double *pgain;   // [tri(p)]    4 pi / (2l+1) N_l^m (l+m)!
  const double *igain;   // [tri(p+1)]  N_l^m / (l-m)!
  const double *xgain;   // [tri(p+1)]  4 pi / (2l+1) N_l^m / (l-m)!
  const double *alpha;   // [tri(kAlphaDeg+1)]  1/sqrt((l-m)!(l+m)!)
  const double *beta;    // [kAlphaDeg+1]       sqrt(4 pi / (2l+1))
  int p, rb;
};
// degree of the alpha / beta tables (every order's translations in any
// flavour read them up to degree 2 kMaxOrder + 1)
constexpr int kAlphaDeg = 2 * kMaxOrder + 2;

struct Stats {
  long zero = 0, p2p = 0, p2m = 0, m2m = 0, m2l = 0, l2l = 0, l2p = 0;
  long p2l = 0, m2p = 0, semifar = 0;  // semifar: (leaf, node) pairs
  int depth = 0;
  long nnodes = 0, nleaves = 0;
  // tiled far-field launch: offset ranges, schedule, tiles
  int m2l_split = 0, m2l_sched = 0, m2l_tiles = 0;
  // phase times in ms (only with Problem::timings)
  float t_tree = 0, t_lists = 0, t_up = 0, t_m2l = 0, t_down = 0, t_near = 0,
        t_total = 0;
  // kernel groups inside the phases: P2M, the M2M levels, the L2L levels,
  // M2P + L2P (after the near field has joined)
  float t_p2m = 0, t_m2m = 0, t_l2l = 0, t_l2p = 0;
};

// Grow-only device buffer.
// While an arena is set on the calling thread (set_arena), buffers are
// carved from it (bump allocation, never freed) instead of hipMalloc: the
// no-allocation solve runs the engine over a fixed device scratch that way.
struct DevArena {
  char *base = nullptr;
  size_t size = 0, used = 0;
};
void set_arena(DevArena *arena);

struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  bool in_arena = false;  // ptr points into an arena (not to be freed)
  hipError_t reserve(size_t nbytes);
  hipError_t grow_keep(size_t nbytes, size_t keep, hipStream_t s);
  void release();
  template <class T> T *as() const { return static_cast<T *>(ptr); }
};

struct Tables {
  int order = -1;
  int ntheta = 0;
  DevBuf rot, eiphi, dir, tm2l, tm2m, tl2l, ynorm, ogain, pgain, igain, xgain,
      dscr, thetas, alpha, beta;
  // tiled far field (mixed form): FP32 copies and octant offset lists
  DevBuf rot32, ephi32, vfac32, oct_codes, code_to_k;
  // MFMA far field: dense BF16 operators of every offset code (built on
  // first use for this order)
  DevBuf mfma_ops;
  int mfma_order = -1;
  // plain form's far field: dense FP64 operators in the MFMA layout
  DevBuf f64_ops;
  int f64_order = -1;
  std::vector<DirEntry> dir_host;
  OpTables view() const;
  void release();
};

// Orders whose tables a workspace keeps besides the current one (switching
// back to a cached order swaps its tables in instead of rebuilding them).
constexpr int kCachedOrders = 3;

struct Workspace {
  DevBuf keys_a, keys_b, idx_a, idx_b, xs, ys, zs, qs, pot_s, grad_s;
  DevBuf packed;     // (x, y, z, q) per particle in input order
  DevBuf lvl_table;  // per level: first node, node count (device)
  DevBuf hist, part;
  DevBuf n_depth, n_ix, n_iy, n_iz, n_pbegin, n_pend, n_parent, n_child,
      n_nchild, n_leaf, n_key, n_cx, n_cy, n_cz;
  DevBuf n_anc;   // ancestor rows (NodeArrays::anc)
  DevBuf n_cmap;  // children by octant (NodeArrays::cmap)
  DevBuf split_ofs, leaf_list;
  DevBuf internal_list;  // internal nodes in node order (level by level)
  DevBuf trans_rows;     // TransRow of each internal node (same order)
  DevBuf v_cnt, v_ofs, v_src, v_code, u_cnt, u_ofs, u_beg, u_end, u_pairs,
      colleagues;
  // semifar lists: per leaf the nodes of its M2P (w_*), per node the leaves
  // of its P2L (x_*); s_part: per-leaf P2L/M2P particle counts
  DevBuf w_cnt, w_ofs, w_node, x_cnt, x_ofs, x_leaf, s_part;
  long semifar_nodes = -1;  // nodes covered by the semifar lists, -1: none
  // autotuned far-field schedules, by (form, order, size bucket)
  std::map<long, int> tuned_sched;
  DevBuf mult, loc, small, pleaf;
  DevBuf tile_tgt, tile_src, node_slot, tile_cnt, tile_rank,
      tile_meta, mult32, loc32, mult_bf, loc_mfma, mult64, loc64;
  DevBuf n_sys, sys_box, sys_off, skey, spos, stop;
  // look-back status words (Lookback)
  DevBuf lb_buf;
  unsigned lb_epoch = 0;
  long node_cap = 0;
  int depth_hint = 0;  // depth of the last tree (levels of the first batch)
  // near-list ranges per leaf of the last solve (0: none yet); the next
  // solve sizes its lists from it and builds them in one launch
  double near_per_leaf = 0.;
  long far_csr_nodes = -1;  // nodes covered by the CSR far lists, -1: none
  void *host_small = nullptr;  // pinned, 8 KiB
  // pinned staging of the host-to-device copies (and the larger reads) of
  // one solve: a bump arena reset when the next solve starts (a solve ends
  // with its stream synchronized, so no copy from it is pending then);
  // blocks it outgrows are kept until then
  char *host_stage = nullptr;
  size_t host_stage_bytes = 0, host_stage_used = 0;
  std::vector<char *> host_stage_old;
  // the staging block is fixed (allocated once; a solve that needs more
  // fails instead of allocating)
  bool fixed_stage = false;
  Tables tab;
  std::vector<Tables> cached_tabs;  // other orders, least recently used first
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // second stream (default priority) for the near field, which runs
  // alongside the upward pass, the far field and the downward pass
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_near = nullptr;
  hipEvent_t ev[8] = {};
  hipEvent_t evk[8] = {};  // kernel-group timers (Stats::t_p2m ...)
  bool events_ready = false;
  int device = -1;
  NodeArrays nodes() const;
  // every device buffer of the workspace and of its tables
  std::vector<DevBuf *> buffers();
  // drops every buffer pointer without freeing it (the arena they came from
  // starts over) and forgets the tables
  void forget_arena_buffers();
  ~Workspace();
};

struct Problem {
  long n = 0;
  int order = 8;
  int leafcap = 20;
  double lb[3] = {0, 0, 0}, ub[3] = {1, 1, 1};
  const double *x = nullptr, *y = nullptr, *z = nullptr, *q = nullptr;
  // independent systems: particles soff[s]..soff[s+1] of the inputs form
  // system s with root bounds lbs/ubs[3s..3s+2] (host arrays); nsys == 0
  // means one system with lb / ub
  int nsys = 0;
  const long *soff = nullptr;
  const double *lbs = nullptr, *ubs = nullptr;
  double *pot = nullptr;   // device, user order, may be null
  double *grad = nullptr;  // device, user order [n][3], may be null
  bool timings = false;
  // 0: plain FP64; 1: mixed (FP32 far field on VALU, FP32 P2P); 2 / 3: far
  // field on MFMA with BF16 operands in 3 / 6 products, FP32 P2P
  int precision = 0;
  // semifar pairs (P2L / M2P) for nodes holding more than this many
  // particles; < 0: none (every leaf / deeper node pair is P2P)
  long semifar = -1;
  // time the far-field schedules on first use (see tune_schedule)
  bool autotune = false;
  // translation flavours of M2M, M2L, L2L (0 rotate, 1 normal, 2 Kylin;
  // the last two in the plain form only)
  int flav_m2m = 0, flav_m2l = 0, flav_l2l = 0;
  // build the tree and its leaf / internal lists only (no expansions)
  bool tree_only = false;
  // nodes that do not split (one system only): (depth, Morton key) pairs
  // sorted by depth then key (host array)
  const unsigned long long *stop = nullptr;  // [2 nstop]: depth, key
  int nstop = 0;
  // plain form, rotate flavour, orders <= 15: the far field as dense FP64
  // products on MFMA (false: the per-pair rotate kernel)
  bool dense_far = true;
};

struct TreeInfo {
  double lo[3] = {0, 0, 0};  // root cube of system 0
  double width = 1;
  int nsys = 1;
  int depth = 0;
  long nnodes = 0, nleaves = 0, nfar = 0, nnear = 0;
  std::vector<int> level_start, level_count;
};

hipError_t solve(Workspace &ws, const Problem &pb, Stats *st, TreeInfo *ti);
// CSR far lists (v_ofs, v_src, v_code) of the last tree, from its colleagues.
hipError_t build_far_lists(Workspace &ws, long nn, int *nfar);
hipError_t ensure_tables(Workspace &ws, int order);
hipError_t ensure_mfma_tables(Workspace &ws, int order);
hipError_t workspace_init(Workspace &ws, hipStream_t stream);
const char *last_error();
void set_error(const std::string &msg);

// ---- tree.hip
hipError_t launch_keys(const double *x, const double *y, const double *z,
                       const double *q, long n, const long *soff, int nsys,
                       const double *box, uint64_t *keys, int *idx,
                       double4 *packed, hipStream_t s);
hipError_t launch_sys_keys(const int *idx, long n, const long *soff, int nsys,
                           uint64_t *skey, int *pos, hipStream_t s);
hipError_t launch_permute_keys(const int *pos, long n, const uint64_t *kin,
                               const int *iin, uint64_t *kout, int *iout,
                               hipStream_t s);
int minmax_blocks(int nsys);
// Extent of each system and its root cube: box[4 s ..] = (lo, width) from
// its bounds (bounds[6 s ..] = lb, ub) and particles (a system with a
// non-finite coordinate sets bit 4 of *err, which starts at zero), and its
// root node s of nd; also sets up the level table dl[0 .. ntab) (level 0:
// the nsys roots).
hipError_t launch_sys_minmax(const double *x, const double *y, const double *z,
                             const long *soff, const double *bounds, int nsys,
                             double *part, double *box, int *dl, int ntab,
                             int *err, const NodeArrays &nd, hipStream_t s);
// Exclusive scan of n <= 2^14 ints in one launch of one workgroup (in and
// out may alias); writes the total to *total.
hipError_t scan_small(const int *in, int *out, long n, int *total,
                      hipStream_t s);
hipError_t radix_sort(uint64_t *keys, int *idx, uint64_t *keys_tmp,
                      int *idx_tmp, long n, DevBuf &hist, int passes,
                      hipStream_t s);
hipError_t launch_gather(const int *idx, long n, const double4 *packed,
                         double *xs, double *ys, double *zs, double *qs,
                         hipStream_t s);
// Status words of a launch that scans across its workgroups by decoupled
// look-back (one word per chunk of work).  The words need no clearing
// between launches: each launch has its own epoch.
struct Lookback {
  unsigned long long *status;
  int *err;        // set when a wait gave up (the result is then wrong)
  unsigned epoch;  // 1 .. 2^30 - 1
};
// Exclusive scan of n ints in one launch of scan_lookback_chunks(n)
// workgroups (in and out may alias); writes the total to *total.
hipError_t scan_lookback(const int *in, int *out, long n, int *total,
                         const Lookback &lb, hipStream_t s);
long scan_lookback_chunks(long n);
// whether n entries fit scan_small
bool scan_fits_single(long n);
// Level table on the device: dl[2 d] first node of level d, dl[2 d + 1] its
// node count; max_count bounds the count for the grid.  One launch splits
// level d and sets dl[2 d + 2] (the next level's first node) and
// dl[2 d + 3] (its count); it uses split_level_chunks(max_count) status
// words.
hipError_t launch_split_level(const NodeArrays &nd, int *dl, int d,
                              long max_count, const uint64_t *keys,
                              int leafcap, const ulonglong2 *stop, int nstop,
                              const Lookback &lb, hipStream_t s);
long split_level_chunks(long max_count);
hipError_t launch_leaf_compact(const NodeArrays &nd, int nnodes,
                               const int *pos, int *leaves, int *internal,
                               hipStream_t s);
// out[0] = leaf count, out[1 + d] = internal nodes of level d (d < nlevels)
hipError_t launch_level_internal(const int *pos, const int *total,
                                 const int *lstart, int nlevels, int nnodes,
                                 int *out, hipStream_t s);
// colleagues of every node (lstart / lcount of the levels on the device)
hipError_t launch_colleagues_all(const NodeArrays &nd, int nnodes,
                                 const int *lstart, const int *lcount,
                                 int *col, hipStream_t s);
hipError_t launch_vlist(const NodeArrays &nd, int nnodes, const int *col,
                        int pass, int *cnt, const int *ofs, int *src,
                        short *code, hipStream_t s);
// Near lists: pass 0 counts each leaf's ranges, pass 1 writes them from the
// scan ofs of the counts; with a capacity cap (ranges the arrays hold;
// LONG_MAX: none) pass 1 stores only the ranges below it, clamps ofs to it
// and writes the true total to *total.
hipError_t launch_ulist(const NodeArrays &nd, int nleaves, const int *leaves,
                        const int *col, int pass, int *cnt, int *ofs,
                        int *ubeg, int *uend, long long *pairs, long cap,
                        int *total, hipStream_t s);
hipError_t launch_sum_ll(const long long *v, int n, long long *out,
                         hipStream_t s);
hipError_t launch_ulist_semifar(const NodeArrays &nd, int nleaves,
                                const int *leaves, const int *col, long thr,
                                int pass, int *cnt, const int *ofs, int *ubeg,
                                int *uend, long long *pairs, int *wcnt,
                                const int *wofs, int *wnode, long long *spart,
                                hipStream_t s);
hipError_t launch_xlist(const NodeArrays &nd, int nnodes, const int *col,
                        long thr, int pass, int *cnt, const int *ofs,
                        int *xleaf, hipStream_t s);

// ---- operators.hip
hipError_t launch_wigner_tables(const double *thetas, int ntheta, int L,
                                double *dscr, double *rot, hipStream_t s);
// low: the FP32 recurrence on width-scaled coordinates (the reduced-
// precision forms, for leaves below the top two levels); FP64 otherwise
hipError_t launch_p2m(const OpTables &tb, const NodeArrays &nd,
                      const int *leaves, int nleaves, const double *xs,
                      const double *ys, const double *zs, const double *qs,
                      double2 *M, hipStream_t s, bool low = false);
hipError_t launch_m2l(const OpTables &tb, const NodeArrays &nd, int nnodes,
                      const int *vofs, const int *vsrc, const short *vcode,
                      const double2 *M, double2 *L, hipStream_t s);
// pleaf[j] = leaf node of sorted particle j (every particle of the leaves)
hipError_t launch_particle_leaf(const NodeArrays &nd, const int *leaves,
                                int nleaves, int *pleaf, hipStream_t s);
// L2P of sorted particles [0, n), adding into pot / grad.  sel 0: every
// particle, FP64; 1: the particles of leaves at depth >= 2, FP32 recurrence
// on width-scaled coordinates (the reduced-precision forms); 2: those of
// leaves at depth < 2, FP64.  With perm, pot / grad stay as they are and
// the sums go to opot[perm[i]] / ograd[perm[i]] instead (the caller's
// order; either may be null).
hipError_t launch_l2p(const OpTables &tb, const NodeArrays &nd,
                      const int *pleaf, int n, const double *xs,
                      const double *ys, const double *zs, const double2 *L,
                      double *pot, double *grad, hipStream_t s, int sel = 0,
                      const int *perm = nullptr, double *opot = nullptr,
                      double *ograd = nullptr);

// M2P: for every leaf, the multipoles of its semifar nodes evaluated at its
// sorted particles, added into pot / grad (potential and gradient).
hipError_t launch_m2p(const OpTables &tb, const NodeArrays &nd,
                      const int *leaves, int nleaves, const int *wofs,
                      const int *wnode, const double *xs, const double *ys,
                      const double *zs, const double2 *M, double *pot,
                      double *grad, hipStream_t s);
// P2L: for every node, the particles of its semifar leaves added into its
// local.
hipError_t launch_p2l(const OpTables &tb, const NodeArrays &nd, int nnodes,
                      const int *xofs, const int *xleaf, const double *xs,
                      const double *ys, const double *zs, const double *qs,
                      double2 *L, hipStream_t s);

// Dense scaled M2L operators O[343][2lc-p-1][2mc-p] in FP64.
hipError_t launch_m2l_dense(const OpTables &tb, double *O, hipStream_t s);

hipError_t launch_translate_one(int kind, const OpTables &tb, double r,
                                const double2 *src, double2 *dst,
                                hipStream_t s);

// ---- p2p.hip
hipError_t launch_p2p(int precision, const int *leaves, int nleaves,
                      const NodeArrays &nd, const int *uofs, const int *ubeg,
                      const int *uend, const double *xs, const double *ys,
                      const double *zs, const double *qs, double *pot,
                      double *grad, hipStream_t s);
hipError_t launch_direct(const double *x, const double *y, const double *z,
                         const double *q, long n, const long *targets, long nt,
                         double *pot, double *grad, hipStream_t s);

}  // namespace aran
