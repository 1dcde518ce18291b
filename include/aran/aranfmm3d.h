/*
 * aranfmm3d.h -- device-resident entry points of the MI355X FMM engine.
 *
 * No counterpart in the reference: this is the engine under
 * aran_solver3d_solve() (aran/aransolver3d.c:1135-1202) for callers whose
 * particles already live in GPU memory (structure-of-arrays FP64 device
 * pointers).  Every call is synchronous with respect to the host unless
 * stated otherwise; all pointers marked "device" are HIP device pointers.
 *
 * Return codes: 0 on success, negative on error (aran_last_error() holds the
 * message).
 */
#ifndef ARAN_FMM3D_H
#define ARAN_FMM3D_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _AranFmm3d AranFmm3d;

#define ARAN_FMM3D_EINVAL (-1)
#define ARAN_FMM3D_EHIP (-2)
#define ARAN_FMM3D_ENOMEM (-3)

/* Context bound to the current HIP device; stream NULL creates its own. */
AranFmm3d *aran_fmm3d_new (void *hip_stream);
void aran_fmm3d_free (AranFmm3d *ctx);

/* One FMM solve: potential phi_i = sum_{j != i} q_j / |x_i - x_j| and its
 * gradient, written (not accumulated) in the input order.  order p:
 * multipole degrees 0..p-1, local degrees 0..p (1 <= p <= 30 in the plain
 * form, <= 15 in the others); leafcap: leaf
 * capacity of the octree; lb/ub: host arrays of 3 doubles giving the root
 * box before growth.  pot / grad (n x 3, row-major) may be NULL. */
int aran_fmm3d_solve (AranFmm3d *ctx, long n, const double *x,
                      const double *y, const double *z, const double *q,
                      const double *lb, const double *ub, int order,
                      int leafcap, double *pot, double *grad);

/* Independent systems in one solve: system s is particles
 * offsets[s] .. offsets[s+1]-1 of the concatenated device arrays (offsets is
 * a host array of nsys+1 values, offsets[0] = 0), with root bounds
 * lbs[3s..3s+2], ubs[3s..3s+2] (host).  All systems share order and leaf
 * capacity; each gets its own octree and its results are those of a solve
 * of that system alone. */
int aran_fmm3d_solve_batch (AranFmm3d *ctx, int nsys, const long *offsets,
                            const double *x, const double *y, const double *z,
                            const double *q, const double *lbs,
                            const double *ubs, int order, int leafcap,
                            double *pot, double *grad);

/* Arithmetic form of the next solves: ARAN_FMM3D_PLAIN (FP64 throughout) or
 * ARAN_FMM3D_MIXED (far field M2L and near field P2P in FP32 on expansions
 * scaled per level by the box width, FP64 tables, tree, P2M, M2M, L2L, L2P
 * and accumulation across staged source passes). */
#define ARAN_FMM3D_PLAIN 0
#define ARAN_FMM3D_MIXED 1
/* Far field on MFMA: per offset code the M2L is one dense real matrix on
 * the scaled expansions, applied with BF16 operands split in parts
 * (x = h + m + l) and FP32 accumulation: 3 products (h h + h m + m h) or
 * 6 products (+ h l + l h + m m, FP32-level accuracy); P2P in FP32 as in
 * ARAN_FMM3D_MIXED. */
#define ARAN_FMM3D_BF16X3 2
#define ARAN_FMM3D_BF16X6 3
int aran_fmm3d_set_precision (AranFmm3d *ctx, int precision);

/* Far-field autotuner: when on, the first solve of each (form, order,
 * problem-size bucket) times every workgroup schedule of the tiled
 * far-field kernel (the order in which its tiles and fixed offset ranges
 * are dealt to the XCDs) on the device, and later solves of that bucket use
 * the fastest (kept in the context).  Every schedule computes the same sums
 * in the same order, so results do not depend on it; nor on the rest of a
 * batch (the offset ranges are fixed per order).  Off by default. */
int aran_fmm3d_set_autotune (AranFmm3d *ctx, int on);

/* Translation flavours of the next solves (plain form only): the operators
 * of record rotate / z-translate / rotate back (ARAN_FMM3D_ROTATE, the
 * default), the O(p^4) "normal" formulas (aran/aransphericalseriesd-
 * translate.c:366-834) or K. Kylin's recurrences (aran/aransphericalseriesd-
 * kkylin.c), for M2M, M2L and L2L each. */
#define ARAN_FMM3D_ROTATE 0
#define ARAN_FMM3D_NORMAL 1
#define ARAN_FMM3D_KYLIN 2
int aran_fmm3d_set_translations (AranFmm3d *ctx, int m2m, int m2l, int l2l);

/* Expansions of the last solve (host arrays, complex FP64 interleaved):
 * multipoles[nnodes * p(p+1)/2], locals[nnodes * (p+1)(p+2)/2] in the node
 * order of export_tree; either may be NULL.  The plain form keeps them in
 * FP64 on the device; after a mixed or MFMA solve the locals of levels >= 2
 * are those of the far field's FP32 accumulation. */
int aran_fmm3d_export_expansions (const AranFmm3d *ctx, double *multipoles,
                                  double *locals);

/* Counters of the last solve: zero, p2p, p2m, m2m, m2l, l2l, l2p (the
 * reference's semantics), then depth, nodes, leaves, far-list entries,
 * near-list entries, then p2l, m2p (particles in semifar pairs, the
 * reference's counters) and the number of semifar (leaf, node) pairs. */
int aran_fmm3d_get_stats (const AranFmm3d *ctx, long *out15);

/* Far-field plan of the last solve: out3 = {tiles of 64 targets (0: the
 * per-pair kernel ran), offset ranges per tile, workgroup schedule}.  No
 * counterpart in the reference (its far field has no tiles); for callers
 * that check which far-field kernel ran. */
int aran_fmm3d_get_far_plan (const AranFmm3d *ctx, int *out3);

/* Semifar pairs of the next solves (aran/aransolver3d.c:357-396): when the
 * near/far recursion pairs a leaf with a deeper node whose box does not
 * touch the leaf's and which holds more than `threshold` particles, the
 * leaf's particles add to the node's local (P2L) and take its multipole's
 * potential and gradient (M2P) instead of P2P with the node's particles.
 * threshold < 0 disables them (the default). */
int aran_fmm3d_set_semifar (AranFmm3d *ctx, long threshold);

/* Threshold at which P2L + M2P per particle cost what P2P with the node's
 * particles costs, from the operators' algorithmic flops at this order. */
long aran_fmm3d_semifar_threshold_auto (int order);

/* Semifar lists of the last solve (host arrays sized from get_stats):
 * w_ofs[nleaves+1], w_node[npairs] (per leaf, in the leaf order of
 * export_lists, the nodes whose multipoles it evaluates), x_ofs[nnodes+1],
 * x_leaf[npairs] (per node, the leaves whose particles it takes). */
int aran_fmm3d_export_semifar (const AranFmm3d *ctx, int *w_ofs, int *w_node,
                               int *x_ofs, int *x_leaf);

/* Phase timers (HIP events) of the next solves when on != 0; read back as
 * tree, lists, up, m2l, down, near, total in ms. */
int aran_fmm3d_set_timings (AranFmm3d *ctx, int on);
int aran_fmm3d_get_timings (const AranFmm3d *ctx, float *out7);
/* Kernel groups inside those phases, in ms: P2M, the M2M levels, the L2L
 * levels, M2P + L2P.  No counterpart in the reference; for per-operator
 * rates. */
int aran_fmm3d_get_kernel_timings (const AranFmm3d *ctx, float *out4);

/* Tree of the last solve, copied to host arrays sized from get_stats:
 * perm[n] (sorted position -> input index), nodes[nnodes*10] with
 * depth, ix, iy, iz, pbegin, pend, parent, isleaf, child_begin, nchild,
 * centers[nnodes*3], lo[3] and width[1]. */
int aran_fmm3d_export_tree (const AranFmm3d *ctx, long *perm, int *nodes,
                            double *centers, double *lo, double *width);

/* Interaction tables of the last solve (host arrays): far_ofs[nnodes+1],
 * far_src[nfar], far_code[nfar] (offset code (dx+3)*49+(dy+3)*7+(dz+3) of
 * target minus source), leaves[nleaves], near_ofs[nleaves+1],
 * near_begin[nnear], near_end[nnear] (particle ranges in sorted order). */
int aran_fmm3d_export_lists (const AranFmm3d *ctx, int *far_ofs,
                             int *far_src, short *far_code, int *leaves,
                             int *near_ofs, int *near_begin, int *near_end);

/* FP64 direct sum on nt targets (device indices), compensated summation. */
int aran_fmm3d_direct (AranFmm3d *ctx, long n, const double *x,
                       const double *y, const double *z, const double *q,
                       long nt, const long *targets, double *pot,
                       double *grad);

/* aran_fmm3d_solve without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same inputs,
 * outputs and results as aran_fmm3d_solve on the same context settings, but
 * every working buffer (tree, lists, expansions, tables of the order) comes
 * from one device scratch array of ARAN_FMM3D_NOALLOC_SCRATCH bytes per
 * device, set up by aran_fmm3d_new, and host staging from a pinned block
 * allocated there too: the call makes no hipMalloc, hipFree or other device
 * or pinned allocation.  Limits: n <= ARAN_FMM3D_NOALLOC_MAX_N, order <=
 * ARAN_FMM3D_NOALLOC_MAX_ORDER, the plain or mixed form, rotate
 * translations; a problem whose working memory exceeds the scratch (small
 * leaf capacities) is refused with ARAN_FMM3D_ENOMEM.  The scratch is
 * shared by every call on a device and keeps the last call's contents: calls
 * that overlap on one device race.  Statistics and exports of the context
 * keep describing its last aran_fmm3d_solve. */
#define ARAN_FMM3D_NOALLOC_SCRATCH (128L << 20)
#define ARAN_FMM3D_NOALLOC_MAX_N 65536
#define ARAN_FMM3D_NOALLOC_MAX_ORDER 10
int aran_fmm3d_solve_noalloc (AranFmm3d *ctx, long n, const double *x,
                              const double *y, const double *z,
                              const double *q, const double *lb,
                              const double *ub, int order, int leafcap,
                              double *pot, double *grad);

/* Stream the context launches on (hipStream_t). */
void *aran_fmm3d_stream (const AranFmm3d *ctx);

#ifdef __cplusplus
}
#endif

#endif /* ARAN_FMM3D_H */
