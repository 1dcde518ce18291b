/*
 * aransolver3d.h -- the AranSolver3d C API (reference:
 * aran/aransolver3d.h:32-214), backed by the MI355X FMM engine.
 *
 * aran_solver3d_solve() gathers the inserted particles, builds the octree
 * and runs P2M, M2M, M2L, L2L, L2P and P2P on the GPU, then adds the results
 * into the caller's particle structs, as the reference's L2P / P2P
 * callbacks do.  The engine takes this path when the registered callbacks
 * are the library's own:
 *   p2p : aran_newton3d_p2p_pot / _field / _potfield
 *   p2m : aran_newton3d_p2m
 *   m2m : aran_development3d_m2m_rotate, _m2m or _m2m_kkylin
 *   m2l : aran_development3d_m2l_rotate, _m2l or _m2l_kkylin
 *   l2l : aran_development3d_l2l_rotate, _l2l or _l2l_kkylin
 *         (the normal and Kylin flavours in the plain FP64 form only)
 *   l2p : aran_newton3d_l2p_pot / _field / _potfield
 * optionally with the semifar pair (aran_solver3d_set_functions_full)
 *   p2l : aran_newton3d_p2l
 *   m2p : aran_newton3d_m2p_pot / _field / _potfield (the outputs of p2p)
 * and a semifar threshold (G_MAXUINT: never semifar, 0: automatic from the
 * operators' algorithmic costs), and the particle layout (offsets of charge, potential, field inside the
 * caller's struct whose first member is a VsgVector3d) is declared with
 * aran_solver3d_set_particle_layout().  Other callbacks are rejected with a
 * critical message; there is no host fallback.
 */
#ifndef ARAN_SOLVER3D_H
#define ARAN_SOLVER3D_H

#include <stdio.h>

#include <aran/aran.h>
#include <aran/arandevelopment3d.h>
#include <aran/arantypes.h>
#include <vsg/vsgd.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _AranSolver3d AranSolver3d;

/* callback types (aran/aransolver3d.h:35-83) */
typedef void (*AranParticleInitFunc3d) (VsgPoint3 particle);
typedef void (*AranParticle2ParticleFunc3d) (VsgPoint3 src, VsgPoint3 dst);
typedef void (*AranParticle2MultipoleFunc3d) (VsgPoint3 src,
                                              const VsgPRTree3dNodeInfo *dst_node,
                                              gpointer dst);
typedef AranParticle2MultipoleFunc3d AranParticle2LocalFunc3d;
typedef void (*AranLocal2ParticleFunc3d) (const VsgPRTree3dNodeInfo *src_node,
                                          gpointer src, VsgPoint3 dst);
typedef AranLocal2ParticleFunc3d AranMultipole2ParticleFunc3d;
typedef void (*AranMultipole2MultipoleFunc3d) (const VsgPRTree3dNodeInfo *src_node,
                                               gpointer src,
                                               const VsgPRTree3dNodeInfo *dst_node,
                                               gpointer dst);
typedef AranMultipole2MultipoleFunc3d AranMultipole2LocalFunc3d;
typedef AranMultipole2MultipoleFunc3d AranLocal2LocalFunc3d;

/* construction (aran/aransolver3d.c:509-585) */
AranSolver3d *aran_solver3d_new (VsgPRTree3d *prtree, GType devel_type,
                                 gpointer devel, AranZeroFunc zero);
void aran_solver3d_free (AranSolver3d *solver);
void aran_solver3d_set_development (AranSolver3d *solver, GType devel_type,
                                    gpointer devel, AranZeroFunc zero);

/* operators (aran/aransolver3d.c:599-685) */
void aran_solver3d_set_functions (AranSolver3d *solver,
                                  AranParticle2ParticleFunc3d p2p,
                                  AranParticle2MultipoleFunc3d p2m,
                                  AranMultipole2MultipoleFunc3d m2m,
                                  AranMultipole2LocalFunc3d m2l,
                                  AranLocal2LocalFunc3d l2l,
                                  AranLocal2ParticleFunc3d l2p);
void aran_solver3d_set_functions_full (AranSolver3d *solver,
                                       AranParticle2ParticleFunc3d p2p,
                                       AranParticle2MultipoleFunc3d p2m,
                                       AranMultipole2MultipoleFunc3d m2m,
                                       AranMultipole2LocalFunc3d m2l,
                                       AranLocal2LocalFunc3d l2l,
                                       AranLocal2ParticleFunc3d l2p,
                                       AranParticle2LocalFunc3d p2l,
                                       AranMultipole2ParticleFunc3d m2p,
                                       guint semifar_threshold);
void aran_solver3d_get_functions_full (AranSolver3d *solver,
                                       AranParticle2ParticleFunc3d *p2p,
                                       AranParticle2MultipoleFunc3d *p2m,
                                       AranMultipole2MultipoleFunc3d *m2m,
                                       AranMultipole2LocalFunc3d *m2l,
                                       AranLocal2LocalFunc3d *l2l,
                                       AranLocal2ParticleFunc3d *l2p,
                                       AranParticle2LocalFunc3d *p2l,
                                       AranMultipole2ParticleFunc3d *m2p,
                                       guint *semifar_threshold);

/* statistics (aran/aransolver3d.c:829-886): counts of the last solve with
 * the reference's semantics (p2p counts unordered pairs, m2l counts both
 * directions). */
void aran_solver3d_reinit_stats (AranSolver3d *solver);
void aran_solver3d_get_stats (AranSolver3d *solver, glong *zero_count,
                              glong *p2p_count, glong *p2m_count,
                              glong *m2m_count, glong *m2l_count,
                              glong *l2l_count, glong *l2p_count,
                              glong *p2l_count, glong *m2p_count,
                              glong *p2p_remote_count,
                              glong *m2l_remote_count);

/* particles (aran/aransolver3d.c:887-1088).  Particles stay owned by the
 * caller; the solver keeps their addresses. */
gdouble aran_solver3d_get_tolerance (AranSolver3d *solver);
void aran_solver3d_set_tolerance (AranSolver3d *solver, gdouble tolerance);
void aran_solver3d_get_bounds (AranSolver3d *solver, VsgVector3d *lbound,
                               VsgVector3d *ubound);
guint aran_solver3d_depth (const AranSolver3d *solver);
guint aran_solver3d_point_count (const AranSolver3d *solver);
void aran_solver3d_insert_point (AranSolver3d *solver, VsgPoint3 point);
/* (aran/aransolver3d.h:161) one process per tree: same as insert_point */
void aran_solver3d_insert_point_local (AranSolver3d *solver,
                                       VsgPoint3 point);
gboolean aran_solver3d_remove_point (AranSolver3d *solver, VsgPoint3 point);
VsgPoint3 aran_solver3d_find_point (AranSolver3d *solver, VsgPoint3 selector);
void aran_solver3d_foreach_point (AranSolver3d *solver, GFunc func,
                                  gpointer user_data);

/* the solve (aran/aransolver3d.c:1135-1202) */
void aran_solver3d_solve (AranSolver3d *solver);

/* (aran/aransolver3d.h:180) Visits the octree of the inserted particles
 * (the last solve's, or one built for the call) in the given order
 * (G_LEVEL_ORDER: breadth first).  Node infos carry bounds, centre, particle
 * count, leaf flag, depth, father and key; point_list is NULL.  After a
 * solve of the same particles user_data is the node's AranDevelopment3d
 * (multipole and local of that solve, freed after the call). */
void aran_solver3d_traverse (AranSolver3d *solver, GTraverseType order,
                             VsgPRTree3dFunc func, gpointer user_data);

/* (aran/aransolver3d.h:185-187) the order of siblings in
 * aran_solver3d_traverse: octant order (default) or along a 3-D Hilbert
 * curve (J. Skilling's index of the children's integer coordinates).  A
 * solve's results do not depend on it. */
void aran_solver3d_set_children_order_hilbert (AranSolver3d *solver);
void aran_solver3d_set_children_order_default (AranSolver3d *solver);

/* (aran/aransolver3d.h:212) Near/far leaf predicate: a node for which
 * isleaf returns TRUE acts as a leaf (its particles interact by P2P with
 * the near leaves, its subtree takes no far pairs), as in the reference's
 * traversal.  The solve builds the octree, asks the predicate of every
 * non-leaf node from the root down (node infos as in traverse, user_data
 * NULL), and solves on the octree where those nodes are leaves; the
 * expansions of a leaf made this way are its P2M (the reference's M2M of
 * its subtree, equal up to rounding: M2M and L2L are exact on truncated
 * expansions).  NULL removes the predicate. */
void aran_solver3d_set_nf_isleaf (AranSolver3d *solver,
                                  VsgPRTree3dNFIsleafFunc isleaf,
                                  gpointer user_data);

/* ---- extensions (no counterpart in the reference) ---- */

/* Status of the last call on this solver: 0 or a negative error code. */
gint aran_solver3d_status (const AranSolver3d *solver);

/* Arithmetic form of the next solves (no counterpart in the reference,
 * which computes in FP64 only; for callers of this interface that accept a
 * far field of lower precision): 0 = ARAN_FMM3D_PLAIN (FP64, the
 * default), 1 = ARAN_FMM3D_MIXED, 2 = ARAN_FMM3D_BF16X3 or
 * 3 = ARAN_FMM3D_BF16X6, the forms of aran_fmm3d_set_precision()
 * (aran/aranfmm3d.h).  Returns 0, or ARAN_FMM3D_EINVAL (-1). */
gint aran_solver3d_set_precision (AranSolver3d *solver, gint precision);

/* Byte offsets inside the caller's particle struct.  off_pot / off_field
 * < 0 disable that output.  pot_is_complex: the potential slot is a
 * gcomplex128 (its real part receives the potential, as in
 * tests/newtonpot3.c:43-54), otherwise a gdouble.  Defaults: charge 24,
 * potential none, field none. */
void aran_solver3d_set_particle_layout (AranSolver3d *solver,
                                        gssize off_charge, gssize off_pot,
                                        gboolean pot_is_complex,
                                        gssize off_field);

/* Library Newton callbacks.  As host functions they perform the
 * reference's symmetric pair update / local evaluation using the layout of
 * the last solver created (tests/newtonpot3.c:61-104,
 * tests/newtonfield3.c:65-153); the solver recognises them by address. */
void aran_newton3d_p2p_pot (VsgPoint3 one, VsgPoint3 other);
void aran_newton3d_p2p_field (VsgPoint3 one, VsgPoint3 other);
void aran_newton3d_p2p_potfield (VsgPoint3 one, VsgPoint3 other);
void aran_newton3d_p2m (VsgPoint3 particle, const VsgPRTree3dNodeInfo *dst_node,
                        gpointer dst);
void aran_newton3d_l2p_pot (const VsgPRTree3dNodeInfo *src_node, gpointer src,
                            VsgPoint3 particle);
void aran_newton3d_l2p_field (const VsgPRTree3dNodeInfo *src_node, gpointer src,
                              VsgPoint3 particle);
void aran_newton3d_l2p_potfield (const VsgPRTree3dNodeInfo *src_node,
                                 gpointer src, VsgPoint3 particle);
/* semifar callbacks (tests/newtonpot3.c:93-104, tests/newtonfield3.c:
 * 139-153) */
void aran_newton3d_p2l (VsgPoint3 particle, const VsgPRTree3dNodeInfo *dst_node,
                        gpointer dst);
void aran_newton3d_m2p_pot (const VsgPRTree3dNodeInfo *src_node, gpointer src,
                            VsgPoint3 particle);
void aran_newton3d_m2p_field (const VsgPRTree3dNodeInfo *src_node, gpointer src,
                              VsgPoint3 particle);
void aran_newton3d_m2p_potfield (const VsgPRTree3dNodeInfo *src_node,
                                 gpointer src, VsgPoint3 particle);

/* Adds nparticles particles given as structure-of-arrays copies (the solver
 * then owns them); results of the next solve are read back with
 * aran_solver3d_get_results_soa(). */
gint aran_solver3d_insert_points_soa (AranSolver3d *solver, glong nparticles,
                                      const gdouble *x, const gdouble *y,
                                      const gdouble *z, const gdouble *charge);
gint aran_solver3d_get_results_soa (AranSolver3d *solver, gdouble *pot,
                                    gdouble *field);

/* FMA particle files, the formats of the reference's parallel Newton test
 * (tests/newtonfield3parallel.c).  aran_solver3d_write_fma writes the
 * header "count half_width cx cy cz" of the solver's bounds, then one line
 * "x y z charge" per particle in insertion order, with %g as the reference
 * does (newtonfield3parallel.c:1186-1212); returns 0 or a negative code.
 * aran_fma_read_points reads lines "index x y z charge" (its -dist load
 * fill, newtonfield3parallel.c:1020-1086) into the first max_points slots of
 * the arrays and returns the number of lines read (which may exceed
 * max_points: call with max_points 0 to size the arrays), or a negative code
 * on a malformed line. */
gint aran_solver3d_write_fma (AranSolver3d *solver, FILE *file);
glong aran_fma_read_points (FILE *file, glong max_points, gdouble *x,
                            gdouble *y, gdouble *z, gdouble *charge);

#ifdef __cplusplus
}
#endif

#endif /* ARAN_SOLVER3D_H */
