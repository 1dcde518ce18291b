/*
 * vsgd.h -- the part of LibVsg's FP64 interface that LibAran's 3D solver
 * API exposes (VsgVector3d, VsgPoint3, VsgPRTree3d, VsgPRTree3dNodeInfo and
 * the tree constructor), declared here because LibVsg is not a dependency of
 * this library.  Call sites in the reference: aran/aransolver3d.h:35-214,
 * tests/newtonpot3.c:602-606.
 *
 * The tree object only records bounds and leaf capacity; the octree itself
 * is built on the GPU by aran_solver3d_solve().
 */
#ifndef ARAN_VSGD_COMPAT_H
#define ARAN_VSGD_COMPAT_H

#include <stdio.h>

#include <aran/arantypes.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _VsgVector3d VsgVector3d;
struct _VsgVector3d
{
  gdouble x, y, z;
};

typedef gpointer VsgPoint3;
typedef gpointer VsgRegion3;

typedef struct _VsgPRTreeKey3d VsgPRTreeKey3d;
struct _VsgPRTreeKey3d
{
  guint64 x, y, z;
  guint8 depth;
};

typedef struct _VsgPRTree3dNodeInfo VsgPRTree3dNodeInfo;
struct _VsgPRTree3dNodeInfo
{
  VsgVector3d lbound;
  VsgVector3d ubound;
  VsgVector3d center;
  guint point_count;
  gpointer point_list;
  gboolean isleaf;
  guint8 depth;
  VsgPRTree3dNodeInfo *father_info;
  gpointer user_data;
  VsgPRTreeKey3d id;
};

typedef struct _VsgPRTree3d VsgPRTree3d;

typedef guint (*VsgPoint3dLocFunc) (const VsgPoint3 candidate,
                                    const VsgVector3d *center);
typedef gdouble (*VsgPoint3dDistFunc) (const VsgPoint3 one,
                                       const VsgPoint3 other);
typedef guint (*VsgRegion3dLocFunc) (const VsgRegion3 region,
                                     const VsgVector3d *center);
typedef guint (*VsgRegion3Point3LocFunc) (const VsgRegion3 region,
                                          const VsgPoint3 point);
typedef void (*VsgPRTree3dFunc) (const VsgPRTree3dNodeInfo *node_info,
                                 gpointer user_data);
typedef gboolean (*VsgPRTree3dNFIsleafFunc) (const VsgPRTree3dNodeInfo *info,
                                             gpointer user_data);

/* Octant of candidate around center (bit 0: x, bit 1: y, bit 2: z, set on
 * the upper side).  Points are given as structs whose first member is a
 * VsgVector3d. */
guint vsg_vector3d_vector3d_locfunc (const VsgPoint3 candidate,
                                     const VsgVector3d *center);
gdouble vsg_vector3d_dist (const VsgPoint3 one, const VsgPoint3 other);

/* Records the root bounds and the leaf capacity maxbox.  Only the two
 * functions above are accepted as point locator / distance. */
VsgPRTree3d *vsg_prtree3d_new_full (const VsgVector3d *lbound,
                                    const VsgVector3d *ubound,
                                    VsgPoint3dLocFunc point_locfunc,
                                    VsgPoint3dDistFunc point_distfunc,
                                    VsgRegion3dLocFunc region_locfunc,
                                    guint maxbox);
void vsg_prtree3d_free (VsgPRTree3d *prtree);

/* Writes the recorded bounds and leaf capacity (the octree itself exists
 * on the GPU during a solve only). */
void vsg_prtree3d_write (VsgPRTree3d *prtree, FILE *file);

/* Vector helpers the reference's harnesses call (tests/newtonpot3.c,
 * tests/newtonfield3.c), written from those call sites: dst may alias an
 * operand. */
#define VSG_V3D_ZERO ((VsgVector3d) {0., 0., 0.})
void vsg_vector3d_set (VsgVector3d *vec, gdouble x, gdouble y, gdouble z);
void vsg_vector3d_copy (const VsgVector3d *src, VsgVector3d *dst);
void vsg_vector3d_add (const VsgVector3d *src, const VsgVector3d *other,
                       VsgVector3d *dst);                 /* src + other */
void vsg_vector3d_sub (const VsgVector3d *src, const VsgVector3d *other,
                       VsgVector3d *dst);                 /* src - other */
void vsg_vector3d_scalp (const VsgVector3d *src, gdouble scal,
                         VsgVector3d *dst);               /* scal * src */
gdouble vsg_vector3d_dotp (const VsgVector3d *one, const VsgVector3d *other);
gdouble vsg_vector3d_norm (const VsgVector3d *vec);
/* "[x, y, z]" */
void vsg_vector3d_write (const VsgVector3d *vec, FILE *file);
/* r, cos theta, sin theta (theta from +z), cos phi, sin phi (phi from +x);
 * on the z axis cos phi = 1, at the origin cos theta = 1. */
void vsg_vector3d_to_spherical_internal (const VsgVector3d *vec, gdouble *r,
                                         gdouble *cost, gdouble *sint,
                                         gdouble *cosp, gdouble *sinp);
void vsg_vector3d_from_spherical_internal (VsgVector3d *vec, gdouble r,
                                           gdouble cost, gdouble sint,
                                           gdouble cosp, gdouble sinp);

/* Message tracing of LibVsg's MPI layer: nothing to trace on one GPU. */
void vsg_packed_msg_trace_set_active (gboolean active);

/* Whether a node is shared between processes of a distributed tree: never
 * here (one process per tree). */
#define VSG_PRTREE3D_NODE_INFO_IS_SHARED(info) ((void) (info), FALSE)

#ifdef __cplusplus
}
#endif

#endif /* ARAN_VSGD_COMPAT_H */
