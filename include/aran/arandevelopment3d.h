/*
 * arandevelopment3d.h -- FMM node data and its operators
 * (reference: aran/arandevelopment3d.h:30-150).
 *
 * An AranDevelopment3d holds two spherical series, as the reference's:
 *   multipole : aran_spherical_seriesd_new (posdeg, negdeg)
 *   local     : aran_spherical_seriesd_new (max (posdeg, negdeg), 0)
 * (aran/arandevelopment3d.c:64-73).  aran_development3d_new (0, p) is the
 * "order p" development of the reference's tests (tests/newtonpot3.c:
 * 608-610): multipole degrees 0..p-1, local degrees 0..p.  Degrees go up to
 * 30.
 *
 * The operator entry points below run one operator as a GPU launch on the
 * current device; inside aran_solver3d_solve() they are recognised by
 * address and replaced by the batched kernels.
 */
#ifndef ARAN_DEVELOPMENT3D_H
#define ARAN_DEVELOPMENT3D_H

#include <stdio.h>

#include <aran/arantypes.h>
#include <aran/aransphericalseriesd.h>
#include <vsg/vsgd.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ARAN_TYPE_DEVELOPMENT3D (aran_development3d_get_type ())

typedef struct _AranDevelopment3d AranDevelopment3d;

/* the reference's layout (aran/arandevelopment3d.h:36-42) */
struct _AranDevelopment3d
{
  AranSphericalSeriesd *multipole;
  AranSphericalSeriesd *local;
};

GType aran_development3d_get_type (void);                 /* :59 */
AranDevelopment3d *aran_development3d_new (guint8 posdeg,  /* :61 */
                                           guint8 negdeg);
void aran_development3d_free (AranDevelopment3d *ad);      /* :63 */
void aran_development3d_copy (const AranDevelopment3d *src,
                              AranDevelopment3d *dst);     /* :65 */
AranDevelopment3d *aran_development3d_clone (AranDevelopment3d *src);
void aran_development3d_set_zero (AranDevelopment3d *ad);  /* :70 */
void aran_development3d_write (AranDevelopment3d *ad, FILE *file);

/* P2M (aran/arandevelopment3d.c:169-216) */
void aran_development3d_p2m (const VsgVector3d *position,
                             const gdouble charge,
                             const VsgPRTree3dNodeInfo *dst_node,
                             AranDevelopment3d *dst);

/* translations in the three flavours: O(p^4) "normal" formulas
 * (aran/arandevelopment3d.c:288-331), K. Kylin's recurrences (:343-390) and
 * rotate / z-translate / rotate-back, the operators of record (:402-451) */
void aran_development3d_m2m (const VsgPRTree3dNodeInfo *src_node,
                             AranDevelopment3d *src,
                             const VsgPRTree3dNodeInfo *dst_node,
                             AranDevelopment3d *dst);
void aran_development3d_m2l (const VsgPRTree3dNodeInfo *src_node,
                             AranDevelopment3d *src,
                             const VsgPRTree3dNodeInfo *dst_node,
                             AranDevelopment3d *dst);
void aran_development3d_l2l (const VsgPRTree3dNodeInfo *src_node,
                             AranDevelopment3d *src,
                             const VsgPRTree3dNodeInfo *dst_node,
                             AranDevelopment3d *dst);
void aran_development3d_m2m_kkylin (const VsgPRTree3dNodeInfo *src_node,
                                    AranDevelopment3d *src,
                                    const VsgPRTree3dNodeInfo *dst_node,
                                    AranDevelopment3d *dst);
void aran_development3d_m2l_kkylin (const VsgPRTree3dNodeInfo *src_node,
                                    AranDevelopment3d *src,
                                    const VsgPRTree3dNodeInfo *dst_node,
                                    AranDevelopment3d *dst);
void aran_development3d_l2l_kkylin (const VsgPRTree3dNodeInfo *src_node,
                                    AranDevelopment3d *src,
                                    const VsgPRTree3dNodeInfo *dst_node,
                                    AranDevelopment3d *dst);
void aran_development3d_m2m_rotate (const VsgPRTree3dNodeInfo *src_node,
                                    AranDevelopment3d *src,
                                    const VsgPRTree3dNodeInfo *dst_node,
                                    AranDevelopment3d *dst);
void aran_development3d_m2l_rotate (const VsgPRTree3dNodeInfo *src_node,
                                    AranDevelopment3d *src,
                                    const VsgPRTree3dNodeInfo *dst_node,
                                    AranDevelopment3d *dst);
void aran_development3d_l2l_rotate (const VsgPRTree3dNodeInfo *src_node,
                                    AranDevelopment3d *src,
                                    const VsgPRTree3dNodeInfo *dst_node,
                                    AranDevelopment3d *dst);

/* evaluation of the local part (aran/arandevelopment3d.c:485-552) */
gcomplex128
aran_development3d_local_evaluate (const VsgPRTree3dNodeInfo *devel_node,
                                   AranDevelopment3d *devel,
                                   const VsgVector3d *pos);
void
aran_development3d_local_gradient_evaluate (const VsgPRTree3dNodeInfo *devel_node,
                                            AranDevelopment3d *devel,
                                            const VsgVector3d *pos,
                                            VsgVector3d *grad);
/* semifar operators (aran/arandevelopment3d.h:64, :84, :92, :112;
 * aran/arandevelopment3d.c:227-277, :485-560) */
void aran_development3d_p2l (const VsgVector3d *position,
                             const gdouble charge,
                             const VsgPRTree3dNodeInfo *dst_node,
                             AranDevelopment3d *dst);
gcomplex128
aran_development3d_multipole_evaluate (const VsgPRTree3dNodeInfo *devel_node,
                                       AranDevelopment3d *devel,
                                       const VsgVector3d *pos);
void
aran_development3d_multipole_gradient_evaluate (const VsgPRTree3dNodeInfo *devel_node,
                                                AranDevelopment3d *devel,
                                                const VsgVector3d *pos,
                                                VsgVector3d *grad);
gcomplex128 aran_development3d_m2p (const VsgPRTree3dNodeInfo *devel_node,
                                    AranDevelopment3d *devel,
                                    const VsgVector3d *pos);
void aran_development3d_m2pv (const VsgPRTree3dNodeInfo *devel_node,
                              AranDevelopment3d *devel,
                              const VsgVector3d *pos, VsgVector3d *grad);

gcomplex128 aran_development3d_l2p (const VsgPRTree3dNodeInfo *devel_node,
                                    AranDevelopment3d *devel,
                                    const VsgVector3d *pos);
void aran_development3d_l2pv (const VsgPRTree3dNodeInfo *devel_node,
                              AranDevelopment3d *devel,
                              const VsgVector3d *pos, VsgVector3d *grad);

#ifdef __cplusplus
}
#endif

#endif /* ARAN_DEVELOPMENT3D_H */
