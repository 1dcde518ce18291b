/*
 * aransphericalseriesd.h -- complex spherical series (reference:
 * aran/aransphericalseriesd.h:36-131).
 *
 * A series holds local (positive-degree) terms c_l^m for l = 0..posdeg and
 * multipole (negative-degree) terms for l = 0..negdeg-1, m >= 0 only
 * (c_l^{-m} = (-1)^m conj(c_l^m)); its value at x (relative to the centre,
 * spherical (r, theta, phi)) is
 *   sum_{l<=posdeg} r^l sum_m c_l^m Y_l^m + sum_{l<negdeg} r^-(l+1) sum_m
 *   d_l^m Y_l^m.
 * The struct is opaque, as in the reference (whose private layout,
 * aransphericalseriesd-private.h:25-99, stores the terms right after a
 * 2-byte header, misaligned); here the terms start 16-byte aligned.  Terms
 * are reached through aran_spherical_seriesd_get_term().
 *
 * The translations run on the current HIP device, one launch each: the
 * rotate / z-translate / rotate-back flavour (_rotate), the O(p^4) "normal"
 * formulas (no suffix) and K. Kylin's recurrences (_kkylin), all in FP64.
 */
#ifndef ARAN_SPHERICAL_SERIESD_H
#define ARAN_SPHERICAL_SERIESD_H

#include <stdio.h>

#include <aran/arantypes.h>
#include <vsg/vsgd.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _AranSphericalSeriesd AranSphericalSeriesd;

/* storage (aran/aransphericalseriesd.c:122-298); degrees up to 30 */
AranSphericalSeriesd *aran_spherical_seriesd_new (guint8 posdeg,
                                                  guint8 negdeg);
void aran_spherical_seriesd_free (AranSphericalSeriesd *ass);
void aran_spherical_seriesd_copy (const AranSphericalSeriesd *src,
                                  AranSphericalSeriesd *dst);
AranSphericalSeriesd *
aran_spherical_seriesd_clone (const AranSphericalSeriesd *src);
/* i >= 0: local term (i, j); i < 0: multipole term (-i-1, j); 0 <= j <= |i|
 * (NULL outside; the reference returns a wrong address for j < 0). */
gcomplex128 *aran_spherical_seriesd_get_term (AranSphericalSeriesd *ass,
                                              gint i, gint j);
guint8 aran_spherical_seriesd_get_posdeg (const AranSphericalSeriesd *ass);
guint8 aran_spherical_seriesd_get_negdeg (const AranSphericalSeriesd *ass);
void aran_spherical_seriesd_set_zero (AranSphericalSeriesd *ass);
void aran_spherical_seriesd_write (const AranSphericalSeriesd *ass,
                                   FILE *file);

/* value and gradient at x (relative to the series' centre)
 * (aran/aransphericalseriesd.c:313-546) */
gcomplex128 aran_spherical_seriesd_evaluate (const AranSphericalSeriesd *ass,
                                             const VsgVector3d *x);
void aran_spherical_seriesd_gradient_evaluate (const AranSphericalSeriesd *ass,
                                               const VsgVector3d *x,
                                               VsgVector3d *grad);

/* dst += T src, T the translation from centre xsrc to centre xdst: local
 * terms to local terms and multipole terms to multipole terms (translate),
 * or multipole terms to local terms (to_local)
 * (aran/aransphericalseriesd-translate.c:770-834,
 *  aran/aransphericalseriesd-kkylin.c:516-579,
 *  aran/aransphericalseriesd-rotate.c:165-263) */
void aran_spherical_seriesd_translate (const AranSphericalSeriesd *src,
                                       const VsgVector3d *xsrc,
                                       AranSphericalSeriesd *dst,
                                       const VsgVector3d *xdst);
void aran_spherical_seriesd_to_local (const AranSphericalSeriesd *src,
                                      const VsgVector3d *xsrc,
                                      AranSphericalSeriesd *dst,
                                      const VsgVector3d *xdst);
void aran_spherical_seriesd_translate_kkylin (const AranSphericalSeriesd *src,
                                              const VsgVector3d *xsrc,
                                              AranSphericalSeriesd *dst,
                                              const VsgVector3d *xdst);
void aran_spherical_seriesd_to_local_kkylin (const AranSphericalSeriesd *src,
                                             const VsgVector3d *xsrc,
                                             AranSphericalSeriesd *dst,
                                             const VsgVector3d *xdst);
void aran_spherical_seriesd_translate_rotate (const AranSphericalSeriesd *src,
                                              const VsgVector3d *xsrc,
                                              AranSphericalSeriesd *dst,
                                              const VsgVector3d *xdst);
void aran_spherical_seriesd_to_local_rotate (const AranSphericalSeriesd *src,
                                             const VsgVector3d *xsrc,
                                             AranSphericalSeriesd *dst,
                                             const VsgVector3d *xdst);

#ifdef __cplusplus
}
#endif

#endif /* ARAN_SPHERICAL_SERIESD_H */
