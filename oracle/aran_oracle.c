/*
 * aran_oracle.c -- FP64 serial CPU restatement of LibAran's 3D Newton FMM.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine and the CPU
 * baseline of the benchmark.  Nothing in the engine links or calls it.
 *
 * Each function cites the reference routine it restates.  Parity pins
 * (tests/test_oracle_kat.py): closed forms for Legendre / Y (tests/legendre.c,
 * tests/sphericalharmonic.c), Wigner d against the factorial sum
 * (tests/wigner.c) and mpmath, rotate/rotate_inverse identity
 * (tests/rotation.c), M2L/M2M/L2L/gradient against 1/r and its gradient
 * (tests/m2l3.c, tests/multipole3.c, tests/taylor3.c, tests/gradient3.c), the
 * whole solve against FP64 direct summation, and the tree / traversal against
 * brute-force pair accounting on small inputs.
 *
 * Tree rules (LibVsg's VsgPRTree3d is not in the reference tree, so these are
 * this project's definition, and this file is where they are fixed):
 *   - root cube: centre (lb+ub)/2, half-width max((ub-lb)/2), doubled about
 *     the same centre until every particle lies inside;
 *   - integer coordinates i = floor((x - lo) * 2^21 / width) clamped to
 *     [0, 2^21-1]; a node splits while it holds more than leafcap particles
 *     and its depth is below 21; a particle goes to the upper child of a
 *     split plane when its integer coordinate has that bit set;
 *   - empty children are not created; children are kept in octant order
 *     (octant = bx | by<<1 | bz<<2), nodes in breadth-first order;
 *   - same-depth nodes are far when their integer coordinates differ by 2 or
 *     more along some axis (M2L both ways); near leaves interact by P2P; a
 *     leaf near a deeper subtree interacts with every particle of it by P2P
 *     (LibVsg's near/far traversal with no semifar function,
 *     aran/aransolver3d.c:1189-1190);
 *   - with a semifar threshold t: when the recursion pairs a leaf with a
 *     deeper node whose box does not touch the leaf's box and which holds
 *     more than t particles, the pair is semifar (P2L of the leaf's
 *     particles into the node's local, M2P of the node's multipole onto
 *     them, aran/aransolver3d.c:357-396) and the recursion stops there.
 */
#include "aran_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

#define TRI(l, m) ((l) * ((l) + 1) / 2 + (m))
#define PHASE(m) ((((m) % 2) == 0) ? 1. : -1.)
#define OR_MAXDEG 80

/* ---------------------------------------------------------------- tables */

/* alpha_l^m = 1/sqrt((l-m)!(l+m)!), built with the recurrences of
 * _alpha_generator (aran/aransphericalseriesd.c:45-59). */
static double alpha_tab[TRI (2 * OR_MAXDEG + 2, 0)];
static int alpha_ready = 0;

static void alpha_init (void)
{
  int l, m;
  for (l = 0; l <= 2 * OR_MAXDEG + 1; l++)
    for (m = 0; m <= l; m++)
      {
        double v;
        if (l == 0)
          v = 1.;
        else if (m == 0)
          v = alpha_tab[TRI (l - 1, 0)] / l;
        else if (l == m)
          v = alpha_tab[TRI (l - 1, l - 1)] / sqrt ((l + l - 1.) * (l + l));
        else
          v = sqrt ((l - m + 1.) / (l + m)) * alpha_tab[TRI (l, m - 1)];
        alpha_tab[TRI (l, m)] = v;
      }
  alpha_ready = 1;
}

double or_alpha (int l, int m)
{
  if (!alpha_ready)
    alpha_init ();
  return alpha_tab[TRI (l, m)];
}

/* beta_l = sqrt(4 pi / (2l+1)) (aran/aransphericalseriesd.c:40-43) */
double or_beta (int l)
{
  return sqrt ((4. * M_PI) / (l + l + 1.));
}

/* Associated Legendre P_l^m with Condon-Shortley phase, all 0<=m<=l<=L
 * (aran/aranlegendre.c:135-169). */
void or_legendre (int L, double x, double s, double *P)
{
  int i, j;
  P[0] = 1.;
  if (L < 1)
    return;
  P[TRI (1, 0)] = x;
  P[TRI (1, 1)] = -s;
  for (i = 2; i <= L; i++)
    {
      double *p = P + TRI (i, 0);
      const double *p1 = P + TRI (i - 1, 0);
      const double *p2 = P + TRI (i - 2, 0);
      for (j = 0; j < i - 1; j++)
        p[j] = ((i + i - 1) * x * p1[j] - (i + j - 1) * p2[j]) / (i - j);
      p[i - 1] = (i + i - 1) * x * p1[i - 1];
      p[i] = -(i + i - 1) * s * p1[i - 1];
    }
}

/* P_l^m / sin(theta) without the division for m >= 1
 * (aran/aranlegendre.c:204-264). */
void or_legendre_special (int L, double x, double s, const double *P,
                          double *S)
{
  int i, j;
  double factor = (s != 0.) ? 1. / s : 1.7976931348623157e308;

  S[0] = P[0] * factor;
  if (L < 1)
    return;
  S[TRI (1, 1)] = -P[0];
  S[TRI (1, 0)] = P[TRI (1, 0)] * factor;
  for (i = 2; i <= L; i++)
    {
      double *sp = S + TRI (i, 0);
      const double *s1 = S + TRI (i - 1, 0);
      const double *s2 = S + TRI (i - 2, 0);
      sp[i] = -(i + i - 1.) * P[TRI (i - 1, i - 1)];
      sp[0] = P[TRI (i, 0)] * factor;
      for (j = 1; j < i - 1; j++)
        sp[j] = ((i + i - 1.) * x * s1[j] - (i + j - 1.) * s2[j]) / (i - j);
      sp[i - 1] = (i + i - 1.) * x * s1[i - 1];
    }
}

/* N_l^m = sqrt((2l+1)/(4pi) (l-m)!/(l+m)!) by the recurrence of
 * _spherical_generator (aran/aransphericalharmonic.c:31-38). */
static double harm_norm (int l, int m)
{
  static double tab[TRI (OR_MAXDEG + 1, 0)];
  static int ready = 0;
  if (!ready)
    {
      int a, b;
      for (a = 0; a <= OR_MAXDEG; a++)
        for (b = 0; b <= a; b++)
          tab[TRI (a, b)] = (b == 0) ? sqrt ((a + a + 1.) / (4. * M_PI))
                                     : tab[TRI (a, b - 1)]
                                         / sqrt ((a - b + 1.) * (a + b));
      ready = 1;
    }
  return tab[TRI (l, m)];
}

/* Y_l^m for 0<=m<=l<=L (aran/aransphericalharmonic.c:149-184) */
void or_harmonics (int L, double cost, double sint, double cosp, double sinp,
                   double *Yd)
{
  cplx *Y = (cplx *) Yd;
  double P[TRI (OR_MAXDEG + 1, 0)];
  cplx expp = cosp + I * sinp, pw = 1., pows[OR_MAXDEG + 1];
  int i, j;
  or_legendre (L, cost, sint, P);
  for (i = 0; i <= L; i++)
    {
      pows[i] = pw;
      for (j = 0; j <= i; j++)
        Y[TRI (i, j)] = pows[j] * P[TRI (i, j)] * harm_norm (i, j);
      pw *= expp;
    }
}

/* Y and (P/sin) * N * e^{im phi} (aran/aransphericalharmonic.c:228-276) */
static void harmonics_pregrad (int L, double cost, double sint, double cosp,
                               double sinp, cplx *Y, cplx *SY)
{
  double P[TRI (OR_MAXDEG + 1, 0)], S[TRI (OR_MAXDEG + 1, 0)];
  cplx expp = cosp + I * sinp, pw = 1., pows[OR_MAXDEG + 1];
  int i, j;
  or_legendre (L, cost, sint, P);
  or_legendre_special (L, cost, sint, P, S);
  for (i = 0; i <= L; i++)
    {
      pows[i] = pw;
      for (j = 0; j <= i; j++)
        {
          cplx t = pows[j] * harm_norm (i, j);
          Y[TRI (i, j)] = t * P[TRI (i, j)];
          SY[TRI (i, j)] = t * S[TRI (i, j)];
        }
      pw *= expp;
    }
}

/* Spherical coordinates of v: theta from +z, phi from +x.  On the z axis
 * phi is taken as 0 (cos 1, sin 0); at the origin theta is 0 too. */
void or_to_spherical (const double *v, double *r, double *cost, double *sint,
                      double *cosp, double *sinp)
{
  double rho2 = v[0] * v[0] + v[1] * v[1];
  double rr = sqrt (rho2 + v[2] * v[2]);
  double rho = sqrt (rho2);
  *r = rr;
  if (rr > 0.)
    {
      *cost = v[2] / rr;
      *sint = rho / rr;
    }
  else
    {
      *cost = 1.;
      *sint = 0.;
    }
  if (rho > 0.)
    {
      *cosp = v[0] / rho;
      *sinp = v[1] / rho;
    }
  else
    {
      *cosp = 1.;
      *sinp = 0.;
    }
}

/* ---------------------------------------------------------------- Wigner */

static long wofs (int l)
{
  return ((long) l * (l + 1) * (4 * l - 1)) / 6;
}

long or_wigner_size (int L)
{
  return ((long) (L + 1) * (L + 2) * (4 * L + 3)) / 6;
}

#define WD(d, l, mp, m) (d)[wofs (l) + (long) (mp) * (2 * (l) + 1) + (l) + (m)]

/* d^l_{m,m'}(beta) for 0<=l<=L, 0<=m'<=l, -l<=m<=l, by the l-recursion of
 * Blanco, Florez and Bermejo (J. Mol. Struct. Theochem 419, 1997) as
 * arranged in aran_wigner_require_d (aran/aranwigner.c:83-246). */
void or_wigner_d (int L, double beta, double *d)
{
  int l, mp, m;
  double cb = cos (beta), sb = sin (beta);
  double cb2 = cos (beta * 0.5), sb2 = sin (beta * 0.5);
  double tb2 = sb2 / cb2;
  double d100, d111, d11m1;

  WD (d, 0, 0, 0) = 1.;
  if (L == 0)
    return;
  WD (d, 1, 0, 0) = cb;
  WD (d, 1, 1, -1) = sb2 * sb2;
  WD (d, 1, 1, 0) = sb / sqrt (2.);
  WD (d, 1, 1, 1) = cb2 * cb2;
  WD (d, 1, 0, -1) = WD (d, 1, 1, 0);
  WD (d, 1, 0, 1) = -WD (d, 1, 1, 0);
  if (L <= 1)
    return;
  d100 = WD (d, 1, 0, 0);
  d111 = WD (d, 1, 1, 1);
  d11m1 = WD (d, 1, 1, -1);

  for (l = 2; l <= L; l++)
    {
      double tl1 = l + l - 1., sql = (double) l * l;
      double sql1 = (l - 1.) * (l - 1.);
      /* three-term recursion for |m| <= m' <= l-2 */
      for (mp = 0; mp <= l - 2; mp++)
        for (m = -mp; m <= mp; m++)
          {
            double a = (l * tl1) / sqrt ((sql - mp * mp) * (sql - m * m));
            double b = d100 - (mp * m) / (l * (l - 1.));
            double c = sqrt ((sql1 - mp * mp) * (sql1 - m * m))
                       / ((l - 1.) * tl1);
            WD (d, l, mp, m) = a * (b * WD (d, l - 1, mp, m)
                                    - c * WD (d, l - 2, mp, m));
          }
      /* corner terms */
      WD (d, l, l, l) = d111 * WD (d, l - 1, l - 1, l - 1);
      WD (d, l, l - 1, l - 1) = (l * d100 - l + 1.) * WD (d, l - 1, l - 1, l - 1);
      WD (d, l, l, -l) = d11m1 * WD (d, l - 1, l - 1, -(l - 1));
      WD (d, l, l - 1, -(l - 1))
        = (l * d100 + l - 1.) * WD (d, l - 1, l - 1, -(l - 1));
      /* last column, descending in m' */
      for (mp = l; mp >= 1; mp--)
        WD (d, l, mp - 1, l) = -sqrt ((l + mp) / (l - mp + 1.)) * tb2
                               * WD (d, l, mp, l);
      /* penultimate column */
      for (mp = l - 1; mp >= 1; mp--)
        {
          double a = sqrt (((double) l + mp) / (((double) l + l) * (l - mp + 1.)));
          WD (d, l, mp - 1, l - 1) = (l * cb - mp + 1.) * a * WD (d, l, mp, l)
                                     / d111;
        }
      /* last two rows by symmetry */
      for (mp = l - 1; mp <= l; mp++)
        for (m = 0; m < mp; m++)
          WD (d, l, mp, m) = PHASE (mp + m) * WD (d, l, m, mp);
      /* negative m of the last row */
      for (m = 0; m < l; m++)
        WD (d, l, l, -m - 1) = sqrt ((l - m) / (l + m + 1.)) * tb2
                               * WD (d, l, l, -m);
      for (m = 0; m < l; m++)
        {
          double a = sqrt (((double) l - m) / (((double) l + l) * (l + m + 1.)));
          WD (d, l, l - 1, -m - 1) = (l * cb + m + 1.) * a * WD (d, l, l, -m)
                                     / d111;
        }
      /* remaining entries by symmetry */
      for (mp = 0; mp <= l; mp++)
        for (m = mp + 1; m <= l; m++)
          {
            WD (d, l, mp, m) = PHASE (m + mp) * WD (d, l, m, mp);
            WD (d, l, mp, -m) = WD (d, l, m, -mp);
          }
    }
}

/* Wigner's explicit sum in log-factorials (tests/wigner.c:104-145):
 * the independent check of or_wigner_d. */
double or_wigner_factorial (int l, int m1, int m2, double beta)
{
  double ret = 0., cb2 = cos (beta * 0.5), sb2 = sin (beta * 0.5), a;
  int k, kmin, kmax;
  int lmm1 = l - m1, lpm1 = l + m1, lmm2 = l - m2, lpm2 = l + m2;
  int m1mm2 = m1 - m2;
  kmin = (0 > -m1mm2) ? 0 : -m1mm2;
  kmax = (lmm1 < lpm2) ? lmm1 : lpm2;
  if (kmin > kmax)
    return 0.;
  a = 0.5 * (lgamma (lpm1 + 1.) + lgamma (lmm1 + 1.) + lgamma (lpm2 + 1.)
             + lgamma (lmm2 + 1.));
  for (k = kmin; k <= kmax; k++)
    {
      double b = lgamma (k + 1.) + lgamma (lmm1 - k + 1.)
                 + lgamma (lpm2 - k + 1.) + lgamma (m1mm2 + k + 1.);
      ret += PHASE (m1mm2 + k) * exp (a - b) * pow (cb2, lmm1 + lpm2 - 2 * k)
             * pow (sb2, m1mm2 + 2 * k);
    }
  return ret;
}

/* D^l_{m,m'} = e^{-i m alpha} d^l_{m,m'}(beta) e^{-i m' gamma}, phases
 * skipped below 1e-5 rad as in aran_wigner_require (aran/aranwigner.c:
 * 256-313). */
void or_wigner_D (int L, double alpha, double beta, double gamma, double *Dd)
{
  cplx *D = (cplx *) Dd;
  long n = or_wigner_size (L), i;
  double *d = malloc (sizeof (double) * n);
  int l, mp, m;
  or_wigner_d (L, beta, d);
  for (i = 0; i < n; i++)
    D[i] = d[i];
  free (d);
  if (fabs (alpha) >= 1.e-5)
    {
      cplx expa = cos (alpha) - I * sin (alpha), e[OR_MAXDEG + 2];
      e[0] = 1.;
      for (m = 1; m <= L; m++)
        e[m] = e[m - 1] * expa;
      for (l = 0; l <= L; l++)
        for (mp = 0; mp <= l; mp++)
          {
            for (m = l; m >= 1; m--)
              WD (D, l, mp, -m) *= conj (e[m]);
            for (m = 0; m <= l; m++)
              WD (D, l, mp, m) *= e[m];
          }
    }
  if (fabs (gamma) >= 1.e-5)
    {
      cplx expg = cos (gamma) - I * sin (gamma), e[OR_MAXDEG + 2];
      e[0] = 1.;
      for (m = 1; m <= L; m++)
        e[m] = e[m - 1] * expg;
      for (l = 0; l <= L; l++)
        for (mp = 0; mp <= l; mp++)
          for (m = -l; m <= l; m++)
            WD (D, l, mp, m) *= e[mp];
    }
}

static inline cplx sph_sym (cplx z, int m)
{
  z = conj (z);
  return (m % 2 != 0) ? -z : z;
}

/* dst[l][m'] += sum_m D^l_{m,m'} src[l][m] for l <= deg
 * (_buffer_rotate, aran/aransphericalseriesd-rotate.c:34-80; arrays are
 * sized l+1 here, where the reference's VLAs are one short). */
static void buffer_rotate (const cplx *D, int deg, const cplx *src, cplx *dst)
{
  int l, mp, m;
  for (l = 0; l <= deg; l++)
    {
      const cplx *s = src + TRI (l, 0);
      for (mp = 0; mp <= l; mp++)
        {
          cplx sum = 0.;
          for (m = -l; m < 0; m++)
            sum += WD (D, l, mp, m) * sph_sym (s[-m], -m);
          for (m = 0; m <= l; m++)
            sum += WD (D, l, mp, m) * s[m];
          dst[TRI (l, mp)] += sum;
        }
    }
}

/* aran_spherical_seriesd_rotate (aran/aransphericalseriesd-rotate.c:93-115) */
void or_rotate (int deg, const double *src, double alpha, double beta,
                double gamma, double *dst)
{
  double *D = malloc (sizeof (double) * 2 * or_wigner_size (deg));
  or_wigner_D (deg, alpha, beta, gamma, D);
  buffer_rotate ((cplx *) D, deg, (const cplx *) src, (cplx *) dst);
  free (D);
}

/* aran_spherical_seriesd_rotate_inverse (rotate.c:130-153): the Wigner
 * entry of (-gamma, -beta, -alpha). */
void or_rotate_inverse (int deg, const double *src, double alpha,
                        double beta, double gamma, double *dst)
{
  or_rotate (deg, src, -gamma, -beta, -alpha, dst);
}

/* ----------------------------------------------------- vertical operators */

/* aran_spherical_seriesd_multipole_to_local_vertical
 * (aran/aransphericalseriesd-translate.c:609-676) with the factor table of
 * _precomputed_translate_vertical_generator (translate.c:257-263). */
static void m2l_vertical (int nm, const cplx *src, int pl, cplx *dst,
                          double r, double c)
{
  int d = pl + nm, l, m, n;
  double rpow[2 * OR_MAXDEG + 4], pw = 1., invr = 1. / r;
  for (l = 0; l <= d; l++)
    {
      rpow[l] = pw;
      if (l % 2 == 0)
        rpow[l] *= c;
      pw *= invr;
    }
  for (l = 0; l <= pl; l++)
    for (m = 0; m <= l; m++)
      {
        cplx sum = 0.;
        for (n = m; n < nm; n++)
          {
            double t = or_beta (l) / or_beta (n) * or_alpha (l, m)
                       * or_alpha (n, m) / or_alpha (l + n, 0);
            sum += src[TRI (n, m)] * (rpow[l + n + 1] * t);
          }
        if ((l + m) % 2 != 0)
          sum = -sum;
        dst[TRI (l, m)] += sum;
      }
}

/* aran_multipole_translate_vertical (translate.c:445-505) */
static void m2m_vertical (int nms, const cplx *src, int nmd, cplx *dst,
                          double r, double cost)
{
  int l, m, n;
  double rpow[OR_MAXDEG + 2], pw = 1.;
  for (l = 0; l < nmd; l++)
    {
      rpow[l] = pw;
      pw *= r;
      for (m = 0; m <= l; m++)
        {
          int nmax = (l < nms - 1) ? l : nms - 1;
          for (n = m; n <= nmax; n++)
            {
              double normaliz = or_beta (l) / or_beta (n);
              double factor = or_alpha (l - n, 0) * or_alpha (n, m)
                              / or_alpha (l, m);
              double h = ((l - n) % 2 == 0) ? 1. : cost;
              dst[TRI (l, m)] += h * src[TRI (n, m)] * factor * normaliz
                                 * rpow[l - n];
            }
        }
    }
}

/* aran_local_translate_vertical (translate.c:301-364) */
static void l2l_vertical (int pls, const cplx *src, int pld, cplx *dst,
                          double r, double cost)
{
  int l, m, n;
  double rpow[OR_MAXDEG + 2], pw = 1.;
  for (l = 0; l <= pls; l++)
    {
      rpow[l] = pw;
      pw *= r;
    }
  for (l = 0; l <= pld; l++)
    for (m = 0; m <= l; m++)
      for (n = l; n <= pls; n++)
        {
          double normaliz = or_beta (l) / or_beta (n);
          double factor = or_alpha (n - l, 0) * or_alpha (l, m)
                          / or_alpha (n, m);
          double h = ((n - l) % 2 == 0) ? 1. : cost;
          dst[TRI (l, m)] += h * src[TRI (n, m)] * factor * normaliz
                             * rpow[n - l];
        }
}

/* translation direction, z flip and rotation angles
 * (aran/aransphericalseriesd-rotate.c:186-197) */
static void direction (const double *xsrc, const double *xdst, double *r,
                       double *theta, double *phi, double *c)
{
  double dir[3], cost, sint, cosp, sinp;
  int k;
  *c = 1.;
  for (k = 0; k < 3; k++)
    dir[k] = xdst[k] - xsrc[k];
  if (dir[2] < 0.)
    {
      *c = -1.;
      for (k = 0; k < 3; k++)
        dir[k] = -dir[k];
    }
  or_to_spherical (dir, r, &cost, &sint, &cosp, &sinp);
  *theta = atan2 (sint, cost);
  *phi = atan2 (sinp, cosp);
}

/* aran_spherical_seriesd_to_local_rotate (rotate.c:165-207) */
void or_m2l_rotate (int nm, const double *M, const double *xsrc, int pl,
                    double *L, const double *xdst)
{
  double r, theta, phi, c;
  int lm = (nm - 1 > pl) ? nm - 1 : pl;
  cplx *rot = calloc (TRI (lm + 1, 0), sizeof (cplx));
  cplx *trans = calloc (TRI (lm + 1, 0), sizeof (cplx));
  direction (xsrc, xdst, &r, &theta, &phi, &c);
  if (nm > 0)
    or_rotate (nm - 1, M, -phi, theta, 0., (double *) rot);
  m2l_vertical (nm, rot, pl, trans, r, c);
  or_rotate_inverse (pl, (double *) trans, -phi, theta, 0., L);
  free (rot);
  free (trans);
}

/* aran_spherical_seriesd_translate_rotate on multipole series
 * (rotate.c:219-263, translate.c:581-594): vertical step with -cost. */
void or_m2m_rotate (int nm, const double *Msrc, const double *xsrc, int nmdst,
                    double *Mdst, const double *xdst)
{
  double r, theta, phi, c;
  int lm = (nm > nmdst) ? nm : nmdst;
  cplx *rot = calloc (TRI (lm + 1, 0), sizeof (cplx));
  cplx *trans = calloc (TRI (lm + 1, 0), sizeof (cplx));
  direction (xsrc, xdst, &r, &theta, &phi, &c);
  if (nm > 0)
    or_rotate (nm - 1, Msrc, -phi, theta, 0., (double *) rot);
  m2m_vertical (nm, rot, nmdst, trans, r, -c);
  if (nmdst > 0)
    or_rotate_inverse (nmdst - 1, (double *) trans, -phi, theta, 0., Mdst);
  free (rot);
  free (trans);
}

/* translate_rotate on local series (rotate.c:219-263, translate.c:301-364) */
void or_l2l_rotate (int pl, const double *Lsrc, const double *xsrc, int pldst,
                    double *Ldst, const double *xdst)
{
  double r, theta, phi, c;
  int lm = (pl > pldst) ? pl : pldst;
  cplx *rot = calloc (TRI (lm + 1, 0), sizeof (cplx));
  cplx *trans = calloc (TRI (lm + 1, 0), sizeof (cplx));
  direction (xsrc, xdst, &r, &theta, &phi, &c);
  or_rotate (pl, Lsrc, -phi, theta, 0., (double *) rot);
  l2l_vertical (pl, rot, pldst, trans, r, c);
  or_rotate_inverse (pldst, (double *) trans, -phi, theta, 0., Ldst);
  free (rot);
  free (trans);
}

/* ---------------------------------------- alternate translation flavours
 *
 * The O(p^4) "normal" translations (aran/aransphericalseriesd-translate.c:
 * 366-443, 507-579, 678-834) and K. Kylin's partial-sum recurrences
 * (aran/aransphericalseriesd-kkylin.c:34-579).  They map the same truncated
 * expansions as the rotate flavour, by other arithmetic.  The translation
 * vector is xdst - xsrc in spherical coordinates (or_to_spherical); the
 * multipole translations take the opposite vector (r, -cost, sint, -cosp,
 * -sinp), as the reference's translate wrappers pass it
 * (translate.c:784-793, kkylin.c:540-549). */

/* Y_l^m for any sign of m from the m >= 0 table (_sph_sym). */
static inline cplx harm_any (const cplx *Y, int l, int m)
{
  return (m >= 0) ? Y[TRI (l, m)] : sph_sym (Y[TRI (l, -m)], -m);
}

/* coefficient (l, m) of a series for any sign of m */
static inline cplx coef_any (const cplx *S, int l, int m)
{
  return (m >= 0) ? S[TRI (l, m)] : sph_sym (S[TRI (l, -m)], -m);
}

/* aran_local_translate (translate.c:366-443) */
static void l2l_normal (int pls, const cplx *src, int pld, cplx *dst,
                        double r, double cost, double sint, double cosp,
                        double sinp)
{
  int d = (pls > pld) ? pls : pld, l, m, n, o;
  cplx *Y = malloc (sizeof (cplx) * TRI (d + 1, 0));
  double rpow[OR_MAXDEG + 2], pw = 1.;
  or_harmonics (pls, cost, sint, cosp, sinp, (double *) Y);
  for (l = 0; l <= pls; l++)
    {
      rpow[l] = pw;
      pw *= r;
    }
  for (l = 0; l <= pld; l++)
    for (m = 0; m <= l; m++)
      for (n = l; n <= pls; n++)
        {
          double normaliz = or_beta (n - l) * (or_beta (l) / or_beta (n));
          cplx sum = 0.;
          for (o = l + m - n; o <= m + n - l; o++)
            {
              double factor = or_alpha (n - l, abs (o - m)) * or_alpha (l, m)
                              / or_alpha (n, abs (o));
              sum += harm_any (Y, n - l, o - m) * coef_any (src, n, o)
                     * factor;
            }
          dst[TRI (l, m)] += sum * (normaliz * rpow[n - l]);
        }
  free (Y);
}

/* aran_multipole_translate (translate.c:507-579); (cost, cosp, sinp) are
 * those of the opposite vector */
static void m2m_normal (int nms, const cplx *src, int nmd, cplx *dst,
                        double r, double cost, double sint, double cosp,
                        double sinp)
{
  int d = ((nms > nmd) ? nms : nmd) - 1, l, m, n, o;
  cplx *Y = malloc (sizeof (cplx) * TRI (d + 2, 0));
  double rpow[OR_MAXDEG + 2], pw = 1.;
  or_harmonics (nmd - 1, cost, sint, cosp, sinp, (double *) Y);
  for (l = 0; l < nmd; l++)
    {
      rpow[l] = pw;
      pw *= r;
    }
  for (l = 0; l < nmd; l++)
    for (m = 0; m <= l; m++)
      {
        int nmax = (l < nms - 1) ? l : nms - 1;
        for (n = 0; n <= nmax; n++)
          {
            double normaliz = or_beta (l - n) * (or_beta (l) / or_beta (n));
            int olo = (-n > m + n - l) ? -n : m + n - l;
            int ohi = (n < l + m - n) ? n : l + m - n;
            cplx sum = 0.;
            for (o = olo; o <= ohi; o++)
              {
                double factor = or_alpha (l - n, abs (m - o))
                                * or_alpha (n, abs (o)) / or_alpha (l, m);
                sum += conj (harm_any (Y, l - n, m - o)) * coef_any (src, n, o)
                       * factor;
              }
            dst[TRI (l, m)] += sum * normaliz * rpow[l - n];
          }
      }
  free (Y);
}

/* aran_multipole_to_local (translate.c:678-768), with the vertical M2L of
 * the reference when |sin theta| < 1e-5 (translate.c:695-701, 609-676:
 * r^-(k) times cos theta for even k, phi unused) */
static void m2l_vertical_cost (int nm, const cplx *src, int pl, cplx *dst,
                               double r, double cost)
{
  m2l_vertical (nm, src, pl, dst, r, cost);
}

static void m2l_normal (int nm, const cplx *src, int pl, cplx *dst, double r,
                        double cost, double sint, double cosp, double sinp)
{
  int d = pl + nm, l, m, n, o;
  cplx *Y;
  double rpow[2 * OR_MAXDEG + 4], pw = 1., invr = 1. / r, sign = 1.;
  if (fabs (sint) < 1.e-5)
    {
      m2l_vertical_cost (nm, src, pl, dst, r, cost);
      return;
    }
  Y = malloc (sizeof (cplx) * TRI (d + 1, 0));
  or_harmonics (d, cost, sint, cosp, sinp, (double *) Y);
  for (l = 0; l <= d; l++)
    {
      rpow[l] = pw;
      pw *= invr;
    }
  for (l = 0; l <= pl; l++)
    {
      for (m = 0; m <= l; m++)
        for (n = 0; n < nm; n++)
          {
            double normaliz = or_beta (l + n) * (or_beta (l) / or_beta (n));
            const cplx *S = src + TRI (n, 0);
            const cplx *h = Y + TRI (l + n, 0);
            cplx sum = or_alpha (l, m) * or_alpha (n, 0)
                       / or_alpha (l + n, m) * h[m] * S[0];
            for (o = 1; o <= n; o++)
              {
                double factor = or_alpha (l, m) * or_alpha (n, o);
                sum += h[m + o] * S[o] * factor / or_alpha (l + n, m + o);
                sum += harm_any (Y, l + n, m - o) * sph_sym (S[o], o) * factor
                       / or_alpha (l + n, abs (m - o));
              }
            dst[TRI (l, m)] += conj (sum) * sign * normaliz * rpow[l + n + 1];
          }
      sign = -sign;
    }
  free (Y);
}

/* aran_multipole_translate_kkylin (kkylin.c:34-205); (cost, cosp, sinp)
 * are those of the opposite vector; degrees of src and dst equal */
static void m2m_kylin (int nms, const cplx *src, int nmd, cplx *dst,
                       double r, double cost, double sint, double cosp,
                       double sinp)
{
  int d = ((nms > nmd) ? nms : nmd) - 1, l, m, n, o;
  int w = d + 1;
  cplx *Y = malloc (sizeof (cplx) * TRI (d + 2, 0));
  cplx *b1 = calloc ((size_t) w * w, sizeof (cplx));
  cplx *b2 = calloc ((size_t) w * w, sizeof (cplx));
  cplx *part = b1, *prev = b2, *t, expp = cosp + I * sinp, fse;
  double rpow[OR_MAXDEG + 2], pw = 1., fact;
#define PT(a, mm, nn) (a)[(mm) * w + (nn)]
  or_harmonics (nmd - 1, cost, sint, cosp, sinp, (double *) Y);
  rpow[0] = pw;
  pw *= r;
  PT (part, 0, 0) = conj (Y[0]) * src[0];
  dst[0] += PT (part, 0, 0) * or_beta (0);
  for (l = 1; l < nmd; l++)
    {
      rpow[l] = pw;
      pw *= r;
      t = part;
      part = prev;
      prev = t;
      for (n = 0; n <= l - 1; n++)
        {
          fact = or_beta (l - n) / or_beta (n);
          PT (part, 0, n) = 0.;
          for (o = (-n > n - l ? -n : n - l); o <= (n < l - n ? n : l - n);
               o++)
            PT (part, 0, n) += (rpow[l - n] * or_alpha (n, abs (o))
                                * or_alpha (l - n, abs (o)))
                               * conj (harm_any (Y, l - n, -o))
                               * coef_any (src, n, o);
          PT (part, 0, n) *= fact;
          if (l > 1)
            {
              PT (part, l - 1, n) = 0.;
              for (o = (-n > n - 1 ? -n : n - 1);
                   o <= (n < l + l - 1 - n ? n : l + l - 1 - n); o++)
                PT (part, l - 1, n) +=
                    (rpow[l - n] * or_alpha (n, abs (o))
                     * or_alpha (l - n, abs (l - o - 1)))
                    * conj (harm_any (Y, l - n, l - o - 1))
                    * coef_any (src, n, o);
              PT (part, l - 1, n) *= fact;
            }
          PT (part, l, n) = 0.;
          for (o = (-n > n ? -n : n); o <= (n < l + l - n ? n : l + l - n);
               o++)
            PT (part, l, n) += (rpow[l - n] * or_alpha (n, abs (o))
                                * or_alpha (l - n, abs (l - o)))
                               * conj (harm_any (Y, l - n, l - o))
                               * coef_any (src, n, o);
          PT (part, l, n) *= fact;
          fact = 0.5 * r / (l - n);
          fse = fact * sint * expp;
          fact *= 2. * cost;
          for (m = 1; m < l - 1; m++)
            PT (part, m, n) = fse * PT (prev, m + 1, n)
                              - conj (fse) * PT (prev, m - 1, n)
                              + fact * PT (prev, m, n);
        }
      fact = or_beta (0) / or_beta (l);
      for (m = 0; m <= l; m++)
        PT (part, m, l) = (or_alpha (l, m) * fact) * conj (Y[0])
                          * src[TRI (l, m)];
      for (m = 0; m <= l; m++)
        {
          cplx sum = 0.;
          for (n = 0; n <= l; n++)
            sum += PT (part, m, n);
          dst[TRI (l, m)] += sum * (or_beta (l) / or_alpha (l, m));
        }
    }
#undef PT
  free (Y);
  free (b1);
  free (b2);
}

/* aran_local_translate_kkylin (kkylin.c:207-364); src degree <= dst */
static void l2l_kylin (int pls, const cplx *src, int pld, cplx *dst,
                       double r, double cost, double sint, double cosp,
                       double sinp)
{
  int d = (pls > pld) ? pls : pld, l, m, n, o;
  int w = d + 1;
  cplx *Y = malloc (sizeof (cplx) * TRI (d + 1, 0));
  cplx *b1 = calloc ((size_t) w * w, sizeof (cplx));
  cplx *b2 = calloc ((size_t) w * w, sizeof (cplx));
  cplx *part = b1, *prev = b2, *t, expp = cosp + I * sinp, fse, y00;
  double rpow[OR_MAXDEG + 2], pw = 1., fact;
#define PT(a, mm, nn) (a)[(mm) * w + (nn)]
  or_harmonics (pls, cost, sint, cosp, sinp, (double *) Y);
  y00 = Y[0];
  for (l = 0; l <= pls; l++)
    {
      rpow[l] = pw;
      pw *= r;
    }
  for (l = pld; l >= 0; l--)
    {
      if (l <= pls)
        {
          fact = or_beta (0) / or_beta (l);
          for (m = 0; m <= l; m++)
            PT (part, m, l) = fact / or_alpha (l, m)
                              * conj (y00 * src[TRI (l, m)]);
        }
      for (n = l + 1; n <= pls; n++)
        {
          fact = or_beta (n - l) / or_beta (n);
          PT (part, 0, n) = 0.;
          for (o = (-n > l - n ? -n : l - n); o <= (n < n - l ? n : n - l);
               o++)
            PT (part, 0, n) += (rpow[n - l] * or_alpha (n - l, abs (o))
                                / or_alpha (n, abs (o)))
                               * conj (harm_any (Y, n - l, o)
                                       * coef_any (src, n, o));
          PT (part, 0, n) *= fact;
          if (l > 0)
            {
              PT (part, l, n) = 0.;
              for (o = (-n > l + l - n ? -n : l + l - n); o <= n; o++)
                PT (part, l, n) += (rpow[n - l]
                                    * or_alpha (n - l, abs (o - l))
                                    / or_alpha (n, abs (o)))
                                   * conj (harm_any (Y, n - l, o - l)
                                           * coef_any (src, n, o));
              PT (part, l, n) *= fact;
            }
          fact = 0.5 * r / (n - l);
          fse = fact * sint * expp;
          fact *= 2. * cost;
          for (m = 1; m < l; m++)
            PT (part, m, n) = fse * PT (prev, m - 1, n)
                              - conj (fse) * PT (prev, m + 1, n)
                              + fact * PT (prev, m, n);
        }
      for (m = 0; m <= l; m++)
        {
          cplx sum = 0.;
          for (n = l; n <= pls; n++)
            sum += PT (part, m, n);
          dst[TRI (l, m)] += conj (sum) * (or_beta (l) * or_alpha (l, m));
        }
      t = part;
      part = prev;
      prev = t;
    }
#undef PT
  free (Y);
  free (b1);
  free (b2);
}

/* aran_multipole_to_local_kkylin (kkylin.c:364-501) */
static void m2l_kylin (int nm, const cplx *src, int pl, cplx *dst, double r,
                       double cost, double sint, double cosp, double sinp)
{
  int d = pl + nm, l, m, n, o;
  int w = d + 1;
  cplx *Y, *b1, *b2, *part, *prev, *t, expp = cosp + I * sinp;
  cplx tce, e2, tl;
  double rpow[2 * OR_MAXDEG + 4], pw = 1., invr = 1. / r, sign = 1., fact;
  if (fabs (sint) < 1.e-5)
    {
      m2l_vertical_cost (nm, src, pl, dst, r, cost);
      return;
    }
  Y = malloc (sizeof (cplx) * TRI (d + 1, 0));
  b1 = calloc ((size_t) w * w, sizeof (cplx));
  b2 = calloc ((size_t) w * w, sizeof (cplx));
  part = b1;
  prev = b2;
#define PT(a, mm, nn) (a)[(mm) * w + (nn)]
  or_harmonics (d, cost, sint, cosp, sinp, (double *) Y);
  for (l = 0; l <= d; l++)
    {
      rpow[l] = pw;
      pw *= invr;
    }
  for (l = 0; l <= pl; l++)
    {
      for (n = 0; n < nm; n++)
        {
          fact = rpow[l + n + 1] * or_beta (l + n) / or_beta (n);
          PT (part, 0, n) = 0.;
          for (o = -n; o <= n; o++)
            PT (part, 0, n) += (or_alpha (n, abs (o))
                                / or_alpha (l + n, abs (o)))
                               * harm_any (Y, l + n, o) * coef_any (src, n, o);
          PT (part, 0, n) *= fact;
          if (l > 0)
            {
              PT (part, 1, n) = 0.;
              for (o = -n; o <= n; o++)
                PT (part, 1, n) += (or_alpha (n, abs (o))
                                    / or_alpha (l + n, abs (o + 1)))
                                   * harm_any (Y, l + n, o + 1)
                                   * coef_any (src, n, o);
              PT (part, 1, n) *= fact;
            }
          tce = 2. * cost / sint * expp;
          e2 = expp * expp;
          tl = 2. * (l + n) * invr / sint * expp;
          for (m = 2; m <= l; m++)
            PT (part, m, n) = tce * PT (part, m - 1, n)
                              + e2 * PT (part, m - 2, n)
                              - tl * PT (prev, m - 1, n);
        }
      fact = or_beta (l) * sign;
      for (m = 0; m <= l; m++)
        {
          cplx sum = 0.;
          for (n = 0; n < nm; n++)
            sum += PT (part, m, n);
          dst[TRI (l, m)] += conj (sum) * (or_alpha (l, m) * fact);
        }
      t = part;
      part = prev;
      prev = t;
      sign = -sign;
    }
#undef PT
  free (Y);
  free (b1);
  free (b2);
}

/* Public entry points: one part of a series, translated by xdst - xsrc and
 * added into dst (flavour 1: normal, 2: Kylin). */
static void flav_vector (const double *xsrc, const double *xdst, int opposite,
                         double *r, double *cost, double *sint, double *cosp,
                         double *sinp)
{
  double v[3];
  int k;
  for (k = 0; k < 3; k++)
    v[k] = xdst[k] - xsrc[k];
  or_to_spherical (v, r, cost, sint, cosp, sinp);
  if (opposite)
    {
      *cost = -*cost;
      *cosp = -*cosp;
      *sinp = -*sinp;
    }
}

void or_m2m_flavour (int flavour, int nm, const double *Msrc,
                     const double *xsrc, int nmdst, double *Mdst,
                     const double *xdst)
{
  double r, ct, st, cp, sp;
  if (nm <= 0)
    return;
  flav_vector (xsrc, xdst, 1, &r, &ct, &st, &cp, &sp);
  if (flavour == 2)
    m2m_kylin (nm, (const cplx *) Msrc, nmdst, (cplx *) Mdst, r, ct, st, cp,
               sp);
  else
    m2m_normal (nm, (const cplx *) Msrc, nmdst, (cplx *) Mdst, r, ct, st, cp,
                sp);
}

void or_l2l_flavour (int flavour, int pl, const double *Lsrc,
                     const double *xsrc, int pldst, double *Ldst,
                     const double *xdst)
{
  double r, ct, st, cp, sp;
  flav_vector (xsrc, xdst, 0, &r, &ct, &st, &cp, &sp);
  if (flavour == 2)
    l2l_kylin (pl, (const cplx *) Lsrc, pldst, (cplx *) Ldst, r, ct, st, cp,
               sp);
  else
    l2l_normal (pl, (const cplx *) Lsrc, pldst, (cplx *) Ldst, r, ct, st, cp,
                sp);
}

void or_m2l_flavour (int flavour, int nm, const double *M, const double *xsrc,
                     int pl, double *L, const double *xdst)
{
  double r, ct, st, cp, sp;
  if (nm <= 0)
    return;
  flav_vector (xsrc, xdst, 0, &r, &ct, &st, &cp, &sp);
  if (flavour == 2)
    m2l_kylin (nm, (const cplx *) M, pl, (cplx *) L, r, ct, st, cp, sp);
  else
    m2l_normal (nm, (const cplx *) M, pl, (cplx *) L, r, ct, st, cp, sp);
}

/* ------------------------------------------------- particle <-> series */

/* aran_development3d_p2m (aran/arandevelopment3d.c:169-216) */
void or_p2m (int nm, const double *pos, double q, const double *center,
             double *Md)
{
  cplx *M = (cplx *) Md;
  cplx Y[TRI (OR_MAXDEG + 1, 0)];
  double v[3], r, cost, sint, cosp, sinp, fact = q;
  int l, m;
  for (l = 0; l < 3; l++)
    v[l] = pos[l] - center[l];
  or_to_spherical (v, &r, &cost, &sint, &cosp, &sinp);
  or_harmonics (nm, cost, sint, cosp, sinp, (double *) Y);
  for (l = 0; l < nm; l++)
    {
      for (m = 0; m <= l; m++)
        {
          cplx term = fact * (4. * M_PI / (l + l + 1.)) * Y[TRI (l, m)];
          M[TRI (l, m)] += conj (term);
        }
      fact *= r;
    }
}

/* aran_development3d_p2l (aran/arandevelopment3d.c:227-277):
 * L_l^m += q (4 pi / (2l+1)) conj (Y_l^m) / r^(l+1) for l = 0..pl, about
 * the local's centre. */
void or_p2l (int pl, const double *pos, double q, const double *center,
             double *Ld)
{
  cplx *L = (cplx *) Ld;
  cplx Y[TRI (OR_MAXDEG + 1, 0)];
  double v[3], r, cost, sint, cosp, sinp, invr, fact;
  int l, m;
  for (l = 0; l < 3; l++)
    v[l] = pos[l] - center[l];
  or_to_spherical (v, &r, &cost, &sint, &cosp, &sinp);
  or_harmonics (pl, cost, sint, cosp, sinp, (double *) Y);
  invr = 1. / r;
  fact = q * invr;
  for (l = 0; l <= pl; l++)
    {
      for (m = 0; m <= l; m++)
        {
          cplx term = fact * (4. * M_PI / (l + l + 1.)) * Y[TRI (l, m)];
          L[TRI (l, m)] += conj (term);
        }
      fact *= invr;
    }
}

/* local part of aran_spherical_seriesd_evaluate_internal
 * (aran/aransphericalseriesd.c:313-346) */
double or_local_evaluate (int pl, const double *Ld, const double *center,
                          const double *pos)
{
  const cplx *L = (const cplx *) Ld;
  cplx Y[TRI (OR_MAXDEG + 1, 0)], res = 0.;
  double v[3], r, cost, sint, cosp, sinp;
  int l, m;
  for (l = 0; l < 3; l++)
    v[l] = pos[l] - center[l];
  or_to_spherical (v, &r, &cost, &sint, &cosp, &sinp);
  or_harmonics (pl, cost, sint, cosp, sinp, (double *) Y);
  for (l = pl; l >= 0; l--)
    {
      cplx sum = 0.;
      for (m = 1; m <= l; m++)
        sum += creal (L[TRI (l, m)] * Y[TRI (l, m)]);
      sum = 2. * sum + L[TRI (l, 0)] * Y[TRI (l, 0)];
      res = res * r + sum;
    }
  return creal (res);
}

/* multipole part of the same routine (aransphericalseriesd.c:348-368) */
double or_multipole_evaluate (int nm, const double *Md, const double *center,
                              const double *pos)
{
  const cplx *M = (const cplx *) Md;
  cplx Y[TRI (OR_MAXDEG + 1, 0)], res = 0.;
  double v[3], r, cost, sint, cosp, sinp, invr;
  int l, m;
  if (nm == 0)
    return 0.;
  for (l = 0; l < 3; l++)
    v[l] = pos[l] - center[l];
  or_to_spherical (v, &r, &cost, &sint, &cosp, &sinp);
  or_harmonics (nm - 1, cost, sint, cosp, sinp, (double *) Y);
  invr = 1. / r;
  for (l = nm - 1; l >= 0; l--)
    {
      cplx sum = 0.;
      for (m = 1; m <= l; m++)
        sum += creal (M[TRI (l, m)] * Y[TRI (l, m)]);
      sum = 2. * sum + M[TRI (l, 0)] * Y[TRI (l, 0)];
      res = (res + sum) * invr;
    }
  return creal (res);
}

static void grad_to_cartesian (double cost, double sint, double cosp,
                               double sinp, double dr, double dt, double dp,
                               double *g)
{
  g[0] = sint * cosp * dr + cost * cosp * dt - sinp * dp;
  g[1] = sint * sinp * dr + cost * sinp * dt + cosp * dp;
  g[2] = cost * dr - sint * dt;
}

/* per-degree sums of aran_spherical_seriesd_gradient_evaluate_internal
 * (aran/aransphericalseriesd.c:434-456 and 474-496) */
static void grad_degree_sums (int l, const cplx *c, const cplx *Y,
                              const cplx *SY, double cost, cplx expp,
                              cplx *sr, cplx *st, cplx *sp)
{
  const cplx *h = Y + TRI (l, 0), *sh = SY + TRI (l, 0);
  cplx sumr = 0., sumt = 0., sump = 0.;
  int m;
  for (m = 1; m <= l; m++)
    {
      sumr += creal (h[m] * c[m]);
      sumt += creal (sh[m] * c[m]) * (cost * m);
      sump += cimag (sh[m] * c[m]) * m;
    }
  for (m = 1; m <= l - 1; m++)
    sumt += creal (h[m + 1] * c[m] * conj (expp))
            * sqrt ((double) (l - m) * (l + m + 1));
  sumr = 2. * sumr + creal (h[0] * c[0]);
  sumt = 2. * sumt + creal (h[1] * c[0] * conj (expp))
                     * sqrt ((double) l * (l + 1));
  sump *= -2.;
  *sr = sumr;
  *st = sumt;
  *sp = sump;
}

void or_local_gradient (int pl, const double *Ld, const double *center,
                        const double *pos, double *grad)
{
  const cplx *L = (const cplx *) Ld;
  cplx Y[TRI (OR_MAXDEG + 1, 0)], SY[TRI (OR_MAXDEG + 1, 0)];
  cplx dr = 0., dt = 0., dp = 0., expp;
  double v[3], r, cost, sint, cosp, sinp;
  int l;
  for (l = 0; l < 3; l++)
    v[l] = pos[l] - center[l];
  or_to_spherical (v, &r, &cost, &sint, &cosp, &sinp);
  expp = cosp + I * sinp;
  harmonics_pregrad (pl, cost, sint, cosp, sinp, Y, SY);
  for (l = pl; l >= 1; l--)
    {
      cplx sr, st, sp;
      grad_degree_sums (l, L + TRI (l, 0), Y, SY, cost, expp, &sr, &st, &sp);
      dr = dr * r + sr * l;
      dt = dt * r + st;
      dp = dp * r + sp;
    }
  grad_to_cartesian (cost, sint, cosp, sinp, creal (dr), creal (dt),
                     creal (dp), grad);
}

void or_multipole_gradient (int nm, const double *Md, const double *center,
                            const double *pos, double *grad)
{
  const cplx *M = (const cplx *) Md;
  cplx Y[TRI (OR_MAXDEG + 1, 0)], SY[TRI (OR_MAXDEG + 1, 0)], expp;
  cplx dr2 = 0., dt2 = 0., dp2 = 0., dr, dt, dp;
  double v[3], r, cost, sint, cosp, sinp, invr;
  int l;
  for (l = 0; l < 3; l++)
    v[l] = pos[l] - center[l];
  or_to_spherical (v, &r, &cost, &sint, &cosp, &sinp);
  expp = cosp + I * sinp;
  harmonics_pregrad (nm, cost, sint, cosp, sinp, Y, SY);
  invr = 1. / r;
  for (l = nm - 1; l >= 1; l--)
    {
      cplx sr, st, sp;
      grad_degree_sums (l, M + TRI (l, 0), Y, SY, cost, expp, &sr, &st, &sp);
      dr2 = dr2 * invr + sr * (l + 1);
      dt2 = dt2 * invr + st;
      dp2 = dp2 * invr + sp;
    }
  dr = -invr * invr * (dr2 * invr + Y[0] * M[0]);
  dt = dt2 * invr * invr * invr;
  dp = dp2 * invr * invr * invr;
  grad_to_cartesian (cost, sint, cosp, sinp, creal (dr), creal (dt),
                     creal (dp), grad);
}

/* ------------------------------------------------------------------ tree */

void or_root_box (const double *lb, const double *ub, long n, const double *x,
                  const double *y, const double *z, double *lo, double *width)
{
  double c[3], h = 0.;
  const double *p[3] = { x, y, z };
  int k, inside;
  long i;
  for (k = 0; k < 3; k++)
    {
      c[k] = 0.5 * (lb[k] + ub[k]);
      if (0.5 * (ub[k] - lb[k]) > h)
        h = 0.5 * (ub[k] - lb[k]);
    }
  if (!(h > 0.))
    h = 1.;
  do
    {
      inside = 1;
      for (k = 0; k < 3 && inside; k++)
        for (i = 0; i < n; i++)
          if (p[k][i] < c[k] - h || p[k][i] > c[k] + h)
            {
              inside = 0;
              break;
            }
      if (!inside)
        h *= 2.;
    }
  while (!inside);
  for (k = 0; k < 3; k++)
    lo[k] = c[k] - h;
  *width = 2. * h;
}

static uint64_t spread3 (uint64_t v)
{
  uint64_t r = 0;
  int b;
  for (b = 0; b < OR_MAXDEPTH; b++)
    r |= ((v >> b) & 1ull) << (3 * b);
  return r;
}

static uint64_t icoord (double x, double lo, double scale)
{
  double f = floor ((x - lo) * scale);
  const double top = (double) ((1 << OR_MAXDEPTH) - 1);
  if (f < 0.)
    f = 0.;
  if (f > top)
    f = top;
  return (uint64_t) f;
}

void or_keys (long n, const double *x, const double *y, const double *z,
              const double *lo, double width, uint64_t *keys)
{
  double scale = (double) (1 << OR_MAXDEPTH) / width;
  long i;
  for (i = 0; i < n; i++)
    keys[i] = spread3 (icoord (x[i], lo[0], scale))
              | (spread3 (icoord (y[i], lo[1], scale)) << 1)
              | (spread3 (icoord (z[i], lo[2], scale)) << 2);
}

static const uint64_t *sort_keys;
static int cmp_perm (const void *a, const void *b)
{
  long i = *(const long *) a, j = *(const long *) b;
  if (sort_keys[i] < sort_keys[j])
    return -1;
  if (sort_keys[i] > sort_keys[j])
    return 1;
  return (i < j) ? -1 : (i > j);
}

long or_tree_build (long n, const double *x, const double *y, const double *z,
                    const double *lb, const double *ub, int leafcap,
                    long *perm, OrNode *nodes, long max_nodes)
{
  double lo[3], width;
  uint64_t *keys = malloc (sizeof (uint64_t) * (n > 0 ? n : 1));
  long i, nn = 1, cur;
  or_root_box (lb, ub, n, x, y, z, lo, &width);
  or_keys (n, x, y, z, lo, width, keys);
  for (i = 0; i < n; i++)
    perm[i] = i;
  sort_keys = keys;
  qsort (perm, n, sizeof (long), cmp_perm);

  memset (&nodes[0], 0, sizeof (OrNode));
  nodes[0].pbegin = 0;
  nodes[0].pend = (int) n;
  nodes[0].parent = -1;
  for (cur = 0; cur < nn; cur++)
    {
      OrNode *nd = &nodes[cur];
      double h = width * ldexp (1., -nd->depth);
      nd->center[0] = fma (nd->ix + 0.5, h, lo[0]);
      nd->center[1] = fma (nd->iy + 0.5, h, lo[1]);
      nd->center[2] = fma (nd->iz + 0.5, h, lo[2]);
      nd->child_begin = -1;
      nd->nchild = 0;
      nd->isleaf = 1;
      if (nd->pend - nd->pbegin > leafcap && nd->depth < OR_MAXDEPTH)
        {
          int shift = 3 * (OR_MAXDEPTH - nd->depth - 1), o;
          long b = nd->pbegin;
          nd->isleaf = 0;
          nd->child_begin = (int) nn;
          for (o = 0; o < 8; o++)
            {
              long e = b;
              while (e < nd->pend && (int) ((keys[perm[e]] >> shift) & 7) == o)
                e++;
              if (e > b)
                {
                  OrNode *ch;
                  if (nn >= max_nodes)
                    {
                      free (keys);
                      return -1;
                    }
                  ch = &nodes[nn++];
                  nd = &nodes[cur];
                  memset (ch, 0, sizeof (OrNode));
                  ch->depth = nd->depth + 1;
                  ch->ix = 2 * nd->ix + (o & 1);
                  ch->iy = 2 * nd->iy + ((o >> 1) & 1);
                  ch->iz = 2 * nd->iz + ((o >> 2) & 1);
                  ch->pbegin = (int) b;
                  ch->pend = (int) e;
                  ch->parent = (int) cur;
                  nd->nchild++;
                }
              b = e;
            }
        }
    }
  free (keys);
  return nn;
}

/* ------------------------------------------------------------ traversal */

typedef struct
{
  const OrNode *nodes;
  void (*far) (void *, long, long);
  void (*near) (void *, long, long);
  /* semifar (leaf, deeper node), or NULL; used when the deeper node does
   * not touch the leaf and holds more than thr particles */
  void (*semifar) (void *, long, long);
  long thr;
  void *ud;
} Trav;

/* Boxes of a and of b (b at least as deep as a) touch or overlap. */
static int touching (const OrNode *a, const OrNode *b)
{
  const int s = b->depth - a->depth;
  const long w = 1L << s;
  const long lo[3] = { (long) a->ix << s, (long) a->iy << s,
                       (long) a->iz << s };
  const long c[3] = { b->ix, b->iy, b->iz };
  int k;
  for (k = 0; k < 3; k++)
    if (c[k] < lo[k] - 1 || c[k] > lo[k] + w)
      return 0;
  return 1;
}

static int adjacent (const OrNode *a, const OrNode *b)
{
  return abs (a->ix - b->ix) <= 1 && abs (a->iy - b->iy) <= 1
         && abs (a->iz - b->iz) <= 1;
}

static void trav_pair (const Trav *t, long a, long b)
{
  const OrNode *A = &t->nodes[a], *B = &t->nodes[b];
  int k;
  if (A->depth == B->depth && !adjacent (A, B))
    {
      t->far (t->ud, a, b);
      return;
    }
  if (t->semifar && A->depth != B->depth)
    {
      /* the shallower side is a leaf (only non-leaves are descended) */
      const long s = A->depth < B->depth ? a : b, d = s == a ? b : a;
      const OrNode *S = &t->nodes[s], *D = &t->nodes[d];
      if (!touching (S, D) && (long) (D->pend - D->pbegin) > t->thr)
        {
          t->semifar (t->ud, s, d);
          return;
        }
    }
  if (A->isleaf && B->isleaf)
    {
      t->near (t->ud, a, b);
      return;
    }
  if (A->isleaf)
    {
      for (k = 0; k < B->nchild; k++)
        trav_pair (t, a, B->child_begin + k);
      return;
    }
  if (B->isleaf)
    {
      for (k = 0; k < A->nchild; k++)
        trav_pair (t, A->child_begin + k, b);
      return;
    }
  {
    int i, j;
    for (i = 0; i < A->nchild; i++)
      for (j = 0; j < B->nchild; j++)
        trav_pair (t, A->child_begin + i, B->child_begin + j);
  }
}

static void trav_self (const Trav *t, long a)
{
  const OrNode *A = &t->nodes[a];
  int i, j;
  if (A->isleaf)
    {
      t->near (t->ud, a, a);
      return;
    }
  for (i = 0; i < A->nchild; i++)
    {
      trav_self (t, A->child_begin + i);
      for (j = i + 1; j < A->nchild; j++)
        trav_pair (t, A->child_begin + i, A->child_begin + j);
    }
}

typedef struct
{
  long *fs, *fd, maxf, nf;
  long *na, *nb, maxn, nn;
  long *sl, *sd, maxs, ns;
} PairOut;

static void po_far (void *ud, long a, long b)
{
  PairOut *o = ud;
  if (o->nf + 2 <= o->maxf)
    {
      o->fs[o->nf] = a;
      o->fd[o->nf] = b;
      o->fs[o->nf + 1] = b;
      o->fd[o->nf + 1] = a;
    }
  o->nf += 2;
}

static void po_near (void *ud, long a, long b)
{
  PairOut *o = ud;
  if (o->nn < o->maxn)
    {
      o->na[o->nn] = a < b ? a : b;
      o->nb[o->nn] = a < b ? b : a;
    }
  o->nn++;
}

static void po_semifar (void *ud, long leaf, long node)
{
  PairOut *o = ud;
  if (o->ns < o->maxs)
    {
      o->sl[o->ns] = leaf;
      o->sd[o->ns] = node;
    }
  o->ns++;
}

void or_traverse_full (const OrNode *nodes, long nnodes, long semifar_thr,
                       long *far_src, long *far_dst, long max_far, long *nfar,
                       long *near_a, long *near_b, long max_near, long *nnear,
                       long *semi_leaf, long *semi_node, long max_semi,
                       long *nsemi)
{
  PairOut o = { far_src, far_dst, max_far, 0, near_a, near_b, max_near, 0,
                semi_leaf, semi_node, max_semi, 0 };
  Trav t = { nodes, po_far, po_near, NULL, 0, &o };
  if (semifar_thr >= 0)
    {
      t.semifar = po_semifar;
      t.thr = semifar_thr;
    }
  if (nnodes > 0)
    trav_self (&t, 0);
  *nfar = o.nf;
  *nnear = o.nn;
  *nsemi = o.ns;
}

void or_traverse (const OrNode *nodes, long nnodes, long *far_src,
                  long *far_dst, long max_far, long *nfar, long *near_a,
                  long *near_b, long max_near, long *nnear)
{
  long ns;
  or_traverse_full (nodes, nnodes, -1, far_src, far_dst, max_far, nfar,
                    near_a, near_b, max_near, nnear, NULL, NULL, 0, &ns);
}

/* ----------------------------------------------------------------- solve */

typedef struct
{
  const OrNode *nodes;
  int nm, pl;
  long mc, lc;
  double *M, *L;
  const double *x, *y, *z, *q; /* key-sorted */
  double *pot, *grad;
  long *stats;
} Solve;

/* Newton P2P, symmetric update (tests/newtonpot3.c:61-78,
 * tests/newtonfield3.c:91-119); only i == j is excluded. */
static void p2p_pair (Solve *s, long i, long j)
{
  double dx, dy, dz, r2, inv_r, inv_r3;
  if (i == j)
    return;
  dx = s->x[i] - s->x[j];
  dy = s->y[i] - s->y[j];
  dz = s->z[i] - s->z[j];
  r2 = dx * dx + dy * dy + dz * dz;
  inv_r = 1. / sqrt (r2);
  inv_r3 = inv_r * inv_r * inv_r;
  s->pot[i] += inv_r * s->q[j];
  s->pot[j] += inv_r * s->q[i];
  s->grad[3 * i + 0] -= inv_r3 * s->q[j] * dx;
  s->grad[3 * i + 1] -= inv_r3 * s->q[j] * dy;
  s->grad[3 * i + 2] -= inv_r3 * s->q[j] * dz;
  s->grad[3 * j + 0] += inv_r3 * s->q[i] * dx;
  s->grad[3 * j + 1] += inv_r3 * s->q[i] * dy;
  s->grad[3 * j + 2] += inv_r3 * s->q[i] * dz;
}

/* near_func_default / near_func_reflexive (aran/aransolver3d.c:254-324) */
static void solve_near (void *ud, long a, long b)
{
  Solve *s = ud;
  const OrNode *A = &s->nodes[a], *B = &s->nodes[b];
  long i, j;
  if (a == b)
    {
      for (i = A->pbegin; i < A->pend; i++)
        for (j = i; j < A->pend; j++)
          p2p_pair (s, i, j);
      s->stats[1] += ((long) (A->pend - A->pbegin)
                      * (A->pend - A->pbegin + 1)) / 2;
    }
  else
    {
      for (i = A->pbegin; i < A->pend; i++)
        for (j = B->pbegin; j < B->pend; j++)
          p2p_pair (s, i, j);
      s->stats[1] += (long) (A->pend - A->pbegin) * (B->pend - B->pbegin);
    }
}

/* far_func: M2L both ways (aran/aransolver3d.c:333-354) */
static void solve_far (void *ud, long a, long b)
{
  Solve *s = ud;
  or_m2l_rotate (s->nm, s->M + 2 * s->mc * a, s->nodes[a].center, s->pl,
                 s->L + 2 * s->lc * b, s->nodes[b].center);
  or_m2l_rotate (s->nm, s->M + 2 * s->mc * b, s->nodes[b].center, s->pl,
                 s->L + 2 * s->lc * a, s->nodes[a].center);
  s->stats[4] += 2;
}

/* semifar_func (aran/aransolver3d.c:357-396) with the newtonfield3
 * callbacks (tests/newtonfield3.c:139-153): every particle of the leaf adds
 * to the deeper node's local (P2L) and takes the potential and gradient of
 * its multipole (M2P). */
static void solve_semifar (void *ud, long leaf, long node)
{
  Solve *s = ud;
  const OrNode *A = &s->nodes[leaf], *D = &s->nodes[node];
  long i;
  for (i = A->pbegin; i < A->pend; i++)
    {
      double pos[3] = { s->x[i], s->y[i], s->z[i] }, g[3];
      or_p2l (s->pl, pos, s->q[i], D->center, s->L + 2 * s->lc * node);
      s->pot[i] += or_multipole_evaluate (s->nm, s->M + 2 * s->mc * node,
                                          D->center, pos);
      or_multipole_gradient (s->nm, s->M + 2 * s->mc * node, D->center, pos,
                             g);
      s->grad[3 * i + 0] += g[0];
      s->grad[3 * i + 1] += g[1];
      s->grad[3 * i + 2] += g[2];
    }
  s->stats[7] += A->pend - A->pbegin;
  s->stats[8] += A->pend - A->pbegin;
}

/* up_func, post-order (aran/aransolver3d.c:412-450) */
static void solve_up (Solve *s, long a)
{
  const OrNode *A = &s->nodes[a];
  long i;
  int k;
  for (k = 0; k < A->nchild; k++)
    solve_up (s, A->child_begin + k);
  if (A->isleaf)
    for (i = A->pbegin; i < A->pend; i++)
      {
        double pos[3] = { s->x[i], s->y[i], s->z[i] };
        or_p2m (s->nm, pos, s->q[i], A->center, s->M + 2 * s->mc * a);
        s->stats[2]++;
      }
  if (A->pend > A->pbegin && A->parent >= 0)
    {
      or_m2m_rotate (s->nm, s->M + 2 * s->mc * a, A->center, s->nm,
                     s->M + 2 * s->mc * A->parent,
                     s->nodes[A->parent].center);
      s->stats[3]++;
    }
}

/* down_func, pre-order (aran/aransolver3d.c:452-490); L2P for potential and
 * gradient as the newtonpot3 / newtonfield3 callbacks do. */
static void solve_down (Solve *s, long a)
{
  const OrNode *A = &s->nodes[a];
  long i;
  int k;
  if (A->pend > A->pbegin && A->parent >= 0)
    {
      or_l2l_rotate (s->pl, s->L + 2 * s->lc * A->parent,
                     s->nodes[A->parent].center, s->pl, s->L + 2 * s->lc * a,
                     A->center);
      s->stats[5]++;
    }
  if (A->isleaf)
    for (i = A->pbegin; i < A->pend; i++)
      {
        double pos[3] = { s->x[i], s->y[i], s->z[i] }, g[3];
        s->pot[i] += or_local_evaluate (s->pl, s->L + 2 * s->lc * a,
                                        A->center, pos);
        or_local_gradient (s->pl, s->L + 2 * s->lc * a, A->center, pos, g);
        s->grad[3 * i + 0] += g[0];
        s->grad[3 * i + 1] += g[1];
        s->grad[3 * i + 2] += g[2];
        s->stats[6]++;
      }
  for (k = 0; k < A->nchild; k++)
    solve_down (s, A->child_begin + k);
}

/* aran_solver3d_solve (aran/aransolver3d.c:1135-1202): clear, up, near/far,
 * down.  order p: multipole degrees 0..p-1, local degrees 0..p
 * (aran_development3d_new (0, p), aran/arandevelopment3d.c:64-73). */
int or_fmm_solve (long n, const double *x, const double *y, const double *z,
                  const double *q, const double *lb, const double *ub,
                  int order, int leafcap, double *pot, double *grad,
                  long *stats)
{
  long st[9];
  int rc = or_fmm_solve_full (n, x, y, z, q, lb, ub, order, leafcap, -1, pot,
                              grad, st);
  memcpy (stats, st, 7 * sizeof (long));
  return rc;
}

int or_fmm_solve_full (long n, const double *x, const double *y,
                       const double *z, const double *q, const double *lb,
                       const double *ub, int order, int leafcap,
                       long semifar_thr, double *pot, double *grad,
                       long *stats)
{
  long max_nodes = 64 + 8 * (n / (leafcap + 1) + 1) * (OR_MAXDEPTH + 1);
  OrNode *nodes;
  long *perm, nn, i;
  double *xs, *ys, *zs, *qs, *ps, *gs;
  Solve s;
  Trav t;

  if (n <= 0 || order < 1 || order > OR_MAXDEG - 2 || leafcap < 1)
    return -1;
  nodes = malloc (sizeof (OrNode) * max_nodes);
  perm = malloc (sizeof (long) * n);
  nn = or_tree_build (n, x, y, z, lb, ub, leafcap, perm, nodes, max_nodes);
  if (nn < 0)
    {
      free (nodes);
      free (perm);
      return -2;
    }
  xs = malloc (sizeof (double) * n);
  ys = malloc (sizeof (double) * n);
  zs = malloc (sizeof (double) * n);
  qs = malloc (sizeof (double) * n);
  ps = calloc (n, sizeof (double));
  gs = calloc (3 * n, sizeof (double));
  for (i = 0; i < n; i++)
    {
      xs[i] = x[perm[i]];
      ys[i] = y[perm[i]];
      zs[i] = z[perm[i]];
      qs[i] = q[perm[i]];
    }
  memset (&s, 0, sizeof (s));
  s.nodes = nodes;
  s.nm = order;
  s.pl = order;
  s.mc = TRI (order, 0);
  s.lc = TRI (order + 1, 0);
  s.M = calloc (2 * s.mc * nn, sizeof (double));
  s.L = calloc (2 * s.lc * nn, sizeof (double));
  s.x = xs;
  s.y = ys;
  s.z = zs;
  s.q = qs;
  s.pot = ps;
  s.grad = gs;
  s.stats = stats;
  for (i = 0; i < 9; i++)
    stats[i] = 0;
  stats[0] = nn; /* clear_func on every node */
  solve_up (&s, 0);
  t.nodes = nodes;
  t.far = solve_far;
  t.near = solve_near;
  t.semifar = semifar_thr >= 0 ? solve_semifar : NULL;
  t.thr = semifar_thr;
  t.ud = &s;
  trav_self (&t, 0);
  solve_down (&s, 0);
  for (i = 0; i < n; i++)
    {
      pot[perm[i]] = ps[i];
      grad[3 * perm[i] + 0] = gs[3 * i + 0];
      grad[3 * perm[i] + 1] = gs[3 * i + 1];
      grad[3 * perm[i] + 2] = gs[3 * i + 2];
    }
  free (s.M);
  free (s.L);
  free (xs);
  free (ys);
  free (zs);
  free (qs);
  free (ps);
  free (gs);
  free (nodes);
  free (perm);
  return 0;
}

/* O(N) per target direct sum (tests/newtonpot3.c:678-716,
 * tests/newtonfield3.c:650-693) with Neumaier compensation. */
void or_direct (long n, const double *x, const double *y, const double *z,
                const double *q, long nt, const long *targets, double *pot,
                double *grad)
{
  long t, j;
  for (t = 0; t < nt; t++)
    {
      long i = targets[t];
      double s[4] = { 0, 0, 0, 0 }, c[4] = { 0, 0, 0, 0 };
      int k;
      for (j = 0; j < n; j++)
        {
          double dx, dy, dz, r2, ir, ir3, v[4];
          if (j == i)
            continue;
          dx = x[i] - x[j];
          dy = y[i] - y[j];
          dz = z[i] - z[j];
          r2 = dx * dx + dy * dy + dz * dz;
          ir = 1. / sqrt (r2);
          ir3 = ir * ir * ir;
          v[0] = q[j] * ir;
          v[1] = -q[j] * ir3 * dx;
          v[2] = -q[j] * ir3 * dy;
          v[3] = -q[j] * ir3 * dz;
          for (k = 0; k < 4; k++)
            {
              double tt = s[k] + v[k];
              if (fabs (s[k]) >= fabs (v[k]))
                c[k] += (s[k] - tt) + v[k];
              else
                c[k] += (v[k] - tt) + s[k];
              s[k] = tt;
            }
        }
      pot[t] = s[0] + c[0];
      grad[3 * t + 0] = s[1] + c[1];
      grad[3 * t + 1] = s[2] + c[2];
      grad[3 * t + 2] = s[3] + c[3];
    }
}
