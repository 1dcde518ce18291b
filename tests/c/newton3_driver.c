/*
 * newton3_driver.c -- a C program that drives the AranSolver3d API the way
 * the reference's solver harnesses do (tests/newtonpot3.c,
 * tests/newtonfield3.c): LibVsg tree and vector calls, a development
 * prototype, the library Newton callbacks with a chosen translation flavour,
 * particle insertion (with tolerance merging and the corner insert/remove of
 * the Plummer fill), an optional near/far leaf predicate, the solve, and an
 * O(N^2) check with the harnesses' pointwise relative error.  It is
 * compiled with cc against include/ and linked to libaran_amd.so.
 *
 *   newton3_driver [-np N] [-pr ORDER] [-s MAXBOX] [-err LIMIT]
 *                  [-dist circle|random|grid|plummer]
 *                  [-translation normal|kkylin|rotate] [-field]
 *                  [-semifar T] [-virtual-maxbox K] [-traverse] [-hilbert]
 *                  [-seed S]
 *
 * Exit status: the number of particles whose error exceeds LIMIT plus the
 * number of failed traversal checks (0 = pass), or 100 + code on an API
 * failure.  Random inputs come from a seeded SplitMix64 stream (the
 * reference's "random" distribution uses GLib's unseeded generator).
 */
#include <complex.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <aran/aran.h>
#include <aran/aransolver3d.h>

#define TR (1.)
#define R (0.999 * TR)

/* the particle of tests/newtonpot3.c:43-54 (complex accumulator) */
typedef struct
{
  VsgVector3d vector;
  gdouble density;
  gcomplex128 accum;
  guint id;
} PointPot;

/* the particle of tests/newtonfield3.c:46-55 (field accumulator) */
typedef struct
{
  VsgVector3d vector;
  gdouble density;
  VsgVector3d field;
  guint id;
} PointField;

static guint np = 12, order = 20, maxbox = 1, virtual_maxbox = 0;
static guint semifar_threshold = 0xffffffffu;
static double err_lim = 1.e-3;
static const char *dist = "circle", *translation = "normal";
static int field_mode = 0, do_traverse = 0, hilbert = 0;
static uint64_t seed = 0;

static uint64_t splitmix (uint64_t *s)
{
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double uniform (uint64_t *s, double a, double b)
{
  return a + (b - a) * ((double) (splitmix (s) >> 11) * 0x1.0p-53);
}

static void parse_args (int argc, char **argv)
{
  int i;
  for (i = 1; i < argc; i++)
    {
      const char *a = argv[i], *v = (i + 1 < argc) ? argv[i + 1] : NULL;
      if (!strcmp (a, "-np") && v) { np = (guint) atoi (v); i++; }
      else if (!strcmp (a, "-pr") && v) { order = (guint) atoi (v); i++; }
      else if (!strcmp (a, "-s") && v) { maxbox = (guint) atoi (v); i++; }
      else if (!strcmp (a, "-err") && v) { err_lim = atof (v); i++; }
      else if (!strcmp (a, "-dist") && v) { dist = v; i++; }
      else if (!strcmp (a, "-translation") && v) { translation = v; i++; }
      else if (!strcmp (a, "-semifar") && v)
        { semifar_threshold = (guint) strtoul (v, NULL, 10); i++; }
      else if (!strcmp (a, "-virtual-maxbox") && v)
        { virtual_maxbox = (guint) atoi (v); i++; }
      else if (!strcmp (a, "-seed") && v)
        { seed = strtoull (v, NULL, 10); i++; }
      else if (!strcmp (a, "-field")) field_mode = 1;
      else if (!strcmp (a, "-traverse")) do_traverse = 1;
      else if (!strcmp (a, "-hilbert")) hilbert = 1;
      else
        {
          fprintf (stderr, "invalid argument \"%s\"\n", a);
          exit (100);
        }
    }
}

/* particle access independent of the layout */
static size_t psize (void)
{
  return field_mode ? sizeof (PointField) : sizeof (PointPot);
}
static VsgVector3d *pvec (void *p) { return (VsgVector3d *) p; }
static double *pdens (void *p)
{
  return field_mode ? &((PointField *) p)->density : &((PointPot *) p)->density;
}

static void *new_point (double x, double y, double z, double q, guint id)
{
  void *p = calloc (1, psize ());
  vsg_vector3d_set (pvec (p), x, y, z);
  *pdens (p) = q;
  if (field_mode)
    ((PointField *) p)->id = id;
  else
    ((PointPot *) p)->id = id;
  return p;
}

/* the distributions of tests/newtonpot3.c:405-542 */
static guint fill (void **pts, AranSolver3d *solver)
{
  guint i, n = 0;
  uint64_t s = seed;
  if (!strcmp (dist, "circle"))
    {
      for (i = 0; i < np; i++)
        {
          pts[n] = new_point (R * cos (2. * M_PI * i / np),
                              R * sin (2. * M_PI * i / np), 0., 1. / np, n);
          aran_solver3d_insert_point (solver, pts[n]);
          n++;
        }
    }
  else if (!strcmp (dist, "random"))
    {
      /* points closer than the tolerance merge into the first one */
      aran_solver3d_set_tolerance (solver, 1.e-2);
      for (i = 0; i < np; i++)
        {
          void *p = new_point (uniform (&s, -R, R), uniform (&s, -R, R),
                               uniform (&s, -R, R), uniform (&s, 0., 2. / np),
                               n);
          void *found = aran_solver3d_find_point (solver, p);
          if (found)
            {
              *pdens (found) += *pdens (p);
              free (p);
            }
          else
            {
              pts[n++] = p;
              aran_solver3d_insert_point_local (solver, p);
            }
        }
    }
  else if (!strcmp (dist, "grid"))
    {
      guint size = (guint) floor (pow (np, 1. / 3.) + 1e-9), j, k;
      double q = 1. / ((double) size * size * size);
      for (i = 0; i < size; i++)
        for (j = 0; j < size; j++)
          for (k = 0; k < size; k++)
            {
              pts[n] = new_point (R * (2. * i / (size - 1.) - 1.),
                                  R * (2. * j / (size - 1.) - 1.),
                                  R * (2. * k / (size - 1.) - 1.), q, n);
              aran_solver3d_insert_point (solver, pts[n]);
              n++;
            }
    }
  else if (!strcmp (dist, "plummer"))
    {
      /* the corners grow the bounds and are removed again
       * (tests/newtonpot3.c:528-542) */
      double rmax = 300.;
      void *ub = new_point (rmax, rmax, rmax, 0., 0);
      void *lb = new_point (-rmax, -rmax, -rmax, 0., 0);
      aran_solver3d_insert_point (solver, ub);
      aran_solver3d_insert_point (solver, lb);
      if (!aran_solver3d_remove_point (solver, lb)
          || !aran_solver3d_remove_point (solver, ub))
        {
          fprintf (stderr, "remove_point failed\n");
          exit (101);
        }
      free (ub);
      free (lb);
      for (i = 0; i < np; i++)
        {
          double x1 = uniform (&s, 0., 1.), x2 = uniform (&s, 0., 1.);
          double x3 = uniform (&s, 0., 1.);
          double r = 1. / sqrt (pow (x1, -2. / 3.) - 1.);
          double z = R * (1. - (x2 + x2)) * r;
          double t = sqrt (r * r - z * z);
          double x = R * t * cos (2. * M_PI * x3), y = R * t * sin (2. * M_PI * x3);
          if (fabs (x) > rmax || fabs (y) > rmax || fabs (z) > rmax)
            continue;
          pts[n] = new_point (x, y, z, 1. / np, n);
          aran_solver3d_insert_point (solver, pts[n]);
          n++;
        }
    }
  else
    {
      fprintf (stderr, "invalid distribution \"%s\"\n", dist);
      exit (100);
    }
  return n;
}

/* tests/newtonfield3.c:585-595: a node holding few particles is a leaf */
static gboolean isleaf_virtual_maxbox (const VsgPRTree3dNodeInfo *info,
                                       gpointer data)
{
  if (VSG_PRTREE3D_NODE_INFO_IS_SHARED (info))
    return FALSE;
  return info->point_count <= *(guint *) data;
}

/* traversal checks: counts, father links, leaf particle total, developments
 * present after a solve */
typedef struct
{
  long visits, leaves, leaf_points, bad_father, with_dev, last_depth;
  long depth_drops;
} TravCount;

static void count_node (const VsgPRTree3dNodeInfo *info, gpointer data)
{
  TravCount *c = (TravCount *) data;
  c->visits++;
  if (info->isleaf)
    {
      c->leaves++;
      c->leaf_points += info->point_count;
    }
  if (info->father_info
      && (info->father_info->depth + 1 != info->depth
          || info->father_info->point_count < info->point_count))
    c->bad_father++;
  if (info->user_data)
    c->with_dev++;
  if ((long) info->depth < c->last_depth)
    c->depth_drops++;
  c->last_depth = info->depth;
}

int main (int argc, char **argv)
{
  VsgVector3d lbound = {-TR, -TR, -TR}, ubound = {TR, TR, TR};
  VsgPRTree3d *prtree;
  AranSolver3d *solver;
  AranMultipole2MultipoleFunc3d m2m;
  AranMultipole2LocalFunc3d m2l;
  AranLocal2LocalFunc3d l2l;
  void **pts;
  guint i, j, n;
  int ret = 0;
  glong stats[11];

  aran_init ();
  vsg_packed_msg_trace_set_active (FALSE);
  parse_args (argc, argv);

  if (!strcmp (translation, "normal"))
    {
      m2m = (AranMultipole2MultipoleFunc3d) aran_development3d_m2m;
      m2l = (AranMultipole2LocalFunc3d) aran_development3d_m2l;
      l2l = (AranLocal2LocalFunc3d) aran_development3d_l2l;
    }
  else if (!strcmp (translation, "kkylin"))
    {
      m2m = (AranMultipole2MultipoleFunc3d) aran_development3d_m2m_kkylin;
      m2l = (AranMultipole2LocalFunc3d) aran_development3d_m2l_kkylin;
      l2l = (AranLocal2LocalFunc3d) aran_development3d_l2l_kkylin;
    }
  else if (!strcmp (translation, "rotate"))
    {
      m2m = (AranMultipole2MultipoleFunc3d) aran_development3d_m2m_rotate;
      m2l = (AranMultipole2LocalFunc3d) aran_development3d_m2l_rotate;
      l2l = (AranLocal2LocalFunc3d) aran_development3d_l2l_rotate;
    }
  else
    {
      fprintf (stderr, "invalid translation \"%s\"\n", translation);
      return 100;
    }

  prtree = vsg_prtree3d_new_full (&lbound, &ubound,
                                  vsg_vector3d_vector3d_locfunc,
                                  vsg_vector3d_dist, NULL, maxbox);
  solver = aran_solver3d_new (prtree, ARAN_TYPE_DEVELOPMENT3D,
                              aran_development3d_new (0, order),
                              (AranZeroFunc) aran_development3d_set_zero);
  if (!solver)
    return 102;
  if (field_mode)
    {
      aran_solver3d_set_particle_layout (solver, offsetof (PointField, density),
                                         -1, FALSE,
                                         offsetof (PointField, field));
      aran_solver3d_set_functions_full (
          solver, aran_newton3d_p2p_field, aran_newton3d_p2m, m2m, m2l, l2l,
          aran_newton3d_l2p_field,
          semifar_threshold != 0xffffffffu ? aran_newton3d_p2l : NULL,
          semifar_threshold != 0xffffffffu ? aran_newton3d_m2p_field : NULL,
          semifar_threshold);
    }
  else
    {
      aran_solver3d_set_particle_layout (solver, offsetof (PointPot, density),
                                         offsetof (PointPot, accum), TRUE,
                                         -1);
      aran_solver3d_set_functions_full (
          solver, aran_newton3d_p2p_pot, aran_newton3d_p2m, m2m, m2l, l2l,
          aran_newton3d_l2p_pot,
          semifar_threshold != 0xffffffffu ? aran_newton3d_p2l : NULL,
          semifar_threshold != 0xffffffffu ? aran_newton3d_m2p_pot : NULL,
          semifar_threshold);
    }
  if (virtual_maxbox != 0)
    aran_solver3d_set_nf_isleaf (solver, isleaf_virtual_maxbox,
                                 &virtual_maxbox);
  if (hilbert)
    aran_solver3d_set_children_order_hilbert (solver);

  pts = calloc (np + 1, sizeof (void *));
  n = fill (pts, solver);
  if (aran_solver3d_point_count (solver) != n)
    {
      fprintf (stderr, "point count %u != %u\n",
               aran_solver3d_point_count (solver), n);
      return 103;
    }

  aran_solver3d_solve (solver);
  if (aran_solver3d_status (solver) != 0)
    {
      fprintf (stderr, "solve failed: %s\n", aran_last_error ());
      return 104;
    }
  aran_solver3d_get_stats (solver, &stats[0], &stats[1], &stats[2], &stats[3],
                           &stats[4], &stats[5], &stats[6], &stats[7],
                           &stats[8], &stats[9], &stats[10]);
  printf ("np %u order %u maxbox %u dist %s translation %s %s depth %u\n", n,
          order, maxbox, dist, translation, field_mode ? "field" : "pot",
          aran_solver3d_depth (solver));
  printf ("stats zero %ld p2p %ld p2m %ld m2m %ld m2l %ld l2l %ld l2p %ld "
          "p2l %ld m2p %ld\n", stats[0], stats[1], stats[2], stats[3],
          stats[4], stats[5], stats[6], stats[7], stats[8]);

  if (do_traverse)
    {
      TravCount pre, post, lvl;
      memset (&pre, 0, sizeof pre);
      memset (&post, 0, sizeof post);
      memset (&lvl, 0, sizeof lvl);
      aran_solver3d_traverse (solver, G_PRE_ORDER, count_node, &pre);
      aran_solver3d_traverse (solver, G_POST_ORDER, count_node, &post);
      aran_solver3d_traverse (solver, G_LEVEL_ORDER, count_node, &lvl);
      printf ("traverse nodes %ld leaves %ld leaf points %ld with dev %ld\n",
              pre.visits, pre.leaves, pre.leaf_points, pre.with_dev);
      if (pre.visits != stats[0] || post.visits != stats[0]
          || lvl.visits != stats[0])
        ret++; /* every node visited once (the zero counter counts nodes) */
      if (pre.leaf_points != (long) n || post.leaf_points != (long) n)
        ret++;
      if (pre.bad_father || post.bad_father)
        ret++;
      if (pre.with_dev != pre.visits)
        ret++;
      if (lvl.depth_drops != 0)
        ret++; /* breadth first: depths never decrease */
    }

  /* O(N^2) check with the harnesses' metrics
   * (tests/newtonpot3.c:678-716, tests/newtonfield3.c:650-693) */
  for (i = 0; i < n; i++)
    {
      double pot = 0., f[3] = {0., 0., 0.}, err;
      for (j = 0; j < n; j++)
        {
          VsgVector3d d;
          double ir;
          if (i == j)
            continue;
          vsg_vector3d_sub (pvec (pts[i]), pvec (pts[j]), &d);
          ir = 1. / vsg_vector3d_norm (&d);
          pot += ir * *pdens (pts[j]);
          vsg_vector3d_scalp (&d, -ir * ir * ir * *pdens (pts[j]), &d);
          f[0] += d.x;
          f[1] += d.y;
          f[2] += d.z;
        }
      if (field_mode)
        {
          const VsgVector3d *g = &((PointField *) pts[i])->field;
          VsgVector3d ref = {f[0], f[1], f[2]}, diff;
          vsg_vector3d_sub (g, &ref, &diff);
          err = vsg_vector3d_norm (&diff) / vsg_vector3d_norm (&ref);
        }
      else
        {
          gcomplex128 a = ((PointPot *) pts[i])->accum;
          err = cabs (a - pot) / fmax (cabs (a), fabs (pot));
        }
      if (err > err_lim || !isfinite (err))
        {
          if (ret < 10)
            {
              fprintf (stderr, "Error: pt%u %e -> ", i, err);
              vsg_vector3d_write (pvec (pts[i]), stderr);
              fprintf (stderr, "\n");
            }
          ret++;
        }
    }
  printf ("failures %d\n", ret);

  aran_solver3d_free (solver);
  for (i = 0; i < n; i++)
    free (pts[i]);
  free (pts);
  return ret > 99 ? 99 : ret;
}
