/*
 * aran_oracle.h -- FP64 serial CPU restatement of LibAran's 3D spherical
 * harmonic FMM (the AranSolver3d path with the rotate / z-translate /
 * rotate-back operators).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only the test-suite, the smoke check and the CPU-baseline leg of the
 * benchmark may load it.  The engine never calls into it and has no CPU
 * fallback.
 *
 * Coefficient layouts (FP64 complex, interleaved re/im):
 *   multipole with degrees l = 0..nm-1 : index l(l+1)/2 + m, 0 <= m <= l
 *   local     with degrees l = 0..pl   : index l(l+1)/2 + m, 0 <= m <= l
 * Only m >= 0 is stored; c_l^{-m} = (-1)^m conj(c_l^m)
 * (aran/aransphericalseriesd-private.h:91-99).
 */
#ifndef ARAN_ORACLE_H
#define ARAN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXDEPTH 21

/* tables */
double or_alpha (int l, int m);
double or_beta (int l);
void or_legendre (int L, double x, double s, double *P);
void or_legendre_special (int L, double x, double s, const double *P,
                          double *S);
void or_harmonics (int L, double cost, double sint, double cosp, double sinp,
                   double *Y /* complex, tri(L) */);
void or_to_spherical (const double *v, double *r, double *cost, double *sint,
                      double *cosp, double *sinp);

/* Wigner: d stored [l][m'][l+m], m' in 0..l, m in -l..l */
long or_wigner_size (int L);
void or_wigner_d (int L, double beta, double *d);
void or_wigner_D (int L, double alpha, double beta, double gamma,
                  double *D /* complex */);
double or_wigner_factorial (int l, int m1, int m2, double beta);

/* series operators (accumulate into dst) */
void or_rotate (int deg, const double *src, double alpha, double beta,
                double gamma, double *dst);
void or_rotate_inverse (int deg, const double *src, double alpha,
                        double beta, double gamma, double *dst);
void or_p2m (int nm, const double *pos, double q, const double *center,
             double *M);
void or_m2l_rotate (int nm, const double *M, const double *xsrc,
                    int pl, double *L, const double *xdst);
void or_m2m_rotate (int nm, const double *Msrc, const double *xsrc,
                    int nmdst, double *Mdst, const double *xdst);
void or_l2l_rotate (int pl, const double *Lsrc, const double *xsrc,
                    int pldst, double *Ldst, const double *xdst);
/* alternate flavours of the same translations (flavour 1: the O(p^4)
 * "normal" formulas, 2: K. Kylin's recurrences); same arguments as the
 * rotate flavour above */
void or_m2m_flavour (int flavour, int nm, const double *Msrc,
                     const double *xsrc, int nmdst, double *Mdst,
                     const double *xdst);
void or_l2l_flavour (int flavour, int pl, const double *Lsrc,
                     const double *xsrc, int pldst, double *Ldst,
                     const double *xdst);
void or_m2l_flavour (int flavour, int nm, const double *M, const double *xsrc,
                     int pl, double *L, const double *xdst);
double or_local_evaluate (int pl, const double *L, const double *center,
                          const double *pos);
void or_local_gradient (int pl, const double *L, const double *center,
                        const double *pos, double *grad);
double or_multipole_evaluate (int nm, const double *M, const double *center,
                              const double *pos);
void or_p2l (int pl, const double *pos, double q, const double *center,
             double *L);
void or_multipole_gradient (int nm, const double *M, const double *center,
                            const double *pos, double *grad);

/* tree: returns node count; arrays sized by caller via or_tree_query */
typedef struct
{
  int depth, ix, iy, iz;
  int pbegin, pend;
  int parent, isleaf;
  int child_begin, nchild;
  double center[3];
} OrNode;

void or_root_box (const double *lb, const double *ub, long n, const double *x,
                  const double *y, const double *z, double *lo, double *width);
void or_keys (long n, const double *x, const double *y, const double *z,
              const double *lo, double width, uint64_t *keys);
long or_tree_build (long n, const double *x, const double *y, const double *z,
                    const double *lb, const double *ub, int leafcap,
                    long *perm /* n */, OrNode *nodes, long max_nodes);

/* near/far traversal: writes ordered far pairs (src,dst) and unordered
 * near pairs (a,b) with a<=b; returns counts through out-params. */
void or_traverse (const OrNode *nodes, long nnodes,
                  long *far_src, long *far_dst, long max_far, long *nfar,
                  long *near_a, long *near_b, long max_near, long *nnear);

/* the same with semifar pairs (leaf, deeper node) when semifar_thr >= 0 */
void or_traverse_full (const OrNode *nodes, long nnodes, long semifar_thr,
                       long *far_src, long *far_dst, long max_far, long *nfar,
                       long *near_a, long *near_b, long max_near, long *nnear,
                       long *semi_leaf, long *semi_node, long max_semi,
                       long *nsemi);

/* whole solve.  stats: zero, p2p, p2m, m2m, m2l, l2l, l2p (7 longs). */
int or_fmm_solve (long n, const double *x, const double *y, const double *z,
                  const double *q, const double *lb, const double *ub,
                  int order, int leafcap, double *pot, double *grad,
                  long *stats);

/* whole solve with semifar pairs when semifar_thr >= 0 (none when < 0).
 * stats: the 7 above, then p2l, m2p (9 longs). */
int or_fmm_solve_full (long n, const double *x, const double *y,
                       const double *z, const double *q, const double *lb,
                       const double *ub, int order, int leafcap,
                       long semifar_thr, double *pot, double *grad,
                       long *stats);

/* FP64 direct sum with compensated accumulation, on chosen targets */
void or_direct (long n, const double *x, const double *y, const double *z,
                const double *q, long nt, const long *targets, double *pot,
                double *grad);

#ifdef __cplusplus
}
#endif

#endif
