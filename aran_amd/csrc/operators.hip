// operators.hip -- FP64 expansion operators of the 3D Newton FMM on gfx950.
//
//   k_wigner_tables : Wigner d^l_{a,b}(theta) by the three-term recurrence
//                     in l for fixed (a, b), FP64, one block per theta;
//                     folded into the real rotation matrices E+-, F+-
//   k_p2m           : aran_development3d_p2m (arandevelopment3d.c:169-216)
//   k_m2l           : aran_development3d_m2l_rotate (arandevelopment3d.c:
//                     423-430, rotate.c:165-207, translate.c:609-676)
//   k_l2p           : local evaluate + gradient (aransphericalseriesd.c:
//                     313-546)
//
// Rotation.  The reference applies D^l_{m,m'} = e^{-im alpha} d^l_{m,m'}
// e^{-im' gamma} to src[l][m] with src[l][-m] = (-1)^m conj(src[l][m])
// (rotate.c:34-80).  With the z-flipped translation angles (alpha, beta,
// gamma) = (-phi, theta, 0) and the inverse (0, -theta, phi), the phases
// factor out and each degree becomes two REAL (l+1)x(l+1) products:
//   forward: u_m = e^{i m phi} s_m;  Re X = E+ Re u,  Im X = E- Im u,
//            E+-[m'][m] = d_{m,m'}(theta) +- [m>0] (-1)^m d_{-m,m'}(theta)
//   inverse: Re W = F+ Re s, Im W = F- Im s, dst_m' = e^{-i m' phi} W_m',
//            F+-[m'][m] = d_{m',m}(theta) +- [m>0] (-1)^m d_{m',-m}(theta)
// (d(-theta) is the transpose of d(theta)).
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <type_traits>

#include "engine.h"
#include "tree_rules.h"

namespace aran {

__device__ inline void lm_of(int k, int &l, int &m) {
  int t = (int)((sqrtf(8.f * k + 1.f) - 1.f) * 0.5f);
  while ((t + 1) * (t + 2) / 2 <= k) t++;
  while (t * (t + 1) / 2 > k) t--;
  l = t;
  m = k - t * (t + 1) / 2;
}

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline double2 cmulc(double2 a, double2 b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// ------------------------------------------------------------ Wigner tables

__device__ double binom_exact(int n, int k) {
  double c = 1.;
  for (int i = 1; i <= k; i++) c = c * (double)(n - k + i) / (double)i;
  return c;
}

__device__ double ipow(double x, int e) {
  double r = 1.;
  for (int i = 0; i < e; i++) r *= x;
  return r;
}

// d^j_{a,b}(beta) for j = max(|a|,|b|) (closed forms of Wigner's sum with a
// single term).
__device__ double wigner_seed(int a, int b, double c, double s) {
  int aa = abs(a), ab = abs(b);
  if (aa >= ab) {
    int j = aa;
    double sq = sqrt(binom_exact(2 * j, j + b));
    if (a == j) {
      double v = sq * ipow(c, j + b) * ipow(s, j - b);
      return ((j - b) & 1) ? -v : v;
    }
    return sq * ipow(c, j - b) * ipow(s, j + b);
  }
  int j = ab;
  double sq = sqrt(binom_exact(2 * j, j + a));
  if (b == j) return sq * ipow(c, j + a) * ipow(s, j - a);
  double v = sq * ipow(c, j - a) * ipow(s, j + a);
  return ((a + j) & 1) ? -v : v;
}

// One block per theta; threads over the (2L+1)^2 pairs (a, b).
__global__ void k_wigner_tables(const double *__restrict__ thetas, int L,
                                double *__restrict__ dscr,
                                double *__restrict__ rot) {
  const int w = 2 * L + 1;
  const int sq = w * w;
  const int th = blockIdx.x;
  double *dth = dscr + (size_t)th * (L + 1) * sq;
  const double beta = thetas[th];
  const double cb = cos(beta), c = cos(0.5 * beta), s = sin(0.5 * beta);
  for (int t = threadIdx.x; t < sq; t += blockDim.x) {
    int a = t / w - L, b = t % w - L;
    int j = max(abs(a), abs(b));
    double dm2 = 0., dm1 = wigner_seed(a, b, c, s);
    dth[(size_t)j * sq + t] = dm1;
    for (int l = j + 1; l <= L; l++) {
      double ll = (double)l * l;
      double l1 = (double)(l - 1) * (l - 1);
      double pre = l * (2. * l - 1.) / sqrt((ll - a * a) * (ll - b * b));
      double mid = (l > 1) ? cb - (double)(a * b) / (l * (l - 1.)) : cb;
      double back = (l > 1) ? sqrt((l1 - a * a) * (l1 - b * b)) /
                                  ((l - 1.) * (2. * l - 1.))
                            : 0.;
      double dl = pre * (mid * dm1 - back * dm2);
      dth[(size_t)l * sq + t] = dl;
      dm2 = dm1;
      dm1 = dl;
    }
  }
  __syncthreads();
  const int rb = rot_block(L);
  double *R = rot + (size_t)th * 4 * rb;
  for (int idx = threadIdx.x; idx < rb; idx += blockDim.x) {
    int l = 0;
    while (rot_ofs(l + 1) <= idx) l++;
    int r = idx - rot_ofs(l);
    int mp = r / (l + 1), m = r % (l + 1);
    const double *dl = dth + (size_t)l * sq;
#define DD(aa, bb) dl[((aa) + L) * w + ((bb) + L)]
    double sgn = (m & 1) ? -1. : 1.;
    double e0 = DD(m, mp), e1 = (m > 0) ? sgn * DD(-m, mp) : 0.;
    double f0 = DD(mp, m), f1 = (m > 0) ? sgn * DD(mp, -m) : 0.;
#undef DD
    R[idx] = e0 + e1;
    R[rb + idx] = e0 - e1;
    R[2 * rb + idx] = f0 + f1;
    R[3 * rb + idx] = f0 - f1;
  }
}

hipError_t launch_wigner_tables(const double *thetas, int ntheta, int L,
                                double *dscr, double *rot, hipStream_t s) {
  int w = 2 * L + 1;
  int nt = std::min(1024, ((w * w + 63) / 64) * 64);
  k_wigner_tables<<<ntheta, nt, 0, s>>>(thetas, L, dscr, rot);
  return hipGetLastError();
}

// ------------------------------------------------------------ shared pieces

constexpr int kMaxMc = kMaxOrder * (kMaxOrder + 1) / 2;        // 465
constexpr int kMaxLc = (kMaxOrder + 1) * (kMaxOrder + 2) / 2;  // 496
constexpr int kOpThreads = 64;
// coefficients per thread of a 64-thread block over a local
constexpr int kLcSlots = (kMaxLc + kOpThreads - 1) / kOpThreads;



// X[l][m'] = rotation (forward) of U over degrees 0..deg.
__device__ inline void rotate_fwd(const double *__restrict__ R, int rb,
                                  int deg, const double2 *U, double2 *X) {
  int n = tri(deg + 1);
  for (int k = threadIdx.x; k < n; k += kOpThreads) {
    int l, mp;
    lm_of(k, l, mp);
    const double *ep = R + rot_ofs(l) + mp * (l + 1);
    const double *em = ep + rb;
    const double2 *u = U + tri(l);
    double re = 0., im = 0.;
    for (int m = 0; m <= l; m++) {
      re += ep[m] * u[m].x;
      im += em[m] * u[m].y;
    }
    X[k] = make_double2(re, im);
  }
}

// out[k] (+)= e^{-i m' phi} * (inverse rotation of Z), thread-strided slots
__device__ inline double2 rotate_inv_one(const double *__restrict__ R, int rb,
                                         const double2 *__restrict__ ph,
                                         const double2 *Z, int k) {
  int l, mp;
  lm_of(k, l, mp);
  const double *fp = R + 2 * rb + rot_ofs(l) + mp * (l + 1);
  const double *fm = fp + rb;
  const double2 *z = Z + tri(l);
  double re = 0., im = 0.;
  for (int m = 0; m <= l; m++) {
    re += fp[m] * z[m].x;
    im += fm[m] * z[m].y;
  }
  return cmulc(ph[mp], make_double2(re, im));
}

// ------------------------------------------------------------ P2M / L2P
//
// Both go through the normalised regular solid harmonics
//   O_l^m = r^l P_l^m(cos theta) e^{i m phi} / (l+m)!
// (P with the Condon-Shortley phase), from
//   O_0^0 = 1,  O_m^m = -(x + iy) / (2m) O_{m-1}^{m-1},
//   O_l^m = ((2l-1) z O_{l-1}^m - r^2 O_{l-2}^m) / (l^2 - m^2)   (l > m),
// which need no square root or angle and are exact at r = 0, and whose
// derivatives are again harmonics of one degree less:
//   d/dz O_l^m = O_{l-1}^m,  (d/dx + i d/dy) O_l^m = O_{l-1}^{m+1},
//   (d/dx - i d/dy) O_l^m = -O_{l-1}^{m-1}.
// With g_l^m = N_l^m (l+m)!, r^l Y_l^m = g_l^m O_l^m.

// 1/n for 0 < n < 1024 (n = l^2 - m^2 <= kMaxOrder^2), each correctly
// rounded (compile-time constants).
struct InvTable {
  double v[1024];
};
static_assert(kMaxOrder * kMaxOrder < 1024, "reciprocal table too short");
constexpr InvTable make_inv_table() {
  InvTable t{};
  for (int i = 1; i < 1024; i++) t.v[i] = 1.0 / i;
  return t;
}
__constant__ InvTable kInv = make_inv_table();

constexpr int kP2mSlots = 1024;  // P2L LDS staging, double2 (16 KiB)

template <int B, int E, class F>
__device__ __forceinline__ void static_loop(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_loop<B + 1, E>(f);
  }
}

// One wave per leaf: M_l^m = (4pi/(2l+1)) g_l^m conj(sum_j q_j O_l^m(x_j))
// (aran_development3d_p2m, arandevelopment3d.c:169-216).  Lane s P + m
// (s < 64 / P particle slots) runs column m of the degree recurrence for the
// leaf's particles s, s + S, ... in registers and accumulates q O_l^m for
// l >= m; the slots are then added in slot order through LDS.  No memory
// access sits in the recurrence (the reciprocals come from a constant
// table), and the LDS is small enough that several leaves share a CU with
// the near field.
//
// R = double: the plain form.  R = float (the reduced-precision forms, for
// leaves below the top two levels): the recurrence runs on coordinates
// scaled by the leaf's width, u = (x - c) / w (|u| <= sqrt(3)/2, so the
// harmonics stay in FP32 range at every order), and each slot's sums are
// scaled back by w^l in FP64; FP32 issues no faster than FP64 on this VALU,
// but the lane needs half the registers and the near field beside it gets
// the issue slots (measured 0.22 -> 0.12 ms at N = 2^20, p = 10).
template <int P, class R>
__device__ __forceinline__ void p2m_slots(
    int b, int e, int s, int m, double cx, double cy, double cz, double w,
    const double *__restrict__ xs, const double *__restrict__ ys,
    const double *__restrict__ zs, const double *__restrict__ qs,
    double2 *__restrict__ Ts) {
  constexpr bool kScaled = !std::is_same<R, double>::value;
  constexpr int S = 64 / P;
  const double iw = kScaled ? 1. / w : 1.;
  R ar[P], ai[P];
  // factors 1 / (l^2 - m^2) of the degree recurrence: m is fixed per lane,
  // so they are read once, not per particle (a per-lane table read inside
  // the recurrence is a dependent memory round trip per degree)
  R gfac[P];
  static_loop<0, P>([&](auto lc) {
    constexpr int l = decltype(lc)::value;
    ar[l] = R(0);
    ai[l] = R(0);
    gfac[l] = (R)kInv.v[l > m ? l * l - m * m : 1];
  });
  // the next particle's coordinates are loaded while this one is summed
  int i = b + s;
  double nx_ = 0., ny_ = 0., nz_ = 0., nq_ = 0.;
  if (i < e) {
    nx_ = xs[i];
    ny_ = ys[i];
    nz_ = zs[i];
    nq_ = qs[i];
  }
  for (; i < e; i += S) {
    R vx, vy, vz;
    if constexpr (kScaled) {
      vx = (R)((nx_ - cx) * iw);
      vy = (R)((ny_ - cy) * iw);
      vz = (R)((nz_ - cz) * iw);
    } else {
      vx = nx_ - cx;
      vy = ny_ - cy;
      vz = nz_ - cz;
    }
    const R qi = (R)nq_;
    if (i + S < e) {
      nx_ = xs[i + S];
      ny_ = ys[i + S];
      nz_ = zs[i + S];
      nq_ = qs[i + S];
    }
    const R r2 = vx * vx + vy * vy + vz * vz;
    R ox = qi, oy = R(0);  // q O_m^m
    static_loop<1, P>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      if (k <= m) {
        const R f = (R)(-0.5 * kInv.v[k]);
        const R ax = f * vx, ay = f * vy;
        const R nx = ox * ax - oy * ay, ny = ox * ay + oy * ax;
        ox = nx;
        oy = ny;
      }
    });
    R px = R(0), py = R(0), qx = R(0), qy = R(0);  // degrees l-1, l-2
    static_loop<0, P>([&](auto lc) {
      constexpr int l = decltype(lc)::value;
      if (l == m) {
        px = ox;
        py = oy;
        ar[l] += ox;
        ai[l] += oy;
      } else if (l > m) {
        const R a = (R)(2. * l - 1.) * vz;
        const R g = gfac[l];
        const R nx = (a * px - r2 * qx) * g, ny = (a * py - r2 * qy) * g;
        qx = px;
        qy = py;
        px = nx;
        py = ny;
        ar[l] += nx;
        ai[l] += ny;
      }
    });
  }
  double wl = 1.;  // w^l of the scaled form
  static_loop<0, P>([&](auto lc) {
    constexpr int l = decltype(lc)::value;
    if (l >= m)
      Ts[tri(l) + m] = make_double2((double)ar[l] * wl, (double)ai[l] * wl);
    if (kScaled) wl *= w;
  });
}

template <int P, class R>
__global__ __launch_bounds__(64) void k_p2m(
    OpTables tb, NodeArrays nd, const int *__restrict__ leaves,
    const double *__restrict__ xs, const double *__restrict__ ys,
    const double *__restrict__ zs, const double *__restrict__ qs,
    double2 *__restrict__ M) {
  // issue priority over the near field's waves (second stream), which
  // would otherwise hold up this latency-bound step of the far-field chain
  __builtin_amdgcn_s_setprio(2);
  constexpr int MC = P * (P + 1) / 2, S = 64 / P;
  __shared__ double2 T[S][MC];
  const int lane = threadIdx.x;
  const int A = leaves[blockIdx.x];
  const int b = nd.pbegin[A], e = nd.pend[A];
  const double cx = nd.cx[A], cy = nd.cy[A], cz = nd.cz[A];
  const int s = lane / P, m = lane % P;
  // the width only for the scaled form (the single-operator entry points
  // describe their one leaf by its centre: no system box)
  if (s < S)
    p2m_slots<P, R>(b, e, s, m, cx, cy, cz,
                    std::is_same<R, float>::value ? node_width(nd, A) : 1.,
                    xs, ys, zs, qs, T[s]);
  __syncthreads();
  for (int k = lane; k < MC; k += 64) {
    double2 acc = make_double2(0., 0.);
    for (int t = 0; t < S; t++) {
      acc.x += T[t][k].x;
      acc.y += T[t][k].y;
    }
    const double g = tb.pgain[k];
    M[(size_t)A * MC + k] = make_double2(g * acc.x, -g * acc.y);
  }
}

// ------------------------------------------------------------ M2L

// One block per target node: L(A) = sum over its far list of
// rotate -> vertical M2L -> rotate back.  Writes L(A) (zero when the list
// is empty).
__global__ __launch_bounds__(kOpThreads) void k_m2l(
    OpTables tb, NodeArrays nd, const int *__restrict__ vofs,
    const int *__restrict__ vsrc, const short *__restrict__ vcode,
    const double2 *__restrict__ M, double2 *__restrict__ L) {
  __shared__ double2 U[kMaxMc], X[kMaxMc], Z[kMaxLc];
  __shared__ double rp[2 * kMaxOrder + 2];
  const int A = blockIdx.x;
  const int p = tb.p, mc = tri(p), lc = tri(p + 1), rb = tb.rb;
  const double w = node_width(nd, A);
  double2 acc[kLcSlots];
  for (int slot = 0; slot < kLcSlots; slot++) acc[slot] = make_double2(0., 0.);
  const int e0 = vofs[A], e1 = vofs[A + 1];
  for (int e = e0; e < e1; e++) {
    const int B = vsrc[e];
    const int code = vcode[e];
    DirEntry de = tb.dir[code];
    const double2 *ph = tb.eiphi + code * (p + 1);
    const double *R = tb.rot + (size_t)de.theta * 4 * rb;
    for (int k = threadIdx.x; k < mc; k += kOpThreads) {
      int l, m;
      lm_of(k, l, m);
      U[k] = cmul(ph[m], M[(size_t)B * mc + k]);
    }
    if (threadIdx.x < 2 * p + 2) {
      double rho = 1. / (de.norm * w);
      int kk = threadIdx.x;
      double v = 1.;
      for (int i = 0; i < kk; i++) v *= rho;
      if (kk > 0 && ((kk - 1) & 1)) v *= de.c;  // c^{l+n} with kk = l+n+1
      rp[kk] = v;
    }
    __syncthreads();
    rotate_fwd(R, rb, p - 1, U, X);
    __syncthreads();
    for (int k = threadIdx.x; k < lc; k += kOpThreads) {
      int l, m;
      lm_of(k, l, m);
      const double *T = tb.tm2l + ((size_t)l * (p + 1) + m) * p;
      double re = 0., im = 0.;
      for (int n = m; n < p; n++) {
        double f = T[n] * rp[l + n + 1];
        re += f * X[tri(n) + m].x;
        im += f * X[tri(n) + m].y;
      }
      if ((l + m) & 1) {
        re = -re;
        im = -im;
      }
      Z[k] = make_double2(re, im);
    }
    __syncthreads();
    for (int slot = 0; slot < kLcSlots; slot++) {
      int k = threadIdx.x + slot * kOpThreads;
      if (k < lc) {
        double2 v = rotate_inv_one(R, rb, ph, Z, k);
        acc[slot].x += v.x;
        acc[slot].y += v.y;
      }
    }
    __syncthreads();
  }
  for (int slot = 0; slot < kLcSlots; slot++) {
    int k = threadIdx.x + slot * kOpThreads;
    if (k < lc) L[(size_t)A * lc + k] = acc[slot];
  }
}

// Dense scaled M2L operators: for offset code `code` (box units, width 1)
// and real basis input k, column k of the real matrix mapping
//   (Re M^_l^m for all (l,m), Im M^_l^m for m >= 1)   [2 mc - p inputs]
// to (Re L^_l^m for all (l,m), Im L^_l^m for m >= 1)  [2 lc - p - 1 outputs]
// through the same rotate / z-translate / rotate-back operator as k_m2l
// (M^_n = M_n / w^n, L^_l = L_l w^(l+1), so w = 1).  Near codes (every
// |component| <= 1) give zero columns.  O is [343][ro][ki] FP64.
__global__ __launch_bounds__(kOpThreads) void k_m2l_dense(OpTables tb,
                                                          double *__restrict__ O) {
  __shared__ double2 U[kMaxMc], X[kMaxMc], Z[kMaxLc], Lo[kMaxLc];
  __shared__ double rp[2 * kMaxOrder + 2];
  const int p = tb.p, mc = tri(p), lc = tri(p + 1), rb = tb.rb;
  const int ki = 2 * mc - p, ro = 2 * lc - p - 1;
  const int code = blockIdx.x / ki, kb = blockIdx.x % ki;
  const int dx = code / 49 - 3, dy = (code / 7) % 7 - 3, dz = code % 7 - 3;
  double *col = O + (size_t)code * ro * ki + kb;
  if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) {
    for (int r = threadIdx.x; r < ro; r += kOpThreads) col[(size_t)r * ki] = 0.;
    return;
  }
  // basis input: Re of coefficient kb (kb < mc) or Im of the (l, m >= 1)
  // coefficient number kb - mc in tri order without m = 0
  int bk = kb, bim = 0;
  if (kb >= mc) {
    int j = kb - mc, l = 1;
    while (tri(l) <= j) l++;  // j = tri(l-1) + m - 1
    bk = tri(l) + (j - tri(l - 1)) + 1;
    bim = 1;
  }
  const DirEntry de = tb.dir[code];
  const double2 *ph = tb.eiphi + code * (p + 1);
  const double *R = tb.rot + (size_t)de.theta * 4 * rb;
  for (int k = threadIdx.x; k < mc; k += kOpThreads) {
    int l, m;
    lm_of(k, l, m);
    const double2 b = make_double2(k == bk && !bim ? 1. : 0.,
                                   k == bk && bim ? 1. : 0.);
    U[k] = cmul(ph[m], b);
  }
  if (threadIdx.x < 2 * p + 2) {
    const double rho = 1. / de.norm;
    const int kk = threadIdx.x;
    double v = 1.;
    for (int i = 0; i < kk; i++) v *= rho;
    if (kk > 0 && ((kk - 1) & 1)) v *= de.c;  // c^{l+n} with kk = l+n+1
    rp[kk] = v;
  }
  __syncthreads();
  rotate_fwd(R, rb, p - 1, U, X);
  __syncthreads();
  for (int k = threadIdx.x; k < lc; k += kOpThreads) {
    int l, m;
    lm_of(k, l, m);
    const double *T = tb.tm2l + ((size_t)l * (p + 1) + m) * p;
    double re = 0., im = 0.;
    for (int n = m; n < p; n++) {
      const double f = T[n] * rp[l + n + 1];
      re += f * X[tri(n) + m].x;
      im += f * X[tri(n) + m].y;
    }
    if ((l + m) & 1) {
      re = -re;
      im = -im;
    }
    Z[k] = make_double2(re, im);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < lc; k += kOpThreads)
    Lo[k] = rotate_inv_one(R, rb, ph, Z, k);
  __syncthreads();
  for (int r = threadIdx.x; r < ro; r += kOpThreads) {
    double v;
    if (r < lc) {
      v = Lo[r].x;
    } else {
      const int j = r - lc;
      int l = 1;
      while (tri(l) <= j) l++;
      v = Lo[tri(l) + (j - tri(l - 1)) + 1].y;
    }
    col[(size_t)r * ki] = v;
  }
}

hipError_t launch_m2l_dense(const OpTables &tb, double *O, hipStream_t s) {
  const int p = tb.p, ki = 2 * tri(p) - p;
  k_m2l_dense<<<kNumCodes * ki, kOpThreads, 0, s>>>(tb, O);
  return hipGetLastError();
}

// ------------------------------------------------------------ L2P

// The terms of one particle at order p (run time), u its offset from the
// leaf centre and La its leaf's row of c_l^m.
template <class R, class R2>
__device__ __forceinline__ void l2p_terms_rt(const R2 *__restrict__ La,
                                           int p, R ux, R uy, R uz,
                                           R &phi, R &dz, R &gx, R &gy) {
  const R r2 = ux * ux + uy * uy + uz * uz;
  R mx = R(1), my = R(0);  // O_m^m
  for (int m = 0; m <= p; m++) {
    if (m > 0) {
      const R f = (R)(-0.5 * kInv.v[m]);
      const R ax = f * ux, ay = f * uy;
      const R nx = mx * ax - my * ay, ny = mx * ay + my * ax;
      mx = nx;
      my = ny;
    }
    const R w = m ? R(2) : R(1);
    R ox = mx, oy = my, qx = R(0), qy = R(0);  // O_j^m, O_{j-1}^m
    R2 c = La[tri(m) + m];
    for (int j = m; j <= p; j++) {
      if (j > m) {
        const R a = (R)(2 * j - 1) * uz, sj = (R)kInv.v[j * j - m * m];
        const R nx = (a * ox - r2 * qx) * sj, ny = (a * oy - r2 * qy) * sj;
        qx = ox;
        qy = oy;
        ox = nx;
        oy = ny;
      }
      phi += w * (c.x * ox - c.y * oy);
      if (j < p) {
        const int k1 = tri(j + 1) + m;
        const R2 c1 = La[k1];  // c_{j+1}^m
        dz += w * (c1.x * ox - c1.y * oy);
        if (m > 0) {  // c_{j+1}^{m-1} O_j^m
          const R2 c2 = La[k1 - 1];
          gx += c2.x * ox - c2.y * oy;
          gy += c2.x * oy + c2.y * ox;
        }
        {  // - conj(c_{j+1}^{m+1} O_j^m)
          const R2 c3 = La[k1 + 1];
          gx -= c3.x * ox - c3.y * oy;
          gy += c3.x * oy + c3.y * ox;
        }
        c = c1;
      }
    }
  }
}

// The same terms at order P (compile time): the recurrence factors and the
// row offsets are constants, so the loop body is multiplies, adds and LDS
// reads with immediate offsets.
template <int P, class R, class R2>
__device__ __forceinline__ void l2p_terms(const R2 *__restrict__ La, R ux,
                                          R uy, R uz, R &phi, R &dz, R &gx,
                                          R &gy) {
  const R r2 = ux * ux + uy * uy + uz * uz;
  R mx = R(1), my = R(0);  // O_m^m
  static_loop<0, P + 1>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if constexpr (m > 0) {
      constexpr R f = (R)(-0.5 * (1.0 / m));
      const R ax = f * ux, ay = f * uy;
      const R nx = mx * ax - my * ay, ny = mx * ay + my * ax;
      mx = nx;
      my = ny;
    }
    constexpr R w = m ? R(2) : R(1);
    R ox = mx, oy = my, qx = R(0), qy = R(0);  // O_j^m, O_{j-1}^m
    static_loop<m, P + 1>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      if constexpr (j > m) {
        constexpr R a0 = (R)(2 * j - 1), sj = (R)(1.0 / (j * j - m * m));
        const R a = a0 * uz;
        const R nx = (a * ox - r2 * qx) * sj, ny = (a * oy - r2 * qy) * sj;
        qx = ox;
        qy = oy;
        ox = nx;
        oy = ny;
      }
      const R2 c = La[j * (j + 1) / 2 + m];
      phi += w * (c.x * ox - c.y * oy);
      if constexpr (j < P) {
        constexpr int k1 = (j + 1) * (j + 2) / 2 + m;
        const R2 c1 = La[k1];  // c_{j+1}^m
        dz += w * (c1.x * ox - c1.y * oy);
        if constexpr (m > 0) {  // c_{j+1}^{m-1} O_j^m
          const R2 c2 = La[k1 - 1];
          gx += c2.x * ox - c2.y * oy;
          gy += c2.x * oy + c2.y * ox;
        }
        {  // - conj(c_{j+1}^{m+1} O_j^m)
          const R2 c3 = La[k1 + 1];
          gx -= c3.x * ox - c3.y * oy;
          gy += c3.x * oy + c3.y * ox;
        }
      }
    });
  });
}

// One lane per particle (aran_development3d_local_evaluate and
// local_gradient_evaluate, arandevelopment3d.c:485-494,541-552): with
// c_l^m = g_l^m L_l^m of the particle's leaf,
//   phi = sum_l [c_l^0 O_l^0 + 2 Re sum_{m>0} c_l^m O_l^m],
//   d/dz phi = same sum with O_{l-1}^m,
//   (d/dx + i d/dy) phi = sum_l [c_l^0 O_{l-1}^1
//       + sum_{m>0} (c_l^m O_{l-1}^{m+1} - conj(c_l^m O_{l-1}^{m-1}))],
// accumulated per harmonic O_j^m as it is generated (column by column).
//
// A block takes 256 consecutive sorted particles; the c rows of their
// leaves (a leaf's particles are consecutive) are staged in LDS, at most
// `rows` leaves per round, so the recurrence reads its coefficients from
// LDS instead of waiting on a global load per degree step.  R = float (the
// reduced-precision forms, particles of leaves below the top two levels):
// the recurrence runs on coordinates scaled by the leaf width,
// u = (x - c) / w, on coefficients c_l^m w^l (staged in FP32), and the
// gradient takes the factor 1 / w at the end.  sel: 0 every particle, 1 the
// particles of leaves at depth >= 2, 2 those of leaves at depth < 2 (the
// choice is per leaf, so a system's result does not depend on its batch).
template <int P, class R>
__global__ __launch_bounds__(256) void k_l2p(
    OpTables tb, NodeArrays nd, const int *__restrict__ pleaf, int n,
    const double *__restrict__ xs, const double *__restrict__ ys,
    const double *__restrict__ zs, const double2 *__restrict__ L,
    double *__restrict__ pot, double *__restrict__ grad, int rows, int sel,
    const int *__restrict__ perm, double *__restrict__ opot,
    double *__restrict__ ograd) {
  using R2 = typename std::conditional<std::is_same<R, float>::value, float2,
                                       double2>::type;
  constexpr bool kScaled = std::is_same<R, float>::value;
  extern __shared__ double2 l2p_shared[];
  R2 *srow = reinterpret_cast<R2 *>(l2p_shared);
  __shared__ int sleaf[256], wbase[4];
  __shared__ double sw[256];
  const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
  const int i = blockIdx.x * 256 + tid;
  const int p = tb.p, lc = tri(p + 1);
  const double *__restrict__ g = tb.ogain;
  int A = i < n ? pleaf[i] : -1;
  if (A >= 0 && sel != 0 && ((nd.depth[A] >= 2) != (sel == 1))) A = -1;
  sleaf[tid] = A;
  __syncthreads();
  // slot of the particle's leaf among the block's distinct leaves
  const bool start = A >= 0 && (tid == 0 || sleaf[tid - 1] != A);
  const unsigned long long bal = __ballot(start);
  const int before = __popcll(bal & ((2ull << lane) - 1ull));  // <= lane
  if (lane == 63) wbase[wv] = __popcll(bal);
  __syncthreads();
  int slot = before - 1, nslots = 0;
  for (int k = 0; k < 4; k++) {
    if (k < wv) slot += wbase[k];
    nslots += wbase[k];
  }
  __syncthreads();  // sleaf is rewritten below
  if (start) sleaf[slot] = A;
  __syncthreads();
  double vx = 0., vy = 0., vz = 0.;
  if (A >= 0) {
    vx = xs[i] - nd.cx[A];
    vy = ys[i] - nd.cy[A];
    vz = zs[i] - nd.cz[A];
  }
  for (int r0 = 0; r0 < nslots; r0 += rows) {
    const int nr = min(rows, nslots - r0);
    if (r0 > 0) __syncthreads();  // the previous round's rows are read
    // leaf widths, for the scaled form only (the single-operator entry
    // points describe their one leaf by its centre: no system box)
    if (kScaled)
      for (int t = tid; t < nr; t += 256)
        sw[t] = node_width(nd, sleaf[r0 + t]);
    __syncthreads();
    for (int e = tid; e < nr * lc; e += 256) {
      const int rr = e / lc, k = e % lc;
      double2 c = L[(size_t)sleaf[r0 + rr] * lc + k];
      double f = g[k];
      if (kScaled) {
        int l = 0;
        while (tri(l + 1) <= k) l++;
        double wl = 1.;
        for (int t = 0; t < l; t++) wl *= sw[rr];
        f *= wl;
      }
      srow[(size_t)rr * lc + k] = R2{(R)(c.x * f), (R)(c.y * f)};
    }
    __syncthreads();
    if (A < 0 || slot < r0 || slot >= r0 + nr) continue;
    const R2 *__restrict__ La = srow + (size_t)(slot - r0) * lc;
    R ux, uy, uz;
    if (kScaled) {
      const double iw = 1. / sw[slot - r0];
      ux = (R)(vx * iw);
      uy = (R)(vy * iw);
      uz = (R)(vz * iw);
    } else {
      ux = vx;
      uy = vy;
      uz = vz;
    }
    R phi = R(0), dz = R(0), gx = R(0), gy = R(0);
    if constexpr (P > 0)
      l2p_terms<P, R, R2>(La, ux, uy, uz, phi, dz, gx, gy);
    else
      l2p_terms_rt<R, R2>(La, p, ux, uy, uz, phi, dz, gx, gy);
    double s = 1.;
    if (kScaled) s = 1. / sw[slot - r0];
    const double rp = pot[i] + (double)phi;
    const double rx = grad[3 * (size_t)i + 0] + (double)gx * s;
    const double ry = grad[3 * (size_t)i + 1] + (double)gy * s;
    const double rz = grad[3 * (size_t)i + 2] + (double)dz * s;
    if (!perm) {
      pot[i] = rp;
      grad[3 * (size_t)i + 0] = rx;
      grad[3 * (size_t)i + 1] = ry;
      grad[3 * (size_t)i + 2] = rz;
    } else {  // the last term of the particle: straight to the caller's order
      const size_t j = (size_t)perm[i];
      if (opot) opot[j] = rp;
      if (ograd) {
        ograd[3 * j + 0] = rx;
        ograd[3 * j + 1] = ry;
        ograd[3 * j + 2] = rz;
      }
    }
  }
}

// ------------------------------------------------------------ M2P / P2L
//
// Semifar operators (aran_development3d_m2pv / multipole_evaluate and
// aran_development3d_p2l, arandevelopment3d.c:227-277,485-560) through the
// irregular solid harmonics
//   I_l^m = (l-m)! P_l^m(cos theta) e^{i m phi} / r^(l+1),
//   I_0^0 = 1/r,  I_m^m = -(2m-1) (x + iy) / r^2 I_{m-1}^{m-1},
//   I_l^m = ((2l-1) z I_{l-1}^m - ((l-1)^2 - m^2) I_{l-2}^m) / r^2,
// whose derivatives are harmonics of one degree more:
//   d/dz I_l^m = -I_{l+1}^m,  (d/dx + i d/dy) I_l^m = I_{l+1}^{m+1},
//   (d/dx - i d/dy) I_l^m = -I_{l+1}^{m-1}.
// With h_l^m = N_l^m / (l-m)!, r^-(l+1) Y_l^m = h_l^m I_l^m.

// One wave per leaf, one lane per particle: for every semifar node D of the
// leaf, with c_l^m = h_l^m M_l^m of D (degrees l < p),
//   phi = sum_l [c_l^0 I_l^0 + 2 Re sum_{m>0} c_l^m I_l^m],
//   d/dz phi = -(same sum with I_{l+1}^m),
//   (d/dx + i d/dy) phi = sum_l [sum_{m>=0} c_l^m I_{l+1}^{m+1}
//                                - sum_{m>0} conj(c_l^m I_{l+1}^{m-1})],
// accumulated per harmonic I_j^m as it is generated (column by column).
__global__ __launch_bounds__(64) void k_m2p(
    OpTables tb, NodeArrays nd, const int *__restrict__ leaves,
    const int *__restrict__ wofs, const int *__restrict__ wnode,
    const double *__restrict__ xs, const double *__restrict__ ys,
    const double *__restrict__ zs, const double2 *__restrict__ M,
    double *__restrict__ pot, double *__restrict__ grad) {
  const int t = blockIdx.x;
  const int w0 = wofs[t], w1 = wofs[t + 1];
  if (w0 == w1) return;
  const int A = leaves[t];
  const int p = tb.p, mc = tri(p);
  const double *__restrict__ g = tb.igain;
  for (int i = nd.pbegin[A] + threadIdx.x; i < nd.pend[A]; i += 64) {
    double phi = 0., dz = 0., gx = 0., gy = 0.;
    for (int w = w0; w < w1; w++) {
      const int D = wnode[w];
      const double2 *__restrict__ Md = M + (size_t)D * mc;
      const double vx = xs[i] - nd.cx[D], vy = ys[i] - nd.cy[D],
                   vz = zs[i] - nd.cz[D];
      const double r2 = vx * vx + vy * vy + vz * vz;
      const double ir2 = 1. / r2;
      double mx = sqrt(ir2), my = 0.;  // I_m^m
      for (int m = 0; m <= p; m++) {
        if (m > 0) {
          const double f = -(2. * m - 1.) * ir2;
          const double ax = f * vx, ay = f * vy;
          const double nx = mx * ax - my * ay, ny = mx * ay + my * ax;
          mx = nx;
          my = ny;
        }
        const double wm = m ? 2. : 1.;
        double ox = mx, oy = my, qx = 0., qy = 0.;  // I_j^m, I_{j-1}^m
        for (int j = m; j <= p; j++) {
          if (j > m) {
            const double a = (2. * j - 1.) * vz;
            const double b = (double)((j - 1) * (j - 1) - m * m);
            const double nx = (a * ox - b * qx) * ir2,
                         ny = (a * oy - b * qy) * ir2;
            qx = ox;
            qy = oy;
            ox = nx;
            oy = ny;
          }
          if (j < p) {  // c_j^m I_j^m
            const int k = tri(j) + m;
            const double cx = g[k] * Md[k].x, cy = g[k] * Md[k].y;
            phi += wm * (cx * ox - cy * oy);
          }
          if (j >= 1 && j - 1 <= p - 1) {
            const int l = j - 1;
            if (l >= m) {  // -c_l^m I_{l+1}^m
              const int k = tri(l) + m;
              const double cx = g[k] * Md[k].x, cy = g[k] * Md[k].y;
              dz -= wm * (cx * ox - cy * oy);
            }
            if (m >= 1) {  // c_l^{m-1} I_{l+1}^m
              const int k = tri(l) + m - 1;
              const double cx = g[k] * Md[k].x, cy = g[k] * Md[k].y;
              gx += cx * ox - cy * oy;
              gy += cx * oy + cy * ox;
            }
            if (l >= m + 1) {  // - conj(c_l^{m+1} I_{l+1}^m)
              const int k = tri(l) + m + 1;
              const double cx = g[k] * Md[k].x, cy = g[k] * Md[k].y;
              gx -= cx * ox - cy * oy;
              gy += cx * oy + cy * ox;
            }
          }
        }
      }
    }
    pot[i] += phi;
    grad[3 * (size_t)i + 0] += gx;
    grad[3 * (size_t)i + 1] += gy;
    grad[3 * (size_t)i + 2] += dz;
  }
}

// One wave per node: L_l^m += (4pi/(2l+1)) h_l^m conj(sum_j q_j I_l^m(x_j))
// over the particles of its semifar leaves (aran_development3d_p2l).  As in
// k_p2m, lanes take (particle, m) columns of a chunk into LDS, then lanes take
// coefficients and add the chunk's particles in order.
__global__ __launch_bounds__(64) void k_p2l(
    OpTables tb, NodeArrays nd, const int *__restrict__ xofs,
    const int *__restrict__ xleaf, const double *__restrict__ xs,
    const double *__restrict__ ys, const double *__restrict__ zs,
    const double *__restrict__ qs, double2 *__restrict__ L) {
  __shared__ double2 T[kP2mSlots];
  const int D = blockIdx.x;
  const int x0 = xofs[D], x1 = xofs[D + 1];
  if (x0 == x1) return;  // uniform per block: no barrier is skipped
  const int p = tb.p, lc = tri(p + 1), np1 = p + 1;
  const int lane = threadIdx.x;
  const double cx = nd.cx[D], cy = nd.cy[D], cz = nd.cz[D];
  const int chunk = min(64, kP2mSlots / lc);
  double2 acc[kLcSlots];
  for (int slot = 0; slot < kLcSlots; slot++) acc[slot] = make_double2(0., 0.);
  for (int x = x0; x < x1; x++) {
    const int J = xleaf[x];
    const int b = nd.pbegin[J], e = nd.pend[J];
    for (int c0 = b; c0 < e; c0 += chunk) {
      const int cnt = min(chunk, e - c0);
      __syncthreads();
      for (int t = lane; t < cnt * np1; t += 64) {
        const int j = t / np1, m = t - j * np1;
        const int i = c0 + j;
        const double vx = xs[i] - cx, vy = ys[i] - cy, vz = zs[i] - cz;
        const double r2 = vx * vx + vy * vy + vz * vz;
        const double ir2 = 1. / r2;
        double ox = qs[i] * sqrt(ir2), oy = 0.;  // q I_m^m
        for (int k = 1; k <= m; k++) {
          const double f = -(2. * k - 1.) * ir2;
          const double ax = f * vx, ay = f * vy;
          const double nx = ox * ax - oy * ay, ny = ox * ay + oy * ax;
          ox = nx;
          oy = ny;
        }
        double2 *row = T + j * lc;
        row[tri(m) + m] = make_double2(ox, oy);
        double px = ox, py = oy, qx = 0., qy = 0.;
        for (int l = m + 1; l <= p; l++) {
          const double a = (2. * l - 1.) * vz;
          const double bb = (double)((l - 1) * (l - 1) - m * m);
          const double nx = (a * px - bb * qx) * ir2,
                       ny = (a * py - bb * qy) * ir2;
          qx = px;
          qy = py;
          px = nx;
          py = ny;
          row[tri(l) + m] = make_double2(nx, ny);
        }
      }
      __syncthreads();
      for (int slot = 0; slot < kLcSlots; slot++) {
        const int k = lane + 64 * slot;
        if (k >= lc) break;
        double2 a = acc[slot];
        for (int j = 0; j < cnt; j++) {
          const double2 v = T[j * lc + k];
          a.x += v.x;
          a.y += v.y;
        }
        acc[slot] = a;
      }
    }
  }
  for (int slot = 0; slot < kLcSlots; slot++) {
    const int k = lane + 64 * slot;
    if (k < lc) {
      const double g = tb.xgain[k];
      double2 v = L[(size_t)D * lc + k];
      v.x += g * acc[slot].x;
      v.y -= g * acc[slot].y;
      L[(size_t)D * lc + k] = v;
    }
  }
}

// pleaf[j] = leaf node of sorted particle j.
__global__ void k_particle_leaf(NodeArrays nd, const int *__restrict__ leaves,
                                int nleaves, int *__restrict__ pleaf) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = t / 8, sub = t % 8;  // 8 lanes per leaf
  if (f >= nleaves) return;
  const int A = leaves[f];
  for (int j = nd.pbegin[A] + sub; j < nd.pend[A]; j += 8) pleaf[j] = A;
}

// P2M above the templated orders: one wave per leaf, lane m runs column m
// of the degree recurrence for the leaf's particles in order and adds
// q O_l^m into the LDS entries of its own column (no two lanes share one).
__global__ __launch_bounds__(64) void k_p2m_generic(
    OpTables tb, NodeArrays nd, const int *__restrict__ leaves,
    const double *__restrict__ xs, const double *__restrict__ ys,
    const double *__restrict__ zs, const double *__restrict__ qs,
    double2 *__restrict__ M) {
  __shared__ double2 T[kMaxMc];
  const int p = tb.p, mc = tri(p);
  const int lane = threadIdx.x;
  const int A = leaves[blockIdx.x];
  const int b = nd.pbegin[A], e = nd.pend[A];
  const double cx = nd.cx[A], cy = nd.cy[A], cz = nd.cz[A];
  for (int k = lane; k < mc; k += 64) T[k] = make_double2(0., 0.);
  __syncthreads();
  if (lane < p) {
    const int m = lane;
    for (int i = b; i < e; i++) {
      const double vx = xs[i] - cx, vy = ys[i] - cy, vz = zs[i] - cz;
      const double r2 = vx * vx + vy * vy + vz * vz;
      double ox = qs[i], oy = 0.;  // q O_m^m
      for (int k = 1; k <= m; k++) {
        const double f = -0.5 * kInv.v[k];
        const double ax = f * vx, ay = f * vy;
        const double nx = ox * ax - oy * ay, ny = ox * ay + oy * ax;
        ox = nx;
        oy = ny;
      }
      T[tri(m) + m].x += ox;
      T[tri(m) + m].y += oy;
      double px = ox, py = oy, qx = 0., qy = 0.;
      for (int l = m + 1; l < p; l++) {
        const double a = (2. * l - 1.) * vz;
        const double g = kInv.v[l * l - m * m];
        const double nx = (a * px - r2 * qx) * g, ny = (a * py - r2 * qy) * g;
        qx = px;
        qy = py;
        px = nx;
        py = ny;
        T[tri(l) + m].x += nx;
        T[tri(l) + m].y += ny;
      }
    }
  }
  __syncthreads();
  for (int k = lane; k < mc; k += 64) {
    const double g = tb.pgain[k];
    M[(size_t)A * mc + k] = make_double2(g * T[k].x, -g * T[k].y);
  }
}

// ------------------------------------------------------------ launchers

template <int P>
static hipError_t p2m_dispatch(const OpTables &tb, const NodeArrays &nd,
                               const int *leaves, int nleaves,
                               const double *xs, const double *ys,
                               const double *zs, const double *qs, double2 *M,
                               int low, hipStream_t s) {
  if constexpr (P > kMaxOrderFast) {
    k_p2m_generic<<<nleaves, 64, 0, s>>>(tb, nd, leaves, xs, ys, zs, qs, M);
    return hipGetLastError();
  } else {
    if (tb.p != P)
      return p2m_dispatch<P + 1>(tb, nd, leaves, nleaves, xs, ys, zs, qs, M,
                                 low, s);
    if (low)
      k_p2m<P, float><<<nleaves, 64, 0, s>>>(tb, nd, leaves, xs, ys, zs, qs,
                                             M);
    else
      k_p2m<P, double><<<nleaves, 64, 0, s>>>(tb, nd, leaves, xs, ys, zs, qs,
                                              M);
    return hipGetLastError();
  }
}

hipError_t launch_p2m(const OpTables &tb, const NodeArrays &nd,
                      const int *leaves, int nleaves, const double *xs,
                      const double *ys, const double *zs, const double *qs,
                      double2 *M, hipStream_t s, bool low) {
  if (nleaves == 0) return hipSuccess;
  return p2m_dispatch<1>(tb, nd, leaves, nleaves, xs, ys, zs, qs, M,
                         low ? 1 : 0, s);
}

hipError_t launch_m2l(const OpTables &tb, const NodeArrays &nd, int nnodes,
                      const int *vofs, const int *vsrc, const short *vcode,
                      const double2 *M, double2 *L, hipStream_t s) {
  if (nnodes == 0) return hipSuccess;
  k_m2l<<<nnodes, kOpThreads, 0, s>>>(tb, nd, vofs, vsrc, vcode, M, L);
  return hipGetLastError();
}

hipError_t launch_m2p(const OpTables &tb, const NodeArrays &nd,
                      const int *leaves, int nleaves, const int *wofs,
                      const int *wnode, const double *xs, const double *ys,
                      const double *zs, const double2 *M, double *pot,
                      double *grad, hipStream_t s) {
  if (nleaves == 0) return hipSuccess;
  k_m2p<<<nleaves, 64, 0, s>>>(tb, nd, leaves, wofs, wnode, xs, ys, zs, M, pot,
                               grad);
  return hipGetLastError();
}

hipError_t launch_p2l(const OpTables &tb, const NodeArrays &nd, int nnodes,
                      const int *xofs, const int *xleaf, const double *xs,
                      const double *ys, const double *zs, const double *qs,
                      double2 *L, hipStream_t s) {
  if (nnodes == 0) return hipSuccess;
  k_p2l<<<nnodes, 64, 0, s>>>(tb, nd, xofs, xleaf, xs, ys, zs, qs, L);
  return hipGetLastError();
}

hipError_t launch_particle_leaf(const NodeArrays &nd, const int *leaves,
                                int nleaves, int *pleaf, hipStream_t s) {
  if (nleaves == 0) return hipSuccess;
  const long t = 8L * nleaves;
  k_particle_leaf<<<(t + 255) / 256, 256, 0, s>>>(nd, leaves, nleaves, pleaf);
  return hipGetLastError();
}

// k_l2p at the table's order: orders up to kMaxOrderFast unrolled, the
// higher ones with the order at run time.
template <class R, int P>
static void l2p_dispatch(const OpTables &tb, const NodeArrays &nd,
                         const int *pleaf, int n, const double *xs,
                         const double *ys, const double *zs, const double2 *L,
                         double *pot, double *grad, int rows, size_t lds,
                         int sel, const int *perm, double *opot, double *ograd,
                         hipStream_t s) {
  if constexpr (P <= kMaxOrderFast) {
    if (tb.p != P)
      return l2p_dispatch<R, P + 1>(tb, nd, pleaf, n, xs, ys, zs, L, pot,
                                    grad, rows, lds, sel, perm, opot, ograd,
                                    s);
  }
  constexpr int PK = P <= kMaxOrderFast ? P : 0;
  k_l2p<PK, R><<<(n + 255) / 256, 256, lds, s>>>(
      tb, nd, pleaf, n, xs, ys, zs, L, pot, grad, rows, sel, perm, opot,
      ograd);
}

hipError_t launch_l2p(const OpTables &tb, const NodeArrays &nd,
                      const int *pleaf, int n, const double *xs,
                      const double *ys, const double *zs, const double2 *L,
                      double *pot, double *grad, hipStream_t s, int sel,
                      const int *perm, double *opot, double *ograd) {
  if (n == 0) return hipSuccess;
  if (sel < 0 || sel > 2) return hipErrorInvalidValue;
  // staged rows per round: up to 16 leaves within 32 KiB of LDS
  const int lc = tri(tb.p + 1);
  const size_t row = (sel == 1 ? sizeof(float2) : sizeof(double2)) * lc;
  const int rows = (int)std::max<size_t>(1, std::min<size_t>(16, 32768 / row));
  if (sel == 1)
    l2p_dispatch<float, 1>(tb, nd, pleaf, n, xs, ys, zs, L, pot, grad, rows,
                           rows * row, sel, perm, opot, ograd, s);
  else
    l2p_dispatch<double, 1>(tb, nd, pleaf, n, xs, ys, zs, L, pot, grad, rows,
                            rows * row, sel, perm, opot, ograd, s);
  return hipGetLastError();
}

}  // namespace aran

namespace aran {

// ------------------------------------------------------------ one operator

// One translation (kind 0: M2M, 1: M2L, 2: L2L) along the single direction
// stored at code 0 of tb, for the development API.  dst is overwritten.
__global__ __launch_bounds__(kOpThreads) void k_translate_one(
    int kind, OpTables tb, double r, const double2 *__restrict__ src,
    double2 *__restrict__ dst) {
  __shared__ double2 U[kMaxLc], X[kMaxLc], Z[kMaxLc];
  __shared__ double rp[2 * kMaxOrder + 2];
  const int p = tb.p, mc = tri(p), lc = tri(p + 1), rb = tb.rb;
  const int nin = (kind == 2) ? lc : mc;
  const int nout = (kind == 0) ? mc : lc;
  const int din = (kind == 2) ? p : p - 1;
  const int dout = (kind == 0) ? p - 1 : p;
  DirEntry de = tb.dir[0];
  const double2 *ph = tb.eiphi;
  const double *R = tb.rot + (size_t)de.theta * 4 * rb;
  for (int k = threadIdx.x; k < nin; k += kOpThreads) {
    int l, m;
    lm_of(k, l, m);
    U[k] = cmul(ph[m], src[k]);
  }
  if (threadIdx.x < 2 * p + 2) {
    int kk = threadIdx.x;
    double v = 1.;
    if (kind == 1) {
      for (int i = 0; i < kk; i++) v /= r;
      if (kk > 0 && ((kk - 1) & 1)) v *= de.c;
    } else {
      double f = (kind == 0) ? -de.c * r : de.c * r;
      for (int i = 0; i < kk; i++) v *= f;
    }
    rp[kk] = v;
  }
  __syncthreads();
  rotate_fwd(R, rb, din, U, X);
  __syncthreads();
  for (int k = threadIdx.x; k < nout; k += kOpThreads) {
    int l, m;
    lm_of(k, l, m);
    double re = 0., im = 0.;
    if (kind == 0) {
      const double *T = tb.tm2m + ((size_t)l * p + m) * p;
      for (int n = m; n <= l; n++) {
        double f = T[n] * rp[l - n];
        re += f * X[tri(n) + m].x;
        im += f * X[tri(n) + m].y;
      }
    } else if (kind == 1) {
      const double *T = tb.tm2l + ((size_t)l * (p + 1) + m) * p;
      for (int n = m; n < p; n++) {
        double f = T[n] * rp[l + n + 1];
        re += f * X[tri(n) + m].x;
        im += f * X[tri(n) + m].y;
      }
      if ((l + m) & 1) {
        re = -re;
        im = -im;
      }
    } else {
      const double *T = tb.tl2l + ((size_t)l * (p + 1) + m) * (p + 1);
      for (int n = l; n <= p; n++) {
        double f = T[n] * rp[n - l];
        re += f * X[tri(n) + m].x;
        im += f * X[tri(n) + m].y;
      }
    }
    Z[k] = make_double2(re, im);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nout; k += kOpThreads)
    dst[k] = rotate_inv_one(R, rb, ph, Z, k);
  (void)dout;
}

hipError_t launch_translate_one(int kind, const OpTables &tb, double r,
                                const double2 *src, double2 *dst,
                                hipStream_t s) {
  k_translate_one<<<1, kOpThreads, 0, s>>>(kind, tb, r, src, dst);
  return hipGetLastError();
}

}  // namespace aran
