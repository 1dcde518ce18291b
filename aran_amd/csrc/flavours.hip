// flavours.hip -- the alternate translation flavours of LibAran on gfx950:
// the O(p^4) "normal" translations (aran/aransphericalseriesd-translate.c:
// 366-443, 507-579, 678-768) and K. Kylin's partial-sum recurrences
// (aran/aransphericalseriesd-kkylin.c:34-501).
//
// They map the same truncated expansions as the rotate / z-translate /
// rotate-back operators (operators.hip, translate.hip), by other
// arithmetic: the harmonics Y_l^m of the translation vector (not rotations)
// and, for Kylin, recurrences in m over partial sums.  They serve
// aran_development3d_{m2m,m2l,l2l}[_kkylin], the spherical series API and
// solves whose registered operators are these flavours.
//
// One workgroup runs one or several translations one after another (the
// sum over a far list, the children of a parent); inside a translation the
// threads take output coefficients (normal) or partial-sum columns (Kylin).
// FP64 throughout, as the reference.
#include <hip/hip_runtime.h>
#include <math.h>

#include "engine.h"
#include "flavours.h"
#include "translate.h"

namespace aran {

namespace {

constexpr int kFlavThreads = 256;
// harmonics of the translation vector: degrees up to 2 kMaxOrder + 1
constexpr int kFlavHarmDeg = 2 * kMaxOrder + 1;
constexpr int kFlavHarm = (kFlavHarmDeg + 1) * (kFlavHarmDeg + 2) / 2;
constexpr int kPartW = kMaxOrder + 1;  // partial sums [m][n], m, n <= 30

__device__ inline int tri_f(int n) { return n * (n + 1) / 2; }

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 conj2(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ inline double2 cscale(double2 a, double f) {
  return make_double2(a.x * f, a.y * f);
}
// c_l^{-m} = (-1)^m conj(c_l^m)
__device__ inline double2 sym(double2 z, int m) {
  return (m & 1) ? make_double2(-z.x, z.y) : make_double2(z.x, -z.y);
}
// coefficient (l, m) of a packed m >= 0 series for either sign of m
__device__ inline double2 coef_any(const double2 *S, int l, int m) {
  return m >= 0 ? S[tri_f(l) + m] : sym(S[tri_f(l) - m], -m);
}

// Y_l^m (N_l^m P_l^m(cos t) e^{i m phi}, Condon-Shortley phase, the
// normalisation of aran/aransphericalharmonic.c:31-38) for l <= D into Y,
// one column m per thread: P by the recurrences of aran/aranlegendre.c:
// 135-169 and N_l^m = N_{l-1}^m sqrt((2l+1)(l-m) / ((2l-1)(l+m))).
__device__ void harmonics(int D, const FlavVec &v, double2 *Y) {
  for (int m = threadIdx.x; m <= D; m += blockDim.x) {
    double er = 1., ei = 0.;
    for (int k = 0; k < m; k++) {
      const double nr = er * v.cosp - ei * v.sinp;
      ei = er * v.sinp + ei * v.cosp;
      er = nr;
    }
    double pmm = 1., nn = (2. * m + 1.) / (4. * M_PI);
    for (int i = 1; i <= m; i++) pmm *= -(2. * i - 1.) * v.sint;
    for (int i = 1; i <= 2 * m; i++) nn /= (double)i;
    double nlm = sqrt(nn);
    double p2 = 0., p1 = pmm;
    Y[tri_f(m) + m] = make_double2(nlm * p1 * er, nlm * p1 * ei);
    for (int l = m + 1; l <= D; l++) {
      const double pl =
          l == m + 1 ? (2. * m + 1.) * v.cost * p1
                     : ((2. * l - 1.) * v.cost * p1 - (l + m - 1.) * p2) /
                           (double)(l - m);
      nlm *= sqrt((2. * l + 1.) * (double)(l - m) /
                  ((2. * l - 1.) * (double)(l + m)));
      p2 = p1;
      p1 = pl;
      Y[tri_f(l) + m] = make_double2(nlm * pl * er, nlm * pl * ei);
    }
  }
}

__device__ inline double2 harm_any(const double2 *Y, int l, int m) {
  return m >= 0 ? Y[tri_f(l) + m] : sym(Y[tri_f(l) - m], -m);
}

struct Tab {
  const double *alpha, *beta;
  __device__ double a(int l, int m) const {
    return alpha[tri_f(l) + (m < 0 ? -m : m)];
  }
  __device__ double b(int l) const { return beta[l]; }
};

// ---- normal flavour

// aran_local_translate (translate.c:366-443)
__device__ void l2l_normal(const Tab &t, int pls, const double2 *src, int pld,
                           double2 *dst, const FlavVec &v, double2 *Y) {
  harmonics(pls, v, Y);
  __syncthreads();
  for (int k = threadIdx.x; k < tri_f(pld + 1); k += blockDim.x) {
    int l = 0;
    while (tri_f(l + 1) <= k) l++;
    const int m = k - tri_f(l);
    double2 acc = make_double2(0., 0.);
    double rp = 1.;
    for (int n = l; n <= pls; n++) {
      const double normaliz = t.b(n - l) * (t.b(l) / t.b(n));
      double2 sum = make_double2(0., 0.);
      for (int o = l + m - n; o <= m + n - l; o++) {
        const double f = t.a(n - l, o - m) * t.a(l, m) / t.a(n, o);
        sum = cadd(sum, cscale(cmul(harm_any(Y, n - l, o - m),
                                    coef_any(src, n, o)), f));
      }
      acc = cadd(acc, cscale(sum, normaliz * rp));
      rp *= v.r;
    }
    dst[k] = cadd(dst[k], acc);
  }
  __syncthreads();
}

// aran_multipole_translate (translate.c:507-579); v is the opposite vector
__device__ void m2m_normal(const Tab &t, int nms, const double2 *src,
                           int nmd, double2 *dst, const FlavVec &v,
                           double2 *Y) {
  harmonics(nmd - 1, v, Y);
  __syncthreads();
  for (int k = threadIdx.x; k < tri_f(nmd); k += blockDim.x) {
    int l = 0;
    while (tri_f(l + 1) <= k) l++;
    const int m = k - tri_f(l);
    double2 acc = make_double2(0., 0.);
    const int nmax = min(l, nms - 1);
    for (int n = 0; n <= nmax; n++) {
      const double normaliz = t.b(l - n) * (t.b(l) / t.b(n));
      double2 sum = make_double2(0., 0.);
      for (int o = max(-n, m + n - l); o <= min(n, l + m - n); o++) {
        const double f = t.a(l - n, m - o) * t.a(n, o) / t.a(l, m);
        sum = cadd(sum, cscale(cmul(conj2(harm_any(Y, l - n, m - o)),
                                    coef_any(src, n, o)), f));
      }
      acc = cadd(acc, cscale(sum, normaliz * pow(v.r, (double)(l - n))));
    }
    dst[k] = cadd(dst[k], acc);
  }
  __syncthreads();
}

// aran_spherical_seriesd_multipole_to_local_vertical (translate.c:609-676)
// as the normal and Kylin flavours call it for |sin t| < 1e-5: r^-k times
// cos t for even k, the azimuth unused.
__device__ void m2l_vertical(const Tab &t, int nm, const double2 *src, int pl,
                             double2 *dst, const FlavVec &v) {
  const double ir = 1. / v.r;
  for (int k = threadIdx.x; k < tri_f(pl + 1); k += blockDim.x) {
    int l = 0;
    while (tri_f(l + 1) <= k) l++;
    const int m = k - tri_f(l);
    double2 sum = make_double2(0., 0.);
    for (int n = m; n < nm; n++) {
      const int e = l + n + 1;
      double rp = pow(ir, (double)e);
      if (e % 2 == 0) rp *= v.cost;
      const double tf =
          t.b(l) / t.b(n) * t.a(l, m) * t.a(n, m) / t.a(l + n, 0);
      sum = cadd(sum, cscale(src[tri_f(n) + m], rp * tf));
    }
    if ((l + m) & 1) sum = make_double2(-sum.x, -sum.y);
    dst[k] = cadd(dst[k], sum);
  }
  __syncthreads();
}

// aran_multipole_to_local (translate.c:678-768)
__device__ void m2l_normal(const Tab &t, int nm, const double2 *src, int pl,
                           double2 *dst, const FlavVec &v, double2 *Y) {
  if (fabs(v.sint) < 1.e-5) {
    m2l_vertical(t, nm, src, pl, dst, v);
    return;
  }
  harmonics(pl + nm, v, Y);
  __syncthreads();
  const double ir = 1. / v.r;
  for (int k = threadIdx.x; k < tri_f(pl + 1); k += blockDim.x) {
    int l = 0;
    while (tri_f(l + 1) <= k) l++;
    const int m = k - tri_f(l);
    const double sign = (l & 1) ? -1. : 1.;
    double2 acc = make_double2(0., 0.);
    double rp = pow(ir, (double)(l + 1));
    for (int n = 0; n < nm; n++) {
      const double normaliz = t.b(l + n) * (t.b(l) / t.b(n));
      const double2 *S = src + tri_f(n);
      const double2 *h = Y + tri_f(l + n);
      double2 sum = cscale(cmul(h[m], S[0]),
                           t.a(l, m) * t.a(n, 0) / t.a(l + n, m));
      for (int o = 1; o <= n; o++) {
        const double f = t.a(l, m) * t.a(n, o);
        sum = cadd(sum, cscale(cmul(h[m + o], S[o]), f / t.a(l + n, m + o)));
        sum = cadd(sum, cscale(cmul(harm_any(Y, l + n, m - o), sym(S[o], o)),
                               f / t.a(l + n, m - o)));
      }
      acc = cadd(acc, cscale(conj2(sum), sign * normaliz * rp));
      rp *= ir;
    }
    dst[k] = cadd(dst[k], acc);
  }
  __syncthreads();
}

// ---- Kylin flavour (partial sums part[m][n] in LDS, two buffers)

#define PT(a, mm, nn) (a)[(mm) * kPartW + (nn)]

// aran_multipole_translate_kkylin (kkylin.c:34-205); v is the opposite
// vector
__device__ void m2m_kylin(const Tab &t, int nms, const double2 *src, int nmd,
                          double2 *dst, const FlavVec &v, double2 *Y,
                          double2 *b1, double2 *b2) {
  (void)nms;  // the caller pads src to nmd degrees (the reference reads
              // source terms up to degree nmd - 1)
  harmonics(nmd - 1, v, Y);
  __syncthreads();
  double2 *part = b1, *prev = b2;
  const double2 ex = make_double2(v.cosp, v.sinp);
  if (threadIdx.x == 0) {
    PT(part, 0, 0) = cmul(conj2(Y[0]), src[0]);
    dst[0] = cadd(dst[0], cscale(PT(part, 0, 0), t.b(0)));
  }
  __syncthreads();
  for (int l = 1; l < nmd; l++) {
    double2 *tmp = part;
    part = prev;
    prev = tmp;
    for (int n = threadIdx.x; n < l; n += blockDim.x) {
      const double rp = pow(v.r, (double)(l - n));
      double fact = t.b(l - n) / t.b(n);
      double2 s0 = make_double2(0., 0.);
      for (int o = max(-n, n - l); o <= min(n, l - n); o++)
        s0 = cadd(s0, cscale(cmul(conj2(harm_any(Y, l - n, -o)),
                                  coef_any(src, n, o)),
                             rp * t.a(n, o) * t.a(l - n, o)));
      PT(part, 0, n) = cscale(s0, fact);
      if (l > 1) {
        double2 s1 = make_double2(0., 0.);
        for (int o = max(-n, n - 1); o <= min(n, l + l - 1 - n); o++)
          s1 = cadd(s1, cscale(cmul(conj2(harm_any(Y, l - n, l - o - 1)),
                                    coef_any(src, n, o)),
                               rp * t.a(n, o) * t.a(l - n, l - o - 1)));
        PT(part, l - 1, n) = cscale(s1, fact);
      }
      double2 s2 = make_double2(0., 0.);
      for (int o = max(-n, n); o <= min(n, l + l - n); o++)
        s2 = cadd(s2, cscale(cmul(conj2(harm_any(Y, l - n, l - o)),
                                  coef_any(src, n, o)),
                             rp * t.a(n, o) * t.a(l - n, l - o)));
      PT(part, l, n) = cscale(s2, fact);
      fact = 0.5 * v.r / (double)(l - n);
      const double2 fse = cscale(ex, fact * v.sint);
      fact *= 2. * v.cost;
      for (int m = 1; m < l - 1; m++)
        PT(part, m, n) = cadd(cadd(cmul(fse, PT(prev, m + 1, n)),
                                   cscale(cmul(conj2(fse), PT(prev, m - 1, n)),
                                          -1.)),
                              cscale(PT(prev, m, n), fact));
    }
    const double fl = t.b(0) / t.b(l);
    for (int m = threadIdx.x; m <= l; m += blockDim.x)
      PT(part, m, l) = cscale(cmul(conj2(Y[0]), src[tri_f(l) + m]),
                              t.a(l, m) * fl);
    __syncthreads();
    for (int m = threadIdx.x; m <= l; m += blockDim.x) {
      double2 sum = make_double2(0., 0.);
      for (int n = 0; n <= l; n++) sum = cadd(sum, PT(part, m, n));
      dst[tri_f(l) + m] =
          cadd(dst[tri_f(l) + m], cscale(sum, t.b(l) / t.a(l, m)));
    }
    __syncthreads();
  }
}

// aran_local_translate_kkylin (kkylin.c:207-362); src degree <= dst degree
__device__ void l2l_kylin(const Tab &t, int pls, const double2 *src, int pld,
                          double2 *dst, const FlavVec &v, double2 *Y,
                          double2 *b1, double2 *b2) {
  harmonics(pls, v, Y);
  __syncthreads();
  double2 *part = b1, *prev = b2;
  const double2 ex = make_double2(v.cosp, v.sinp);
  const double2 y00 = Y[0];
  for (int l = pld; l >= 0; l--) {
    if (l <= pls) {
      const double f = t.b(0) / t.b(l);
      for (int m = threadIdx.x; m <= l; m += blockDim.x)
        PT(part, m, l) =
            cscale(conj2(cmul(y00, src[tri_f(l) + m])), f / t.a(l, m));
    }
    for (int n = l + 1 + (int)threadIdx.x; n <= pls; n += blockDim.x) {
      const double rp = pow(v.r, (double)(n - l));
      double fact = t.b(n - l) / t.b(n);
      double2 s0 = make_double2(0., 0.);
      for (int o = max(-n, l - n); o <= min(n, n - l); o++)
        s0 = cadd(s0, cscale(conj2(cmul(harm_any(Y, n - l, o),
                                        coef_any(src, n, o))),
                             rp * t.a(n - l, o) / t.a(n, o)));
      PT(part, 0, n) = cscale(s0, fact);
      if (l > 0) {
        double2 s1 = make_double2(0., 0.);
        for (int o = max(-n, l + l - n); o <= n; o++)
          s1 = cadd(s1, cscale(conj2(cmul(harm_any(Y, n - l, o - l),
                                          coef_any(src, n, o))),
                               rp * t.a(n - l, o - l) / t.a(n, o)));
        PT(part, l, n) = cscale(s1, fact);
      }
      fact = 0.5 * v.r / (double)(n - l);
      const double2 fse = cscale(ex, fact * v.sint);
      fact *= 2. * v.cost;
      for (int m = 1; m < l; m++)
        PT(part, m, n) = cadd(cadd(cmul(fse, PT(prev, m - 1, n)),
                                   cscale(cmul(conj2(fse), PT(prev, m + 1, n)),
                                          -1.)),
                              cscale(PT(prev, m, n), fact));
    }
    __syncthreads();
    for (int m = threadIdx.x; m <= l; m += blockDim.x) {
      double2 sum = make_double2(0., 0.);
      for (int n = l; n <= pls; n++) sum = cadd(sum, PT(part, m, n));
      dst[tri_f(l) + m] =
          cadd(dst[tri_f(l) + m], cscale(conj2(sum), t.b(l) * t.a(l, m)));
    }
    __syncthreads();
    double2 *tmp = part;
    part = prev;
    prev = tmp;
  }
}

// aran_multipole_to_local_kkylin (kkylin.c:364-501)
__device__ void m2l_kylin(const Tab &t, int nm, const double2 *src, int pl,
                          double2 *dst, const FlavVec &v, double2 *Y,
                          double2 *b1, double2 *b2) {
  if (fabs(v.sint) < 1.e-5) {
    m2l_vertical(t, nm, src, pl, dst, v);
    return;
  }
  harmonics(pl + nm, v, Y);
  __syncthreads();
  double2 *part = b1, *prev = b2;
  const double2 ex = make_double2(v.cosp, v.sinp);
  const double ir = 1. / v.r;
  const double2 tce = cscale(ex, 2. * v.cost / v.sint);
  const double2 e2 = cmul(ex, ex);
  double sign = 1.;
  for (int l = 0; l <= pl; l++) {
    for (int n = threadIdx.x; n < nm; n += blockDim.x) {
      const double fact = pow(ir, (double)(l + n + 1)) * t.b(l + n) / t.b(n);
      double2 s0 = make_double2(0., 0.);
      for (int o = -n; o <= n; o++)
        s0 = cadd(s0, cscale(cmul(harm_any(Y, l + n, o), coef_any(src, n, o)),
                             t.a(n, o) / t.a(l + n, o)));
      PT(part, 0, n) = cscale(s0, fact);
      if (l > 0) {
        double2 s1 = make_double2(0., 0.);
        for (int o = -n; o <= n; o++)
          s1 = cadd(s1, cscale(cmul(harm_any(Y, l + n, o + 1),
                                    coef_any(src, n, o)),
                               t.a(n, o) / t.a(l + n, o + 1)));
        PT(part, 1, n) = cscale(s1, fact);
      }
      const double2 tl = cscale(ex, 2. * (l + n) * ir / v.sint);
      for (int m = 2; m <= l; m++)
        PT(part, m, n) = cadd(cadd(cmul(tce, PT(part, m - 1, n)),
                                   cmul(e2, PT(part, m - 2, n))),
                              cscale(cmul(tl, PT(prev, m - 1, n)), -1.));
    }
    __syncthreads();
    const double f = t.b(l) * sign;
    for (int m = threadIdx.x; m <= l; m += blockDim.x) {
      double2 sum = make_double2(0., 0.);
      for (int n = 0; n < nm; n++) sum = cadd(sum, PT(part, m, n));
      dst[tri_f(l) + m] =
          cadd(dst[tri_f(l) + m], cscale(conj2(sum), t.a(l, m) * f));
    }
    __syncthreads();
    double2 *tmp = part;
    part = prev;
    prev = tmp;
    sign = -sign;
  }
}

#undef PT

// One translation of the given kind and flavour, added into dst.
__device__ void flavour_one(const Tab &t, int kind, int flavour, int sdeg,
                            const double2 *src, int ddeg, double2 *dst,
                            const FlavVec &v, double2 *Y, double2 *b1,
                            double2 *b2) {
  if (kind == kFlavM2M) {
    FlavVec w = v;  // the multipole translations take the opposite vector
    w.cost = -v.cost;
    w.cosp = -v.cosp;
    w.sinp = -v.sinp;
    if (flavour == kFlavKylin)
      m2m_kylin(t, sdeg, src, ddeg, dst, w, Y, b1, b2);
    else
      m2m_normal(t, sdeg, src, ddeg, dst, w, Y);
  } else if (kind == kFlavL2L) {
    if (flavour == kFlavKylin)
      l2l_kylin(t, sdeg, src, ddeg, dst, v, Y, b1, b2);
    else
      l2l_normal(t, sdeg, src, ddeg, dst, v, Y);
  } else {
    if (flavour == kFlavKylin)
      m2l_kylin(t, sdeg, src, ddeg, dst, v, Y, b1, b2);
    else
      m2l_normal(t, sdeg, src, ddeg, dst, v, Y);
  }
}

// spherical coordinates of d (the conventions of the oracle's
// or_to_spherical: theta from +z, phi from +x; on the axis cos phi = 1)
__device__ FlavVec to_vec(double dx, double dy, double dz) {
  FlavVec v;
  const double rho2 = dx * dx + dy * dy;
  const double rho = sqrt(rho2);
  v.r = sqrt(rho2 + dz * dz);
  if (v.r > 0.) {
    v.cost = dz / v.r;
    v.sint = rho / v.r;
  } else {
    v.cost = 1.;
    v.sint = 0.;
  }
  if (rho > 0.) {
    v.cosp = dx / rho;
    v.sinp = dy / rho;
  } else {
    v.cosp = 1.;
    v.sinp = 0.;
  }
  return v;
}

struct FlavLds {
  double2 Y[kFlavHarm];
  double2 b1[kPartW * kPartW], b2[kPartW * kPartW];
};

}  // namespace

// One translation (the single-operator entry points).
__global__ __launch_bounds__(kFlavThreads) void k_flavour_one(
    const double *__restrict__ alpha, const double *__restrict__ beta,
    int kind, int flavour, int sdeg, const double2 *src, int ddeg,
    double2 *dst, FlavVec v) {
  __shared__ FlavLds s;
  const Tab t{alpha, beta};
  flavour_one(t, kind, flavour, sdeg, src, ddeg, dst, v, s.Y, s.b1, s.b2);
}

// Far field of a solve: one workgroup per target node, L(A) = sum over its
// far list (source node B, vector c(A) - c(B)) in list order.
__global__ __launch_bounds__(kFlavThreads) void k_flavour_m2l(
    const double *__restrict__ alpha, const double *__restrict__ beta,
    int flavour, int p, NodeArrays nd, const int *__restrict__ vofs,
    const int *__restrict__ vsrc, const double2 *__restrict__ M,
    double2 *__restrict__ L) {
  __shared__ FlavLds s;
  const Tab t{alpha, beta};
  const int A = blockIdx.x, mc = tri_f(p), lc = tri_f(p + 1);
  double2 *dst = L + (size_t)A * lc;
  for (int k = threadIdx.x; k < lc; k += blockDim.x)
    dst[k] = make_double2(0., 0.);
  __syncthreads();
  for (int e = vofs[A]; e < vofs[A + 1]; e++) {
    const int B = vsrc[e];
    const FlavVec v =
        to_vec(nd.cx[A] - nd.cx[B], nd.cy[A] - nd.cy[B], nd.cz[A] - nd.cz[B]);
    flavour_one(t, kFlavM2L, flavour, p, M + (size_t)B * mc, p, dst, v, s.Y,
                s.b1, s.b2);
  }
}

// M2M / L2L of one tree level: one workgroup per parent; M2M sums the
// children (octant order) into the parent, L2L adds the parent's local into
// each child.
__global__ __launch_bounds__(kFlavThreads) void k_flavour_level(
    const double *__restrict__ alpha, const double *__restrict__ beta,
    int kind, int flavour, int p, NodeArrays nd,
    const TransRow *__restrict__ rows, double2 *__restrict__ X) {
  __shared__ FlavLds s;
  const Tab t{alpha, beta};
  const TransRow &row = rows[blockIdx.x];
  const int P = row.node;
  if (kind == kFlavM2M) {
    const int mc = tri_f(p);
    double2 *dst = X + (size_t)P * mc;
    for (int k = threadIdx.x; k < mc; k += blockDim.x)
      dst[k] = make_double2(0., 0.);
    __syncthreads();
    for (int o = 0; o < 8; o++) {
      const int C = row.child[o];
      if (C < 0) continue;  // uniform over the workgroup
      const FlavVec v = to_vec(nd.cx[P] - nd.cx[C], nd.cy[P] - nd.cy[C],
                               nd.cz[P] - nd.cz[C]);
      flavour_one(t, kFlavM2M, flavour, p, X + (size_t)C * mc, p, dst, v, s.Y,
                  s.b1, s.b2);
    }
  } else {
    const int lc = tri_f(p + 1);
    for (int o = 0; o < 8; o++) {
      const int C = row.child[o];
      if (C < 0) continue;
      const FlavVec v = to_vec(nd.cx[C] - nd.cx[P], nd.cy[C] - nd.cy[P],
                               nd.cz[C] - nd.cz[P]);
      flavour_one(t, kFlavL2L, flavour, p, X + (size_t)P * lc, p,
                  X + (size_t)C * lc, v, s.Y, s.b1, s.b2);
    }
  }
}

hipError_t launch_flavour_one(const OpTables &tb, int kind, int flavour,
                              int sdeg, const double2 *src, int ddeg,
                              double2 *dst, const FlavVec &v, hipStream_t s) {
  if (sdeg < 0 || ddeg < 0 || sdeg > kMaxOrder || ddeg > kMaxOrder ||
      (kind == kFlavM2M && flavour == kFlavKylin && sdeg != ddeg))
    return hipErrorInvalidValue;
  k_flavour_one<<<1, kFlavThreads, 0, s>>>(tb.alpha, tb.beta, kind, flavour,
                                           sdeg, src, ddeg, dst, v);
  return hipGetLastError();
}

hipError_t launch_flavour_m2l(const OpTables &tb, int flavour,
                              const NodeArrays &nd, int nnodes,
                              const int *vofs, const int *vsrc,
                              const double2 *M, double2 *L, hipStream_t s) {
  if (nnodes <= 0) return hipSuccess;
  k_flavour_m2l<<<nnodes, kFlavThreads, 0, s>>>(tb.alpha, tb.beta, flavour,
                                                tb.p, nd, vofs, vsrc, M, L);
  return hipGetLastError();
}

hipError_t launch_flavour_level(const OpTables &tb, int kind, int flavour,
                                const NodeArrays &nd, const TransRow *rows,
                                int nrows, double2 *X, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  k_flavour_level<<<nrows, kFlavThreads, 0, s>>>(tb.alpha, tb.beta, kind,
                                                 flavour, tb.p, nd, rows, X);
  return hipGetLastError();
}

}  // namespace aran
