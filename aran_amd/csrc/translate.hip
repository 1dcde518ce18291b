// translate.hip -- M2M and L2L of one tree level.
//
// Same operators as aran_development3d_m2m_rotate / l2l_rotate
// (arandevelopment3d.c:402-409, 444-451; aransphericalseriesd-rotate.c:
// 219-263; translate.c:301-364): rotate the expansion so the translation
// runs along z, translate along z, rotate back, in FP64.  One workgroup per
// parent, one wave per child octant, lanes over the coefficients.  A level
// is a short kernel on the critical path, so everything is shaped for
// latency: the order is a template parameter and every loop over degrees is
// unrolled with predicates, so each phase issues all of its table loads at
// once (one memory round trip per phase instead of one per loop step); the
// parent's record (children by octant, width) is one load; the per-wave
// buffers are two (LDS stays small enough to run beside the near field).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "engine.h"
#include "translate.h"

namespace aran {

namespace {

template <int B, int E, class F>
__device__ __forceinline__ void static_loop(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_loop<B + 1, E>(f);
  }
}

__device__ inline double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// (l, m) of coefficient k in (l, m) order, k < tri(kMaxOrder + 2)
struct LmTab {
  unsigned char l[(kMaxOrder + 2) * (kMaxOrder + 3) / 2];
  unsigned char m[(kMaxOrder + 2) * (kMaxOrder + 3) / 2];
};
constexpr LmTab make_lm() {
  LmTab t{};
  int k = 0;
  for (int l = 0; l <= kMaxOrder + 1; l++)
    for (int m = 0; m <= l; m++) {
      t.l[k] = (unsigned char)l;
      t.m[k] = (unsigned char)m;
      k++;
    }
  return t;
}
__constant__ LmTab kLm = make_lm();

constexpr __host__ __device__ int tri_c(int n) { return n * (n + 1) / 2; }
constexpr __host__ __device__ int rofs_c(int l) {
  return l * (l + 1) * (2 * l + 1) / 6;
}

// Forward rotation of one output coefficient (l, mp) over degrees <= D:
// Re X = E+ Re U, Im X = E- Im U (U already phased).
template <int D>
__device__ inline double2 rot_fwd(const double *__restrict__ R, int rb, int l,
                                  int mp, const double2 *U) {
  const double *ep = R + rofs_c(l) + mp * (l + 1);
  const double *em = ep + rb;
  const double2 *u = U + tri_c(l);
  double re = 0., im = 0.;
  static_loop<0, D + 1>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if (m <= l) {
      re += ep[m] * u[m].x;
      im += em[m] * u[m].y;
    }
  });
  return make_double2(re, im);
}

// Inverse rotation of one output coefficient and the phase back:
// e^{-i mp phi} (F+ Re Z, F- Im Z).
template <int D>
__device__ inline double2 rot_inv(const double *__restrict__ R, int rb,
                                  const double2 *__restrict__ ph, int l,
                                  int mp, const double2 *Z) {
  const double *fp = R + 2 * rb + rofs_c(l) + mp * (l + 1);
  const double *fm = fp + rb;
  const double2 *z = Z + tri_c(l);
  double re = 0., im = 0.;
  static_loop<0, D + 1>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if (m <= l) {
      re += fp[m] * z[m].x;
      im += fm[m] * z[m].y;
    }
  });
  const double2 e = ph[mp];
  return make_double2(e.x * re + e.y * im, e.x * im - e.y * re);
}

}  // namespace

__global__ void k_trans_rows(NodeArrays nd, const int *__restrict__ internal,
                             int ninternal, TransRow *__restrict__ rows) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ninternal) return;
  const int v = internal[j];
  TransRow r;
  r.node = v;
  // children by octant, as the level launches recorded them
  const int4 *cm = reinterpret_cast<const int4 *>(nd.cmap + 8 * (size_t)v);
  const int4 c0 = cm[0], c1 = cm[1];
  r.child[0] = c0.x;
  r.child[1] = c0.y;
  r.child[2] = c0.z;
  r.child[3] = c0.w;
  r.child[4] = c1.x;
  r.child[5] = c1.y;
  r.child[6] = c1.z;
  r.child[7] = c1.w;
  r.pad = 0;
  r.width = node_width(nd, v);
  rows[j] = r;
}

hipError_t launch_trans_rows(const NodeArrays &nd, const int *internal,
                             int ninternal, TransRow *rows, hipStream_t s) {
  if (ninternal <= 0) return hipSuccess;
  k_trans_rows<<<(ninternal + 255) / 256, 256, 0, s>>>(nd, internal,
                                                       ninternal, rows);
  return hipGetLastError();
}

// M2M of one parent per workgroup: wave o translates the child in octant o
// (rotate, vertical M2M, rotate back) and the workgroup adds the children in
// octant order (the order of a serial loop over the children).
template <int P>
__global__ __launch_bounds__(512) void k_m2m_level(
    OpTables tb, const TransRow *__restrict__ rows,
    double2 *__restrict__ M) {
  // issue priority over the near field's waves (second stream), which
  // would otherwise hold up this latency-bound step of the far-field chain
  __builtin_amdgcn_s_setprio(2);
  constexpr int MC = tri_c(P), NS = (MC + 63) / 64;
  __shared__ double2 U[8][MC], X[8][MC];
  __shared__ double rp[8][2 * P + 2];
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int rb = tb.rb;
  const TransRow &row = rows[blockIdx.x];
  const int C = row.child[w];
  const double half = row.width * 0.25;  // half a child width
  const int code =
      code_of(1 - 2 * (w & 1), 1 - 2 * ((w >> 1) & 1), 1 - 2 * (w >> 2));
  const DirEntry de = tb.dir[code];
  const double2 *ph = tb.eiphi + code * (P + 1);
  const double *R = tb.rot + (size_t)de.theta * 4 * rb;
  if (C >= 0) {
    static_loop<0, NS>([&](auto sc) {
      const int k = lane + 64 * decltype(sc)::value;
      if (k < MC) U[w][k] = cmul2(ph[kLm.m[k]], M[(size_t)C * MC + k]);
    });
    if (lane < 2 * P + 2) {
      const double f = -de.c * de.norm * half;
      double v = 1.;
      for (int i = 0; i < lane; i++) v *= f;
      rp[w][lane] = v;  // ((-c) r)^k
    }
  }
  __syncthreads();
  if (C >= 0)  // forward rotation, degrees < P
    static_loop<0, NS>([&](auto sc) {
      const int k = lane + 64 * decltype(sc)::value;
      if (k < MC) X[w][k] = rot_fwd<P - 1>(R, rb, kLm.l[k], kLm.m[k], U[w]);
    });
  __syncthreads();
  if (C >= 0)  // vertical M2M into U
    static_loop<0, NS>([&](auto sc) {
      const int k = lane + 64 * decltype(sc)::value;
      if (k < MC) {
        const int l = kLm.l[k], m = kLm.m[k];
        const double *T = tb.tm2m + ((size_t)l * P + m) * P;
        double re = 0., im = 0.;
        static_loop<0, P>([&](auto nc) {
          constexpr int n = decltype(nc)::value;
          if (n >= m && n <= l) {
            const double f = T[n] * rp[w][l - n];
            re += f * X[w][tri_c(n) + m].x;
            im += f * X[w][tri_c(n) + m].y;
          }
        });
        U[w][k] = make_double2(re, im);
      }
    });
  __syncthreads();
  if (C >= 0)  // rotate back into X: the child's contribution
    static_loop<0, NS>([&](auto sc) {
      const int k = lane + 64 * decltype(sc)::value;
      if (k < MC) X[w][k] = rot_inv<P - 1>(R, rb, ph, kLm.l[k], kLm.m[k], U[w]);
    });
  __syncthreads();
  for (int k = threadIdx.x; k < MC; k += 512) {
    double2 acc = make_double2(0., 0.);
    for (int o = 0; o < 8; o++)
      if (row.child[o] >= 0) {
        acc.x += X[o][k].x;
        acc.y += X[o][k].y;
      }
    M[(size_t)row.node * MC + k] = acc;
  }
}

// Rotation rows of degree l, order mp in an LDS copy of the E+ E- F+ F-
// tables (stride RB between the four).
template <int D, int RB>
__device__ inline double2 rot_fwd_s(const double *Rs, int l, int mp,
                                    const double2 *U) {
  const double *ep = Rs + rofs_c(l) + mp * (l + 1);
  const double *em = ep + RB;
  const double2 *u = U + tri_c(l);
  double re = 0., im = 0.;
  static_loop<0, D + 1>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if (m <= l) {
      re += ep[m] * u[m].x;
      im += em[m] * u[m].y;
    }
  });
  return make_double2(re, im);
}

template <int D, int RB>
__device__ inline double2 rot_inv_s(const double *Rs, double2 e, int l,
                                    int mp, const double2 *Z) {
  const double *fp = Rs + 2 * RB + rofs_c(l) + mp * (l + 1);
  const double *fm = fp + RB;
  const double2 *z = Z + tri_c(l);
  double re = 0., im = 0.;
  static_loop<0, D + 1>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if (m <= l) {
      re += fp[m] * z[m].x;
      im += fm[m] * z[m].y;
    }
  });
  return make_double2(e.x * re + e.y * im, e.x * im - e.y * re);
}

// L2L of one parent per workgroup: wave o translates the parent's local to
// the child in octant o and adds it there.  The eight offsets (+-1, +-1,
// +-1) share one rotation angle, so the workgroup stages that angle's
// rotation tables and the eight phases in LDS once; the vertical L2L
// factors are read from L2 (in LDS too they would leave room for one
// workgroup less per CU).
template <int P>
__global__ __launch_bounds__(512) void k_l2l_level(
    OpTables tb, const TransRow *__restrict__ rows,
    double2 *__restrict__ L) {
  constexpr int LC = tri_c(P + 1), NS = (LC + 63) / 64, RB = rofs_c(P + 1);
  __shared__ double Rs[4 * RB];
  __shared__ double2 Ph[8][P + 1];
  __shared__ double2 L0[LC], U[8][LC], X[8][LC];
  __shared__ double rp[8][2 * P + 2];
  const int tid = threadIdx.x, w = tid / 64, lane = tid % 64;
  const TransRow &row = rows[blockIdx.x];
  const int C = row.child[w];
  const double half = row.width * 0.25;  // half a child width
  {
    const double *R = tb.rot + (size_t)tb.dir[code_of(-1, -1, -1)].theta * 4 * RB;
    for (int i = tid; i < 4 * RB; i += 512) Rs[i] = R[i];
    for (int i = tid; i < 8 * (P + 1); i += 512) {
      const int o = i / (P + 1), m = i % (P + 1);
      const int code = code_of(2 * (o & 1) - 1, 2 * ((o >> 1) & 1) - 1,
                               2 * (o >> 2) - 1);
      Ph[o][m] = tb.eiphi[code * (P + 1) + m];
    }
  }
  for (int k = tid; k < LC; k += 512) L0[k] = L[(size_t)row.node * LC + k];
  if (C >= 0 && lane < 2 * P + 2) {
    const DirEntry de = tb.dir[code_of(2 * (w & 1) - 1, 2 * ((w >> 1) & 1) - 1,
                                       2 * (w >> 2) - 1)];
    const double f = de.c * de.norm * half;
    double v = 1.;
    for (int i = 0; i < lane; i++) v *= f;
    rp[w][lane] = v;  // (c r)^k
  }
  __syncthreads();
  if (C >= 0)
    static_loop<0, NS>([&](auto sc) {
      const int k = lane + 64 * decltype(sc)::value;
      if (k < LC) U[w][k] = cmul2(Ph[w][kLm.m[k]], L0[k]);
    });
  __syncthreads();
  if (C >= 0)  // forward rotation, degrees <= P
    static_loop<0, NS>([&](auto sc) {
      const int k = lane + 64 * decltype(sc)::value;
      if (k < LC) X[w][k] = rot_fwd_s<P, RB>(Rs, kLm.l[k], kLm.m[k], U[w]);
    });
  __syncthreads();
  if (C >= 0)  // vertical L2L into U
    static_loop<0, NS>([&](auto sc) {
      const int k = lane + 64 * decltype(sc)::value;
      if (k < LC) {
        const int l = kLm.l[k], m = kLm.m[k];
        const double *T = tb.tl2l + (l * (P + 1) + m) * (P + 1);
        double re = 0., im = 0.;
        static_loop<0, P + 1>([&](auto nc) {
          constexpr int n = decltype(nc)::value;
          if (n >= l) {
            const double f = T[n] * rp[w][n - l];
            re += f * X[w][tri_c(n) + m].x;
            im += f * X[w][tri_c(n) + m].y;
          }
        });
        U[w][k] = make_double2(re, im);
      }
    });
  __syncthreads();
  if (C >= 0)  // rotate back and add into the child's local
    static_loop<0, NS>([&](auto sc) {
      const int k = lane + 64 * decltype(sc)::value;
      if (k < LC) {
        const int mp = kLm.m[k];
        const double2 v = rot_inv_s<P, RB>(Rs, Ph[w][mp], kLm.l[k], mp, U[w]);
        double2 *dst = L + (size_t)C * LC + k;
        const double2 o = *dst;
        *dst = make_double2(o.x + v.x, o.y + v.y);
      }
    });
}

// ---- orders above the templated kernels: one workgroup per parent, the
// children one after another in octant order (the sum order of the
// templated kernels), tables from global memory.

constexpr int kGenThreads = 256;
constexpr int kGenLc = tri_c(kMaxOrder + 1);

// X[k] = forward rotation of U over degrees <= deg (threads over k).
__device__ inline void gen_rot_fwd(const double *__restrict__ R, int rb,
                                   int deg, const double2 *U, double2 *X) {
  for (int k = threadIdx.x; k < tri_c(deg + 1); k += kGenThreads) {
    const int l = kLm.l[k], mp = kLm.m[k];
    const double *ep = R + rofs_c(l) + mp * (l + 1);
    const double *em = ep + rb;
    const double2 *u = U + tri_c(l);
    double re = 0., im = 0.;
    for (int m = 0; m <= l; m++) {
      re += ep[m] * u[m].x;
      im += em[m] * u[m].y;
    }
    X[k] = make_double2(re, im);
  }
}

// Inverse rotation and phase back of coefficient k of Z.
__device__ inline double2 gen_rot_inv(const double *__restrict__ R, int rb,
                                      const double2 *__restrict__ ph, int k,
                                      const double2 *Z) {
  const int l = kLm.l[k], mp = kLm.m[k];
  const double *fp = R + 2 * rb + rofs_c(l) + mp * (l + 1);
  const double *fm = fp + rb;
  const double2 *z = Z + tri_c(l);
  double re = 0., im = 0.;
  for (int m = 0; m <= l; m++) {
    re += fp[m] * z[m].x;
    im += fm[m] * z[m].y;
  }
  const double2 e = ph[mp];
  return make_double2(e.x * re + e.y * im, e.x * im - e.y * re);
}

__global__ __launch_bounds__(kGenThreads) void k_m2m_generic(
    OpTables tb, const TransRow *__restrict__ rows, double2 *__restrict__ M) {
  __shared__ double2 U[kGenLc], X[kGenLc];
  __shared__ double rp[2 * kMaxOrder + 2];
  constexpr int NS = (kGenLc + kGenThreads - 1) / kGenThreads;
  const int p = tb.p, mc = tri_c(p), rb = tb.rb;
  const TransRow &row = rows[blockIdx.x];
  const double half = row.width * 0.25;  // half a child width
  double2 acc[NS];
  for (int j = 0; j < NS; j++) acc[j] = make_double2(0., 0.);
  for (int o = 0; o < 8; o++) {
    const int C = row.child[o];
    if (C < 0) continue;  // uniform over the workgroup
    const int code =
        code_of(1 - 2 * (o & 1), 1 - 2 * ((o >> 1) & 1), 1 - 2 * (o >> 2));
    const DirEntry de = tb.dir[code];
    const double2 *ph = tb.eiphi + code * (p + 1);
    const double *R = tb.rot + (size_t)de.theta * 4 * rb;
    __syncthreads();
    for (int k = threadIdx.x; k < mc; k += kGenThreads)
      U[k] = cmul2(ph[kLm.m[k]], M[(size_t)C * mc + k]);
    if (threadIdx.x < 2 * p + 2) {
      const double f = -de.c * de.norm * half;
      double v = 1.;
      for (int i = 0; i < (int)threadIdx.x; i++) v *= f;
      rp[threadIdx.x] = v;  // ((-c) r)^k
    }
    __syncthreads();
    gen_rot_fwd(R, rb, p - 1, U, X);
    __syncthreads();
    for (int k = threadIdx.x; k < mc; k += kGenThreads) {
      const int l = kLm.l[k], m = kLm.m[k];
      const double *T = tb.tm2m + ((size_t)l * p + m) * p;
      double re = 0., im = 0.;
      for (int n = m; n <= l; n++) {
        const double f = T[n] * rp[l - n];
        re += f * X[tri_c(n) + m].x;
        im += f * X[tri_c(n) + m].y;
      }
      U[k] = make_double2(re, im);
    }
    __syncthreads();
    for (int j = 0; j < NS; j++) {
      const int k = threadIdx.x + j * kGenThreads;
      if (k < mc) {
        const double2 v = gen_rot_inv(R, rb, ph, k, U);
        acc[j].x += v.x;
        acc[j].y += v.y;
      }
    }
  }
  for (int j = 0; j < NS; j++) {
    const int k = threadIdx.x + j * kGenThreads;
    if (k < mc) M[(size_t)row.node * mc + k] = acc[j];
  }
}

__global__ __launch_bounds__(kGenThreads) void k_l2l_generic(
    OpTables tb, const TransRow *__restrict__ rows, double2 *__restrict__ L) {
  __shared__ double2 L0[kGenLc], U[kGenLc], X[kGenLc];
  __shared__ double rp[2 * kMaxOrder + 2];
  const int p = tb.p, lc = tri_c(p + 1), rb = tb.rb;
  const TransRow &row = rows[blockIdx.x];
  const double half = row.width * 0.25;  // half a child width
  for (int k = threadIdx.x; k < lc; k += kGenThreads)
    L0[k] = L[(size_t)row.node * lc + k];
  for (int o = 0; o < 8; o++) {
    const int C = row.child[o];
    if (C < 0) continue;  // uniform over the workgroup
    const int code = code_of(2 * (o & 1) - 1, 2 * ((o >> 1) & 1) - 1,
                             2 * (o >> 2) - 1);
    const DirEntry de = tb.dir[code];
    const double2 *ph = tb.eiphi + code * (p + 1);
    const double *R = tb.rot + (size_t)de.theta * 4 * rb;
    __syncthreads();
    for (int k = threadIdx.x; k < lc; k += kGenThreads)
      U[k] = cmul2(ph[kLm.m[k]], L0[k]);
    if (threadIdx.x < 2 * p + 2) {
      const double f = de.c * de.norm * half;
      double v = 1.;
      for (int i = 0; i < (int)threadIdx.x; i++) v *= f;
      rp[threadIdx.x] = v;  // (c r)^k
    }
    __syncthreads();
    gen_rot_fwd(R, rb, p, U, X);
    __syncthreads();
    for (int k = threadIdx.x; k < lc; k += kGenThreads) {
      const int l = kLm.l[k], m = kLm.m[k];
      const double *T = tb.tl2l + ((size_t)l * (p + 1) + m) * (p + 1);
      double re = 0., im = 0.;
      for (int n = l; n <= p; n++) {
        const double f = T[n] * rp[n - l];
        re += f * X[tri_c(n) + m].x;
        im += f * X[tri_c(n) + m].y;
      }
      U[k] = make_double2(re, im);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < lc; k += kGenThreads) {
      const double2 v = gen_rot_inv(R, rb, ph, k, U);
      double2 *dst = L + (size_t)C * lc + k;
      const double2 d = *dst;
      *dst = make_double2(d.x + v.x, d.y + v.y);
    }
  }
}

template <int P>
static hipError_t level_dispatch(int kind, const OpTables &tb,
                                 const TransRow *rows, int nrows, double2 *X,
                                 hipStream_t s) {
  if constexpr (P > kMaxOrderFast) {
    if (kind == 0)
      k_m2m_generic<<<nrows, kGenThreads, 0, s>>>(tb, rows, X);
    else
      k_l2l_generic<<<nrows, kGenThreads, 0, s>>>(tb, rows, X);
    return hipGetLastError();
  } else {
    if (tb.p != P)
      return level_dispatch<P + 1>(kind, tb, rows, nrows, X, s);
    if (kind == 0)
      k_m2m_level<P><<<nrows, 512, 0, s>>>(tb, rows, X);
    else
      k_l2l_level<P><<<nrows, 512, 0, s>>>(tb, rows, X);
    return hipGetLastError();
  }
}

hipError_t launch_m2m_level(const OpTables &tb, const TransRow *rows,
                            int nrows, double2 *M, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  return level_dispatch<1>(0, tb, rows, nrows, M, s);
}

hipError_t launch_l2l_level(const OpTables &tb, const TransRow *rows,
                            int nrows, double2 *L, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  return level_dispatch<1>(1, tb, rows, nrows, L, s);
}

}  // namespace aran
