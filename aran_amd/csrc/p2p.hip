// p2p.hip -- near-field Newton interactions and the FP64 direct sum.
//
// P2P replaces near_func_default / near_func_reflexive
// (aran/aransolver3d.c:254-324) with the Newton callbacks of the reference
// harness (tests/newtonpot3.c:61-78, tests/newtonfield3.c:91-119): each
// target leaf gathers from its list of source particle ranges, so every
// result is written by exactly one thread (no atomics, fixed order).  Only
// i == j is excluded.
#include <hip/hip_runtime.h>

#include "engine.h"

namespace aran {

constexpr int kP2PThreads = 64;

__global__ __launch_bounds__(kP2PThreads) void k_p2p(
    const int *__restrict__ leaves, const int *__restrict__ lpb,
    const int *__restrict__ lpe, const int *__restrict__ uofs,
    const int *__restrict__ ubeg, const int *__restrict__ uend,
    const double *__restrict__ xs, const double *__restrict__ ys,
    const double *__restrict__ zs, const double *__restrict__ qs,
    double *__restrict__ pot, double *__restrict__ grad) {
  __shared__ double sx[kP2PThreads], sy[kP2PThreads], sz[kP2PThreads],
      sq[kP2PThreads];
  const int A = leaves[blockIdx.x];
  const int tb = lpb[A], te = lpe[A];
  const int e0 = uofs[blockIdx.x], e1 = uofs[blockIdx.x + 1];
  for (int t0 = tb; t0 < te; t0 += kP2PThreads) {
    const int i = t0 + threadIdx.x;
    const bool act = i < te;
    const double xi = act ? xs[i] : 0., yi = act ? ys[i] : 0.,
                 zi = act ? zs[i] : 0.;
    double ph = 0., gx = 0., gy = 0., gz = 0.;
    for (int e = e0; e < e1; e++) {
      const int sb = ubeg[e], se = uend[e];
      for (int s0 = sb; s0 < se; s0 += kP2PThreads) {
        int j = s0 + threadIdx.x;
        if (j < se) {
          sx[threadIdx.x] = xs[j];
          sy[threadIdx.x] = ys[j];
          sz[threadIdx.x] = zs[j];
          sq[threadIdx.x] = qs[j];
        }
        __syncthreads();
        const int cnt = min(kP2PThreads, se - s0);
        if (act) {
          for (int jj = 0; jj < cnt; jj++) {
            if (s0 + jj == i) continue;
            double dx = xi - sx[jj], dy = yi - sy[jj], dz = zi - sz[jj];
            double r2 = dx * dx + dy * dy + dz * dz;
            double ir = 1. / sqrt(r2);
            double qr = sq[jj] * ir;
            double qr3 = qr * ir * ir;
            ph += qr;
            gx -= qr3 * dx;
            gy -= qr3 * dy;
            gz -= qr3 * dz;
          }
        }
        __syncthreads();
      }
    }
    if (act) {
      pot[i] = ph;
      grad[3 * (size_t)i + 0] = gx;
      grad[3 * (size_t)i + 1] = gy;
      grad[3 * (size_t)i + 2] = gz;
    }
  }
}

hipError_t launch_p2p(const int *leaves, int nleaves, const int *lpb,
                      const int *lpe, const int *uofs, const int *ubeg,
                      const int *uend, const double *xs, const double *ys,
                      const double *zs, const double *qs, double *pot,
                      double *grad, hipStream_t s) {
  if (nleaves == 0) return hipSuccess;
  k_p2p<<<nleaves, kP2PThreads, 0, s>>>(leaves, lpb, lpe, uofs, ubeg, uend, xs,
                                        ys, zs, qs, pot, grad);
  return hipGetLastError();
}

// FP64 direct sum for chosen targets (tests/newtonpot3.c:678-716) with
// Neumaier-compensated accumulation; targets are indices in user order.
constexpr int kDirThreads = 256;

__device__ inline void neumaier(double &s, double &c, double v) {
  double t = s + v;
  if (fabs(s) >= fabs(v))
    c += (s - t) + v;
  else
    c += (v - t) + s;
  s = t;
}

__global__ __launch_bounds__(kDirThreads) void k_direct(
    const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ q, long n,
    const long *__restrict__ targets, long nt, double *__restrict__ pot,
    double *__restrict__ grad) {
  __shared__ double sx[kDirThreads], sy[kDirThreads], sz[kDirThreads],
      sq[kDirThreads];
  const long t = blockIdx.x * (long)kDirThreads + threadIdx.x;
  const bool act = t < nt;
  const long i = act ? targets[t] : -1;
  const double xi = act ? x[i] : 0., yi = act ? y[i] : 0., zi = act ? z[i] : 0.;
  double s[4] = {0., 0., 0., 0.}, c[4] = {0., 0., 0., 0.};
  for (long s0 = 0; s0 < n; s0 += kDirThreads) {
    long j = s0 + threadIdx.x;
    if (j < n) {
      sx[threadIdx.x] = x[j];
      sy[threadIdx.x] = y[j];
      sz[threadIdx.x] = z[j];
      sq[threadIdx.x] = q[j];
    }
    __syncthreads();
    const int cnt = (int)min((long)kDirThreads, n - s0);
    if (act) {
      for (int jj = 0; jj < cnt; jj++) {
        if (s0 + jj == i) continue;
        double dx = xi - sx[jj], dy = yi - sy[jj], dz = zi - sz[jj];
        double ir = 1. / sqrt(dx * dx + dy * dy + dz * dz);
        double qr = sq[jj] * ir, qr3 = qr * ir * ir;
        neumaier(s[0], c[0], qr);
        neumaier(s[1], c[1], -qr3 * dx);
        neumaier(s[2], c[2], -qr3 * dy);
        neumaier(s[3], c[3], -qr3 * dz);
      }
    }
    __syncthreads();
  }
  if (act) {
    pot[t] = s[0] + c[0];
    grad[3 * t + 0] = s[1] + c[1];
    grad[3 * t + 1] = s[2] + c[2];
    grad[3 * t + 2] = s[3] + c[3];
  }
}

hipError_t launch_direct(const double *x, const double *y, const double *z,
                         const double *q, long n, const long *targets, long nt,
                         double *pot, double *grad, hipStream_t s) {
  if (nt == 0) return hipSuccess;
  k_direct<<<(nt + kDirThreads - 1) / kDirThreads, kDirThreads, 0, s>>>(
      x, y, z, q, n, targets, nt, pot, grad);
  return hipGetLastError();
}

}  // namespace aran
