// p2p.hip -- near-field Newton interactions and the FP64 direct sum.
//
// P2P replaces near_func_default / near_func_reflexive
// (aran/aransolver3d.c:254-324) with the Newton callbacks of the reference
// harness (tests/newtonpot3.c:61-78, tests/newtonfield3.c:91-119).  Each
// target leaf gathers from its list of source particle ranges, so every
// result is written once, in a fixed order (no atomics).  Only i == j is
// excluded.
//
// One 256-thread block per target leaf: the leaf's targets are spread over
// the block as (target, slice) pairs, the near sources are staged in LDS in
// passes of kStage particles (coordinates relative to the leaf centre), and
// the slices of each target are summed in a fixed order at the end.  T is
// the arithmetic type of one interaction (double: plain form; float: mixed
// form, where each staged pass is summed in FP32 and passes in FP64).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "engine.h"

namespace aran {

constexpr int kNearThreads = 256;
constexpr int kStage = 1024;
constexpr int kListChunk = 256;

template <typename T> __device__ inline T rsqrt_t(T x);
template <> __device__ inline double rsqrt_t<double>(double x) {
  return 1.0 / sqrt(x);
}
// v_rsq_f32 directly: rsqrtf adds a rescaling for denormal arguments, which
// a squared distance here never is (two particles that close have equal
// FP32 coordinates relative to the leaf centre, r2 == 0: k_near then sums
// that target again in FP64)
template <> __device__ inline float rsqrt_t<float>(float x) {
  return __builtin_amdgcn_rsqf(x);
}

template <typename T> struct Src4;
template <> struct Src4<float> {
  typedef float4 type;
};
template <> struct Src4<double> {
  typedef double4 type;
};

// One source against the lane's target; f == fself is the target itself.
template <typename T>
__device__ __forceinline__ void near_one(const typename Src4<T>::type &s,
                                         T xi, T yi, T zi, bool self, T &p,
                                         T &gx, T &gy, T &gz) {
  const T dx = xi - s.x, dy = yi - s.y, dz = zi - s.z;
  const T r2 = dx * dx + dy * dy + dz * dz;
  // computed for every source and then selected (a branch around it costs
  // more than the reciprocal square root)
  const T ir0 = rsqrt_t<T>(r2);
  const T ir = self ? (T)0 : ir0;
  const T qr = s.w * ir;
  const T qr3 = qr * (ir * ir);
  p += qr;
  gx -= qr3 * dx;
  gy -= qr3 * dy;
  gz -= qr3 * dz;
}

// Near-field sum of target i over its near ranges e0..e1 in FP64 from
// global memory (the recomputation of a target whose FP32 sum overflowed).
__device__ __noinline__ void near_fp64_one(
    int i, int e0, int e1, const int *__restrict__ ubeg,
    const int *__restrict__ uend, const double *__restrict__ xs,
    const double *__restrict__ ys, const double *__restrict__ zs,
    const double *__restrict__ qs, double &p, double &gx, double &gy,
    double &gz) {
  p = gx = gy = gz = 0.;
  const double xi = xs[i], yi = ys[i], zi = zs[i];
  for (int e = e0; e < e1; e++)
    for (int j = ubeg[e]; j < uend[e]; j++) {
      if (j == i) continue;
      const double dx = xi - xs[j], dy = yi - ys[j], dz = zi - zs[j];
      const double ir = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
      const double qr = qs[j] * ir, qr3 = qr * (ir * ir);
      p += qr;
      gx -= qr3 * dx;
      gy -= qr3 * dy;
      gz -= qr3 * dz;
    }
}

template <typename T>
// seven blocks per CU fit the LDS; the register cap keeps them resident
__global__ __launch_bounds__(kNearThreads) __attribute__((
    amdgpu_waves_per_eu(7))) void k_near(
    const int *__restrict__ leaves, NodeArrays nd, const int *__restrict__ uofs,
    const int *__restrict__ ubeg, const int *__restrict__ uend,
    const double *__restrict__ xs, const double *__restrict__ ys,
    const double *__restrict__ zs, const double *__restrict__ qs,
    double *__restrict__ pot, double *__restrict__ grad) {
  typedef typename Src4<T>::type S4;
  // staged sources (x, y, z relative to the leaf centre; q), and after the
  // last pass of a target chunk the per-slice partial sums in the same LDS
  // (18 KB in FP32, so that more leaves run per CU)
  constexpr size_t kSrcBytes = sizeof(S4) * kStage,
                   kRedBytes = sizeof(double) * 4 * kNearThreads;
  __shared__ __align__(16) unsigned char
      smem[kSrcBytes > kRedBytes ? kSrcBytes : kRedBytes];
  S4 *src = reinterpret_cast<S4 *>(smem);
  double(*red)[kNearThreads] = reinterpret_cast<double(*)[kNearThreads]>(smem);
  __shared__ int lbeg[kListChunk], lpre[kListChunk + 1];
  __shared__ int own;  // staged position of the target leaf's first particle
  const int tid = threadIdx.x;
  const int A = leaves[blockIdx.x];
  const int tb = nd.pbegin[A], te = nd.pend[A];
  const double ox = nd.cx[A], oy = nd.cy[A], oz = nd.cz[A];
  const int e0 = uofs[blockIdx.x], e1 = uofs[blockIdx.x + 1];
  for (int t0 = tb; t0 < te; t0 += kNearThreads) {
    const int ntc = min(kNearThreads, te - t0);
    const int nsl = kNearThreads / ntc;
    const int my_t = tid % ntc, my_s = tid / ntc;
    const bool act = my_s < nsl;
    const int i = t0 + my_t;
    const T xi = (T)(xs[i] - ox), yi = (T)(ys[i] - oy), zi = (T)(zs[i] - oz);
    double ap = 0., ax = 0., ay = 0., az = 0.;
    for (int c0 = e0; c0 < e1; c0 += kListChunk) {
      const int nc = min(kListChunk, e1 - c0);
      __syncthreads();
      if (tid < 64) {
        // wave 0: staged position of each range (prefix of the lengths, by
        // shuffles, 64 ranges at a time) and of the target leaf's own range
        if (tid == 0) own = -(1 << 30);
        int carry = 0;
        for (int b0 = 0; b0 < nc; b0 += 64) {
          const int k = b0 + tid;
          int beg = 0, len = 0;
          if (k < nc) {
            beg = ubeg[c0 + k];
            len = uend[c0 + k] - beg;
          }
          int x = len;  // inclusive scan over the wave
          for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off);
            if (tid >= off) x += y;
          }
          if (k < nc) {
            lbeg[k] = beg;
            lpre[k] = carry + x - len;
            if (beg == tb) own = carry + x - len;
          }
          carry += __shfl(x, 63);
        }
        if (tid == 0) lpre[nc] = carry;
      }
      __syncthreads();
      const int total = lpre[nc];
      for (int s0 = 0; s0 < total; s0 += kStage) {
        const int ns = min(kStage, total - s0);
        for (int f = tid; f < ns; f += kNearThreads) {
          const int g = s0 + f;
          int lo = 0, hi = nc - 1;
          while (lo < hi) {  // last k with lpre[k] <= g
            int mid = (lo + hi + 1) >> 1;
            if (lpre[mid] <= g)
              lo = mid;
            else
              hi = mid - 1;
          }
          const int j = lbeg[lo] + (g - lpre[lo]);
          S4 v;
          v.x = (T)(xs[j] - ox);
          v.y = (T)(ys[j] - oy);
          v.z = (T)(zs[j] - oz);
          v.w = (T)qs[j];
          src[f] = v;
        }
        __syncthreads();
        if (act) {
          // the target's own staged position in this pass (never matches
          // when the target leaf is not in it)
          const int fself = own - s0 + (i - tb);
          // the target leaf's own particles are staged at [o0, o1) of this
          // pass: only that range needs the self test.  The ranges are taken
          // in order, so each lane still sums its sources by increasing f.
          const int o0 = min(max(own - s0, 0), ns);
          const int o1 = min(max(own - s0 + (te - tb), 0), ns);
          T p = 0, gx = 0, gy = 0, gz = 0;
          auto range = [&](int lo, int hi, auto chk)
              __attribute__((always_inline)) {
            constexpr bool C = decltype(chk)::value;
            // first f >= lo with f = my_s (mod nsl)
            int f = lo + ((my_s - lo) % nsl + nsl) % nsl;
            for (; f + 3 * nsl < hi; f += 4 * nsl) {
              const S4 a = src[f], b = src[f + nsl], c = src[f + 2 * nsl],
                       d = src[f + 3 * nsl];
              near_one<T>(a, xi, yi, zi, C && f == fself, p, gx, gy, gz);
              near_one<T>(b, xi, yi, zi, C && f + nsl == fself, p, gx, gy, gz);
              near_one<T>(c, xi, yi, zi, C && f + 2 * nsl == fself, p, gx, gy,
                          gz);
              near_one<T>(d, xi, yi, zi, C && f + 3 * nsl == fself, p, gx, gy,
                          gz);
            }
            for (; f < hi; f += nsl)
              near_one<T>(src[f], xi, yi, zi, C && f == fself, p, gx, gy, gz);
          };
          range(0, o0, std::false_type());
          range(o0, o1, std::true_type());
          range(o1, ns, std::false_type());
          ap += (double)p;
          ax += (double)gx;
          ay += (double)gy;
          az += (double)gz;
        }
        __syncthreads();
      }
    }
    red[0][tid] = ap;
    red[1][tid] = ax;
    red[2][tid] = ay;
    red[3][tid] = az;
    __syncthreads();
    if (tid < ntc) {
      double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
      for (int s = 0; s < nsl; s++) {
        const int k = s * ntc + tid;
        s0 += red[0][k];
        s1 += red[1][k];
        s2 += red[2][k];
        s3 += red[3][k];
      }
      if constexpr (std::is_same<T, float>::value) {
        // Two distinct particles whose FP32 leaf-relative coordinates
        // coincide give r2 == 0 and an infinite FP32 term (and a NaN
        // gradient).  Such a target is summed again in FP64, as the plain
        // form sums it; no other target pays for the check.
        if (!(isfinite(s0) && isfinite(s1) && isfinite(s2) && isfinite(s3)))
          near_fp64_one(i, e0, e1, ubeg, uend, xs, ys, zs, qs, s0, s1, s2,
                        s3);
      }
      pot[i] = s0;
      grad[3 * (size_t)i + 0] = s1;
      grad[3 * (size_t)i + 1] = s2;
      grad[3 * (size_t)i + 2] = s3;
    }
    __syncthreads();
  }
}

hipError_t launch_p2p(int precision, const int *leaves, int nleaves,
                      const NodeArrays &nd, const int *uofs, const int *ubeg,
                      const int *uend, const double *xs, const double *ys,
                      const double *zs, const double *qs, double *pot,
                      double *grad, hipStream_t s) {
  if (nleaves == 0) return hipSuccess;
  if (precision == 0)
    k_near<double><<<nleaves, kNearThreads, 0, s>>>(leaves, nd, uofs, ubeg,
                                                    uend, xs, ys, zs, qs, pot,
                                                    grad);
  else
    k_near<float><<<nleaves, kNearThreads, 0, s>>>(leaves, nd, uofs, ubeg,
                                                   uend, xs, ys, zs, qs, pot,
                                                   grad);
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
