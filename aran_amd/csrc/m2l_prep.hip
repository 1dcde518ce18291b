// m2l_prep.hip -- preparation kernels of the tiled far field: tiles of
// same-octant targets, their dense source table, the per-level scaling of
// expansions and the FP32 tables.
#include <hip/hip_runtime.h>

#include "engine.h"
#include "m2l_tile.h"

namespace aran {

template <typename T> struct V2;
template <> struct V2<float> {
  typedef float2 type;
};
template <> struct V2<double> {
  typedef double2 type;
};

constexpr __host__ __device__ int tri_p(int n) { return n * (n + 1) / 2; }

hipError_t launch_m2l_tile_1(int p, int ntiles, const M2LTileArgs &a,
                             const void *Ms, void *Ls, hipStream_t s);
hipError_t launch_m2l_tile_7(int p, int ntiles, const M2LTileArgs &a,
                             const void *Ms, void *Ls, hipStream_t s);
hipError_t launch_m2l_tile_10(int p, int ntiles, const M2LTileArgs &a,
                              const void *Ms, void *Ls, hipStream_t s);
hipError_t launch_m2l_tile_12(int p, int ntiles, const M2LTileArgs &a,
                              const void *Ms, void *Ls, hipStream_t s);
hipError_t launch_m2l_tile_14(int p, int ntiles, const M2LTileArgs &a,
                              const void *Ms, void *Ls, hipStream_t s);

hipError_t launch_m2l_tile(int precision, int p, int ntiles,
                           const M2LTileArgs &a, const void *Ms, void *Ls,
                           hipStream_t s) {
  if (ntiles == 0) return hipSuccess;
  if (precision == 0) return hipErrorInvalidValue;
  if (p <= 6) return launch_m2l_tile_1(p, ntiles, a, Ms, Ls, s);
  if (p <= 9) return launch_m2l_tile_7(p, ntiles, a, Ms, Ls, s);
  if (p <= 11) return launch_m2l_tile_10(p, ntiles, a, Ms, Ls, s);
  if (p <= 13) return launch_m2l_tile_12(p, ntiles, a, Ms, Ls, s);
  return launch_m2l_tile_14(p, ntiles, a, Ms, Ls, s);
}

// ------------------------------------------------------------ preparation

// Per-level octant counts of the nodes of depth >= 2.
__global__ void k_tile_count(NodeArrays nd, int first, int count,
                             int *__restrict__ cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  int node = first + i;
  int oct = (nd.ix[node] & 1) | ((nd.iy[node] & 1) << 1) | ((nd.iz[node] & 1) << 2);
  atomicAdd(&cnt[nd.depth[node] * 8 + oct], 1);
}

// Slot of each node inside the tiles of its (level, octant) group.  The
// slot only chooses the lane; a target's result does not depend on it.
__global__ void k_tile_fill(NodeArrays nd, int first, int count,
                            const int *__restrict__ group_base,
                            int *__restrict__ fill, int *__restrict__ tile_tgt,
                            int *__restrict__ node_slot) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  int node = first + i;
  int oct = (nd.ix[node] & 1) | ((nd.iy[node] & 1) << 1) | ((nd.iz[node] & 1) << 2);
  int g = nd.depth[node] * 8 + oct;
  int slot = group_base[g] * 64 + atomicAdd(&fill[g], 1);
  tile_tgt[slot] = node;
  node_slot[node] = slot;
}

// Dense source table of the tiles from the far (M2L) lists.
__global__ void k_tile_src(NodeArrays nd, int first, int count,
                           const int *__restrict__ vofs,
                           const int *__restrict__ vsrc,
                           const short *__restrict__ vcode,
                           const short *__restrict__ code_to_k,
                           const int *__restrict__ node_slot,
                           int *__restrict__ tile_src) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  int node = first + i;
  int oct = (nd.ix[node] & 1) | ((nd.iy[node] & 1) << 1) | ((nd.iz[node] & 1) << 2);
  int slot = node_slot[node];
  int tile = slot / 64, lane = slot % 64;
  for (int e = vofs[node]; e < vofs[node + 1]; e++) {
    int k = code_to_k[oct * kNumCodes + vcode[e]];
    if (k < 0) continue;  // not a far offset of this octant (cannot happen
                          // for lists built by k_vlist)
    tile_src[((size_t)tile * kOctOffsets + k) * 64 + lane] = vsrc[e];
  }
}

__global__ void k_fill_int(int *__restrict__ p, long n, int v) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_tile_count(const NodeArrays &nd, int first, int count,
                             int *cnt, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  k_tile_count<<<(count + 255) / 256, 256, 0, s>>>(nd, first, count, cnt);
  return hipGetLastError();
}

hipError_t launch_tile_fill(const NodeArrays &nd, int first, int count,
                            const int *group_base, int *fill, int *tile_tgt,
                            int *node_slot, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  k_tile_fill<<<(count + 255) / 256, 256, 0, s>>>(nd, first, count, group_base,
                                                  fill, tile_tgt, node_slot);
  return hipGetLastError();
}

hipError_t launch_tile_src(const NodeArrays &nd, int first, int count,
                           const int *vofs, const int *vsrc, const short *vcode,
                           const short *code_to_k, const int *node_slot,
                           int *tile_src, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  k_tile_src<<<(count + 255) / 256, 256, 0, s>>>(nd, first, count, vofs, vsrc,
                                                 vcode, code_to_k, node_slot,
                                                 tile_src);
  return hipGetLastError();
}

hipError_t launch_fill_int(int *p, long n, int v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_fill_int<<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

// ------------------------------------------------------------ scaling

// Ms = M / w^l (type T), for every node.
template <typename T>
__global__ void k_scale_mult(NodeArrays nd, int nnodes, int p,
                             const double2 *__restrict__ M,
                             typename V2<T>::type *__restrict__ Ms) {
  const int mc = tri_p(p);
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= (long)nnodes * mc) return;
  int node = (int)(idx / mc), k = (int)(idx % mc);
  int l = 0;
  while (tri_p(l + 1) <= k) l++;
  double w = node_width(nd, node);
  double f = 1.;
  for (int i = 0; i < l; i++) f /= w;
  double2 v = M[idx];
  typename V2<T>::type o;
  o.x = (T)(v.x * f);
  o.y = (T)(v.y * f);
  Ms[idx] = o;
}

// L = (sum of the nsplit partial slices of Ls, in slice order) / w^(l+1)
// for nodes of depth >= 2, zero above.
template <typename T>
__global__ void k_unscale_local(NodeArrays nd, int nnodes, int p,
                                const typename V2<T>::type *__restrict__ Ls,
                                int nsplit, long split_stride,
                                double2 *__restrict__ L) {
  const int lc = tri_p(p + 1);
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= (long)nnodes * lc) return;
  int node = (int)(idx / lc), k = (int)(idx % lc);
  if (nd.depth[node] < 2) {
    L[idx] = make_double2(0., 0.);
    return;
  }
  int l = 0;
  while (tri_p(l + 1) <= k) l++;
  double w = node_width(nd, node);
  double f = 1. / w;
  for (int i = 0; i < l; i++) f /= w;
  double sx = 0., sy = 0.;
  for (int sl = 0; sl < nsplit; sl++) {
    typename V2<T>::type v = Ls[sl * split_stride + idx];
    sx += (double)v.x;
    sy += (double)v.y;
  }
  L[idx] = make_double2(sx * f, sy * f);
}

hipError_t launch_scale_mult(int precision, const NodeArrays &nd, int nnodes,
                             int p, const double2 *M, void *Ms, hipStream_t s) {
  long n = (long)nnodes * tri_p(p);
  if (n == 0) return hipSuccess;
  if (precision == 0)
    k_scale_mult<double><<<(n + 255) / 256, 256, 0, s>>>(
        nd, nnodes, p, M, static_cast<double2 *>(Ms));
  else
    k_scale_mult<float><<<(n + 255) / 256, 256, 0, s>>>(
        nd, nnodes, p, M, static_cast<float2 *>(Ms));
  return hipGetLastError();
}

hipError_t launch_unscale_local(int precision, const NodeArrays &nd, int nnodes,
                                int p, const void *Ls, int nsplit,
                                long split_stride, double2 *L, hipStream_t s) {
  long n = (long)nnodes * tri_p(p + 1);
  if (n == 0) return hipSuccess;
  if (precision == 0)
    k_unscale_local<double><<<(n + 255) / 256, 256, 0, s>>>(
        nd, nnodes, p, static_cast<const double2 *>(Ls), nsplit,
        split_stride, L);
  else
    k_unscale_local<float><<<(n + 255) / 256, 256, 0, s>>>(
        nd, nnodes, p, static_cast<const float2 *>(Ls), nsplit,
        split_stride, L);
  return hipGetLastError();
}
// ------------------------------------------------------------ tables

// Per offset code: (-1)^(l+m) c^(l+n) T(l,m,n) / |o|^(l+n+1) in type T.
template <typename T>
__global__ void k_vfac(int p, const DirEntry *__restrict__ dir,
                       const double *__restrict__ tm2l, T *__restrict__ out) {
  const int code = blockIdx.x;
  const int nv = (p + 1) * (p + 1) * p;
  const DirEntry de = dir[code];
  for (int i = threadIdx.x; i < nv; i += blockDim.x) {
    int l = i / ((p + 1) * p), m = (i / p) % (p + 1), n = i % p;
    double v = 0.;
    if (de.norm > 0. && m <= l && n >= m) {
      double rho = 1. / de.norm, f = 1.;
      for (int k = 0; k < l + n + 1; k++) f *= rho;
      v = tm2l[i] * f;
      if ((l + m) & 1) v = -v;
      if ((l + n) & 1) v *= de.c;
    }
    out[(size_t)code * nv + i] = (T)v;
  }
}

template <typename T>
__global__ void k_convert(const double *__restrict__ in, T *__restrict__ out,
                          long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) out[i] = (T)in[i];
}

hipError_t launch_m2l_tables(int precision, int p, const DirEntry *dir,
                             const double *tm2l, const double *rot,
                             long nrot, const double2 *eiphi, void *vfac,
                             void *rotT, void *ephi, hipStream_t s) {
  long ne = (long)kNumCodes * 2 * (p + 1);
  if (precision == 0) {
    k_vfac<double><<<kNumCodes, 256, 0, s>>>(p, dir, tm2l,
                                             static_cast<double *>(vfac));
    k_convert<double><<<(nrot + 255) / 256, 256, 0, s>>>(
        rot, static_cast<double *>(rotT), nrot);
    k_convert<double><<<(ne + 255) / 256, 256, 0, s>>>(
        reinterpret_cast<const double *>(eiphi), static_cast<double *>(ephi),
        ne);
  } else {
    k_vfac<float><<<kNumCodes, 256, 0, s>>>(p, dir, tm2l,
                                            static_cast<float *>(vfac));
    k_convert<float><<<(nrot + 255) / 256, 256, 0, s>>>(
        rot, static_cast<float *>(rotT), nrot);
    k_convert<float><<<(ne + 255) / 256, 256, 0, s>>>(
        reinterpret_cast<const double *>(eiphi), static_cast<float *>(ephi), ne);
  }
  return hipGetLastError();
}

}  // namespace aran
