// m2l_prep.hip -- preparation kernels of the tiled far field: tiles of
// same-octant targets, their dense source table, the per-level scaling of
// expansions and the FP32 tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "engine.h"
#include "m2l_mfma.h"
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

#define ARAN_MFMA_DECL(n)                                                    \
  hipError_t launch_m2l_mfma_##n(int nprod, int p, int nwg, const int *wgt,  \
                                 int ntiles, const int *tsrc,                \
                                 const int *toct, const short *codes,        \
                                 const void *ops, const void *mb,            \
                                 int zero_node, int nsplit, int sched,       \
                                 float *Lsm, hipStream_t s);
ARAN_MFMA_DECL(1)
ARAN_MFMA_DECL(7)
ARAN_MFMA_DECL(10)
ARAN_MFMA_DECL(12)
ARAN_MFMA_DECL(14)

hipError_t launch_m2l_mfma(int nprod, int p, int nwg, const int *wg_tiles,
                           int ntiles, const int *tsrc, const int *toct,
                           const short *codes, const void *ops,
                           const void *mb, int zero_node, int nsplit,
                           int sched, float *Lsm, hipStream_t s) {
  if (nwg <= 0 || ntiles <= 0) return hipSuccess;
  if (nprod != 3 && nprod != 6) return hipErrorInvalidValue;
  auto f = p <= 6 ? launch_m2l_mfma_1 : p <= 9 ? launch_m2l_mfma_7
           : p <= 11 ? launch_m2l_mfma_10 : p <= 13 ? launch_m2l_mfma_12
                                                    : launch_m2l_mfma_14;
  return f(nprod, p, nwg, wg_tiles, ntiles, tsrc, toct, codes, ops, mb,
           zero_node, nsplit, sched, Lsm, s);
}

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

// Node count of each (level, octant) group g = 8 d + o, d >= 2, from the
// exclusive scan of the octant flags (rank_scan[8 * count] is the total).
__global__ void k_group_counts(int first, int count, int depth,
                               const int *__restrict__ lstart,
                               const int *__restrict__ rank_scan,
                               int *__restrict__ cnt) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (kMaxDepth + 1) * 8) return;
  const int d = g / 8, o = g % 8;
  int c = 0;
  if (d >= 2 && d <= depth) {
    const long row = (long)o * count;
    const int a = lstart[d] - first;
    const int b = d < depth ? lstart[d + 1] - first : count;
    c = rank_scan[row + b] - rank_scan[row + a];
  }
  cnt[g] = c;
}

// flags[o * count + i] = 1 if node first + i is child o of its parent.
__global__ void k_octant_flags(NodeArrays nd, int first, int count,
                               int *__restrict__ flags) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= 8L * count) return;
  const int o = (int)(idx / count), node = first + (int)(idx % count);
  const int oct = (nd.ix[node] & 1) | ((nd.iy[node] & 1) << 1) |
                  ((nd.iz[node] & 1) << 2);
  flags[idx] = oct == o;
}

// Slot of each node: its rank r among the nodes of its (level, octant)
// group (Morton order; from the exclusive scan of the octant flags) picks
// chunk r / 64 of the group, whose tile is tile_of[gofs[group] + r / 64],
// and lane r % 64.
__global__ void k_tile_fill(NodeArrays nd, int first, int count,
                            const int *__restrict__ rank_scan,
                            const int *__restrict__ lstart,
                            const int *__restrict__ gofs,
                            const int *__restrict__ tile_of,
                            int *__restrict__ tile_tgt,
                            int *__restrict__ node_slot) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int node = first + i;
  const int oct = (nd.ix[node] & 1) | ((nd.iy[node] & 1) << 1) |
                  ((nd.iz[node] & 1) << 2);
  const int d = nd.depth[node];
  const long row = (long)oct * count;
  const int r = rank_scan[row + i] - rank_scan[row + (lstart[d] - first)];
  const int slot = tile_of[gofs[d * 8 + oct] + r / 64] * 64 + r % 64;
  tile_tgt[slot] = node;
  node_slot[node] = slot;
}

// Dense source table of the tiles, straight from the colleagues: for a
// target A (octant b in its parent P) and far offset o (one of the 189 of
// octant b, target minus source), the source sits at A - o, whose parent is
// the colleague of P at floor((b - o) / 2) and whose octant there is
// (b - o) mod 2; it is a source when that colleague is not a leaf and has
// that child.  Entries with no source, and lanes with no target, are -1.
// One workgroup per tile: the children of the 27 parent colleagues of each
// lane's target are first summarised in LDS (first child, octant mask; the
// children of a node are stored in octant order), then the 189 x 64 entries
// are written with coalesced stores.  *nfar receives the number of sources
// (the far-field operation count).
// childmap[8 node + o]: child of node in octant o (bit 0: x, 1: y, 2: z),
// -1 when absent (every octant of a leaf).
// Sources of every (tile, offset slot, lane): the child in octant
// (b - offset) mod 2 of the parent's colleague at slot floor((b - offset) /
// 2), from the child map; the parents' colleagues of the tile go through LDS
// first, and so do the tile's offset slots (decoded once per tile).  Every
// load of a phase is independent of the others: a thread takes its entries
// eight at a time, all eight child-map loads in flight before the stores.
__global__ __launch_bounds__(256) void k_tile_src(
    NodeArrays nd, int ntiles, const int *__restrict__ tile_tgt,
    const int *__restrict__ tile_oct, const int *__restrict__ col,
    const int *__restrict__ childmap, const short *__restrict__ oct_codes,
    int *__restrict__ tile_src, int *__restrict__ nfar) {
  // issue priority over the near field's waves (second stream), which
  // would otherwise hold up this latency-bound step of the far-field chain
  __builtin_amdgcn_s_setprio(2);
  __shared__ int scol[27 * 64];
  __shared__ int slot[kOctOffsets];  // (colleague slot << 3 | octant), -1
  __shared__ int wcount[4];
  const int tile = blockIdx.x;
  const int b = tile_oct[tile];
  {  // thread: one target (lane), colleague slots ps = wave, wave + 4, ...
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const int A = tile_tgt[tile * 64 + lane];
    const int P = A >= 0 ? nd.parent[A] : -1;
    int c[7];
#pragma unroll
    for (int i = 0; i < 7; i++) {
      const int ps = w + 4 * i;
      c[i] = (P >= 0 && ps < 27) ? col[(size_t)P * 27 + ps] : -1;
    }
#pragma unroll
    for (int i = 0; i < 7; i++)
      if (w + 4 * i < 27) scol[(w + 4 * i) * 64 + lane] = c[i];
  }
  for (int k = threadIdx.x; k < kOctOffsets; k += blockDim.x) {
    const int bx = b & 1, by = (b >> 1) & 1, bz = b >> 2;
    const int code = oct_codes[b * kOctOffsets + k];
    const int ox = code / 49 - 3, oy = (code / 7) % 7 - 3, oz = code % 7 - 3;
    // (in [-3, 4]; the parent's colleagues cover [-2, 3])
    const int ux = bx - ox, uy = by - oy, uz = bz - oz;
    int v = -1;
    if (ux >= -2 && ux <= 3 && uy >= -2 && uy <= 3 && uz >= -2 && uz <= 3) {
      const int ps =
          ((ux >> 1) + 1) * 9 + ((uy >> 1) + 1) * 3 + ((uz >> 1) + 1);
      v = ps << 3 | ((ux & 1) | (uy & 1) << 1 | (uz & 1) << 2);
    }
    slot[k] = v;
  }
  __syncthreads();
  int found = 0;
  constexpr int kU = 8;
  for (int e0 = threadIdx.x; e0 < kOctOffsets * 64; e0 += kU * 256) {
    int res[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int e = e0 + u * 256;
      res[u] = -1;
      if (e < kOctOffsets * 64) {
        const int v = slot[e / 64];
        const int C = v >= 0 ? scol[(v >> 3) * 64 + e % 64] : -1;
        if (C >= 0) res[u] = childmap[8 * (size_t)C + (v & 7)];
      }
    }
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int e = e0 + u * 256;
      if (e < kOctOffsets * 64) {
        tile_src[(size_t)tile * kOctOffsets * 64 + e] = res[u];
        found += res[u] >= 0;
      }
    }
  }
  // per-wave then per-block count, one atomic per block (exact integers)
  for (int off = 32; off > 0; off >>= 1) found += __shfl_down(found, off);
  if (threadIdx.x % 64 == 0) wcount[threadIdx.x / 64] = found;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicAdd(nfar, wcount[0] + wcount[1] + wcount[2] + wcount[3]);
}

__global__ void k_fill_int(int *__restrict__ p, long n, int v) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_group_counts(int first, int count, int depth,
                               const int *lstart, const int *rank_scan,
                               int *cnt, hipStream_t s) {
  k_group_counts<<<1, (kMaxDepth + 1) * 8, 0, s>>>(first, count, depth,
                                                   lstart, rank_scan, cnt);
  return hipGetLastError();
}

hipError_t launch_octant_flags(const NodeArrays &nd, int first, int count,
                               int *flags, hipStream_t s) {
  const long n = 8L * count;
  if (n <= 0) return hipSuccess;
  k_octant_flags<<<(n + 255) / 256, 256, 0, s>>>(nd, first, count, flags);
  return hipGetLastError();
}

hipError_t launch_tile_fill(const NodeArrays &nd, int first, int count,
                            const int *rank_scan, const int *lstart,
                            const int *gofs, const int *tile_of, int *tile_tgt,
                            int *node_slot, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  k_tile_fill<<<(count + 255) / 256, 256, 0, s>>>(
      nd, first, count, rank_scan, lstart, gofs, tile_of, tile_tgt, node_slot);
  return hipGetLastError();
}

hipError_t launch_tile_src(const NodeArrays &nd, int ntiles,
                           const int *tile_tgt, const int *tile_oct,
                           const int *col, const int *childmap,
                           const short *oct_codes, int *tile_src, int *nfar,
                           hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  k_tile_src<<<ntiles, 256, 0, s>>>(nd, ntiles, tile_tgt, tile_oct, col,
                                    childmap, oct_codes, tile_src, nfar);
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
  const double iw = 1. / node_width(nd, node);
  double f = 1.;
  for (int i = 0; i < l; i++) f *= iw;
  double2 v = M[idx];
  typename V2<T>::type o;
  o.x = (T)(v.x * f);
  o.y = (T)(v.y * f);
  Ms[idx] = o;
}

// One workgroup per tile: sum the nsplit partial slices [split][tile][k][lane]
// in slice order (coalesced reads), transpose through LDS, and write the 64
// target rows L[node][k] = sum / w^(l+1) (coalesced writes), kUnscaleRows
// coefficients at a time.
constexpr int kUnscaleRows = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_unscale_local(
    NodeArrays nd, int p, int ntiles, const int *__restrict__ tile_tgt,
    const typename V2<T>::type *__restrict__ Ls, int nsplit,
    double2 *__restrict__ L) {
  __shared__ double2 sl[64 * (kUnscaleRows + 1)];
  __shared__ double spw[64][kMaxOrderFast + 2];  // 1 / w^(l+1)
  __shared__ int snode[64];
  __shared__ unsigned char sdeg[(kMaxOrderFast + 2) * (kMaxOrderFast + 3) / 2];
  const int lc = tri_p(p + 1);
  const int row = kUnscaleRows + 1;  // odd row stride in LDS
  const int tile = blockIdx.x;
  const long slice = (long)ntiles * lc * 64;
  const typename V2<T>::type *src = Ls + (long)tile * lc * 64;
  if (threadIdx.x < 64) {
    const int node = tile_tgt[tile * 64 + threadIdx.x];
    snode[threadIdx.x] = node;
    const double iw = node >= 0 ? 1. / node_width(nd, node) : 1.;
    double f = iw;
    for (int l = 0; l <= p; l++) {
      spw[threadIdx.x][l] = f;
      f *= iw;
    }
  }
  for (int k = threadIdx.x; k < lc; k += blockDim.x) {
    int l = 0;
    while (tri_p(l + 1) <= k) l++;
    sdeg[k] = (unsigned char)l;
  }
  for (int k0 = 0; k0 < lc; k0 += kUnscaleRows) {
    const int nk = min(kUnscaleRows, lc - k0);
    __syncthreads();
    for (int e = threadIdx.x; e < nk * 64; e += blockDim.x) {
      const int kk = e / 64, lane = e % 64;
      double sx = 0., sy = 0.;
      for (int s = 0; s < nsplit; s++) {
        typename V2<T>::type v = src[s * slice + (long)(k0 + kk) * 64 + lane];
        sx += (double)v.x;
        sy += (double)v.y;
      }
      sl[lane * row + kk] = make_double2(sx, sy);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nk * 64; e += blockDim.x) {
      const int lane = e / nk, kk = e % nk, k = k0 + kk;
      const int node = snode[lane];
      if (node < 0) continue;
      const double f = spw[lane][sdeg[k]];
      const double2 v = sl[lane * row + kk];
      L[(long)node * lc + k] = make_double2(v.x * f, v.y * f);
    }
  }
}

__global__ void k_zero_double2(double2 *__restrict__ p, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_double2(0., 0.);
}

int m2l_tile_major() {
  if (const char *e = getenv("ARAN_M2L_TILE_MAJOR")) return atoi(e) != 0;
  return 1;
}

// Offset ranges per tile of the FP32 far field: a fixed number per order,
// never derived from the size of the solve, so that a target's partial sums
// (and the order they are added in) are the same whatever else is solved
// with it (a batch of systems, a served mix).  ARAN_M2L_SPLIT overrides it
// (1..189) for experiments; results then follow that split.
int m2l_choose_split(int p) {
  (void)p;
  if (const char *e = getenv("ARAN_M2L_SPLIT")) {
    int v = atoi(e);
    if (v >= 1 && v <= kOctOffsets) return v;
  }
  return kFarRangesTile;
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

hipError_t launch_unscale_local(int precision, const NodeArrays &nd, int p,
                                int first, int ntiles, const int *tile_tgt,
                                const void *Ls, int nsplit, double2 *L,
                                hipStream_t s) {
  const int lc = tri_p(p + 1);
  const long nz = (long)first * lc;
  if (nz > 0) k_zero_double2<<<(nz + 255) / 256, 256, 0, s>>>(L, nz);
  if (ntiles <= 0) return hipGetLastError();
  if (precision == 0)
    k_unscale_local<double><<<ntiles, 256, 0, s>>>(
        nd, p, ntiles, tile_tgt, static_cast<const double2 *>(Ls), nsplit, L);
  else
    k_unscale_local<float><<<ntiles, 256, 0, s>>>(
        nd, p, ntiles, tile_tgt, static_cast<const float2 *>(Ls), nsplit, L);
  return hipGetLastError();
}
// ------------------------------------------------------------ MFMA form

__device__ inline unsigned bf16_rn(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

// x = h + m + l in BF16 (after rounding x to FP32; each residual is exact)
__device__ inline void split3(double x, unsigned &h, unsigned &m,
                              unsigned &l) {
  const float f = (float)x;
  h = bf16_rn(f);
  const float r1 = f - __uint_as_float(h << 16);
  m = bf16_rn(r1);
  const float r2 = r1 - __uint_as_float(m << 16);
  l = bf16_rn(r2);
}

// 8 values -> three 16-byte BF16 parts
__device__ inline void pack8(const double *v, uint4 *dst) {
  unsigned hw[4] = {0, 0, 0, 0}, mw[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0};
  for (int j = 0; j < 8; j++) {
    unsigned h, m, l;
    split3(v[j], h, m, l);
    const int sh = (j & 1) * 16;
    hw[j / 2] |= h << sh;
    mw[j / 2] |= m << sh;
    lw[j / 2] |= l << sh;
  }
  dst[0] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
  dst[1] = make_uint4(mw[0], mw[1], mw[2], mw[3]);
  dst[2] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
}

// ------------------------------------------------------------ tables

// One thread per (code, ks, rb, lane): 8 operator entries of row
// 32 rb + lane % 32, columns 16 ks + 8 (lane / 32) + j; table layout
// [code][ks][rb][part][lane] (16-byte units), so that one k step of one
// code is contiguous.
__global__ void k_mfma_pack_ops(int p, const double *__restrict__ O,
                                uint4 *__restrict__ ops) {
  const MfmaDims D = mfma_dims(p);
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long total = (long)kNumCodes * D.rb * D.ks * 64;
  if (t >= total) return;
  const int lane = (int)(t % 64);
  const int rb = (int)((t / 64) % D.rb);
  const int s = (int)((t / 64 / D.rb) % D.ks);
  const int code = (int)(t / 64 / D.rb / D.ks);
  const int row = 32 * rb + lane % 32;
  double v[8];
  for (int j = 0; j < 8; j++) {
    const int k = 16 * s + 8 * (lane / 32) + j;
    v[j] = (row < D.ro && k < D.ki)
               ? O[((size_t)code * D.ro + row) * D.ki + k]
               : 0.;
  }
  uint4 parts[3];
  pack8(v, parts);
  uint4 *dst = ops + (((size_t)code * D.ks + s) * D.rb + rb) * 3 * 64 + lane;
  for (int q = 0; q < 3; q++) dst[q * 64] = parts[q];
}

// Nodes per workgroup of k_mfma_pack_mult, and the longest real input
// vector its tables hold (2 mc - p = 225 at order 15).
constexpr int kPackNodes = 16;
constexpr int kPackMaxKi = 256;

// One thread per (node, ks, half): 8 scaled multipole entries 16 ks +
// 8 half + j of the real input vector, as the first nparts BF16 parts (the
// parts the form's products read, so a source's bytes are all used).  A
// workgroup takes kPackNodes consecutive nodes: their multipole rows (one
// contiguous stretch of M) are staged in LDS by coalesced loads, and the
// degree and source slot of each entry k come from a table built once per
// workgroup.  The scale w^-l is the product of l factors 1/w in sequence,
// as the entries of a section advance in degree.
__global__ void k_mfma_pack_mult(NodeArrays nd, int nnodes, int p, int nparts,
                                 const double2 *__restrict__ M,
                                 uint4 *__restrict__ mb) {
  // issue priority over the near field's waves (second stream), which
  // would otherwise hold up this latency-bound step of the far-field chain
  __builtin_amdgcn_s_setprio(2);
  extern __shared__ double2 pack_rows[];  // [kPackNodes][mc]
  __shared__ short sidx[kPackMaxKi];
  __shared__ unsigned char sdeg[kPackMaxKi], sim[kPackMaxKi];
  const MfmaDims D = mfma_dims(p);
  const int tid = threadIdx.x;
  const int n0 = blockIdx.x * kPackNodes;
  const int nreal = min(kPackNodes, nnodes - n0);  // may be <= 0
  for (int k = tid; k < D.ki; k += blockDim.x) {
    const int ks = k < D.mc ? 0 : 1;
    const int kk = ks ? k - D.mc : k;  // Im part: kk = tri(l-1) + m - 1
    int l = ks;
    while (ks ? tri(l) <= kk : tri(l + 1) <= kk) l++;
    sidx[k] = (short)(ks ? tri(l) + (kk - tri(l - 1)) + 1 : kk);
    sdeg[k] = (unsigned char)l;
    sim[k] = (unsigned char)ks;
  }
  if (nreal > 0) {
    const double2 *src = M + (size_t)n0 * D.mc;
    for (int e = tid; e < nreal * D.mc; e += blockDim.x) pack_rows[e] = src[e];
  }
  __syncthreads();
  const int nl = tid / (2 * D.ks), rest = tid % (2 * D.ks);
  const int node = n0 + nl;
  if (nl >= kPackNodes || node > nnodes) return;
  const long t = (long)node * 2 * D.ks + rest;
  if (node == nnodes) {  // the all-zero node that stands in for no source
    const uint4 z = make_uint4(0, 0, 0, 0);
    for (int q = 0; q < nparts; q++) mb[t * nparts + q] = z;
    return;
  }
  const int s = rest / 2, g = rest % 2;
  const double iw = 1. / node_width(nd, node);
  const double2 *Mn = pack_rows + (size_t)nl * D.mc;
  double v[8];
  int cl = 0;
  double f = 1.;  // iw^cl
  for (int j = 0; j < 8; j++) {
    const int k = 16 * s + 8 * g + j;
    double x = 0.;
    if (k < D.ki) {
      const int l = sdeg[k];
      if (l < cl) {  // the Im section starts again at degree 1
        cl = 0;
        f = 1.;
      }
      while (cl < l) {
        f *= iw;
        cl++;
      }
      const double2 c = Mn[sidx[k]];
      x = (sim[k] ? c.y : c.x) * f;
    }
    v[j] = x;
  }
  uint4 parts[3];
  pack8(v, parts);
  // constant indices keep parts[] in registers
  mb[t * nparts] = parts[0];
  mb[t * nparts + 1] = parts[1];
  if (nparts == 3) mb[t * nparts + 2] = parts[2];
}

hipError_t launch_mfma_pack_ops(int p, const double *O, void *ops,
                                hipStream_t s) {
  const MfmaDims D = mfma_dims(p);
  const long total = (long)kNumCodes * D.rb * D.ks * 64;
  k_mfma_pack_ops<<<(total + 255) / 256, 256, 0, s>>>(
      p, O, static_cast<uint4 *>(ops));
  return hipGetLastError();
}

hipError_t launch_mfma_pack_mult(const NodeArrays &nd, int nnodes, int p,
                                 int nparts, const double2 *M, void *mb,
                                 hipStream_t s) {
  if (nparts < 2 || nparts > 3) return hipErrorInvalidValue;
  const MfmaDims D = mfma_dims(p);
  if (D.ki > kPackMaxKi) return hipErrorInvalidValue;
  const int threads = kPackNodes * 2 * D.ks;  // <= 512 at order 15
  const int blocks = (nnodes + 1 + kPackNodes - 1) / kPackNodes;
  k_mfma_pack_mult<<<blocks, threads, sizeof(double2) * kPackNodes * D.mc,
                     s>>>(nd, nnodes, p, nparts, M,
                          static_cast<uint4 *>(mb));
  return hipGetLastError();
}


// One workgroup per tile: sum the nsplit FP32 slices [split][tile][row][64]
// in slice order, transpose through LDS and write L[node][k] =
// (Re row k, Im row of (l, m >= 1)) / w^(l+1), kUnscaleRows coefficients at
// a time.
template <class T>
__global__ __launch_bounds__(256) void k_unscale_mfma(
    NodeArrays nd, int p, int ntiles, const int *__restrict__ tile_tgt,
    const T *__restrict__ Lsm, int nsplit, double2 *__restrict__ L) {
  __shared__ double2 sl[64 * (kUnscaleRows + 1)];
  __shared__ double spw[64][kMaxOrderFast + 2];  // 1 / w^(l+1)
  __shared__ int snode[64];
  __shared__ short srim[(kMaxOrderFast + 2) * (kMaxOrderFast + 3) / 2];
  __shared__ unsigned char sdeg[(kMaxOrderFast + 2) * (kMaxOrderFast + 3) / 2];
  const MfmaDims D = mfma_dims(p);
  const int lc = D.lc;
  const int row = kUnscaleRows + 1;  // odd row stride in LDS
  const int tile = blockIdx.x;
  const long slice = (long)ntiles * D.ro * 64;
  const T *src = Lsm + (long)tile * D.ro * 64;
  if (threadIdx.x < 64) {
    const int node = tile_tgt[tile * 64 + threadIdx.x];
    snode[threadIdx.x] = node;
    const double iw = node >= 0 ? 1. / node_width(nd, node) : 1.;
    double f = iw;
    for (int l = 0; l <= p; l++) {
      spw[threadIdx.x][l] = f;
      f *= iw;
    }
  }
  // per coefficient: degree, and the row of its imaginary part (-1: m = 0)
  for (int k = threadIdx.x; k < lc; k += blockDim.x) {
    int l = 0;
    while (tri_p(l + 1) <= k) l++;
    const int m = k - tri_p(l);
    sdeg[k] = (unsigned char)l;
    srim[k] = (short)(m > 0 ? lc + tri_p(l - 1) + m - 1 : -1);
  }
  // blockIdx.y takes one chunk of kUnscaleRows coefficients (the chunks of
  // a tile run in parallel)
  for (int k0 = blockIdx.y * kUnscaleRows; k0 < lc;
       k0 += gridDim.y * kUnscaleRows) {
    const int nk = min(kUnscaleRows, lc - k0);
    __syncthreads();
    for (int e = threadIdx.x; e < nk * 64; e += blockDim.x) {
      const int kk = e / 64, lane = e % 64, k = k0 + kk;
      const int rim = srim[k];
      double sx = 0., sy = 0.;
      for (int s = 0; s < nsplit; s++) {
        sx += (double)src[s * slice + (long)k * 64 + lane];
        if (rim >= 0) sy += (double)src[s * slice + (long)rim * 64 + lane];
      }
      sl[lane * row + kk] = make_double2(sx, sy);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nk * 64; e += blockDim.x) {
      const int lane = e / nk, kk = e % nk, k = k0 + kk;
      const int node = snode[lane];
      if (node < 0) continue;
      const double f = spw[lane][sdeg[k]];
      const double2 v = sl[lane * row + kk];
      L[(long)node * lc + k] = make_double2(v.x * f, v.y * f);
    }
  }
}

template <class T>
static hipError_t unscale_slices(const NodeArrays &nd, int p, int first,
                                 int ntiles, const int *tile_tgt, const T *Lsm,
                                 int nsplit, double2 *L, hipStream_t s) {
  const int lc = tri_p(p + 1);
  const long nz = (long)first * lc;
  if (nz > 0) k_zero_double2<<<(nz + 255) / 256, 256, 0, s>>>(L, nz);
  if (ntiles <= 0) return hipGetLastError();
  const dim3 grid(ntiles, (lc + kUnscaleRows - 1) / kUnscaleRows);
  k_unscale_mfma<T><<<grid, 256, 0, s>>>(nd, p, ntiles, tile_tgt, Lsm, nsplit,
                                         L);
  return hipGetLastError();
}

hipError_t launch_unscale_mfma(const NodeArrays &nd, int p, int first,
                               int ntiles, const int *tile_tgt,
                               const float *Lsm, int nsplit, double2 *L,
                               hipStream_t s) {
  return unscale_slices(nd, p, first, ntiles, tile_tgt, Lsm, nsplit, L, s);
}

hipError_t launch_unscale_mfma(const NodeArrays &nd, int p, int first,
                               int ntiles, const int *tile_tgt,
                               const double *Lsm, int nsplit, double2 *L,
                               hipStream_t s) {
  return unscale_slices(nd, p, first, ntiles, tile_tgt, Lsm, nsplit, L, s);
}

// Offset ranges per tile of the MFMA far field: fixed per order, as for the
// FP32 far field (results do not depend on the batch a system is solved
// in); ARAN_M2L_SPLIT overrides it.
int m2l_mfma_split(int p) {
  (void)p;
  if (const char *e = getenv("ARAN_M2L_SPLIT")) {
    int v = atoi(e);
    if (v >= 1 && v <= kOctOffsets) return v;
  }
  return kFarRangesMfma;
}

int mfma_default_schedule(int fallback) {
  if (const char *e = getenv("ARAN_MFMA_SCHED")) {
    const int v = atoi(e);
    if (v >= 0 && v < kMfmaSchedules) return v;
  }
  return fallback;
}

// ------------------------------------------------------------ tables

// Per offset code: (-1)^(l+m) c^(l+n) T(l,m,n) / |o|^(l+n+1) in type T.
template <typename T>
__global__ void k_vfac(int p, const DirEntry *__restrict__ dir,
                       const double *__restrict__ tm2l, T *__restrict__ out) {
  const int code = blockIdx.x;
  const int vs = vfac_size(p);
  const DirEntry de = dir[code];
  for (int l = 0; l <= p; l++)
    for (int m = 0; m <= min(l, p - 1); m++)
      for (int n = m + threadIdx.x; n < p; n += blockDim.x) {
        const int i = (l * (p + 1) + m) * p + n;  // [l][m][n] of tm2l
        double v = 0.;
        if (de.norm > 0.) {
          double rho = 1. / de.norm, f = 1.;
          for (int k = 0; k < l + n + 1; k++) f *= rho;
          v = tm2l[i] * f;
          if ((l + m) & 1) v = -v;
          if ((l + n) & 1) v *= de.c;
        }
        out[(size_t)code * vs + vfac_ofs(p, l, m) + (n - m)] = (T)v;
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
