// m2l_f64.hip -- the plain form's far field as dense FP64 products on MFMA.
//
// Same operator as aran_development3d_m2l_rotate (arandevelopment3d.c:
// 423-430; rotate.c:165-207; translate.c:609-676) on expansions scaled per
// level by the box width (M^_n = M_n / w^n, L^_l = L_l w^(l+1)): for every
// offset code one real matrix, built in FP64 by k_m2l_dense.
//
// Work: like the BF16 kernel, a workgroup takes tiles of 64 same-octant
// targets (one (level, octant) group, neighbours in Morton order) and one
// fixed range of their 189 offset slots; per slot the B operand is the 64
// sources' scaled multipoles (k x 64, FP64, each lane one 16-byte load per
// step) and the A operand the slot's operator, staged through LDS for the
// whole workgroup.  v_mfma_f64_16x16x4_f64: lane i holds A[row i % 16][k
// i / 16] and B[k i / 16][col i % 16]; C[i / 16 + 4 r][i % 16] in
// accumulator r (measured: unlike the FP32 16x16 forms, the four rows of a
// lane are four apart).  Sums over a range stay in FP64 accumulators (fixed order,
// no atomics); the ranges are added in range order by the unscale kernel.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "engine.h"
#include "m2l_f64.h"
#include "m2l_tile.h"
#include "tree_rules.h"

namespace aran {

typedef double f64x4_t __attribute__((ext_vector_type(4)));

constexpr int kF64Waves = 8;
constexpr int kF64Threads = 64 * kF64Waves;

template <int B, int E, class F>
__device__ __forceinline__ void f64_static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    f64_static_for<B + 1, E>(f);
  }
}

// Steps per LDS group: a multiple of the three B buffers, two groups in at
// most 128 KB of LDS.
constexpr __host__ __device__ int f64_group_steps(int step_bytes) {
  return 65536 / step_bytes >= 6 ? 6 : 3;
}

template <int P>
__global__ __launch_bounds__(kF64Threads) __attribute__((
    amdgpu_waves_per_eu(2, 2))) void k_m2l_f64(
    const int *__restrict__ wg_tiles, int nwg, const int *__restrict__ tile_src,
    const int *__restrict__ tile_oct, const short *__restrict__ oct_codes,
    const uint4 *__restrict__ ops, const double2 *__restrict__ mb,
    int zero_node, double *__restrict__ Lsm, int ntiles, int nsplit,
    int sched) {
  constexpr F64Dims D = f64_dims(P);
  constexpr int KS = D.ks2, RBW = D.rbw, RBP = D.rs * D.rbw, CBW = D.cbw;
  constexpr int WPT = D.wpt, TPW = D.tpw, RO = D.ro, KP2 = D.kp / 2;
  constexpr int NCG = 4 / CBW;             // column groups per tile
  constexpr int CHUNK = RBP * 64;          // 16-byte units of one step
  constexpr int NT = kF64Threads;
  constexpr int NB = 3;                    // B buffers (two steps ahead)
  constexpr int G = f64_group_steps(CHUNK * 16);
  constexpr int GU = G * CHUNK;
  constexpr int PER = (GU + NT - 1) / NT;
  static_assert(PER <= 8, "operator group larger than eight units/thread");
  static_assert(G % NB == 0, "group steps must be a multiple of NB");
  __shared__ uint4 ring[2][GU];
  __shared__ short codes[kOctOffsets];
  // workgroups dealt round-robin to the 8 XCDs; each XCD takes one
  // contiguous range of the logical order (schedule 0: tile groups in
  // space order with their offset ranges side by side; 1: range major)
  const int per_xcd = gridDim.x / 8;
  const int logical = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (logical >= nwg * nsplit) return;  // whole workgroup
  const int split = sched ? logical / nwg : logical % nsplit;
  const int wgi = sched ? logical % nwg : logical / nsplit;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const int tid = threadIdx.x, lane = tid % 64;
  const int kq = lane / 16, c16 = lane % 16;
  const int *wt = wg_tiles + (size_t)wgi * TPW;
  const int tile = __builtin_amdgcn_readfirstlane(wt[wave / WPT]);
  const int wpos = wave % WPT;
  const int rsi = wpos / NCG;                 // row split of this wave
  const int colbase = 16 * CBW * (wpos % NCG);
  const int oct = __builtin_amdgcn_readfirstlane(tile_oct[wt[0]]);
  const int chunk = (kOctOffsets + nsplit - 1) / nsplit;
  const int k0 = split * chunk, k1 = min(kOctOffsets, k0 + chunk);
  const int nsteps = max(0, k1 - k0) * KS;
  for (int k = tid; k < kOctOffsets; k += NT)
    codes[k] = oct_codes[oct * kOctOffsets + k];
  __syncthreads();
  // operator unit u of step group grp (steps past the end read the last
  // step: a valid address whose value is never used); table layout
  // [code][step][half][row block][lane] in doubles = [code][step][CHUNK]
  auto group_unit = [&](int grp, int u) __attribute__((always_inline)) -> const uint4 * {
    const int t = min(grp * G + u / CHUNK, nsteps - 1);
    const int k = k0 + t / KS, s = t % KS;
    return ops + ((size_t)codes[k] * KS + s) * CHUNK + u % CHUNK;
  };
  const int ngroups = (nsteps + G - 1) / G;
  if (nsteps > 0)
    for (int u = tid; u < GU; u += NT) ring[0][u] = *group_unit(0, u);
  uint4 p0, p1, p2, p3, p4, p5, p6, p7;
#define ARAN_GFETCH(i, v)                                     \
  if constexpr (PER > i)                                      \
    if (tid + i * NT < GU) v = *group_unit(grp, tid + i * NT);
#define ARAN_GSTORE(i, v)                                     \
  if constexpr (PER > i)                                      \
    if (tid + i * NT < GU) slot[tid + i * NT] = v;
  auto gfetch = [&](int grp) __attribute__((always_inline)) {
    ARAN_GFETCH(0, p0) ARAN_GFETCH(1, p1) ARAN_GFETCH(2, p2)
    ARAN_GFETCH(3, p3) ARAN_GFETCH(4, p4) ARAN_GFETCH(5, p5)
    ARAN_GFETCH(6, p6) ARAN_GFETCH(7, p7)
  };
  auto gstore = [&](uint4 *slot) __attribute__((always_inline)) {
    ARAN_GSTORE(0, p0) ARAN_GSTORE(1, p1) ARAN_GSTORE(2, p2)
    ARAN_GSTORE(3, p3) ARAN_GSTORE(4, p4) ARAN_GSTORE(5, p5)
    ARAN_GSTORE(6, p6) ARAN_GSTORE(7, p7)
  };
#undef ARAN_GFETCH
#undef ARAN_GSTORE
  f64x4_t acc[RBW][CBW];
#pragma unroll
  for (int rb = 0; rb < RBW; rb++)
#pragma unroll
    for (int cb = 0; cb < CBW; cb++) acc[rb][cb] = f64x4_t{0., 0., 0., 0.};
  const int tclamp = max(tile, 0);
  const bool has_tile = tile >= 0;
  // source nodes of offset slot k for this wave's columns (-1: none)
  auto load_src = [&](int k, int *dst) __attribute__((always_inline)) {
    const int kc = min(k, kOctOffsets - 1);
#pragma unroll
    for (int cb = 0; cb < CBW; cb++)
      dst[cb] = tile_src[((size_t)tclamp * kOctOffsets + kc) * 64 + colbase +
                         16 * cb + c16];
  };
  // B fragments of step s (two MFMA k steps) of the sources srcv; a
  // missing source reads the all-zero node zero_node
  auto load_b = [&](const int *srcv, bool ok, int s, double2 (&bb)[CBW])
      __attribute__((always_inline)) {
#pragma unroll
    for (int cb = 0; cb < CBW; cb++) {
      const int node = (ok && srcv[cb] >= 0) ? srcv[cb] : zero_node;
      bb[cb] = mb[((size_t)node * 4 + kq) * KP2 + s];
    }
  };
  auto alive = [&](const int *srcv, bool ok) __attribute__((always_inline)) -> bool {
    bool any = false;
#pragma unroll
    for (int cb = 0; cb < CBW; cb++) any |= ok && srcv[cb] >= 0;
    return __ballot(any) != 0ull;
  };
  int src_c[CBW], src_n[CBW], src_nn[CBW];
  double2 bq[NB][CBW];
  load_src(k0, src_c);
  load_src(k0 + 1, src_n);
  load_src(k0 + 2, src_nn);
  auto load_ahead = [&](int k, int s, int d, double2 (&dst)[CBW])
      __attribute__((always_inline)) {
    const int sa = s + d;
    if (sa < KS)
      load_b(src_c, has_tile && k < k1, sa, dst);
    else if (sa < 2 * KS)
      load_b(src_n, has_tile && k + 1 < k1, sa - KS, dst);
    else
      load_b(src_nn, has_tile && k + 2 < k1, sa - 2 * KS, dst);
  };
  load_ahead(k0, 0, 0, bq[0]);
  load_ahead(k0, 0, 1, bq[1]);
  bool live = alive(src_c, has_tile && k0 < k1);
  auto step = [&](int t, const uint4 *ra, double2 (&cur)[CBW],
                  double2 (&nxt)[CBW]) __attribute__((always_inline)) {
    const int k = k0 + t / KS, s = t % KS;
    load_ahead(k, s, NB - 1, nxt);
    const bool live_t = live;
    if (s + 1 == KS) {
#pragma unroll
      for (int cb = 0; cb < CBW; cb++) {
        src_c[cb] = src_n[cb];
        src_n[cb] = src_nn[cb];
      }
      load_src(k + 3, src_nn);
      live = alive(src_c, has_tile && k + 1 < k1);
    }
    if (!live_t) return;
    const double *A = reinterpret_cast<const double *>(ra);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      double a[RBW];
#pragma unroll
      for (int rb = 0; rb < RBW; rb++)
        a[rb] = A[(h * RBP + rsi * RBW + rb) * 64 + lane];
#pragma unroll
      for (int rb = 0; rb < RBW; rb++)
#pragma unroll
        for (int cb = 0; cb < CBW; cb++)
          acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              a[rb], h ? cur[cb].y : cur[cb].x, acc[rb][cb], 0, 0, 0);
    }
  };
  __syncthreads();  // group 0 in the ring
  for (int grp = 0; grp < ngroups; grp++) {
    if (grp + 1 < ngroups) gfetch(grp + 1);
    const uint4 *rg = ring[grp & 1];
    f64_static_for<0, G>([&](auto gc) {
      constexpr int gg = decltype(gc)::value;
      const int t = grp * G + gg;
      if (t < nsteps)
        step(t, rg + gg * CHUNK, bq[gg % NB], bq[(gg + NB - 1) % NB]);
    });
    if (grp + 1 < ngroups) gstore(ring[(grp + 1) & 1]);
    __syncthreads();
  }
  if (tile < 0) return;
  // slice layout [split][tile][row][64]
  double *out = Lsm + ((size_t)split * ntiles + tile) * RO * 64;
#pragma unroll
  for (int rb = 0; rb < RBW; rb++)
#pragma unroll
    for (int cb = 0; cb < CBW; cb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = 16 * (rsi * RBW + rb) + kq + 4 * r;
        if (row < RO) out[(size_t)row * 64 + colbase + 16 * cb + c16] =
            acc[rb][cb][r];
      }
}

template <int P>
static hipError_t f64_range(int p, int nwg, const int *wgt, int ntiles,
                            const int *tsrc, const int *toct,
                            const short *codes, const uint4 *ops,
                            const double2 *mb, int zero_node, int nsplit,
                            int sched, double *Lsm, hipStream_t s) {
  if constexpr (P > kMaxOrderFast) {
    return hipErrorInvalidValue;
  } else {
    if (p == P) {
      const long wgs = (long)nwg * nsplit;
      const int nblocks = (int)((wgs + 7) / 8 * 8);
      k_m2l_f64<P><<<nblocks, kF64Threads, 0, s>>>(
          wgt, nwg, tsrc, toct, codes, ops, mb, zero_node, Lsm, ntiles,
          nsplit, sched);
      return hipGetLastError();
    }
    return f64_range<P + 1>(p, nwg, wgt, ntiles, tsrc, toct, codes, ops, mb,
                            zero_node, nsplit, sched, Lsm, s);
  }
}

hipError_t launch_m2l_f64(int p, int nwg, const int *wg_tiles, int ntiles,
                          const int *tile_src, const int *tile_oct,
                          const short *oct_codes, const double *ops,
                          const double *mb, int zero_node, int nsplit,
                          int sched, double *Lsm, hipStream_t s) {
  if (p < 1 || p > kMaxOrderFast || nwg <= 0) return hipErrorInvalidValue;
  return f64_range<1>(p, nwg, wg_tiles, ntiles, tile_src, tile_oct, oct_codes,
                      reinterpret_cast<const uint4 *>(ops),
                      reinterpret_cast<const double2 *>(mb), zero_node, nsplit,
                      sched, Lsm, s);
}

// ------------------------------------------------------------ tables

// One thread per operator entry of the table: code, step s, half h, row
// block rb, lane (row 16 rb + lane % 16, k = (lane / 16) kp + 2 s + h).
__global__ void k_f64_pack_ops(int p, const double *__restrict__ O,
                               double *__restrict__ ops) {
  const F64Dims D = f64_dims(p);
  const int rbp = D.rs * D.rbw;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long per_code = (long)D.ks2 * 2 * rbp * 64;
  if (t >= kNumCodes * per_code) return;
  const int code = (int)(t / per_code);
  long r = t % per_code;
  const int lane = (int)(r % 64);
  r /= 64;
  const int rb = (int)(r % rbp);
  r /= rbp;
  const int h = (int)(r % 2);
  const int s = (int)(r / 2);
  const int row = 16 * rb + lane % 16;
  const int k = (lane / 16) * D.kp + 2 * s + h;
  ops[t] = (row < D.ro && k < D.ki) ? O[((size_t)code * D.ro + row) * D.ki + k]
                                    : 0.;
}

hipError_t launch_f64_pack_ops(int p, const double *O, double *ops,
                               hipStream_t s) {
  const long total = (long)kNumCodes * f64_ops_doubles(p);
  k_f64_pack_ops<<<(total + 255) / 256, 256, 0, s>>>(p, O, ops);
  return hipGetLastError();
}

// One thread per (node, quarter, chunk of 8 positions): the scaled real
// input entries k = quarter * kp + position of the node's multipole (zero
// past ki and for the all-zero node nnodes).  The entries of a chunk are
// consecutive in k, so the degree l and the scale w^-l advance
// incrementally.
__global__ void k_f64_pack_mult(NodeArrays nd, int nnodes, int p,
                                const double2 *__restrict__ M,
                                double *__restrict__ mb) {
  // issue priority over the near field's waves (second stream)
  __builtin_amdgcn_s_setprio(2);
  const F64Dims D = f64_dims(p);
  const int nch = (D.kp + 7) / 8;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)(nnodes + 1) * 4 * nch) return;
  const int node = (int)(t / (4 * nch));
  const int q = (int)(t / nch % 4), c = (int)(t % nch);
  double *dst = mb + ((size_t)node * 4 + q) * D.kp + 8 * c;
  const int npos = min(8, D.kp - 8 * c);
  if (node == nnodes) {
    for (int j = 0; j < npos; j++) dst[j] = 0.;
    return;
  }
  const double iw = 1. / node_width(nd, node);
  const double2 *Mn = M + (size_t)node * D.mc;
  int l = -1, sec = -1;  // sec 0: Re part (k < mc), 1: Im part
  double f = 1.;
  for (int j = 0; j < npos; j++) {
    const int k = q * D.kp + 8 * c + j;
    double x = 0.;
    if (k < D.ki) {
      const int ks = k < D.mc ? 0 : 1;
      const int kk = ks ? k - D.mc : k;  // Im part: kk = tri(l-1) + m - 1
      if (ks != sec) {
        sec = ks;
        l = ks;
        f = ks ? iw : 1.;
      }
      while (ks ? tri(l) <= kk : tri(l + 1) <= kk) {
        l++;
        f *= iw;
      }
      const int idx = ks ? tri(l) + (kk - tri(l - 1)) + 1 : kk;
      x = (ks ? Mn[idx].y : Mn[idx].x) * f;
    }
    dst[j] = x;
  }
}

hipError_t launch_f64_pack_mult(const NodeArrays &nd, int nnodes, int p,
                                const double2 *M, double *mb, hipStream_t s) {
  const long total = (long)(nnodes + 1) * 4 * ((f64_dims(p).kp + 7) / 8);
  k_f64_pack_mult<<<(total + 255) / 256, 256, 0, s>>>(nd, nnodes, p, M, mb);
  return hipGetLastError();
}

int m2l_f64_split(int p) {
  (void)p;
  if (const char *e = getenv("ARAN_M2L_SPLIT")) {
    const int v = atoi(e);
    if (v >= 1 && v <= kOctOffsets) return v;
  }
  return kFarRangesF64;
}

bool m2l_f64_enabled(int p) {
  if (p < 1 || p > kMaxOrderFast) return false;
  const char *e = getenv("ARAN_PLAIN_FAR");
  return !(e && strcmp(e, "rotate") == 0);
}

}  // namespace aran
