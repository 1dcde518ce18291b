// m2l_mfma.hip -- far field (M2L) as dense BF16 products on MFMA (gfx950).
//
// Same operator as aran_development3d_m2l_rotate (arandevelopment3d.c:
// 423-430; rotate.c:165-207; translate.c:609-676) on expansions scaled per
// level by the box width (M^_n = M_n / w^n, L^_l = L_l w^(l+1)), where the
// rotate / z-translate / rotate-back map of an offset code is one real
// matrix independent of the level (built in FP64 by k_m2l_dense).
//
// Work: a wave takes a tile of 64 same-octant targets (or one 32-target half
// of it at p >= 11) and a range of its 189 offset slots; per slot the B
// operand is the 64 sources' multipoles (k x 64, each lane loading 16-byte
// pieces of two sources, from L2: the tiles of a workgroup and of an XCD are
// neighbours in space) and the A operand the slot's operator (rows x k,
// shared by the workgroup through LDS).
// v_mfma_f32_32x32x16_bf16: lane i holds A[row = i % 32][k-chunk i / 32] and
// B[k-chunk i / 32][col = i % 32] (8 values each) and C[row 8(r/4) + 4(i/32)
// + r%4][col i % 32] in accumulator r.  Sums over offsets stay in the
// accumulators (fixed order, no atomics); the split over offset ranges is
// summed in slice order by k_unscale_mfma.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "engine.h"
#include "m2l_mfma.h"
#include "m2l_tile.h"

namespace aran {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16_t;

// ------------------------------------------------------------ M2L

__device__ inline bf16x8_t as_bf16x8(uint4 u) {
  return __builtin_bit_cast(bf16x8_t, u);
}

// Workgroup: kMfmaWaves waves on kMfmaWaves / HALVES tiles of one (level,
// octant) group, consecutive in Morton order, so they share the offset codes
// of every step.  The operator slices (one 16-wide k step of every row block
// per step) go through LDS in groups of G steps, double buffered: the whole
// workgroup loads group j + 1 into registers while it multiplies group j,
// stores it, and meets at one barrier per group.  B fragments (the sources)
// are loaded one step ahead, and the first step of the next offset slot PF
// steps ahead (a slot's first step gathers new source nodes; its later steps
// reread the same nodes' next bytes).
constexpr int kMfmaWaves = 8;
constexpr int kMfmaRingUnits = 8192;  // 16-byte LDS units of the two buffers

template <int B, int E, class F>
__device__ __forceinline__ void mfma_static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    mfma_static_for<B + 1, E>(f);
  }
}

constexpr __host__ __device__ int mfma_group_steps(int chunk) {
  const int g = 2 * (kMfmaRingUnits / (4 * chunk));
  return g < 2 ? 2 : (g > 4 ? 4 : g);
}

template <int P, int NPROD>
__global__ __launch_bounds__(64 * kMfmaWaves) __attribute__((
    amdgpu_waves_per_eu(2, 2))) void k_m2l_mfma(
    const int *__restrict__ wg_tiles, int nwg, const int *__restrict__ tile_src,
    const int *__restrict__ tile_oct, const short *__restrict__ oct_codes,
    const uint4 *__restrict__ ops, const uint4 *__restrict__ mb,
    int zero_node, float *__restrict__ Lsm, int ntiles, int nsplit,
    int sched) {
  constexpr MfmaDims D = mfma_dims(P);
  constexpr int KS = D.ks, RB = D.rb, RO = D.ro;
  constexpr int CB = mfma_col_blocks(P);
  constexpr int HALVES = 2 / CB;
  constexpr int NPART = NPROD == 6 ? 3 : 2;
  constexpr int CHUNK = RB * NPART * 64;  // 16-byte units of one step
  constexpr int NT = 64 * kMfmaWaves;
  // B buffers: three for the three-product form (B loaded two steps
  // ahead), two for the six-product form (one step ahead; its registers
  // allow no more); a group holds a multiple of NB steps
  constexpr int NB = NPROD == 3 ? 3 : 2;
  constexpr int G = NB == 3 ? (2 * 6 * CHUNK <= kMfmaRingUnits ? 6 : 3) : 2;
  constexpr int GU = G * CHUNK;
  constexpr int PER = (GU + NT - 1) / NT;
  __shared__ uint4 ring[2][GU];
  // Workgroups are dealt round-robin to the 8 XCDs; each XCD takes one
  // contiguous range of the logical order.  Schedule 0: tile groups in
  // space order with all their offset ranges side by side (the sources a
  // group gathers stay in that XCD's L2).  Schedule 1: offset range major
  // over an octant-major group list (the XCD's workgroups share the
  // operators of one octant's range in L2).  Both give the same sums: a
  // target's partial sum over one fixed offset range does not depend on
  // which workgroup computes it or when.
  const int per_xcd = gridDim.x / 8;
  const int logical = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (logical >= nwg * nsplit) return;  // whole workgroup
  const int split = sched ? logical / nwg : logical % nsplit;
  const int wgi = sched ? logical % nwg : logical / nsplit;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const int tid = threadIdx.x, lane = tid % 64;
  const int g = lane / 32, c32 = lane % 32;
  const int *wt = wg_tiles + (size_t)wgi * (kMfmaWaves / HALVES);
  const int tile = __builtin_amdgcn_readfirstlane(wt[wave / HALVES]);
  const int half = wave % HALVES;
  const int colbase = CB == 1 ? 32 * half : 0;
  const int oct = __builtin_amdgcn_readfirstlane(tile_oct[wt[0]]);
  const int chunk = (kOctOffsets + nsplit - 1) / nsplit;
  const int k0 = split * chunk, k1 = min(kOctOffsets, k0 + chunk);
  const int nsteps = max(0, k1 - k0) * KS;
  // offset codes of this octant, in LDS (read by every step's prefetch)
  __shared__ short codes[kOctOffsets];
  for (int k = tid; k < kOctOffsets; k += 64 * kMfmaWaves)
    codes[k] = oct_codes[oct * kOctOffsets + k];
  __syncthreads();
  // global 16-byte unit of step t for copy slot u (layout of the table:
  // [code][ks][rb][part 3][lane])
  auto src_unit = [&](int t, int u) __attribute__((always_inline)) -> const uint4 * {
    const int k = k0 + t / KS, s = t % KS;
    const int code = codes[k];
    const int rb = u / (NPART * 64), part = (u / 64) % NPART, ln = u % 64;
    return ops + ((((size_t)code * KS + s) * RB + rb) * 3 + part) * 64 + ln;
  };
  // operator unit u of step group grp (steps past the end read the last
  // step: a valid address whose value is never used)
  auto group_unit = [&](int grp, int u) __attribute__((always_inline)) -> const uint4 * {
    const int t = min(grp * G + u / CHUNK, nsteps - 1);
    return src_unit(t, u % CHUNK);
  };
  const int ngroups = (nsteps + G - 1) / G;
  if (nsteps > 0)
    for (int u = tid; u < GU; u += NT) ring[0][u] = *group_unit(0, u);
  // staging registers of the next group (named variables: an array here
  // is moved to LDS by the compiler)
  static_assert(PER <= 8, "operator group larger than eight units/thread");
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
  f32x16_t acc[RB][CB];
#pragma unroll
  for (int rb = 0; rb < RB; rb++)
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[rb][cb][r] = 0.f;
  // Raw source entries of offset slot k for this wave's columns (-1: none;
  // loads are unconditional from a clamped valid address, so nothing waits
  // for them until they are used a whole slot later).
  const int tclamp = max(tile, 0);
  const bool has_tile = tile >= 0;
  auto load_src = [&](int k, int *dst) __attribute__((always_inline)) {
    const int kc = min(k, kOctOffsets - 1);
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
      dst[cb] = tile_src[((size_t)tclamp * kOctOffsets + kc) * 64 + colbase +
                         32 * cb + c32];
  };
  // B fragments (k step s) of the sources srcv; a missing source reads the
  // all-zero node zero_node
  auto load_b = [&](const int *srcv, bool ok, int s,
                    bf16x8_t (&bb)[CB][NPART]) __attribute__((always_inline)) {
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
      const int node = (ok && srcv[cb] >= 0) ? srcv[cb] : zero_node;
      const uint4 *pb = mb + ((size_t)node * KS + s) * (2 * NPART) + g * NPART;
#pragma unroll
      for (int q = 0; q < NPART; q++) bb[cb][q] = as_bf16x8(pb[q]);
    }
  };
  // whether any column of the wave has a source at this slot (wave-uniform)
  auto alive = [&](const int *srcv, bool ok) __attribute__((always_inline)) -> bool {
    bool any = false;
#pragma unroll
    for (int cb = 0; cb < CB; cb++) any |= ok && srcv[cb] >= 0;
    return __ballot(any) != 0ull;
  };
  // sources of the current slot and of the next two
  int src_c[CB], src_n[CB], src_nn[CB];
  bf16x8_t bq[NB][CB][NPART];
  load_src(k0, src_c);
  load_src(k0 + 1, src_n);
  load_src(k0 + 2, src_nn);
  // B fragments of step t + d (d < 3) from the step t of slot k = k0 + t / KS
  auto load_ahead = [&](int k, int s, int d, bf16x8_t (&dst)[CB][NPART])
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
  if constexpr (NB == 3) load_ahead(k0, 0, 1, bq[1]);
  bool live = alive(src_c, has_tile && k0 < k1);
  // one step: MFMAs of step t from `cur` and the LDS slice ra, B of step
  // t + NB - 1 into `nxt`
  auto step = [&](int t, const uint4 *ra, bf16x8_t (&cur)[CB][NPART],
                  bf16x8_t (&nxt)[CB][NPART]) __attribute__((always_inline)) {
    const int k = k0 + t / KS, s = t % KS;
    load_ahead(k, s, NB - 1, nxt);
    const bool next_slot = s + 1 == KS;
    const bool live_t = live;
    if (next_slot) {
#pragma unroll
      for (int cb = 0; cb < CB; cb++) {
        src_c[cb] = src_n[cb];
        src_n[cb] = src_nn[cb];
      }
      load_src(k + 3, src_nn);
      live = alive(src_c, has_tile && k + 1 < k1);
    }
    if (!live_t) return;  // no source in any column: skip the products
    bf16x8_t a[RB][NPART];
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
      for (int q = 0; q < NPART; q++)
        a[rb][q] = as_bf16x8(ra[(rb * NPART + q) * 64 + lane]);
    // products, smallest first ((l,h) (h,l) (m,m) with 6 products, then
    // (m,h) (h,m) (h,h)); the inner loops walk independent accumulators
    constexpr int NQ = NPROD == 6 ? 6 : 3;
    constexpr int QA[6] = {2, 0, 1, 1, 0, 0}, QB[6] = {0, 2, 1, 0, 1, 0};
#pragma unroll
    for (int qq = 0; qq < NQ; qq++) {
      const int q = NPROD == 6 ? qq : qq + 3;
#pragma unroll
      for (int rb = 0; rb < RB; rb++)
#pragma unroll
        for (int cb = 0; cb < CB; cb++)
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
              a[rb][QA[q]], cur[cb][QB[q]], acc[rb][cb], 0, 0, 0);
    }
  };
  __syncthreads();  // group 0 in the ring
  // G is a multiple of NB, so the B buffer of step t is fixed by its place
  // in the group (no register copies between buffers).  Steps past the end
  // would fall in the slot after the range, whose products are skipped (no
  // live source), with operator units repeating the last step, so they may
  // run; the three-product form skips them (the register allocation of the
  // two forms prefers opposite choices).
  for (int grp = 0; grp < ngroups; grp++) {
    if (grp + 1 < ngroups) gfetch(grp + 1);
    const uint4 *rg = ring[grp & 1];
    mfma_static_for<0, G>([&](auto gc) {
      constexpr int gg = decltype(gc)::value;
      const int t = grp * G + gg;
      if (NPROD == 6 || t < nsteps)
        step(t, rg + gg * CHUNK, bq[gg % NB], bq[(gg + NB - 1) % NB]);
    });
    if (grp + 1 < ngroups) gstore(ring[(grp + 1) & 1]);
    __syncthreads();
  }
  if (tile < 0) return;
  // slice layout [split][tile][row][64]: each store is two 128-byte rows
  float *out = Lsm + ((size_t)split * ntiles + tile) * RO * 64;
#pragma unroll
  for (int rb = 0; rb < RB; rb++)
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int row = 32 * rb + 8 * (r / 4) + 4 * g + (r % 4);
        if (row < RO) out[(size_t)row * 64 + colbase + 32 * cb + c32] =
            acc[rb][cb][r];
      }
}

#ifndef ARAN_MFMA_PMIN
#define ARAN_MFMA_PMIN 1
#define ARAN_MFMA_PMAX 15
#endif

template <int P>
static hipError_t mfma_range(int nprod, int p, int nwg, const int *wgt,
                             int ntiles, const int *tsrc, const int *toct,
                             const short *codes, const uint4 *ops,
                             const uint4 *mb, int zero_node, int nsplit,
                             int sched, float *Lsm, hipStream_t s) {
  if constexpr (P > ARAN_MFMA_PMAX) {
    return hipErrorInvalidValue;
  } else {
    if (p == P) {
      const long wgs = (long)nwg * nsplit;
      const int nblocks = (int)((wgs + 7) / 8 * 8);
      if (nprod == 6)
        k_m2l_mfma<P, 6><<<nblocks, 64 * kMfmaWaves, 0, s>>>(
            wgt, nwg, tsrc, toct, codes, ops, mb, zero_node, Lsm, ntiles,
            nsplit, sched);
      else
        k_m2l_mfma<P, 3><<<nblocks, 64 * kMfmaWaves, 0, s>>>(
            wgt, nwg, tsrc, toct, codes, ops, mb, zero_node, Lsm, ntiles,
            nsplit, sched);
      return hipGetLastError();
    }
    return mfma_range<P + 1>(nprod, p, nwg, wgt, ntiles, tsrc, toct, codes,
                             ops, mb, zero_node, nsplit, sched, Lsm, s);
  }
}

#define ARAN_MCAT2(a, b) a##b
#define ARAN_MCAT(a, b) ARAN_MCAT2(a, b)
// launch_m2l_mfma_<PMIN>: orders PMIN..PMAX of this object
hipError_t ARAN_MCAT(launch_m2l_mfma_, ARAN_MFMA_PMIN)(
    int nprod, int p, int nwg, const int *wgt, int ntiles, const int *tsrc,
    const int *toct, const short *codes, const void *ops, const void *mb,
    int zero_node, int nsplit, int sched, float *Lsm, hipStream_t s) {
  if (p < ARAN_MFMA_PMIN || p > ARAN_MFMA_PMAX) return hipErrorInvalidValue;
  return mfma_range<ARAN_MFMA_PMIN>(
      nprod, p, nwg, wgt, ntiles, tsrc, toct, codes,
      static_cast<const uint4 *>(ops), static_cast<const uint4 *>(mb),
      zero_node, nsplit, sched, Lsm, s);
}

}  // namespace aran
