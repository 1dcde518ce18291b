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
// of every step; the operator slice of a step (one 16-wide k step of every
// row block) goes through a ring of kMfmaStages LDS buffers, loaded by the
// whole workgroup kMfmaStages - 1 steps ahead.
constexpr int kMfmaWaves = 8;
constexpr int kMfmaStages = 3;

template <int P, int NPROD>
__global__ __launch_bounds__(64 * kMfmaWaves) __attribute__((
    amdgpu_waves_per_eu(2, 2))) void k_m2l_mfma(
    const int *__restrict__ wg_tiles, int nwg, const int *__restrict__ tile_src,
    const int *__restrict__ tile_oct, const short *__restrict__ oct_codes,
    const uint4 *__restrict__ ops, const uint4 *__restrict__ mb,
    int zero_node, float *__restrict__ Lsm, int ntiles, int nsplit) {
  constexpr MfmaDims D = mfma_dims(P);
  constexpr int KS = D.ks, RB = D.rb, RO = D.ro;
  constexpr int CB = mfma_col_blocks(P);
  constexpr int HALVES = 2 / CB;
  constexpr int NPART = NPROD == 6 ? 3 : 2;
  constexpr int CHUNK = RB * NPART * 64;  // 16-byte units of one step
  constexpr int PER_T = (CHUNK + 64 * kMfmaWaves - 1) / (64 * kMfmaWaves);
  __shared__ uint4 ring[kMfmaStages][CHUNK];
  // workgroups are dealt round-robin to the 8 XCDs: give each XCD one
  // contiguous range of tile groups, with all their offset ranges side by
  // side, so that the sources a group gathers stay in that XCD's L2
  const int per_xcd = gridDim.x / 8;
  const int logical = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (logical >= nwg * nsplit) return;  // whole workgroup
  const int split = logical % nsplit, wgi = logical / nsplit;
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
  auto src_unit = [&](int t, int u) -> const uint4 * {
    const int k = k0 + t / KS, s = t % KS;
    const int code = codes[k];
    const int rb = u / (NPART * 64), part = (u / 64) % NPART, ln = u % 64;
    return ops + ((((size_t)code * KS + s) * RB + rb) * 3 + part) * 64 + ln;
  };
  // prologue: steps 0 .. kMfmaStages - 2 straight into the ring
  for (int t = 0; t < kMfmaStages - 1 && t < nsteps; t++)
    for (int u = tid; u < CHUNK; u += 64 * kMfmaWaves)
      ring[t][u] = *src_unit(t, u);
  // the next ring slot is staged in registers (up to three units per thread)
  static_assert(PER_T <= 3, "operator chunk larger than three units/thread");
  constexpr int NT = 64 * kMfmaWaves;
  const uint4 zero4 = make_uint4(0, 0, 0, 0);
  uint4 pre0 = zero4, pre1 = zero4, pre2 = zero4;
  auto fetch = [&](int t) {
    if (CHUNK >= NT || tid < CHUNK) pre0 = *src_unit(t, tid);
    if constexpr (PER_T > 1)
      if (tid + NT < CHUNK) pre1 = *src_unit(t, tid + NT);
    if constexpr (PER_T > 2)
      if (tid + 2 * NT < CHUNK) pre2 = *src_unit(t, tid + 2 * NT);
  };
  auto store = [&](uint4 *slot) {
    if (CHUNK >= NT || tid < CHUNK) slot[tid] = pre0;
    if constexpr (PER_T > 1)
      if (tid + NT < CHUNK) slot[tid + NT] = pre1;
    if constexpr (PER_T > 2)
      if (tid + 2 * NT < CHUNK) slot[tid + 2 * NT] = pre2;
  };
  if (kMfmaStages - 1 < nsteps) fetch(kMfmaStages - 1);
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
  auto load_src = [&](int k, int *dst) {
    const int kc = min(k, kOctOffsets - 1);
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
      dst[cb] = tile_src[((size_t)tclamp * kOctOffsets + kc) * 64 + colbase +
                         32 * cb + c32];
  };
  // B fragments (k step s) of the sources srcv; a missing source reads the
  // all-zero node zero_node
  auto load_b = [&](const int *srcv, bool ok, int s,
                    bf16x8_t (&bb)[CB][NPART]) {
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
      const int node = (ok && srcv[cb] >= 0) ? srcv[cb] : zero_node;
      const uint4 *pb = mb + ((size_t)node * KS + s) * 6 + g * 3;
#pragma unroll
      for (int q = 0; q < NPART; q++) bb[cb][q] = as_bf16x8(pb[q]);
    }
  };
  // whether any column of the wave has a source at this slot (wave-uniform)
  auto alive = [&](const int *srcv, bool ok) -> bool {
    bool any = false;
#pragma unroll
    for (int cb = 0; cb < CB; cb++) any |= ok && srcv[cb] >= 0;
    return __ballot(any) != 0ull;
  };
  int src_c[CB], src_n[CB];
  bf16x8_t bA[CB][NPART], bB[CB][NPART];
  load_src(k0, src_c);
  load_src(k0 + 1, src_n);
  load_b(src_c, has_tile, 0, bA);
  bool live = alive(src_c, has_tile && k0 < k1);
  // one step: MFMAs of step t from `cur`, B of step t + 1 into `nxt`
  auto step = [&](int t, bf16x8_t (&cur)[CB][NPART],
                  bf16x8_t (&nxt)[CB][NPART]) {
    __syncthreads();
    // ring slot of step t + kMfmaStages - 1 (read last at step t - 1)
    const int tw = t + kMfmaStages - 1;
    if (tw < nsteps) store(ring[tw % kMfmaStages]);
    fetch(min(tw + 1, nsteps - 1));
    const int k = k0 + t / KS, s = t % KS;
    const bool next_slot = s + 1 == KS;
    if (next_slot) {
      load_b(src_n, has_tile && k + 1 < k1, 0, nxt);
#pragma unroll
      for (int cb = 0; cb < CB; cb++) src_c[cb] = src_n[cb];
      load_src(k + 2, src_n);
    } else {
      load_b(src_c, has_tile, s + 1, nxt);
    }
    const bool live_t = live;
    if (next_slot) live = alive(src_c, has_tile && k + 1 < k1);
    if (!live_t) return;  // no source in any column: skip the products
    const uint4 *ra = ring[t % kMfmaStages];
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
  // two steps per iteration with ping-pong B buffers (no register copies)
  int t = 0;
  for (; t + 1 < nsteps; t += 2) {
    step(t, bA, bB);
    step(t + 1, bB, bA);
  }
  if (t < nsteps) step(t, bA, bB);
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
                             float *Lsm, hipStream_t s) {
  if constexpr (P > ARAN_MFMA_PMAX) {
    return hipErrorInvalidValue;
  } else {
    if (p == P) {
      const long wgs = (long)nwg * nsplit;
      const int nblocks = (int)((wgs + 7) / 8 * 8);
      if (nprod == 6)
        k_m2l_mfma<P, 6><<<nblocks, 64 * kMfmaWaves, 0, s>>>(
            wgt, nwg, tsrc, toct, codes, ops, mb, zero_node, Lsm, ntiles,
            nsplit);
      else
        k_m2l_mfma<P, 3><<<nblocks, 64 * kMfmaWaves, 0, s>>>(
            wgt, nwg, tsrc, toct, codes, ops, mb, zero_node, Lsm, ntiles,
            nsplit);
      return hipGetLastError();
    }
    return mfma_range<P + 1>(nprod, p, nwg, wgt, ntiles, tsrc, toct, codes,
                             ops, mb, zero_node, nsplit, Lsm, s);
  }
}

#define ARAN_MCAT2(a, b) a##b
#define ARAN_MCAT(a, b) ARAN_MCAT2(a, b)
// launch_m2l_mfma_<PMIN>: orders PMIN..PMAX of this object
hipError_t ARAN_MCAT(launch_m2l_mfma_, ARAN_MFMA_PMIN)(
    int nprod, int p, int nwg, const int *wgt, int ntiles, const int *tsrc,
    const int *toct, const short *codes, const void *ops, const void *mb,
    int zero_node, int nsplit, float *Lsm, hipStream_t s) {
  if (p < ARAN_MFMA_PMIN || p > ARAN_MFMA_PMAX) return hipErrorInvalidValue;
  return mfma_range<ARAN_MFMA_PMIN>(
      nprod, p, nwg, wgt, ntiles, tsrc, toct, codes,
      static_cast<const uint4 *>(ops), static_cast<const uint4 *>(mb),
      zero_node, nsplit, Lsm, s);
}

}  // namespace aran
