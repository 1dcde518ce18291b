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

#include <cstdlib>
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


// ------------------------------------------------------- M2L, one wave per SIMD
//
// The three-product form with 256 FP32 accumulators per wave (in AGPRs):
// a wave holds every row block of a step and W1Cfg::CB 32-column blocks
// (two 64-target tiles at p <= 10, one above), four waves per workgroup on
// the same tiles of wg_tiles as k_m2l_mfma.  Each accumulator receives the
// same sequence of MFMAs as in k_m2l_mfma, so the two kernels give the same
// bits.
//
// Vector-memory loads retire in order: a wait for one load also waits for
// every load issued before it, so a load with a short lead would cut the
// lead of all older ones.  The loop therefore issues only two kinds of
// global loads, both with the same lead: source fragments W1_NB - 1 steps
// ahead of their products (a register ring), and operator units
// W1Cfg::G + W1_L steps ahead (staging registers, stored into a two-group
// LDS ring W1_L steps after their load).  The source node indices of the
// workgroup's tiles over its offset range are copied into LDS once, as byte
// offsets of the packed multipoles, so no index load sits in the loop.
// Loads go through raw buffer descriptors with 32-bit offsets (no 64-bit
// address arithmetic); a missing source carries an offset past the end of
// the table, which the buffer unit answers with zeros.
constexpr int kW1Waves = 4;
constexpr int W1_NB = 4;  // source buffers (loads W1_NB - 1 steps ahead)
constexpr int W1_L = 2;   // steps between an operator unit's load and store
constexpr int kW1MinRanges = 5;
constexpr int kW1Rows = 40;  // index rows per tile: ceil(189 / 5) rounded up
static_assert((kOctOffsets + kW1MinRanges - 1) / kW1MinRanges < kW1Rows, "");
constexpr unsigned kW1Missing = 0x80000000u;  // byte offset past every table

typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

template <int P>
struct W1Cfg {
  static constexpr MfmaDims D = mfma_dims(P);
  static constexpr int KS = D.ks, RB = D.rb, RO = D.ro;
  static constexpr int CB = D.rb <= 4 ? 4 : 2;  // 32-column blocks per wave
  static constexpr int TPW = CB / 2;             // tiles per wave
  static constexpr int NPART = 2;
  static constexpr int CHUNK = RB * NPART * 64;  // 16-byte units of a step
  static constexpr int G = D.rb <= 4 ? 4 : 2;    // steps per operator group
  static constexpr int NT = 64 * kW1Waves;
  static constexpr int SPT = (CHUNK + NT - 1) / NT;  // units / thread / step
  static constexpr unsigned kNodeBytes = KS * 2 * NPART * 16;
  static_assert(W1_NB % G == 0 && W1_NB % W1_L == 0, "loop body");
  static_assert(kW1Waves * TPW == 8 / (2 / mfma_col_blocks(P)),
                "tiles per workgroup differ from the plan's");
};

template <int P>
__global__ __launch_bounds__(64 * kW1Waves) __attribute__((
    amdgpu_waves_per_eu(1, 1))) void k_m2l_mfma1w(
    const int *__restrict__ wg_tiles, int nwg, const int *__restrict__ tile_src,
    const int *__restrict__ tile_oct, const short *__restrict__ oct_codes,
    const uint4 *__restrict__ ops, const uint4 *__restrict__ mb,
    int mb_bytes, float *__restrict__ Lsm, int ntiles, int nsplit,
    int sched) {
  using C = W1Cfg<P>;
  constexpr int KS = C::KS, RB = C::RB, RO = C::RO, CB = C::CB, TPW = C::TPW;
  constexpr int NPART = C::NPART, CHUNK = C::CHUNK, G = C::G, NT = C::NT;
  constexpr int SPT = C::SPT, NB = W1_NB, DL = W1_NB - 1;
  __shared__ u32x4_t ring[2][G * CHUNK];
  __shared__ int idx_tab[kW1Waves][TPW][kW1Rows * 64];
  __shared__ short codes[kOctOffsets];
  // workgroup order over the XCDs as in k_m2l_mfma
  const int per_xcd = gridDim.x / 8;
  const int logical = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (logical >= nwg * nsplit) return;  // whole workgroup
  const int split = sched ? logical / nwg : logical % nsplit;
  const int wgi = sched ? logical % nwg : logical / nsplit;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const int tid = threadIdx.x, lane = tid % 64;
  const int g = lane / 32, c32 = lane % 32;
  const int *wt = wg_tiles + (size_t)wgi * (kW1Waves * TPW);
  int tl[TPW];
#pragma unroll
  for (int h = 0; h < TPW; h++)
    tl[h] = __builtin_amdgcn_readfirstlane(wt[TPW * wave + h]);
  const int oct = __builtin_amdgcn_readfirstlane(tile_oct[wt[0]]);
  const int chunk = (kOctOffsets + nsplit - 1) / nsplit;
  const int k0 = split * chunk, k1 = min(kOctOffsets, k0 + chunk);
  const int nrange = max(0, k1 - k0);
  __shared__ unsigned live_bits[2];
  if (tid < 2) live_bits[tid] = 0u;
  __syncthreads();
  // index tables of the wave's tiles: rows k0 .. k1 - 1 (64 sources each,
  // one 16-byte load per lane and row quarter) as byte offsets; a missing
  // source or tile holds kW1Missing.  Row j is live for the wave when one
  // of its tiles has a source at slot k0 + j.
  // the range's codes (one per thread, written compacted below), loaded
  // beside the tables
  const int my_code = tid < nrange ? oct_codes[oct * kOctOffsets + k0 + tid] : 0;
  unsigned long long wlive = 0;
  {
    // every load first (one round trip), then the stores
    constexpr int NPASS = kW1Rows * 16 / 64;
    int4 raw[NPASS][TPW];
#pragma unroll
    for (int i = 0; i < NPASS; i++)
#pragma unroll
      for (int h = 0; h < TPW; h++) {
        const int e = lane + 64 * i;
        raw[i][h] = (tl[h] >= 0 && e < nrange * 16)
                        ? ((const int4 *)(tile_src + ((size_t)tl[h] * kOctOffsets + k0) * 64))[e]
                        : make_int4(-1, -1, -1, -1);
      }
#pragma unroll
    for (int i = 0; i < NPASS; i++) {
      unsigned long long any = 0;
#pragma unroll
      for (int h = 0; h < TPW; h++) {
        const int4 r4 = raw[i][h];
        u32x4_t v;
        v[0] = r4.x >= 0 ? (unsigned)r4.x * C::kNodeBytes : kW1Missing;
        v[1] = r4.y >= 0 ? (unsigned)r4.y * C::kNodeBytes : kW1Missing;
        v[2] = r4.z >= 0 ? (unsigned)r4.z * C::kNodeBytes : kW1Missing;
        v[3] = r4.w >= 0 ? (unsigned)r4.w * C::kNodeBytes : kW1Missing;
        ((u32x4_t *)&idx_tab[wave][h][0])[lane + 64 * i] = v;
        any |= __ballot(r4.x >= 0 || r4.y >= 0 || r4.z >= 0 || r4.w >= 0);
      }
      // rows 4i .. 4i + 3 of this pass: 16 lanes each
#pragma unroll
      for (int r = 0; r < 4; r++)
        if ((any >> (16 * r)) & 0xffffull) wlive |= 1ull << (4 * i + r);
    }
  }
  if (lane == 0) {
    atomicOr(&live_bits[0], (unsigned)wlive);
    atomicOr(&live_bits[1], (unsigned)(wlive >> 32));
  }
  __syncthreads();  // index tables, liveness of every wave
  // the workgroup runs only the slots live for one of its waves: rows and
  // codes of those slots, compacted in slot order (a row moves down only)
  const unsigned long long live =
      ((unsigned long long)live_bits[1] << 32) | live_bits[0];
  const int nslots = __builtin_popcountll(live);
  const int nsteps = nslots * KS;
  {
    unsigned long long rest = live;
    int pos = 0;
    while (rest) {
      const int j = __builtin_ctzll(rest);
      rest &= rest - 1;
      if (j != pos) {
#pragma unroll
        for (int h = 0; h < TPW; h++)
          idx_tab[wave][h][pos * 64 + lane] = idx_tab[wave][h][j * 64 + lane];
      }
      pos++;
    }
    // row nslots: read by the steps past the range (zero sources)
#pragma unroll
    for (int h = 0; h < TPW; h++)
      idx_tab[wave][h][nslots * 64 + lane] = (int)kW1Missing;
    if (tid < nrange && ((live >> tid) & 1ull))
      codes[__builtin_popcountll(live & ((1ull << tid) - 1))] = (short)my_code;
  }
  __syncthreads();  // compacted codes
  const __amdgpu_buffer_rsrc_t ops_rs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)ops, (short)0, (int)((long)kNumCodes * KS * RB * 3 * 64 * 16),
      0x00020000);
  const __amdgpu_buffer_rsrc_t mb_rs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)mb, (short)0, mb_bytes, 0x00020000);
  // operator units: this thread's part of a unit's offset, and the step's
  // part (uniform, the instruction's scalar offset)
  unsigned uoff[SPT];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int u = min(tid + i * NT, CHUNK - 1);
    const int rb = u / (NPART * 64), part = (u / 64) % NPART, ln = u % 64;
    uoff[i] = (unsigned)(((rb * 3 + part) * 64 + ln) * 16);
  }
  auto stage_load = [&](int code, int s, u32x4_t (&r)[SPT])
      __attribute__((always_inline)) {
    const int cs = __builtin_amdgcn_readfirstlane((code * KS + s) *
                                                  (RB * 3 * 64 * 16));
#pragma unroll
    for (int i = 0; i < SPT; i++)
      r[i] = __builtin_amdgcn_raw_buffer_load_b128(ops_rs, (int)uoff[i], cs, 0);
  };
  auto stage_store = [&](int t, const u32x4_t (&r)[SPT])
      __attribute__((always_inline)) {
    u32x4_t *dst = ring[(t / G) % 2] + (t % G) * CHUNK;
#pragma unroll
    for (int i = 0; i < SPT; i++)
      if (CHUNK % NT == 0 || tid + i * NT < CHUNK) dst[tid + i * NT] = r[i];
  };
  // group 0 in LDS; steps G .. G + L - 1 staged in registers (steps past
  // the range reread the last step: valid units, never multiplied with a
  // source)
  u32x4_t stg[W1_L][SPT];
  int cbuf[2];
  if (nsteps > 0) {
    u32x4_t g0[G][SPT];
#pragma unroll
    for (int j = 0; j < G; j++) {
      const int tc = min(j, nsteps - 1);
      stage_load(codes[tc / KS], tc % KS, g0[j]);
    }
#pragma unroll
    for (int j = 0; j < W1_L; j++) {
      const int tc = min(G + j, nsteps - 1);
      stage_load(codes[tc / KS], tc % KS, stg[j]);
    }
#pragma unroll
    for (int j = 0; j < G; j++) stage_store(j, g0[j]);
    cbuf[0] = codes[min(G + W1_L, nsteps - 1) / KS];
  }
  f32x16_t acc[RB][CB];
#pragma unroll
  for (int rb = 0; rb < RB; rb++)
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[rb][cb][r] = 0.f;
  // byte offsets of the sources of index row `row` (read one step before
  // the loads that use them)
  auto read_idx = [&](int row, int (&iv)[CB]) __attribute__((always_inline)) {
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
      iv[cb] = idx_tab[wave][cb / 2][row * 64 + 32 * (cb % 2) + c32];
  };
  // B fragments of k step s of the sources iv
  const unsigned goff = (unsigned)(g * NPART * 16);
  auto load_src = [&](int s, const int (&iv)[CB], bf16x8_t (&dst)[CB][NPART])
      __attribute__((always_inline)) {
    const unsigned so = (unsigned)(s * 2 * NPART * 16) + goff;
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
      const unsigned ob = (unsigned)iv[cb] + so;
#pragma unroll
      for (int q = 0; q < NPART; q++)
        dst[cb][q] = __builtin_bit_cast(
            bf16x8_t, __builtin_amdgcn_raw_buffer_load_b128(
                          mb_rs, (int)(ob + 16 * q), 0, 0));
    }
  };
  bf16x8_t bq[NB][CB][NPART];
  int ib[2][CB];
  // sources of steps 0 .. DL - 1, in step order as the loop issues them
  // (the loop's waits count the loads issued after each step's)
#pragma unroll
  for (int d = 0; d < DL; d++) {
    read_idx(d < nsteps ? d / KS : nslots, ib[0]);
    load_src(d % KS, ib[0], bq[d]);
    __builtin_amdgcn_sched_barrier(0);
  }
  read_idx(DL < nsteps ? DL / KS : nslots, ib[0]);
  __syncthreads();  // operator group 0
  auto read_a = [&](const u32x4_t *ra, int rb, bf16x8_t (&a)[NPART])
      __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < NPART; q++)
      a[q] = __builtin_bit_cast(bf16x8_t, ra[(rb * NPART + q) * 64 + lane]);
  };
  // streams as (row, k step): sources of step t + DL, operator units of
  // step t + G + L
  int rB = DL / KS, sB = DL % KS;
  int rO = (G + W1_L) / KS, sO = (G + W1_L) % KS;
  const int nbody = (nsteps + NB - 1) / NB;
  for (int it = 0; it < nbody; it++) {
    mfma_static_for<0, NB>([&](auto gc) {
      constexpr int gg = decltype(gc)::value;
      const int t = it * NB + gg;
      bf16x8_t (&cur)[CB][NPART] = bq[gg % NB];
      bf16x8_t (&nxt)[CB][NPART] = bq[(gg + DL) % NB];
      // operator units of step t + G (loaded L steps ago) into the ring
      // slot of the next group, units of step t + G + L into the same
      // registers, the code of step t + G + L + 1
      stage_store(t + G, stg[gg % W1_L]);
      stage_load(cbuf[gg % 2], sO, stg[gg % W1_L]);
      {
        int s1 = sO + 1, r1 = rO;
        if (s1 == KS) {
          s1 = 0;
          r1++;
        }
        rO = r1;
        sO = s1;
        cbuf[(gg + 1) % 2] = codes[min(r1, nslots - 1)];
      }
      // sources of step t + DL, indices of step t + DL + 1
      {
        load_src(sB, ib[gg % 2], nxt);
        int s1 = sB + 1, r1 = rB;
        if (s1 == KS) {
          s1 = 0;
          r1++;
        }
        rB = r1;
        sB = s1;
        read_idx(min(r1, nslots), ib[(gg + 1) % 2]);
      }
      // products of step t: row block by row block, each block's operator
      // fragments read one block ahead; per accumulator the products in
      // the order (m,h) (h,m) (h,h)
      const u32x4_t *ra = ring[(t / G) % 2] + (t % G) * CHUNK;
      bf16x8_t a[2][NPART];
      read_a(ra, 0, a[0]);
      mfma_static_for<0, RB>([&](auto rc) {
        constexpr int rb = decltype(rc)::value;
        if constexpr (rb + 1 < RB) read_a(ra, rb + 1, a[(rb + 1) % 2]);
        constexpr int QA[3] = {1, 0, 0}, QB[3] = {0, 1, 0};
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
          for (int cb = 0; cb < CB; cb++)
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                a[rb % 2][QA[q]], cur[cb][QB[q]], acc[rb][cb], 0, 0, 0);
      });
      // issue order: index reads and the first two row blocks' fragments,
      // source loads between the first products, each later row block's
      // fragments one block ahead of its products
      __builtin_amdgcn_sched_group_barrier(0x100, 2 + 2 * NPART, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
      __builtin_amdgcn_sched_group_barrier(0x020, CB * NPART / 2, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
      __builtin_amdgcn_sched_group_barrier(0x020, CB * NPART / 2, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 3 * CB - 8, 0);
      mfma_static_for<1, RB>([&](auto rc) {
        constexpr int rb = decltype(rc)::value;
        if constexpr (rb + 1 < RB)
          __builtin_amdgcn_sched_group_barrier(0x100, NPART, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 3 * CB, 0);
      });
      if constexpr (gg % G == G - 1) __syncthreads();
      // each step is scheduled on its own (loads stay in their step)
      __builtin_amdgcn_sched_barrier(0);
    });
  }
  // steps past the range multiplied zero sources: every accumulator holds
  // the sum over the range
#pragma unroll
  for (int h = 0; h < TPW; h++) {
    if (tl[h] < 0) continue;
    float *out = Lsm + ((size_t)split * ntiles + tl[h]) * RO * 64;
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
      for (int cc = 0; cc < 2; cc++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
          const int row = 32 * rb + 8 * (r / 4) + 4 * g + (r % 4);
          if (row < RO)
            out[(size_t)row * 64 + 32 * cc + c32] = acc[rb][2 * h + cc][r];
        }
  }
}

#ifndef ARAN_MFMA_PMIN
#define ARAN_MFMA_PMIN 1
#define ARAN_MFMA_PMAX 15
#endif

// ARAN_MFMA_KERNEL=8w selects k_m2l_mfma for the three-product form (the
// same sums; for comparisons)
static bool mfma_one_wave() {
  const char *e = getenv("ARAN_MFMA_KERNEL");
  return !(e && e[0] == '8');
}

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
      // the one-wave kernel where its index tables and 32-bit offsets fit
      const long mb_bytes = ((long)zero_node + 1) * W1Cfg<P>::kNodeBytes;
      if (nprod == 3 && nsplit >= kW1MinRanges && mb_bytes < (long)kW1Missing &&
          mfma_one_wave()) {
        k_m2l_mfma1w<P><<<nblocks, 64 * kW1Waves, 0, s>>>(
            wgt, nwg, tsrc, toct, codes, ops, mb, (int)mb_bytes, Lsm, ntiles,
            nsplit, sched);
        return hipGetLastError();
      }
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
