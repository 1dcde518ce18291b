// tree.hip -- Morton keys, LSD radix sort, level-synchronous octree build
// and the device-side interaction tables (M2L and P2P lists).
//
// Replaces the particle filtering of LibVsg's VsgPRTree3d
// (aran/aransolver3d.c:972-978) and its near/far traversal
// (aran/aransolver3d.c:1189-1190) for the FMM hot path.  The tree rules are
// spelled out in tree_rules.h and are shared with the FP64 oracle.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <climits>

#include "engine.h"
#include "tree_rules.h"

namespace aran {

// ------------------------------------------------------------ keys

// System of particle i (systems are contiguous ranges soff[s]..soff[s+1]).
__device__ inline int system_of(const long *__restrict__ soff, int nsys,
                                long i) {
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {  // last s with soff[s] <= i
    int mid = (lo + hi + 1) >> 1;
    if (soff[mid] <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Morton keys and identity indices; also packs (x, y, z, q) per particle
// so that the gather after the sort reads one 32-byte record per particle
// instead of four scattered words.
__global__ void k_keys(const double *__restrict__ x,
                       const double *__restrict__ y,
                       const double *__restrict__ z,
                       const double *__restrict__ q, long n,
                       const long *__restrict__ soff, int nsys,
                       const double *__restrict__ box,
                       uint64_t *__restrict__ keys, int *__restrict__ idx,
                       double4 *__restrict__ packed) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sy = system_of(soff, nsys, i);
  const double *b = box + 4 * sy;
  const double scale = 2097152.0 / b[3];
  const double xi = x[i], yi = y[i], zi = z[i];
  keys[i] = morton_key(icoord(xi, b[0], scale), icoord(yi, b[1], scale),
                       icoord(zi, b[2], scale));
  idx[i] = (int)i;
  packed[i] = make_double4(xi, yi, zi, q[i]);
}

hipError_t launch_keys(const double *x, const double *y, const double *z,
                       const double *q, long n, const long *soff, int nsys,
                       const double *box, uint64_t *keys, int *idx,
                       double4 *packed, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_keys<<<(n + 255) / 256, 256, 0, s>>>(x, y, z, q, n, soff, nsys, box, keys,
                                         idx, packed);
  return hipGetLastError();
}

// Second sort key: the system of each (already Morton-sorted) particle.
__global__ void k_sys_keys(const int *__restrict__ idx, long n,
                           const long *__restrict__ soff, int nsys,
                           uint64_t *__restrict__ skey, int *__restrict__ pos) {
  long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (j >= n) return;
  skey[j] = (uint64_t)system_of(soff, nsys, idx[j]);
  pos[j] = (int)j;
}

__global__ void k_permute_keys(const int *__restrict__ pos, long n,
                               const uint64_t *__restrict__ kin,
                               const int *__restrict__ iin,
                               uint64_t *__restrict__ kout,
                               int *__restrict__ iout) {
  long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (j >= n) return;
  kout[j] = kin[pos[j]];
  iout[j] = iin[pos[j]];
}

hipError_t launch_sys_keys(const int *idx, long n, const long *soff, int nsys,
                           uint64_t *skey, int *pos, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_sys_keys<<<(n + 255) / 256, 256, 0, s>>>(idx, n, soff, nsys, skey, pos);
  return hipGetLastError();
}

hipError_t launch_permute_keys(const int *pos, long n, const uint64_t *kin,
                               const int *iin, uint64_t *kout, int *iout,
                               hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_permute_keys<<<(n + 255) / 256, 256, 0, s>>>(pos, n, kin, iin, kout, iout);
  return hipGetLastError();
}

// Per-system extent: grid (systems, blocks per system); partials are then
// combined per system.
__global__ void k_sys_minmax(const double *__restrict__ x,
                             const double *__restrict__ y,
                             const double *__restrict__ z,
                             const long *__restrict__ soff,
                             double *__restrict__ part, int *__restrict__ err) {
  __shared__ double red[6][256];
  const int sy = blockIdx.x, nb = gridDim.y;
  if (sy == 0 && blockIdx.y == 0 && threadIdx.x == 0) *err = 0;
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY,
                 -INFINITY};
  bool bad = false;
  for (long i = soff[sy] + blockIdx.y * 256L + threadIdx.x; i < soff[sy + 1];
       i += 256L * nb) {
    double a = x[i], b = y[i], c = z[i];
    bad |= !(isfinite(a) && isfinite(b) && isfinite(c));
    v[0] = fmin(v[0], a);
    v[1] = fmin(v[1], b);
    v[2] = fmin(v[2], c);
    v[3] = fmax(v[3], a);
    v[4] = fmax(v[4], b);
    v[5] = fmax(v[5], c);
  }
  if (bad) v[0] = NAN;
  for (int k = 0; k < 6; k++) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      for (int k = 0; k < 3; k++) {
        double a = red[k][threadIdx.x], b = red[k][threadIdx.x + st];
        red[k][threadIdx.x] = (isnan(a) || isnan(b)) ? NAN : fmin(a, b);
      }
      for (int k = 3; k < 6; k++)
        red[k][threadIdx.x] = fmax(red[k][threadIdx.x], red[k][threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 6)
    part[((size_t)sy * nb + blockIdx.y) * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// One block per system: its nb partial extents combined (a NaN minimum
// marks a non-finite coordinate and survives), then the system's root cube
// from its bounds (lb, ub at bounds[6 sy ..]): box[4 sy ..] = (lo, width),
// and its root node sy, covering the system's particle range.  A system
// with a non-finite coordinate sets bit 4 of *err.  Block 0 also sets up
// the level table: level 0 holds the nsys roots.
__global__ void k_sys_minmax_final(const double *__restrict__ part, int nsys,
                                   int nb, const long *__restrict__ soff,
                                   const double *__restrict__ bounds,
                                   double *__restrict__ box,
                                   int *__restrict__ dl, int ntab,
                                   int *__restrict__ err, NodeArrays nd) {
  __shared__ double red[6][256];
  const int sy = blockIdx.x;
  if (sy == 0)
    for (int k = threadIdx.x; k < ntab; k += 256)
      dl[k] = k == 1 ? nsys : 0;
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY,
                 -INFINITY};
  bool bad = false;
  for (int b = threadIdx.x; b < nb; b += 256) {
    const double *p = part + ((size_t)sy * nb + b) * 6;
    for (int k = 0; k < 3; k++) {
      bad |= isnan(p[k]);
      v[k] = fmin(v[k], p[k]);
    }
    for (int k = 3; k < 6; k++) v[k] = fmax(v[k], p[k]);
  }
  if (bad) v[0] = NAN;
  for (int k = 0; k < 6; k++) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      for (int k = 0; k < 3; k++) {
        double a = red[k][threadIdx.x], b = red[k][threadIdx.x + st];
        red[k][threadIdx.x] = (isnan(a) || isnan(b)) ? NAN : fmin(a, b);
      }
      for (int k = 3; k < 6; k++)
        red[k][threadIdx.x] = fmax(red[k][threadIdx.x], red[k][threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double *lb = bounds + 6 * sy, *ub = lb + 3;
  double pmin[3], pmax[3];
  bool finite = true;
  for (int k = 0; k < 3; k++) {
    pmin[k] = red[k][0];
    pmax[k] = red[3 + k][0];
    finite = finite && isfinite(pmin[k]) && isfinite(pmax[k]);
  }
  double lo[3] = {0., 0., 0.}, width = 1.;
  if (soff[sy + 1] == soff[sy]) {  // empty: the cube of its bounds
    for (int k = 0; k < 3; k++) pmin[k] = pmax[k] = lb[k];
    root_cube(lb, ub, pmin, pmax, lo, &width);
  } else if (!finite) {  // (the solve fails at its first level read)
    atomicOr(err, 4);
  } else {
    root_cube(lb, ub, pmin, pmax, lo, &width);
  }
  for (int k = 0; k < 3; k++) box[4 * sy + k] = lo[k];
  box[4 * sy + 3] = width;
  nd.depth[sy] = 0;
  nd.ix[sy] = nd.iy[sy] = nd.iz[sy] = 0;
  nd.key[sy] = 0;
  nd.sys[sy] = sy;
  nd.anc[(size_t)sy * kAncRow] = sy;
  nd.pbegin[sy] = (int)soff[sy];
  nd.pend[sy] = (int)soff[sy + 1];
  nd.parent[sy] = -1;
  nd.child[sy] = -1;
  nd.nchild[sy] = 0;
  nd.leaf[sy] = 1;
  nd.cx[sy] = fma(0.5, width, lo[0]);
  nd.cy[sy] = fma(0.5, width, lo[1]);
  nd.cz[sy] = fma(0.5, width, lo[2]);
}

// part must hold 6 * nsys * minmax_blocks(nsys) doubles.
int minmax_blocks(int nsys) {
  return std::max(1, std::min(1024, 8192 / std::max(nsys, 1)));
}

hipError_t launch_sys_minmax(const double *x, const double *y, const double *z,
                             const long *soff, const double *bounds, int nsys,
                             double *part, double *box, int *dl, int ntab,
                             int *err, const NodeArrays &nd, hipStream_t s) {
  if (nsys == 0) return hipSuccess;
  const int nb = minmax_blocks(nsys);
  k_sys_minmax<<<dim3(nsys, nb), 256, 0, s>>>(x, y, z, soff, part, err);
  k_sys_minmax_final<<<nsys, 256, 0, s>>>(part, nsys, nb, soff, bounds, box,
                                          dl, ntab, err, nd);
  return hipGetLastError();
}

// ------------------------------------------------------------ look-back
//
// A launch whose workgroups each own one chunk of a sequence and need the
// sum of the totals of the chunks before theirs (a scan across workgroups)
// gets it in the same launch: a workgroup publishes its chunk's total in the
// chunk's status word, then reads the words of the chunks below, 64 at a
// time, down to the nearest one that already holds its inclusive total
// (decoupled look-back).  Workgroup b owns chunk b: workgroups are
// dispatched in index order, so the chunks a resident workgroup waits for
// are held by workgroups that were dispatched before it (a ticket taken
// with an atomic would not need that order, but same-address atomics
// serialize at about 11 ns each, 25 us per 1000 workgroups).  Every wait is
// bounded: past kLbMaxPolls polls it sets *err and goes on with a wrong
// prefix, which the host reports as an error.
//
// Status word: bits 63..34 the launch's epoch, 33..32 the state (1: the
// chunk's own total, 2: the total of chunks 0..c), 31..0 the value; a word
// of another launch (another epoch, or the zero of the clear at allocation)
// reads as not yet published, so the words need no clearing between
// launches.  Every access to them is a single 64-bit atomic at agent scope.

constexpr unsigned kLbAgg = 1u, kLbIncl = 2u;
constexpr int kLbMaxPolls = 1 << 20;

__device__ inline unsigned long long lb_word(unsigned epoch, unsigned st,
                                             unsigned v) {
  return ((unsigned long long)((epoch << 2) | st) << 32) | v;
}

// Exclusive prefix of chunk c whose own total is 'total'; called by one
// whole wave of the workgroup, every lane returns the prefix.
__device__ inline unsigned lb_prefix_wave(const Lookback &lb, int c,
                                          unsigned total) {
  const int lane = threadIdx.x & 63;
  if (lane == 0)
    __hip_atomic_store(&lb.status[c],
                       lb_word(lb.epoch, c == 0 ? kLbIncl : kLbAgg, total),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned pre = 0;
  if (c == 0) return 0;
  const unsigned tag_agg = (lb.epoch << 2) | kLbAgg;
  const unsigned tag_incl = (lb.epoch << 2) | kLbIncl;
  int top = c - 1, polls = 0;
  while (true) {
    const int k = top - lane;  // lane 0 reads the nearest chunk
    unsigned long long w = 0;
    if (k >= 0)
      w = __hip_atomic_load(&lb.status[k], __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
    const unsigned tag = (unsigned)(w >> 32);
    const bool incl = k < 0 || tag == tag_incl;  // below chunk 0: zero
    const bool ready = incl || tag == tag_agg;
    const unsigned long long im = __ballot(incl);
    const int first = im ? __builtin_ctzll(im) : 64;
    // the lanes up to the first inclusive word must all be published
    const unsigned long long need = first >= 63 ? ~0ull : (2ull << first) - 1;
    if (__ballot(!ready) & need) {
      if (++polls > kLbMaxPolls) {
        if (lane == 0) atomicOr(lb.err, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    unsigned v = (lane <= first && k >= 0) ? (unsigned)w : 0u;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    pre += v;
    if (first < 64) break;
    top -= 64;
  }
  if (lane == 0)
    __hip_atomic_store(&lb.status[c], lb_word(lb.epoch, kLbIncl, pre + total),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return pre;
}

// ------------------------------------------------------------ scan

// One-workgroup exclusive scan (in and out may alias); writes the grand
// total to *total.  Tiles of kScan1Max entries in turn: thread t scans
// entries [kScan1Per t, kScan1Per (t + 1)) of the tile held in registers
// (all of its loads in flight at once), so a tile costs two memory round
// trips; the chunk is short because a single CU serves every access.
constexpr int kScan1Threads = 1024;
constexpr int kScan1Per = 16;
constexpr long kScan1Max = (long)kScan1Threads * kScan1Per;

__global__ __launch_bounds__(kScan1Threads) void k_scan_single(
    const int *in, int *out, int n, int *__restrict__ total) {
  __shared__ int ws[kScan1Threads / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int carry = 0;
  for (int base = 0; base < n; base += (int)kScan1Max) {
    const int b = base + t * kScan1Per;
    const bool full = b + kScan1Per <= n;  // 16-byte loads and stores
    int v[kScan1Per];
    int sum = 0;
    if (full) {
#pragma unroll
      for (int k = 0; k < kScan1Per; k += 4) {
        const int4 q = *reinterpret_cast<const int4 *>(in + b + k);
        v[k] = q.x;
        v[k + 1] = q.y;
        v[k + 2] = q.z;
        v[k + 3] = q.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kScan1Per; k++) v[k] = b + k < n ? in[b + k] : 0;
    }
#pragma unroll
    for (int k = 0; k < kScan1Per; k++) sum += v[k];
    int x = sum;  // inclusive scan over the wave
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    if (w == 0) {
      int u = lane < kScan1Threads / 64 ? ws[lane] : 0;
      for (int off = 1; off < kScan1Threads / 64; off <<= 1) {
        const int y = __shfl_up(u, off);
        if (lane >= off) u += y;
      }
      if (lane < kScan1Threads / 64) ws[lane] = u;
    }
    __syncthreads();
    int run = carry + x - sum + (w > 0 ? ws[w - 1] : 0);
    if (full) {
#pragma unroll
      for (int k = 0; k < kScan1Per; k += 4) {
        int4 q;
        q.x = run;
        q.y = q.x + v[k];
        q.z = q.y + v[k + 1];
        q.w = q.z + v[k + 2];
        run = q.w + v[k + 3];
        *reinterpret_cast<int4 *>(out + b + k) = q;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kScan1Per; k++) {
        if (b + k < n) out[b + k] = run;
        run += v[k];
      }
    }
    carry += ws[kScan1Threads / 64 - 1];
    __syncthreads();
  }
  if (t == 0) *total = carry;
}

// Exclusive scan in one launch of one workgroup (n <= kScan1Max).
hipError_t scan_small(const int *in, int *out, long n, int *total,
                      hipStream_t s) {
  if (n > kScan1Max) return hipErrorInvalidValue;
  k_scan_single<<<1, kScan1Threads, 0, s>>>(in, out, (int)n, total);
  return hipGetLastError();
}

bool scan_fits_single(long n) { return n <= kScan1Max; }

// ------------------------------------------------------------ look-back scan

// Exclusive scan of n ints in one launch: workgroup c owns entries
// [kLsTile c, kLsTile (c + 1)) (16 per thread, read with 16-byte loads when
// whole), scans them and takes the total of the entries before its tile by
// look-back.  Each thread reads its entries before it writes them, so in
// and out may alias.  The workgroup of the last tile writes the total.
constexpr int kLsThreads = 256;
constexpr int kLsPer = 16;
constexpr long kLsTile = (long)kLsThreads * kLsPer;

__global__ __launch_bounds__(kLsThreads) void k_scan_lookback(
    const int *in, int *out, long n, int *__restrict__ total, Lookback lb,
    int nwork) {
  __shared__ int ws[kLsThreads / 64];
  __shared__ unsigned sh_pre;
  const int c = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const long b = c * kLsTile + (long)t * kLsPer;
  const bool full = b + kLsPer <= n;
  int v[kLsPer];
  if (full) {
#pragma unroll
    for (int k = 0; k < kLsPer; k += 4) {
      const int4 q = *reinterpret_cast<const int4 *>(in + b + k);
      v[k] = q.x;
      v[k + 1] = q.y;
      v[k + 2] = q.z;
      v[k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kLsPer; k++) v[k] = b + k < n ? in[b + k] : 0;
  }
  int sum = 0;
#pragma unroll
  for (int k = 0; k < kLsPer; k++) sum += v[k];
  int x = sum;  // inclusive scan over the wave
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) ws[w] = x;
  __syncthreads();
  int wpre = 0, tot = 0;
  for (int k = 0; k < kLsThreads / 64; k++) {
    if (k < w) wpre += ws[k];
    tot += ws[k];
  }
  if (w == 0) {
    const unsigned p = lb_prefix_wave(lb, c, (unsigned)tot);
    if (lane == 0) sh_pre = p;
  }
  __syncthreads();
  int run = (int)sh_pre + wpre + x - sum;
  if (full) {
#pragma unroll
    for (int k = 0; k < kLsPer; k += 4) {
      int4 q;
      q.x = run;
      q.y = q.x + v[k];
      q.z = q.y + v[k + 1];
      q.w = q.z + v[k + 2];
      run = q.w + v[k + 3];
      *reinterpret_cast<int4 *>(out + b + k) = q;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kLsPer; k++) {
      if (b + k < n) out[b + k] = run;
      run += v[k];
    }
  }
  if (c == nwork - 1 && t == 0) *total = (int)sh_pre + tot;
}

long scan_lookback_chunks(long n) {
  return std::max(1L, (n + kLsTile - 1) / kLsTile);
}

hipError_t scan_lookback(const int *in, int *out, long n, int *total,
                         const Lookback &lb, hipStream_t s) {
  const long nb = scan_lookback_chunks(n);
  k_scan_lookback<<<nb, kLsThreads, 0, s>>>(in, out, n, total, lb, (int)nb);
  return hipGetLastError();
}

// ------------------------------------------------------------ radix sort

constexpr int kRsThreads = 256;
constexpr int kRsWaves = kRsThreads / 64;
constexpr int kRsItems = 8;
constexpr int kRsTile = kRsThreads * kRsItems;

__global__ __launch_bounds__(kRsThreads) void k_rs_hist(
    const uint64_t *__restrict__ keys, long n, int shift,
    int *__restrict__ hist, int nblocks) {
  __shared__ int cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  long base = blockIdx.x * (long)kRsTile;
  for (int k = 0; k < kRsItems; k++) {
    long i = base + k * kRsThreads + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255], 1);
  }
  __syncthreads();
  hist[threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Block-wide exclusive scan of one int per thread (kRsThreads threads);
// returns the thread's prefix, *tot the block total.
__device__ inline int rs_block_scan(int v, int *wsum, int *tot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int pre = 0, all = 0;
  for (int k = 0; k < kRsWaves; k++) {
    if (k < w) pre += wsum[k];
    all += wsum[k];
  }
  __syncthreads();
  *tot = all;
  return x - v + pre;
}

// One block per digit: exclusive scan of the digit's per-block counts
// (hist[d][b] -> offs[d][b]) and the digit total dtot[d].
__global__ __launch_bounds__(kRsThreads) void k_rs_scan_digits(
    const int *__restrict__ hist, int *__restrict__ offs, int nblocks,
    int *__restrict__ dtot) {
  __shared__ int wsum[kRsWaves];
  const int d = blockIdx.x;
  int carry = 0;
  for (int b0 = 0; b0 < nblocks; b0 += kRsThreads) {
    const int b = b0 + threadIdx.x;
    const int v = b < nblocks ? hist[(size_t)d * nblocks + b] : 0;
    int tot;
    const int pre = rs_block_scan(v, wsum, &tot);
    if (b < nblocks) offs[(size_t)d * nblocks + b] = carry + pre;
    carry += tot;
  }
  if (threadIdx.x == 0) dtot[d] = carry;
}

// Stable scatter: elements keep their input order inside each digit.  The
// block's base of digit d is the total of the digits below d plus its
// offset among the blocks.  Wave w of the block owns the w-th quarter of the
// block's tile (its items in rounds of 64): it ranks them with its own
// per-digit counters (wave-synchronous, no block barrier); one pass over
// (digit, wave) turns the counters into block-local run starts, the items
// are placed in LDS in block-local sorted order, and the block then writes
// each digit's run to consecutive addresses (a direct scatter would write
// 64 different buckets per wave store).  Every position stays below the
// digit's base plus the block's count of it, the count the histogram kernel
// took over the same items.
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(
    const uint64_t *__restrict__ kin, const int *__restrict__ vin,
    uint64_t *__restrict__ kout, int *__restrict__ vout, long n, int shift,
    const int *__restrict__ offs, const int *__restrict__ dtot,
    int nblocks) {
  __shared__ int base[256], lstart[256];
  __shared__ int wcnt[kRsWaves][256];
  __shared__ int wsum[kRsWaves];
  __shared__ uint64_t lkey[kRsTile];
  __shared__ int lval[kRsTile];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  {
    int tot;
    const int below = rs_block_scan(dtot[threadIdx.x], wsum, &tot);
    base[threadIdx.x] = below + offs[threadIdx.x * nblocks + blockIdx.x];
  }
  for (int w = 0; w < kRsWaves; w++) wcnt[w][threadIdx.x] = 0;
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const long first =
      blockIdx.x * (long)kRsTile + (long)wave * (kRsItems * 64) + lane;
  uint64_t key[kRsItems];
  int val[kRsItems], loc[kRsItems];
#pragma unroll
  for (int k = 0; k < kRsItems; k++) {
    const long i = first + 64L * k;
    key[k] = i < n ? kin[i] : 0ull;
    val[k] = i < n ? vin[i] : 0;
  }
#pragma unroll
  for (int k = 0; k < kRsItems; k++) {
    const bool valid = first + 64L * k < n;
    const int d = (int)((key[k] >> shift) & 255);
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < 8; b++) {
      const uint64_t m = __ballot(valid && ((d >> b) & 1));
      peers &= ((d >> b) & 1) ? m : ~m;
    }
    const int rank = __popcll(peers & lt);
    const int old = valid ? wcnt[wave][d] : 0;
    // every lane has read its digit's count before the lowest lane of the
    // digit advances it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (valid && rank == 0) wcnt[wave][d] = old + __popcll(peers);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    loc[k] = old + rank;
  }
  __syncthreads();
  {  // block-local start of digit d, and of each wave's run inside it
    const int d = threadIdx.x;
    int c = 0;
    for (int w = 0; w < kRsWaves; w++) c += wcnt[w][d];
    int tot;
    int run = rs_block_scan(c, wsum, &tot);
    lstart[d] = run;
    for (int w = 0; w < kRsWaves; w++) {
      const int cw = wcnt[w][d];
      wcnt[w][d] = run;
      run += cw;
    }
  }
  __syncthreads();
  // the block's items in block-local sorted order, in LDS
#pragma unroll
  for (int k = 0; k < kRsItems; k++) {
    if (first + 64L * k < n) {
      const int d = (int)((key[k] >> shift) & 255);
      const int lp = wcnt[wave][d] + loc[k];
      lkey[lp] = key[k];
      lval[lp] = val[k];
    }
  }
  __syncthreads();
  // runs of one digit go to consecutive addresses: consecutive threads
  // write consecutive words
  const long tile = blockIdx.x * (long)kRsTile;
  const int nvalid = (int)min((long)kRsTile, n - tile);
  for (int t = threadIdx.x; t < nvalid; t += kRsThreads) {
    const uint64_t kk = lkey[t];
    const int d = (int)((kk >> shift) & 255);
    const int pos = base[d] + (t - lstart[d]);
    kout[pos] = kk;
    vout[pos] = lval[t];
  }
}

// Sorts (keys, idx) by the low 8*passes bits of key; ties keep the input
// order.  passes must be even: the result then ends in keys / idx.
hipError_t radix_sort(uint64_t *keys, int *idx, uint64_t *keys_tmp,
                      int *idx_tmp, long n, DevBuf &hist, int passes,
                      hipStream_t s) {
  if (n <= 1) return hipSuccess;
  int nb = (int)((n + kRsTile - 1) / kRsTile);
  hipError_t e = hist.reserve(sizeof(int) * (256 * (size_t)nb * 2 + 256));
  if (e != hipSuccess) return e;
  int *h = hist.as<int>();
  int *o = h + 256 * (size_t)nb;
  int *dt = o + 256 * (size_t)nb;
  uint64_t *ka = keys, *kb = keys_tmp;
  int *va = idx, *vb = idx_tmp;
  for (int pass = 0; pass < passes; pass++) {
    int shift = 8 * pass;
    k_rs_hist<<<nb, kRsThreads, 0, s>>>(ka, n, shift, h, nb);
    k_rs_scan_digits<<<256, kRsThreads, 0, s>>>(h, o, nb, dt);
    k_rs_scatter<<<nb, kRsThreads, 0, s>>>(ka, va, kb, vb, n, shift, o, dt,
                                           nb);
    uint64_t *tk = ka;
    ka = kb;
    kb = tk;
    int *tv = va;
    va = vb;
    vb = tv;
  }
  return hipGetLastError();
}

// ------------------------------------------------------------ gather

__global__ void k_gather(const int *__restrict__ idx, long n,
                         const double4 *__restrict__ packed,
                         double *__restrict__ xs, double *__restrict__ ys,
                         double *__restrict__ zs, double *__restrict__ qs) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 v = packed[idx[i]];
  xs[i] = v.x;
  ys[i] = v.y;
  zs[i] = v.z;
  qs[i] = v.w;
}

hipError_t launch_gather(const int *idx, long n, const double4 *packed,
                         double *xs, double *ys, double *zs, double *qs,
                         hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_gather<<<(n + 255) / 256, 256, 0, s>>>(idx, n, packed, xs, ys, zs, qs);
  return hipGetLastError();
}

// ------------------------------------------------------------ level build

// First index in [b, e) whose octant digit at 'shift' is >= o.
__device__ inline int digit_lower_bound(const uint64_t *__restrict__ keys,
                                        int b, int e, int shift, int o) {
  while (b < e) {
    int mid = b + ((e - b) >> 1);
    if ((int)((keys[mid] >> shift) & 7) < o)
      b = mid + 1;
    else
      e = mid;
  }
  return b;
}

// Level table dl[2 d] = first node of level d, dl[2 d + 1] = its node count
// (device), so that consecutive levels launch without a host round trip;
// grids are sized by host upper bounds and threads beyond the count exit.

// Eight lanes per node of level d: lane o finds where octant digit o starts
// inside the node's sorted key range (lane 0: the range start); the node
// splits when it holds more than leafcap particles below the maximum depth,
// and cnt counts its non-empty octants.  Also sets the next level's first
// node.
// Whether (depth, key) is in the sorted stop list.
__device__ inline bool stopped(const ulonglong2 *__restrict__ stop, int nstop,
                               unsigned long long dep, uint64_t key) {
  int lo = 0, hi = nstop;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const ulonglong2 v = stop[mid];
    if (v.x < dep || (v.x == dep && v.y < key))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < nstop && stop[lo].x == dep && stop[lo].y == key;
}

// One launch per level: eight lanes per node of level d (32 nodes per
// workgroup); lane o finds where octant digit o starts inside the node's
// sorted key range (lane 0: the range start).  The node splits when it
// holds more than leafcap particles below the maximum depth and is not in
// the stop list; each non-empty octant of a split node becomes a child.
// Children are numbered in (node, octant) order from the next level's first
// node: the workgroup ranks its own children with ballots and takes the
// count of the children of the nodes before its chunk by look-back.  The
// chunk holding the level's last node writes the next level's first node
// and count into the level table.
constexpr int kLvNodes = 32;

__global__ __launch_bounds__(256) void k_split_level(
    NodeArrays nd, int *__restrict__ dl, int d,
    const uint64_t *__restrict__ keys, int leafcap,
    const ulonglong2 *__restrict__ stop, int nstop, Lookback lb) {
  __shared__ unsigned sh_wcnt[4], sh_pre;
  const int lvl_begin = dl[2 * d], lvl_count = dl[2 * d + 1];
  // a level without nodes still has one chunk (it writes the zero count)
  const int nwork = max(1, (lvl_count + kLvNodes - 1) / kLvNodes);
  if ((int)gridDim.x < nwork) {  // the bound was wrong: report, touch nothing
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(lb.err, 2);
    return;
  }
  if ((int)blockIdx.x >= nwork) return;  // grid sized by an upper bound
  const int c = blockIdx.x;
  const int next_begin = lvl_begin + lvl_count;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = c * kLvNodes + (threadIdx.x >> 3), o = threadIdx.x & 7;
  const bool valid = i < lvl_count;
  const int node = lvl_begin + i;
  int b = 0, e = 0, dep = 0;
  bool split = false;
  if (valid) {
    b = nd.pbegin[node];
    e = nd.pend[node];
    dep = nd.depth[node];
    split = e - b > leafcap && dep < kMaxDepth &&
            !(nstop > 0 && stopped(stop, nstop, dep, nd.key[node]));
  }
  int x = b;
  if (split && o > 0)
    x = digit_lower_bound(keys, b, e, 3 * (kMaxDepth - dep - 1), o);
  // next boundary: the lane above in the group of 8, the range end for o = 7
  int nx = __shfl_down(x, 1, 8);
  if (o == 7) nx = e;
  const bool ex = split && nx > x;
  const unsigned long long m = __ballot(ex);
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  if (lane == 0) sh_wcnt[wave] = (unsigned)__popcll(m);
  __syncthreads();
  unsigned wpre = 0, tot = 0;
  for (int w = 0; w < 4; w++) {
    if (w < wave) wpre += sh_wcnt[w];
    tot += sh_wcnt[w];
  }
  if (wave == 0) {
    const unsigned p = lb_prefix_wave(lb, c, tot);
    if (lane == 0) sh_pre = p;
  }
  __syncthreads();
  const unsigned pre = sh_pre;
  // children of this block before this lane (octant o of node i)
  const int rank = (int)(pre + wpre) + __popcll(m & lt);
  const int nch = __popcll((m >> (lane & ~7)) & 0xffull);
  if (valid && o == 0) {
    nd.leaf[node] = nch == 0 ? 1 : 0;
    nd.child[node] = nch == 0 ? -1 : next_begin + rank;
    nd.nchild[node] = nch;
  }
  if (valid) nd.cmap[(size_t)node * 8 + o] = ex ? next_begin + rank : -1;
  if (ex) {
    const int ch = next_begin + rank;
    const int dd = dep + 1;
    const int sy = nd.sys[node];
    const double *bx = nd.box + 4 * sy;
    const double h = ldexp(bx[3], -dd);
    const int ix = 2 * nd.ix[node] + (o & 1);
    const int iy = 2 * nd.iy[node] + ((o >> 1) & 1);
    const int iz = 2 * nd.iz[node] + ((o >> 2) & 1);
    nd.depth[ch] = dd;
    nd.ix[ch] = ix;
    nd.iy[ch] = iy;
    nd.iz[ch] = iz;
    nd.key[ch] = (nd.key[node] << 3) | (uint64_t)o;
    nd.pbegin[ch] = x;
    nd.pend[ch] = nx;
    nd.sys[ch] = sy;
    nd.parent[ch] = node;
    nd.child[ch] = -1;
    nd.nchild[ch] = 0;
    nd.leaf[ch] = 1;
    nd.cx[ch] = fma(ix + 0.5, h, bx[0]);
    nd.cy[ch] = fma(iy + 0.5, h, bx[1]);
    nd.cz[ch] = fma(iz + 0.5, h, bx[2]);
    // ancestor row: the parent's, then the child itself
    const int *pa = nd.anc + (size_t)node * kAncRow;
    int *ca = nd.anc + (size_t)ch * kAncRow;
    for (int k = 0; k < dd; k++) ca[k] = pa[k];
    ca[dd] = ch;
  }
  if (c == nwork - 1 && threadIdx.x == 0) {
    dl[2 * d + 2] = next_begin;
    dl[2 * d + 3] = (int)(pre + tot);
  }
}

hipError_t launch_split_level(const NodeArrays &nd, int *dl, int d,
                              long max_count, const uint64_t *keys,
                              int leafcap, const ulonglong2 *stop, int nstop,
                              const Lookback &lb, hipStream_t s) {
  const long nb = std::max(1L, (max_count + kLvNodes - 1) / kLvNodes);
  k_split_level<<<nb, 256, 0, s>>>(nd, dl, d, keys, leafcap, stop, nstop, lb);
  return hipGetLastError();
}

long split_level_chunks(long max_count) {
  return std::max(1L, (max_count + kLvNodes - 1) / kLvNodes);
}

// ------------------------------------------------------------ leaves

// leaves[pos[i]] = i for leaves; internal[i - pos[i]] = i for the others
// (both in node order, so grouped by level)
__global__ void k_leaf_compact(NodeArrays nd, int nnodes,
                               const int *__restrict__ pos,
                               int *__restrict__ leaves,
                               int *__restrict__ internal) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  if (nd.leaf[i])
    leaves[pos[i]] = i;
  else
    internal[i - pos[i]] = i;
}

// out[0] = leaves; out[1 + d] = internal nodes of level d, d < nlevels
// (pos: exclusive scan of the leaf flags, total: its sum)
__global__ void k_level_internal(const int *__restrict__ pos,
                                 const int *__restrict__ total,
                                 const int *__restrict__ lstart, int nlevels,
                                 int nnodes, int *__restrict__ out) {
  const int d = threadIdx.x;
  if (d == 0) out[0] = *total;
  if (d >= nlevels) return;
  const int a = lstart[d];
  const int b = d + 1 < nlevels ? lstart[d + 1] : nnodes;
  const int la = pos[a];
  const int lb = b < nnodes ? pos[b] : *total;
  out[1 + d] = (b - a) - (lb - la);
}

hipError_t launch_leaf_compact(const NodeArrays &nd, int nnodes,
                               const int *pos, int *leaves, int *internal,
                               hipStream_t s) {
  k_leaf_compact<<<(nnodes + 255) / 256, 256, 0, s>>>(nd, nnodes, pos, leaves,
                                                      internal);
  return hipGetLastError();
}

hipError_t launch_level_internal(const int *pos, const int *total,
                                 const int *lstart, int nlevels, int nnodes,
                                 int *out, hipStream_t s) {
  if (nlevels > 64) return hipErrorInvalidValue;
  k_level_internal<<<1, 64, 0, s>>>(pos, total, lstart, nlevels, nnodes, out);
  return hipGetLastError();
}

// ------------------------------------------------------------ lists

// Colleagues: the 27 same-depth nodes adjacent to (or equal to) a node,
// stored by offset slot (dx+1)*9 + (dy+1)*3 + (dz+1), -1 when absent,
// for every node in one launch: the neighbour at offset d has the node's
// level, system and coordinates plus d, and a level's nodes are ordered by
// (system, level key), so each slot is one binary search in that level
// (lstart / lcount: the levels' first node and node count).
__global__ void k_colleagues_all(NodeArrays nd, int nnodes,
                                 const int *__restrict__ lstart,
                                 const int *__restrict__ lcount,
                                 int *__restrict__ col) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= 27L * nnodes) return;
  const int A = (int)(t / 27), slot = (int)(t % 27);
  const int d = nd.depth[A];
  const int dx = slot / 9 - 1, dy = (slot / 3) % 3 - 1, dz = slot % 3 - 1;
  const int nx = nd.ix[A] + dx, ny = nd.iy[A] + dy, nz = nd.iz[A] + dz;
  const int side = 1 << d;
  int res = -1;
  if (slot == 13) {
    res = A;
  } else if (nx >= 0 && ny >= 0 && nz >= 0 && nx < side && ny < side &&
             nz < side) {
    const uint64_t key = morton_key(nx, ny, nz);
    const int sy = nd.sys[A];
    int lo = lstart[d], hi = lstart[d] + lcount[d];
    while (lo < hi) {  // first node not below (sy, key)
      const int mid = lo + ((hi - lo) >> 1);
      const int ms = nd.sys[mid];
      if (ms < sy || (ms == sy && nd.key[mid] < key))
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo < lstart[d] + lcount[d] && nd.sys[lo] == sy && nd.key[lo] == key)
      res = lo;
  }
  col[(size_t)A * 27 + slot] = res;
}

hipError_t launch_colleagues_all(const NodeArrays &nd, int nnodes,
                                 const int *lstart, const int *lcount,
                                 int *col, hipStream_t s) {
  if (nnodes == 0) return hipSuccess;
  const long t = 27L * nnodes;
  k_colleagues_all<<<(t + 255) / 256, 256, 0, s>>>(nd, nnodes, lstart, lcount,
                                                   col);
  return hipGetLastError();
}

// M2L (far) list of every node: children of the parent's colleagues that
// are not adjacent to the node.  pass 0 counts, pass 1 fills.
__global__ void k_vlist(NodeArrays nd, int nnodes, const int *__restrict__ col,
                        int pass, int *__restrict__ cnt,
                        const int *__restrict__ ofs, int *__restrict__ src,
                        short *__restrict__ code) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  int c = 0;
  int pos = pass ? ofs[i] : 0;
  if (nd.depth[i] >= 1) {
    const int P = nd.parent[i];
    const int ax = nd.ix[i], ay = nd.iy[i], az = nd.iz[i];
    for (int s = 0; s < 27; s++) {
      int C = col[(size_t)P * 27 + s];
      if (C < 0 || nd.leaf[C]) continue;
      for (int b = nd.child[C]; b < nd.child[C] + nd.nchild[C]; b++) {
        int ox = ax - nd.ix[b], oy = ay - nd.iy[b], oz = az - nd.iz[b];
        if (abs(ox) <= 1 && abs(oy) <= 1 && abs(oz) <= 1) continue;
        if (pass) {
          src[pos] = b;
          code[pos] = (short)code_of(ox, oy, oz);
          pos++;
        } else {
          c++;
        }
      }
    }
  }
  if (!pass) cnt[i] = c;
}

// P2P (near) list of every leaf as particle ranges: its colleagues (whole
// subtrees for non-leaf ones) and the leaf colleagues of its ancestors.
// Near list of every leaf: the same-depth colleagues of the leaf and the
// leaf colleagues of its ancestors (not the ancestors themselves), in the
// order (depth, slot); pass 0 counts, pass 1 fills.  One wave per leaf:
// lanes 0-26 take the slots of one ancestor level and lanes 32-58 those of
// the next, so a level costs one round of loads; positions come from a
// ballot prefix, which keeps the (depth, slot) order.  pairs[t]: reference
// pair counter (n(n+1)/2 for the leaf itself, half of the ordered count
// otherwise: each unordered pair is seen twice).
constexpr int kUlistWaves = 4;

// Level pairs of a near list walk (depths 2 it and 2 it + 1).
constexpr int kUlIter = (kAncRow + 1) / 2;

// This lane's slots of leaf A's near list: in pair it, lanes 0-26 take the
// 27 colleague slots of A's ancestor at depth 2 it and lanes 32-58 those at
// depth 2 it + 1.  Every load of a round is issued at once: the ancestor
// row, then the colleagues of every level, then their leaf flags.
struct UlistSlots {
  int j[kUlIter];                 // the slot's node, -1: not in the list
  unsigned long long m[kUlIter];  // the wave's taken slots
};

// Fills u (one whole wave) and returns the length of A's near list.
__device__ inline int ulist_slots(const NodeArrays &nd,
                                  const int *__restrict__ col, int A, int dA,
                                  UlistSlots &u) {
  const int lane = threadIdx.x & 63, half = lane >> 5, s = lane & 31;
  const int my_anc = lane <= dA ? nd.anc[(size_t)A * kAncRow + lane] : 0;
#pragma unroll
  for (int it = 0; it < kUlIter; it++) {
    const int d = 2 * it + half;
    const int a = __shfl(my_anc, d);
    u.j[it] = (s < 27 && d <= dA) ? col[(size_t)a * 27 + s] : -1;
  }
  int c = 0;
#pragma unroll
  for (int it = 0; it < kUlIter; it++) {
    const int d = 2 * it + half, j = u.j[it];
    // ancestors' slots: leaf colleagues only, not the ancestor itself
    const bool take = j >= 0 && (d == dA || (s != 13 && nd.leaf[j]));
    u.j[it] = take ? j : -1;
    u.m[it] = __ballot(take);
    c += __popcll(u.m[it]);
  }
  return c;
}

// Stores the ranges of u that fall below cap from pos on, in (depth, slot)
// order, and returns A's pair count (every lane).
__device__ inline long long ulist_write(const NodeArrays &nd,
                                        const UlistSlots &u, int A, long pos,
                                        long cap, int *__restrict__ ubeg,
                                        int *__restrict__ uend) {
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = lane == 0 ? 0ull : ~0ull >> (64 - lane);
  const int nA = nd.pend[A] - nd.pbegin[A];
  long long pp = 0;
#pragma unroll
  for (int it = 0; it < kUlIter; it++) {
    const int j = u.j[it];
    if (j >= 0) {
      const int b = nd.pbegin[j], e = nd.pend[j];
      const long k = pos + __popcll(u.m[it] & lt);
      if (k < cap) {
        ubeg[k] = b;
        uend[k] = e;
      }
      pp += j == A ? 2LL * ((long long)nA * (nA + 1) / 2)
                   : (long long)nA * (e - b);
    }
    pos += __popcll(u.m[it]);
  }
  for (int off = 32; off > 0; off >>= 1) pp += __shfl_xor(pp, off);
  return pp;
}

// pass 0 counts each leaf's ranges into cnt; pass 1 writes them from
// ofs[t] on (ofs: the exclusive scan of cnt, ofs[nleaves] its total).  A
// caller that sized the range arrays before knowing the total passes their
// capacity cap: the ranges at cap and beyond are not stored and pass 1
// clamps ofs to cap, so a reader of [ofs[t], ofs[t + 1]) reads stored ranges
// only; *total then receives the true total (above cap: incomplete lists).
// Without a capacity (cap = LONG_MAX) ofs stays as it is.
__global__ __launch_bounds__(64 * kUlistWaves) void k_ulist(
    NodeArrays nd, int nleaves, const int *__restrict__ leaves,
    const int *__restrict__ col, int pass, int *__restrict__ cnt,
    int *__restrict__ ofs, int *__restrict__ ubeg, int *__restrict__ uend,
    long long *__restrict__ pairs, long cap, int *__restrict__ total) {
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int t = blockIdx.x * kUlistWaves + w;
  if (t >= nleaves) return;  // whole wave; no barrier below
  const int A = leaves[t];
  const int dA = nd.depth[A];
  UlistSlots u;
  const int c = ulist_slots(nd, col, A, dA, u);
  if (!pass) {
    if (lane == 0) cnt[t] = c;
    return;
  }
  const long pos = ofs[t];
  const long long np =
      pos < cap ? ulist_write(nd, u, A, pos, cap, ubeg, uend) : 0;
  if (lane == 0) {
    pairs[t] = np;
    if (cap != LONG_MAX) {
      ofs[t] = (int)min(pos, cap);
      if (t == nleaves - 1) {
        const long all = ofs[nleaves];
        *total = (int)min(all, (long)INT32_MAX);
        ofs[nleaves] = (int)min(all, cap);
      }
    }
  }
}

// ---- semifar lists (leaf / deeper node pairs, aran/aransolver3d.c:357-396)
//
// The recursion of the near/far traversal pairs a leaf A with the children
// of a same-depth non-leaf neighbour, and goes on down while the deeper node
// touches A.  A deeper node D that does not touch A and holds more than thr
// particles forms a semifar pair (A, D): P2L of A's particles into D's local
// and M2P of D's multipole onto them; the recursion stops there.  Any other
// non-touching node, and every touching leaf, is near: P2P with all of its
// particles.

// Boxes of node a and of node b (b at least as deep as a) touch or overlap.
__device__ inline bool touching(const NodeArrays &nd, int a, int b) {
  const int s = nd.depth[b] - nd.depth[a];
  const int w = 1 << s;
  const int ax = nd.ix[a] << s, ay = nd.iy[a] << s, az = nd.iz[a] << s;
  return nd.ix[b] >= ax - 1 && nd.ix[b] <= ax + w && nd.iy[b] >= ay - 1 &&
         nd.iy[b] <= ay + w && nd.iz[b] >= az - 1 && nd.iz[b] <= az + w;
}

__device__ inline int node_count(const NodeArrays &nd, int a) {
  return nd.pend[a] - nd.pbegin[a];
}

// Near list (particle ranges) and semifar list (deeper nodes, M2P onto this
// leaf) of every leaf; pass 0 counts, pass 1 fills.  pairs[t]: ordered P2P
// pair count (as k_ulist); spart[t]: particles x semifar nodes of the leaf.
__global__ void k_ulist_semifar(NodeArrays nd, int nleaves,
                                const int *__restrict__ leaves,
                                const int *__restrict__ col, long long thr,
                                int pass, int *__restrict__ cnt,
                                const int *__restrict__ ofs,
                                int *__restrict__ ubeg, int *__restrict__ uend,
                                long long *__restrict__ pairs,
                                int *__restrict__ wcnt,
                                const int *__restrict__ wofs,
                                int *__restrict__ wnode,
                                long long *__restrict__ spart) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nleaves) return;
  const int A = leaves[t];
  const int dA = nd.depth[A];
  const int nA = node_count(nd, A);
  int anc[kMaxDepth + 1];
  anc[dA] = A;
  for (int d = dA - 1; d >= 0; d--) anc[d] = nd.parent[anc[d + 1]];
  int c = 0, cw = 0;
  int pos = pass ? ofs[t] : 0, wpos = pass ? wofs[t] : 0;
  long long np = 0;
  auto near = [&](int j) {
    if (pass) {
      const int b = nd.pbegin[j], e = nd.pend[j];
      ubeg[pos] = b;
      uend[pos] = e;
      pos++;
      if (j == A)
        np += 2LL * ((long long)nA * (nA + 1) / 2);
      else
        np += (long long)nA * (e - b);
    } else {
      c++;
    }
  };
  // same-depth neighbours, descending into non-leaf ones (no stack: the
  // children of a node are contiguous, so the walk moves to the next
  // sibling or back up)
  const int *cA = col + (size_t)A * 27;
  for (int s = 0; s < 27; s++) {
    const int j = cA[s];
    if (j < 0) continue;
    if (nd.leaf[j]) {
      near(j);
      continue;
    }
    int C = nd.child[j];
    for (;;) {
      bool down = false;
      if (!touching(nd, A, C)) {
        if (node_count(nd, C) > thr) {
          if (pass) wnode[wpos++] = C;
          else cw++;
        } else {
          near(C);
        }
      } else if (nd.leaf[C]) {
        near(C);
      } else {
        down = true;
      }
      if (down) {
        C = nd.child[C];
        continue;
      }
      for (;;) {
        const int P = nd.parent[C];
        if (C + 1 < nd.child[P] + nd.nchild[P]) {
          C++;
          break;
        }
        C = P;
        if (C == j) break;
      }
      if (C == j) break;
    }
  }
  // shallower leaves next to an ancestor: near unless the first ancestor
  // that does not touch them took them as a semifar pair
  for (int d = 0; d < dA; d++) {
    const int *cx = col + (size_t)anc[d] * 27;
    for (int s = 0; s < 27; s++) {
      const int J = cx[s];
      if (J < 0 || s == 13 || !nd.leaf[J]) continue;
      int k = d + 1;
      while (k <= dA && touching(nd, J, anc[k])) k++;
      if (k > dA || node_count(nd, anc[k]) <= thr) near(J);
    }
  }
  if (!pass) {
    cnt[t] = c;
    wcnt[t] = cw;
  } else {
    pairs[t] = np;
    spart[t] = (long long)nA * (wpos - wofs[t]);
  }
}

// Semifar list of every node holding more than thr particles: the leaves
// whose particles add to its local by P2L.  pass 0 counts, pass 1 fills.
__global__ void k_xlist(NodeArrays nd, int nnodes, const int *__restrict__ col,
                        long long thr, int pass, int *__restrict__ cnt,
                        const int *__restrict__ ofs,
                        int *__restrict__ xleaf) {
  const int C = blockIdx.x * blockDim.x + threadIdx.x;
  if (C >= nnodes) return;
  const int k = nd.depth[C];
  int c = 0, pos = pass ? ofs[C] : 0;
  if (k >= 1 && node_count(nd, C) > thr) {
    int anc[kMaxDepth + 1];
    anc[k] = C;
    for (int d = k - 1; d >= 0; d--) anc[d] = nd.parent[anc[d + 1]];
    for (int d = 0; d < k; d++) {
      const int *cx = col + (size_t)anc[d] * 27;
      for (int s = 0; s < 27; s++) {
        const int J = cx[s];
        if (J < 0 || s == 13 || !nd.leaf[J]) continue;
        int m = d + 1;
        while (m < k && touching(nd, J, anc[m])) m++;
        if (m == k && !touching(nd, J, C)) {
          if (pass) xleaf[pos++] = J;
          else c++;
        }
      }
    }
  }
  if (!pass) cnt[C] = c;
}

hipError_t launch_ulist_semifar(const NodeArrays &nd, int nleaves,
                                const int *leaves, const int *col, long thr,
                                int pass, int *cnt, const int *ofs, int *ubeg,
                                int *uend, long long *pairs, int *wcnt,
                                const int *wofs, int *wnode, long long *spart,
                                hipStream_t s) {
  if (nleaves == 0) return hipSuccess;
  k_ulist_semifar<<<(nleaves + 127) / 128, 128, 0, s>>>(
      nd, nleaves, leaves, col, thr, pass, cnt, ofs, ubeg, uend, pairs, wcnt,
      wofs, wnode, spart);
  return hipGetLastError();
}

hipError_t launch_xlist(const NodeArrays &nd, int nnodes, const int *col,
                        long thr, int pass, int *cnt, const int *ofs,
                        int *xleaf, hipStream_t s) {
  if (nnodes == 0) return hipSuccess;
  k_xlist<<<(nnodes + 127) / 128, 128, 0, s>>>(nd, nnodes, col, thr, pass, cnt,
                                               ofs, xleaf);
  return hipGetLastError();
}

hipError_t launch_vlist(const NodeArrays &nd, int nnodes, const int *col,
                        int pass, int *cnt, const int *ofs, int *src,
                        short *code, hipStream_t s) {
  if (nnodes == 0) return hipSuccess;
  k_vlist<<<(nnodes + 127) / 128, 128, 0, s>>>(nd, nnodes, col, pass, cnt, ofs,
                                                src, code);
  return hipGetLastError();
}

hipError_t launch_ulist(const NodeArrays &nd, int nleaves, const int *leaves,
                        const int *col, int pass, int *cnt, int *ofs,
                        int *ubeg, int *uend, long long *pairs, long cap,
                        int *total, hipStream_t s) {
  if (nleaves == 0) return hipSuccess;
  k_ulist<<<(nleaves + kUlistWaves - 1) / kUlistWaves, 64 * kUlistWaves, 0,
            s>>>(nd, nleaves, leaves, col, pass, cnt, ofs, ubeg, uend, pairs,
                 cap, total);
  return hipGetLastError();
}

// Sum of per-leaf pair counts: kSumBlocks blocks sum fixed chunks into
// partials, then one block adds the partials (integers: exact).
constexpr int kSumBlocks = 64;

__global__ void k_sum_ll(const long long *__restrict__ v, int n,
                         long long *__restrict__ part) {
  __shared__ long long sh[256];
  const int chunk = (n + kSumBlocks - 1) / kSumBlocks;
  const int b = blockIdx.x * chunk, e = min(n, b + chunk);
  long long s = 0;
  for (int i = b + threadIdx.x; i < e; i += 256) s += v[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_sum_parts(const long long *__restrict__ part,
                            long long *__restrict__ out) {
  if (threadIdx.x == 0) {
    long long s = 0;
    for (int i = 0; i < kSumBlocks; i++) s += part[i];
    *out = s;
  }
}

hipError_t launch_sum_ll(const long long *v, int n, long long *out,
                         hipStream_t s) {
  // partials live right after *out (the caller's buffer holds 1 + 64 values)
  k_sum_ll<<<kSumBlocks, 256, 0, s>>>(v, n, out + 1);
  k_sum_parts<<<1, 64, 0, s>>>(out + 1, out);
  return hipGetLastError();
}

}  // namespace aran
