// engine.hip -- host orchestration of one FMM solve on one GPU.
//
// Order of work follows aran_solver3d_solve (aran/aransolver3d.c:1135-1202):
// tree (replaces VsgPRTree3d insertion), upward pass (P2M at leaves, M2M
// child -> father, level by level from the deepest), far field (M2L gathered
// per target node), downward pass (L2L father -> child, level by level) and
// near field + L2P at leaves.  Every accumulation is a gather with a fixed
// order, so a solve is deterministic.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <complex>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "engine.h"
#include "flavours.h"
#include "m2l_f64.h"
#include "m2l_mfma.h"
#include "translate.h"
#include "m2l_tile.h"

namespace aran {

static thread_local std::string g_err;
const char *last_error() { return g_err.c_str(); }
void set_error(const std::string &msg) { g_err = msg; }

// On failure: the expression, the HIP error and the detail a callee set
// (a solve clears the detail when it starts).
#define CK(expr)                                                        \
  do {                                                                  \
    hipError_t _e = (expr);                                             \
    if (_e != hipSuccess) {                                             \
      const std::string _d = g_err;                                     \
      set_error(std::string(#expr) + ": " + hipGetErrorString(_e) +     \
                (_d.empty() ? std::string() : " (" + _d + ")"));        \
      return _e;                                                        \
    }                                                                   \
  } while (0)

static thread_local DevArena *g_arena = nullptr;
void set_arena(DevArena *arena) { g_arena = arena; }

// A region of the current arena (256-byte aligned).
static hipError_t arena_take(size_t nbytes, void **out) {
  const size_t off = (g_arena->used + 255) & ~(size_t)255;
  const size_t want = std::max<size_t>(nbytes, 256);
  if (off > g_arena->size || want > g_arena->size - off) {
    set_error("the problem needs more working memory than the no-allocation "
              "scratch holds");
    return hipErrorOutOfMemory;
  }
  *out = g_arena->base + off;
  g_arena->used = off + want;
  return hipSuccess;
}

hipError_t DevBuf::reserve(size_t nbytes) {
  if (nbytes <= bytes && ptr) return hipSuccess;
  if (g_arena) {
    if (ptr && !in_arena) {
      set_error("internal error: heap buffer in no-allocation mode");
      return hipErrorUnknown;
    }
    void *np = nullptr;
    hipError_t e = arena_take(nbytes, &np);
    if (e != hipSuccess) return e;
    ptr = np;
    bytes = std::max<size_t>(nbytes, 256);
    in_arena = true;
    return hipSuccess;
  }
  if (ptr) {
    hipError_t e = hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    if (e != hipSuccess) return e;
  }
  size_t want = std::max<size_t>(nbytes + nbytes / 4, 256);
  hipError_t e = hipMalloc(&ptr, want);
  if (e != hipSuccess) {
    ptr = nullptr;
    return e;
  }
  bytes = want;
  return hipSuccess;
}

hipError_t DevBuf::grow_keep(size_t nbytes, size_t keep, hipStream_t s) {
  if (nbytes <= bytes && ptr) return hipSuccess;
  void *np = nullptr;
  size_t want = std::max<size_t>(nbytes + nbytes / 2, 256);
  hipError_t e;
  if (g_arena) {
    if (ptr && !in_arena) {
      set_error("internal error: heap buffer in no-allocation mode");
      return hipErrorUnknown;
    }
    want = std::max<size_t>(nbytes, 256);
    e = arena_take(want, &np);
  } else {
    e = hipMalloc(&np, want);
  }
  if (e != hipSuccess) return e;
  if (ptr && keep) {
    e = hipMemcpyAsync(np, ptr, std::min(keep, bytes), hipMemcpyDeviceToDevice,
                       s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
  }
  if (ptr && !in_arena) (void)hipFree(ptr);
  ptr = np;
  bytes = want;
  in_arena = g_arena != nullptr;
  return hipSuccess;
}

void DevBuf::release() {
  if (ptr && !in_arena) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
  in_arena = false;
}

void Tables::release() {
  DevBuf *all[] = {&rot,    &eiphi,  &dir,       &tm2l,    &tm2m,  &tl2l,
                   &ynorm,  &ogain,  &pgain,     &igain,   &xgain, &dscr,
                   &thetas, &rot32,  &ephi32,    &vfac32,  &oct_codes,
                   &code_to_k, &mfma_ops, &f64_ops, &alpha, &beta};
  for (DevBuf *b : all) b->release();
  order = -1;
  mfma_order = -1;
  f64_order = -1;
  dir_host.clear();
}

OpTables Tables::view() const {
  OpTables t;
  t.rot = rot.as<double>();
  t.eiphi = eiphi.as<double2>();
  t.dir = dir.as<DirEntry>();
  t.tm2l = tm2l.as<double>();
  t.tm2m = tm2m.as<double>();
  t.tl2l = tl2l.as<double>();
  t.ynorm = ynorm.as<double>();
  t.ogain = ogain.as<double>();
  t.pgain = pgain.as<double>();
  t.igain = igain.as<double>();
  t.xgain = xgain.as<double>();
  t.alpha = alpha.as<double>();
  t.beta = beta.as<double>();
  t.p = order;
  t.rb = rot_block(order);
  return t;
}

NodeArrays Workspace::nodes() const {
  NodeArrays a;
  a.depth = n_depth.as<int>();
  a.ix = n_ix.as<int>();
  a.iy = n_iy.as<int>();
  a.iz = n_iz.as<int>();
  a.pbegin = n_pbegin.as<int>();
  a.pend = n_pend.as<int>();
  a.parent = n_parent.as<int>();
  a.child = n_child.as<int>();
  a.nchild = n_nchild.as<int>();
  a.leaf = n_leaf.as<int>();
  a.sys = n_sys.as<int>();
  a.anc = n_anc.as<int>();
  a.cmap = n_cmap.as<int>();
  a.box = sys_box.as<double>();
  a.key = n_key.as<uint64_t>();
  a.cx = n_cx.as<double>();
  a.cy = n_cy.as<double>();
  a.cz = n_cz.as<double>();
  return a;
}

std::vector<DevBuf *> Workspace::buffers() {
  std::vector<DevBuf *> all = {&keys_a, &keys_b, &idx_a, &idx_b, &xs, &ys, &zs, &qs,
                   &pot_s, &grad_s, &hist, &part, &n_depth, &n_ix,
                   &n_iy, &n_iz, &n_pbegin, &n_pend, &n_parent, &n_child,
                   &n_nchild, &n_leaf, &n_key, &n_cx, &n_cy, &n_cz, &n_anc, &n_cmap,
                   &split_ofs, &leaf_list,
                   &v_cnt, &v_ofs, &v_src, &v_code, &u_cnt,
                   &u_ofs, &u_beg, &u_end, &u_pairs, &colleagues, &w_cnt, &w_ofs,
                   &w_node, &x_cnt, &x_ofs, &x_leaf, &s_part, &mult, &loc,
                   &small, &pleaf, &tile_tgt, &tile_src, &node_slot, &tile_cnt,
                   &tile_rank, &tile_meta, &mult32, &loc32,
                   &mult_bf, &loc_mfma, &mult64, &loc64, &n_sys, &sys_box,
                   &sys_off,
                   &skey, &spos, &stop, &internal_list,
                   &trans_rows, &packed,
                   &lvl_table, &lb_buf};
  Tables *ts[1] = {&tab};
  for (Tables *t : ts) {
    DevBuf *tb[] = {&t->rot,    &t->eiphi,  &t->dir,    &t->tm2l,   &t->tm2m,
                    &t->tl2l,   &t->ynorm,  &t->ogain,  &t->pgain,  &t->igain,
                    &t->xgain,  &t->dscr,   &t->thetas, &t->rot32,  &t->ephi32,
                    &t->vfac32, &t->oct_codes, &t->code_to_k, &t->mfma_ops,
                    &t->f64_ops, &t->alpha,  &t->beta};
    all.insert(all.end(), tb, tb + sizeof tb / sizeof tb[0]);
  }
  return all;
}

void Workspace::forget_arena_buffers() {
  for (DevBuf *b : buffers()) {
    b->ptr = nullptr;
    b->bytes = 0;
    b->in_arena = false;
  }
  // capacities recorded beside the buffers go with them
  node_cap = 0;
  far_csr_nodes = -1;
  semifar_nodes = -1;
  tab.order = -1;
  tab.mfma_order = -1;
  tab.f64_order = -1;
  cached_tabs.clear();
}

Workspace::~Workspace() {
  std::vector<DevBuf *> all = buffers();
  for (DevBuf *b : all) b->release();
  tab.release();
  for (Tables &t : cached_tabs) t.release();
  if (host_small) (void)hipHostFree(host_small);
  if (host_stage) (void)hipHostFree(host_stage);
  for (char *b : host_stage_old) (void)hipHostFree(b);
  if (events_ready) {
    for (auto &e : ev) (void)hipEventDestroy(e);
    for (auto &e : evk) (void)hipEventDestroy(e);
    (void)hipEventDestroy(ev_fork);
    (void)hipEventDestroy(ev_join);
    (void)hipEventDestroy(ev_near);
  }
  if (aux) (void)hipStreamDestroy(aux);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

hipError_t workspace_init(Workspace &ws, hipStream_t stream) {
  if (ws.device < 0) CK(hipGetDevice(&ws.device));
  if (!ws.host_small) CK(hipHostMalloc(&ws.host_small, 8192));
  if (stream) {
    ws.stream = stream;
  } else if (!ws.stream) {
    CK(hipStreamCreate(&ws.stream));
    ws.own_stream = true;
  }
  if (!ws.events_ready) {
    for (auto &e : ws.ev) CK(hipEventCreate(&e));
    for (auto &e : ws.evk) CK(hipEventCreate(&e));
    CK(hipEventCreateWithFlags(&ws.ev_fork, hipEventDisableTiming));
    CK(hipEventCreateWithFlags(&ws.ev_join, hipEventDisableTiming));
    CK(hipEventCreate(&ws.ev_near));
    ws.events_ready = true;
  }
  if (!ws.aux) CK(hipStreamCreateWithFlags(&ws.aux, hipStreamNonBlocking));
  const void *old_small = ws.small.ptr;
  CK(ws.small.reserve(4096));
  // (zero once: the look-back error word in it is sticky)
  if (ws.small.ptr != old_small)
    CK(launch_fill_int(ws.small.as<int>(), 1024, 0, ws.stream));
  return hipSuccess;
}

// Error word of the look-back launches outside the level build (set when a
// wait gave up); read back with the pair count at the end of each solve.
static int *lb_err_word(Workspace &ws) {
  return reinterpret_cast<int *>(ws.small.as<char>() + 1016);
}

// Pinned host memory for one copy of this solve (see Workspace::host_stage).
static hipError_t stage_alloc(Workspace &ws, size_t bytes, void **out) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (ws.host_stage_used + bytes > ws.host_stage_bytes && ws.fixed_stage) {
    set_error("the problem needs more pinned staging than the no-allocation "
              "context holds");
    return hipErrorOutOfMemory;
  }
  if (ws.host_stage_used + bytes > ws.host_stage_bytes) {
    if (ws.host_stage) ws.host_stage_old.push_back(ws.host_stage);
    const size_t want =
        std::max(std::max<size_t>(65536, 2 * ws.host_stage_bytes), 2 * bytes);
    ws.host_stage = nullptr;
    ws.host_stage_bytes = ws.host_stage_used = 0;
    CK(hipHostMalloc(&ws.host_stage, want));
    ws.host_stage_bytes = want;
  }
  *out = ws.host_stage + ws.host_stage_used;
  ws.host_stage_used += bytes;
  return hipSuccess;
}

// Host array to device memory through the pinned staging (so the copy is
// asynchronous: a copy from pageable memory makes the host wait).
static hipError_t stage_copy(Workspace &ws, void *dst, const void *src,
                             size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  void *h = nullptr;
  CK(stage_alloc(ws, bytes, &h));
  memcpy(h, src, bytes);
  return hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, s);
}

// Look-back state of one launch over nchunks chunks.  The status words
// are cleared when their buffer is (re)allocated or the epoch wraps.
static hipError_t lookback_take(Workspace &ws, long nchunks, int *err,
                                Lookback *lb, hipStream_t s) {
  const void *old = ws.lb_buf.ptr;
  const size_t old_bytes = ws.lb_buf.bytes;
  CK(ws.lb_buf.reserve(sizeof(unsigned long long) * (size_t)nchunks));
  bool clear = ws.lb_buf.ptr != old || ws.lb_buf.bytes != old_bytes;
  if (++ws.lb_epoch >= (1u << 30)) {
    ws.lb_epoch = 1;
    clear = true;
  }
  if (clear)
    CK(launch_fill_int(ws.lb_buf.as<int>(), (long)(ws.lb_buf.bytes / 4), 0, s));
  lb->status = ws.lb_buf.as<unsigned long long>();
  lb->err = err;
  lb->epoch = ws.lb_epoch;
  return hipSuccess;
}

// Exclusive scan in one launch whatever n: one workgroup for short
// arrays, look-back across workgroups for the others.
static hipError_t scan_one_launch(Workspace &ws, const int *in, int *out,
                                  long n, int *total, hipStream_t s) {
  if (scan_fits_single(n))
    return scan_small(in, out, n, total, s);
  Lookback lb;
  CK(lookback_take(ws, scan_lookback_chunks(n), lb_err_word(ws), &lb, s));
  return scan_lookback(in, out, n, total, lb, s);
}

static void stage_reset(Workspace &ws) {
  // (a solve that returned an error may have left copies in flight)
  (void)hipStreamSynchronize(ws.stream);
  for (char *b : ws.host_stage_old) (void)hipHostFree(b);
  ws.host_stage_old.clear();
  ws.host_stage_used = 0;
}

// ------------------------------------------------------------ tables

// alpha_l^m = 1/sqrt((l-m)!(l+m)!) by the recurrences of _alpha_generator
// (aran/aransphericalseriesd.c:45-59).
static std::vector<double> alpha_table(int deg) {
  std::vector<double> a(tri(deg + 1));
  for (int l = 0; l <= deg; l++)
    for (int m = 0; m <= l; m++) {
      double v;
      if (l == 0)
        v = 1.;
      else if (m == 0)
        v = a[tri(l - 1)] / l;
      else if (l == m)
        v = a[tri(l - 1) + l - 1] / sqrt((l + l - 1.) * (l + l));
      else
        v = sqrt((l - m + 1.) / (l + m)) * a[tri(l) + m - 1];
      a[tri(l) + m] = v;
    }
  return a;
}

static inline double beta_l(int l) { return sqrt((4. * M_PI) / (l + l + 1.)); }

void octant_offsets(const std::vector<DirEntry> &dirs, std::vector<short> &codes,
                    std::vector<short> &code_to_k) {
  codes.assign(8 * kOctOffsets, 0);
  code_to_k.assign(8 * kNumCodes, -1);
  for (int oct = 0; oct < 8; oct++) {
    int lo[3];
    for (int a = 0; a < 3; a++) lo[a] = ((oct >> a) & 1) ? -2 : -3;
    std::vector<int> list;
    for (int dx = lo[0]; dx < lo[0] + 6; dx++)
      for (int dy = lo[1]; dy < lo[1] + 6; dy++)
        for (int dz = lo[2]; dz < lo[2] + 6; dz++) {
          if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) continue;
          list.push_back(code_of(dx, dy, dz));
        }
    std::stable_sort(list.begin(), list.end(), [&](int a, int b) {
      return dirs[a].theta < dirs[b].theta;
    });
    for (int k = 0; k < (int)list.size(); k++) {
      codes[oct * kOctOffsets + k] = (short)list[k];
      code_to_k[oct * kNumCodes + list[k]] = (short)k;
    }
  }
}

// Waves per far-field tile: the 189 offsets of a tile are split in this many
// ranges so that a solve has several waves per SIMD (ARAN_M2L_SPLIT
// overrides, 1..16).
hipError_t ensure_tables(Workspace &ws, int p) {
  if (ws.tab.order == p) return hipSuccess;
  // the current order's tables join the cache; a cached order is swapped in
  // (Tables hold plain device pointers: a moved-from Tables is reset so
  // that each buffer has one owner)
  if (ws.tab.order >= 0) {
    ws.cached_tabs.push_back(std::move(ws.tab));
    ws.tab = Tables();
  }
  for (size_t k = 0; k < ws.cached_tabs.size(); k++)
    if (ws.cached_tabs[k].order == p) {
      ws.tab = std::move(ws.cached_tabs[k]);
      ws.cached_tabs.erase(ws.cached_tabs.begin() + k);
      return hipSuccess;
    }
  while ((int)ws.cached_tabs.size() > kCachedOrders) {
    ws.cached_tabs.front().release();
    ws.cached_tabs.erase(ws.cached_tabs.begin());
  }
  hipStream_t s = ws.stream;
  Tables &T = ws.tab;
  // directions and theta values
  std::map<std::pair<int, int>, int> theta_idx;
  std::vector<double> thetas;
  std::vector<DirEntry> dirs(kNumCodes);
  std::vector<std::complex<double>> eiphi((size_t)kNumCodes * (p + 1));
  for (int dx = -3; dx <= 3; dx++)
    for (int dy = -3; dy <= 3; dy++)
      for (int dz = -3; dz <= 3; dz++) {
        int code = code_of(dx, dy, dz);
        int vx = dx, vy = dy, vz = dz;
        double c = 1.;
        if (vz < 0) {
          c = -1.;
          vx = -vx;
          vy = -vy;
          vz = -vz;
        }
        int rho2 = vx * vx + vy * vy;
        DirEntry de;
        de.c = c;
        de.norm = sqrt((double)(rho2 + vz * vz));
        if (rho2 + vz * vz == 0) {
          de.theta = 0;
          de.norm = 0.;
        } else {
          auto key = std::make_pair(vz, rho2);
          auto it = theta_idx.find(key);
          if (it == theta_idx.end()) {
            int id = (int)thetas.size();
            theta_idx[key] = id;
            thetas.push_back(atan2(sqrt((double)rho2), (double)vz));
            de.theta = id;
          } else {
            de.theta = it->second;
          }
        }
        dirs[code] = de;
        std::complex<double> e1(1., 0.), e = 1.;
        if (rho2 > 0) e1 = std::complex<double>(vx, vy) / sqrt((double)rho2);
        for (int m = 0; m <= p; m++) {
          eiphi[(size_t)code * (p + 1) + m] = e;
          e *= e1;
        }
      }
  if (thetas.empty()) thetas.push_back(0.);
  // the level L2L stages one rotation table for its eight offsets
  // (+-1, +-1, +-1): after the z flip they share one angle
  for (int o = 1; o < 8; o++)
    if (dirs[code_of(2 * (o & 1) - 1, 2 * ((o >> 1) & 1) - 1,
                     2 * (o >> 2) - 1)].theta != dirs[code_of(-1, -1, -1)].theta) {
      set_error("internal error: diagonal offsets with different angles");
      return hipErrorUnknown;
    }
  int nth = (int)thetas.size();
  // translation factor tables
  std::vector<double> al = alpha_table(2 * p + 2);
  auto A = [&](int l, int m) { return al[tri(l) + m]; };
  std::vector<double> tm2l((size_t)(p + 1) * (p + 1) * p, 0.);
  for (int l = 0; l <= p; l++)
    for (int m = 0; m <= l; m++)
      for (int n = m; n < p; n++)
        tm2l[((size_t)l * (p + 1) + m) * p + n] =
            beta_l(l) / beta_l(n) * A(l, m) * A(n, m) / A(l + n, 0);
  std::vector<double> tm2m((size_t)p * p * p, 0.);
  for (int l = 0; l < p; l++)
    for (int m = 0; m <= l; m++)
      for (int n = m; n <= l; n++)
        tm2m[((size_t)l * p + m) * p + n] =
            A(l - n, 0) * A(n, m) / A(l, m) * (beta_l(l) / beta_l(n));
  std::vector<double> tl2l((size_t)(p + 1) * (p + 1) * (p + 1), 0.);
  for (int l = 0; l <= p; l++)
    for (int m = 0; m <= l; m++)
      for (int n = l; n <= p; n++)
        tl2l[((size_t)l * (p + 1) + m) * (p + 1) + n] =
            A(n - l, 0) * A(l, m) / A(n, m) * (beta_l(l) / beta_l(n));
  // N_l^m (aran/aransphericalharmonic.c:31-38)
  std::vector<double> yn(tri(p + 2));
  for (int l = 0; l <= p + 1; l++)
    for (int m = 0; m <= l; m++)
      yn[tri(l) + m] = (m == 0) ? sqrt((l + l + 1.) / (4. * M_PI))
                                : yn[tri(l) + m - 1] / sqrt((l - m + 1.) * (l + m));

  // gains of the solid harmonics: g_l^m = N_l^m (l+m)!, and for P2M
  // 4 pi / (2l+1) g_l^m
  std::vector<double> og(tri(p + 1)), pg(std::max(1, tri(p)));
  for (int l = 0; l <= p; l++)
    for (int m = 0; m <= l; m++) {
      double f = 1.;
      for (int i = 2; i <= l + m; i++) f *= i;
      og[tri(l) + m] = yn[tri(l) + m] * f;
      if (l < p) pg[tri(l) + m] = 4. * M_PI / (l + l + 1.) * og[tri(l) + m];
    }
  // irregular solid harmonics (semifar M2P / P2L): r^-(l+1) Y_l^m =
  // N_l^m / (l-m)! I_l^m with I_l^m = (l-m)! P_l^m e^{i m phi} / r^(l+1)
  std::vector<double> ig(tri(p + 1)), xg(tri(p + 1));
  for (int l = 0; l <= p; l++)
    for (int m = 0; m <= l; m++) {
      double f = 1.;
      for (int i = 2; i <= l - m; i++) f *= i;
      ig[tri(l) + m] = yn[tri(l) + m] / f;
      xg[tri(l) + m] = 4. * M_PI / (l + l + 1.) * ig[tri(l) + m];
    }

  int rb = rot_block(p);
  int w = 2 * p + 1;
  CK(T.thetas.reserve(sizeof(double) * nth));
  CK(T.dscr.reserve(sizeof(double) * (size_t)nth * (p + 1) * w * w));
  CK(T.rot.reserve(sizeof(double) * (size_t)nth * 4 * rb));
  CK(T.eiphi.reserve(sizeof(double) * 2 * eiphi.size()));
  CK(T.dir.reserve(sizeof(DirEntry) * dirs.size()));
  CK(T.tm2l.reserve(sizeof(double) * std::max<size_t>(1, tm2l.size())));
  CK(T.tm2m.reserve(sizeof(double) * std::max<size_t>(1, tm2m.size())));
  CK(T.tl2l.reserve(sizeof(double) * tl2l.size()));
  CK(T.ynorm.reserve(sizeof(double) * yn.size()));
  CK(T.ogain.reserve(sizeof(double) * og.size()));
  CK(T.pgain.reserve(sizeof(double) * pg.size()));
  CK(T.igain.reserve(sizeof(double) * ig.size()));
  CK(T.xgain.reserve(sizeof(double) * xg.size()));
  CK(hipMemcpyAsync(T.thetas.ptr, thetas.data(), sizeof(double) * nth,
                    hipMemcpyHostToDevice, s));
  CK(hipMemsetAsync(T.dscr.ptr, 0, sizeof(double) * (size_t)nth * (p + 1) * w * w,
                    s));
  CK(hipMemcpyAsync(T.eiphi.ptr, eiphi.data(), sizeof(double) * 2 * eiphi.size(),
                    hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(T.dir.ptr, dirs.data(), sizeof(DirEntry) * dirs.size(),
                    hipMemcpyHostToDevice, s));
  if (!tm2l.empty())
    CK(hipMemcpyAsync(T.tm2l.ptr, tm2l.data(), sizeof(double) * tm2l.size(),
                      hipMemcpyHostToDevice, s));
  if (!tm2m.empty())
    CK(hipMemcpyAsync(T.tm2m.ptr, tm2m.data(), sizeof(double) * tm2m.size(),
                      hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(T.tl2l.ptr, tl2l.data(), sizeof(double) * tl2l.size(),
                    hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(T.ynorm.ptr, yn.data(), sizeof(double) * yn.size(),
                    hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(T.ogain.ptr, og.data(), sizeof(double) * og.size(),
                    hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(T.pgain.ptr, pg.data(), sizeof(double) * pg.size(),
                    hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(T.igain.ptr, ig.data(), sizeof(double) * ig.size(),
                    hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(T.xgain.ptr, xg.data(), sizeof(double) * xg.size(),
                    hipMemcpyHostToDevice, s));
  // alpha / beta to degree kAlphaDeg (the normal and Kylin translations)
  {
    std::vector<double> ak = alpha_table(kAlphaDeg), bk(kAlphaDeg + 1);
    for (int l = 0; l <= kAlphaDeg; l++) bk[l] = beta_l(l);
    CK(T.alpha.reserve(sizeof(double) * ak.size()));
    CK(T.beta.reserve(sizeof(double) * bk.size()));
    CK(hipMemcpyAsync(T.alpha.ptr, ak.data(), sizeof(double) * ak.size(),
                      hipMemcpyHostToDevice, s));
    CK(hipMemcpyAsync(T.beta.ptr, bk.data(), sizeof(double) * bk.size(),
                      hipMemcpyHostToDevice, s));
  }
  CK(launch_wigner_tables(T.thetas.as<double>(), nth, p, T.dscr.as<double>(),
                          T.rot.as<double>(), s));
  // tiled far field: octant offset lists and FP32 tables
  std::vector<short> oct_codes, code_to_k;
  octant_offsets(dirs, oct_codes, code_to_k);
  CK(T.oct_codes.reserve(sizeof(short) * oct_codes.size()));
  CK(T.code_to_k.reserve(sizeof(short) * code_to_k.size()));
  CK(hipMemcpyAsync(T.oct_codes.ptr, oct_codes.data(),
                    sizeof(short) * oct_codes.size(), hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(T.code_to_k.ptr, code_to_k.data(),
                    sizeof(short) * code_to_k.size(), hipMemcpyHostToDevice, s));
  const long nrot = (long)nth * 4 * rb;
  // FP32 tables of the tiled far field (mixed form: orders <= 15)
  if (p <= kMaxOrderFast) {
    CK(T.rot32.reserve(sizeof(float) * nrot));
    CK(T.ephi32.reserve(sizeof(float) * 2 * eiphi.size()));
    CK(T.vfac32.reserve(sizeof(float) * (size_t)kNumCodes *
                        std::max(1, vfac_size(p))));
    CK(launch_m2l_tables(1, p, T.dir.as<DirEntry>(), T.tm2l.as<double>(),
                         T.rot.as<double>(), nrot, T.eiphi.as<double2>(),
                         T.vfac32.ptr, T.rot32.ptr, T.ephi32.ptr, s));
  }
  T.dir_host = dirs;
  CK(hipStreamSynchronize(s));
  T.ntheta = nth;
  T.order = p;
  return hipSuccess;
}

// ------------------------------------------------------------ solve

static hipError_t read_int(Workspace &ws, const int *dev, int *out) {
  int *h = static_cast<int *>(ws.host_small);
  CK(hipMemcpyAsync(h, dev, sizeof(int), hipMemcpyDeviceToHost, ws.stream));
  CK(hipStreamSynchronize(ws.stream));
  *out = *h;
  return hipSuccess;
}

// Root cube of the first system (lo, width), read back for the tree info.
static hipError_t read_root_cube(Workspace &ws, double *lo, double *width) {
  double *h = reinterpret_cast<double *>(static_cast<char *>(ws.host_small) +
                                         4096 + 2048);
  CK(hipMemcpyAsync(h, ws.sys_box.ptr, 4 * sizeof(double),
                    hipMemcpyDeviceToHost, ws.stream));
  CK(hipStreamSynchronize(ws.stream));
  for (int k = 0; k < 3; k++) lo[k] = h[k];
  *width = h[3];
  return hipSuccess;
}

static hipError_t reserve_nodes(Workspace &ws, long cap, long keep) {
  hipStream_t s = ws.stream;
  DevBuf *ints[] = {&ws.n_depth, &ws.n_ix, &ws.n_iy, &ws.n_iz, &ws.n_pbegin,
                    &ws.n_pend, &ws.n_parent, &ws.n_child, &ws.n_nchild,
                    &ws.n_leaf, &ws.n_sys};
  for (DevBuf *b : ints) CK(b->grow_keep(sizeof(int) * cap, sizeof(int) * keep, s));
  CK(ws.n_key.grow_keep(sizeof(uint64_t) * cap, sizeof(uint64_t) * keep, s));
  CK(ws.n_anc.grow_keep(sizeof(int) * kAncRow * cap,
                        sizeof(int) * kAncRow * keep, s));
  CK(ws.n_cmap.grow_keep(sizeof(int) * 8 * cap, sizeof(int) * 8 * keep, s));
  DevBuf *dbl[] = {&ws.n_cx, &ws.n_cy, &ws.n_cz};
  for (DevBuf *b : dbl)
    CK(b->grow_keep(sizeof(double) * cap, sizeof(double) * keep, s));
  ws.node_cap = std::min(ws.n_depth.bytes / sizeof(int),
                         ws.n_cx.bytes / sizeof(double));
  ws.node_cap = std::min<long>(ws.node_cap, ws.n_key.bytes / sizeof(uint64_t));
  ws.node_cap =
      std::min<long>(ws.node_cap, ws.n_anc.bytes / (sizeof(int) * kAncRow));
  ws.node_cap = std::min<long>(ws.node_cap, ws.n_cmap.bytes / (sizeof(int) * 8));
  return hipSuccess;
}

// Far (M2L) lists in CSR form from the colleagues of the last tree: the
// children of the parent's colleagues that are not adjacent to the node.
hipError_t build_far_lists(Workspace &ws, long nn, int *nfar) {
  hipStream_t s = ws.stream;
  NodeArrays nd = ws.nodes();
  const int *col = ws.colleagues.as<int>();
  CK(ws.v_cnt.reserve(sizeof(int) * nn));
  CK(ws.v_ofs.reserve(sizeof(int) * (nn + 1)));
  CK(launch_vlist(nd, (int)nn, col, 0, ws.v_cnt.as<int>(), nullptr, nullptr,
                  nullptr, s));
  CK(scan_one_launch(ws, ws.v_cnt.as<int>(), ws.v_ofs.as<int>(), nn,
                     ws.v_ofs.as<int>() + nn, s));
  int nf = 0;
  CK(read_int(ws, ws.v_ofs.as<int>() + nn, &nf));
  CK(ws.v_src.reserve(sizeof(int) * std::max(1, nf)));
  CK(ws.v_code.reserve(sizeof(short) * std::max(1, nf)));
  CK(launch_vlist(nd, (int)nn, col, 1, ws.v_cnt.as<int>(), ws.v_ofs.as<int>(),
                  ws.v_src.as<int>(), ws.v_code.as<short>(), s));
  ws.far_csr_nodes = nn;
  *nfar = nf;
  return hipSuccess;
}

// ---- far-field schedule autotuner
//
// The tiled far-field kernels can walk their work in several orders over the
// XCDs (which tiles and offset ranges share an XCD's L2 at a time); every
// order computes the same partial sums over the same fixed offset ranges, so
// the choice changes the speed, never the result.  With Problem::autotune
// the first solve of each (form, order, size bucket) times every schedule on
// the device (HIP events on the engine stream, far field plus its slice
// reduction) and keeps the fastest for later solves of that bucket.
static long tune_key(int precision, int p, int units) {
  int bucket = 0;  // units rounded down to a power of two
  while ((2 << bucket) <= units) bucket++;
  return ((long)precision * 64 + p) * 64 + bucket;
}

template <class Run>
static hipError_t tune_schedule(Workspace &ws, long key, int nchoices, Run run,
                                int *choice) {
  auto it = ws.tuned_sched.find(key);
  if (it != ws.tuned_sched.end()) {
    *choice = it->second;
    return hipSuccess;
  }
  hipStream_t s = ws.stream;
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  float best_ms = -1.f;
  int best = *choice;
  hipError_t err = run(best);  // warm the caches once
  for (int c = 0; c < nchoices; c++) {
    if (err != hipSuccess) break;
    float ms = 0.f;
    if ((err = hipEventRecord(e0, s)) != hipSuccess) break;
    if ((err = run(c)) != hipSuccess) break;
    if ((err = hipEventRecord(e1, s)) != hipSuccess) break;
    if ((err = hipEventSynchronize(e1)) != hipSuccess) break;
    if ((err = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
    if (best_ms < 0.f || ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (err != hipSuccess) return err;
  ws.tuned_sched[key] = best;
  *choice = best;
  return hipSuccess;
}

hipError_t ensure_mfma_tables(Workspace &ws, int order) {
  CK(ensure_tables(ws, order));
  Tables &T = ws.tab;
  if (T.mfma_order == order) return hipSuccess;
  hipStream_t s = ws.stream;
  const MfmaDims D = mfma_dims(order);
  DevBuf tmp;
  hipError_t e = tmp.reserve(sizeof(double) * (size_t)kNumCodes * D.ro * D.ki);
  if (e == hipSuccess) e = launch_m2l_dense(T.view(), tmp.as<double>(), s);
  if (e == hipSuccess)
    e = T.mfma_ops.reserve(16 * (size_t)kNumCodes * mfma_ops_units(order));
  if (e == hipSuccess)
    e = launch_mfma_pack_ops(order, tmp.as<double>(), T.mfma_ops.ptr, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  tmp.release();
  if (e != hipSuccess) return e;
  T.mfma_order = order;
  return hipSuccess;
}

hipError_t ensure_f64_tables(Workspace &ws, int order) {
  CK(ensure_tables(ws, order));
  Tables &T = ws.tab;
  if (T.f64_order == order) return hipSuccess;
  hipStream_t s = ws.stream;
  const MfmaDims D = mfma_dims(order);
  DevBuf tmp;
  hipError_t e = tmp.reserve(sizeof(double) * (size_t)kNumCodes * D.ro * D.ki);
  if (e == hipSuccess) e = launch_m2l_dense(T.view(), tmp.as<double>(), s);
  if (e == hipSuccess)
    e = T.f64_ops.reserve(sizeof(double) * (size_t)kNumCodes *
                          f64_ops_doubles(order));
  if (e == hipSuccess)
    e = launch_f64_pack_ops(order, tmp.as<double>(), T.f64_ops.as<double>(),
                            s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  tmp.release();
  if (e != hipSuccess) return e;
  T.f64_order = order;
  return hipSuccess;
}

hipError_t solve(Workspace &ws, const Problem &pb, Stats *st, TreeInfo *ti) {
  set_error("");
  const long n = pb.n;
  const int p = pb.order;
  if (p < 1 || p > kMaxOrder) {
    set_error("order must be in [1, 30]");
    return hipErrorInvalidValue;
  }
  if (pb.precision != 0 && p > kMaxOrderFast) {
    set_error("orders above 15 run in the plain FP64 form only");
    return hipErrorInvalidValue;
  }
  const int flav[3] = {pb.flav_m2m, pb.flav_m2l, pb.flav_l2l};
  for (int f : flav)
    if (f < kFlavRotate || f > kFlavKylin ||
        (f != kFlavRotate && pb.precision != 0)) {
      set_error("the normal and Kylin translation flavours run in the plain "
                "FP64 form only");
      return hipErrorInvalidValue;
    }
  if (pb.nstop > 0 && (pb.nsys > 1 || !pb.stop)) {
    set_error("a stop list needs one system");
    return hipErrorInvalidValue;
  }
  if (pb.leafcap < 1 || n < 0 || n > (1L << 30)) {
    set_error("leaf capacity must be >= 1 and 0 <= n <= 2^30");
    return hipErrorInvalidValue;
  }
  if (pb.precision < 0 || pb.precision > 3) {
    set_error("precision must be 0 (plain), 1 (mixed), 2 (BF16x3) or 3 "
              "(BF16x6)");
    return hipErrorInvalidValue;
  }
  CK(workspace_init(ws, ws.stream));
  CK(ensure_tables(ws, p));
  if (pb.precision >= 2) CK(ensure_mfma_tables(ws, p));
  // plain form: dense FP64 far field on MFMA where it applies
  const bool dense64 = pb.precision == 0 && pb.dense_far &&
                       pb.flav_m2l == kFlavRotate && m2l_f64_enabled(p);
  if (dense64) CK(ensure_f64_tables(ws, p));
  hipStream_t s = ws.stream;
  Stats local_st;
  Stats &S = st ? *st : local_st;
  S = Stats();
  if (n == 0) return hipSuccess;
  // a near field left running by a failed solve must end before this one
  // reuses its buffers (no-op when the event was never recorded)
  CK(hipStreamWaitEvent(s, ws.ev_join, 0));
  if (pb.timings) CK(hipEventRecord(ws.ev[0], s));

  // ---- systems: particle ranges and root boxes from their extents
  const int nsys = pb.nsys > 0 ? pb.nsys : 1;
  std::vector<long> soff(nsys + 1);
  std::vector<double> lbs(3 * nsys), ubs(3 * nsys);
  if (pb.nsys > 0) {
    for (int k = 0; k <= nsys; k++) soff[k] = pb.soff[k];
    for (int k = 0; k < 3 * nsys; k++) {
      lbs[k] = pb.lbs[k];
      ubs[k] = pb.ubs[k];
    }
    for (int k = 0; k < nsys; k++)
      if (soff[k + 1] < soff[k]) {
        set_error("system offsets must be non-decreasing");
        return hipErrorInvalidValue;
      }
    if (soff[0] != 0 || soff[nsys] != n) {
      set_error("system offsets must start at 0 and end at n");
      return hipErrorInvalidValue;
    }
  } else {
    soff[0] = 0;
    soff[1] = n;
    for (int k = 0; k < 3; k++) {
      lbs[k] = pb.lb[k];
      ubs[k] = pb.ub[k];
    }
  }
  // system offsets, then each system's bounds (lb, ub): one copy
  const size_t nsoff = (size_t)(nsys + 1);
  CK(ws.sys_off.reserve(sizeof(long) * nsoff + sizeof(double) * 6 * nsys));
  CK(ws.sys_box.reserve(sizeof(double) * 4 * nsys));
  stage_reset(ws);
  {
    std::vector<char> h(sizeof(long) * nsoff + sizeof(double) * 6 * nsys);
    memcpy(h.data(), soff.data(), sizeof(long) * nsoff);
    double *hb = reinterpret_cast<double *>(h.data() + sizeof(long) * nsoff);
    for (int k = 0; k < nsys; k++)
      for (int j = 0; j < 3; j++) {
        hb[6 * k + j] = lbs[3 * k + j];
        hb[6 * k + 3 + j] = ubs[3 * k + j];
      }
    CK(stage_copy(ws, ws.sys_off.ptr, h.data(), h.size(), s));
  }
  const double *d_bounds =
      reinterpret_cast<const double *>(ws.sys_off.as<long>() + nsoff);
  // Level table on the device: per level its first node and count, then
  // the error word of the tree launches (bit 4: a non-finite coordinate,
  // others: a level launch gave up); set up by the extent launch.
  constexpr int kTab = 2 * (kMaxDepth + 3);
  CK(ws.lvl_table.reserve(sizeof(int) * (kTab + 1)));
  int *d_lvl_tab = ws.lvl_table.as<int>();
  // root cubes on the device (no host round trip): a non-finite coordinate
  // is reported with the first level counts
  CK(ws.part.reserve(sizeof(double) * 6 * (size_t)nsys * minmax_blocks(nsys)));
  // node arrays (grown by the level batches when needed); the extent
  // launch writes the roots
  const long cap0 = std::max<long>(1024, 4 * (n / pb.leafcap + 1) + 64 + nsys);
  if (ws.node_cap < cap0 || !ws.n_depth.ptr) CK(reserve_nodes(ws, cap0, 0));
  NodeArrays nd = ws.nodes();
  CK(launch_sys_minmax(pb.x, pb.y, pb.z, ws.sys_off.as<long>(), d_bounds,
                       nsys, ws.part.as<double>(), ws.sys_box.as<double>(),
                       d_lvl_tab, kTab, d_lvl_tab + kTab, nd, s));

  // ---- keys and sort (by system, then Morton key)
  CK(ws.keys_a.reserve(sizeof(uint64_t) * n));
  CK(ws.keys_b.reserve(sizeof(uint64_t) * n));
  CK(ws.idx_a.reserve(sizeof(int) * n));
  CK(ws.idx_b.reserve(sizeof(int) * n));
  DevBuf *pbufs[] = {&ws.xs, &ws.ys, &ws.zs, &ws.qs, &ws.pot_s};
  for (DevBuf *b : pbufs) CK(b->reserve(sizeof(double) * n));
  CK(ws.grad_s.reserve(sizeof(double) * 3 * n));
  int *d_total = ws.small.as<int>() + 64;
  CK(ws.packed.reserve(sizeof(double4) * n));
  CK(launch_keys(pb.x, pb.y, pb.z, pb.q, n, ws.sys_off.as<long>(), nsys,
                 ws.sys_box.as<double>(), ws.keys_a.as<uint64_t>(),
                 ws.idx_a.as<int>(), ws.packed.as<double4>(), s));
  CK(radix_sort(ws.keys_a.as<uint64_t>(), ws.idx_a.as<int>(),
                ws.keys_b.as<uint64_t>(), ws.idx_b.as<int>(), n, ws.hist, 8,
                s));
  if (nsys > 1) {
    // stable sort of the Morton order by system (2 or 4 byte passes)
    const int passes = nsys <= 65536 ? 2 : 4;
    CK(ws.skey.reserve(sizeof(uint64_t) * 2 * n));
    CK(ws.spos.reserve(sizeof(int) * 2 * n));
    uint64_t *sk = ws.skey.as<uint64_t>();
    int *sp = ws.spos.as<int>();
    CK(launch_sys_keys(ws.idx_a.as<int>(), n, ws.sys_off.as<long>(), nsys, sk,
                       sp, s));
    CK(radix_sort(sk, sp, sk + n, sp + n, n, ws.hist, passes, s));
    CK(launch_permute_keys(sp, n, ws.keys_a.as<uint64_t>(), ws.idx_a.as<int>(),
                           ws.keys_b.as<uint64_t>(), ws.idx_b.as<int>(), s));
    std::swap(ws.keys_a, ws.keys_b);
    std::swap(ws.idx_a, ws.idx_b);
  }
  CK(launch_gather(ws.idx_a.as<int>(), n, ws.packed.as<double4>(),
                   ws.xs.as<double>(), ws.ys.as<double>(), ws.zs.as<double>(),
                   ws.qs.as<double>(), s));
  const uint64_t *keys = ws.keys_a.as<uint64_t>();

  // ---- levels (level 0: one root per system)
  const ulonglong2 *d_stop = nullptr;
  if (pb.nstop > 0) {
    CK(ws.stop.reserve(sizeof(ulonglong2) * pb.nstop));
    CK(stage_copy(ws, ws.stop.ptr, pb.stop, sizeof(ulonglong2) * pb.nstop, s));
    d_stop = ws.stop.as<ulonglong2>();
  }
  // Levels in batches of kLevelBatch without a host round trip: the level
  // table lives on the device and grids are sized by upper bounds (a level
  // has at most 8 nodes per split node of the level above, and a split node
  // holds more than leafcap particles); one read per batch brings the
  // counts back.
  std::vector<int> lstart = {0}, lcount = {nsys};
  long nnodes = nsys;
  {
    constexpr int kLevelBatch = 4;
    const long max_split = n / ((long)pb.leafcap + 1);
    int *d_lvl = d_lvl_tab;
    long ub = nsys;  // bound on the count of level d
    bool done = false;
    // the first batch reaches the previous solve's depth (repeated solves of
    // similar systems then read the level counts back once)
    const int first_batch =
        std::max(kLevelBatch, std::min(kMaxDepth, ws.depth_hint) + 1);
    for (int d0 = 0; !done && d0 <= kMaxDepth;
         d0 += d0 == 0 ? first_batch : kLevelBatch) {
      const int d1 = std::min(
          kMaxDepth, d0 + (d0 == 0 ? first_batch : kLevelBatch) - 1);
      // bounds of the levels in this batch and of the nodes they add
      std::vector<long> ubs;
      long b = ub, add = 0, bmax = 1;  // bmax: the largest level bound
      for (int d = d0; d <= d1; d++) {
        ubs.push_back(b);
        bmax = std::max(bmax, b);
        b = std::min(8 * b, 8 * max_split);
        add += b;
      }
      if (nnodes + add > ws.node_cap) {
        CK(reserve_nodes(ws, std::max(2 * ws.node_cap, nnodes + add), nnodes));
        nd = ws.nodes();
      }
      for (int d = d0; d <= d1; d++) {
        const long ubd = ubs[d - d0];
        Lookback lb;  // (sized for the batch: one allocation at most)
        CK(lookback_take(ws, split_level_chunks(bmax), d_lvl + kTab, &lb, s));
        CK(launch_split_level(nd, d_lvl, d, ubd, keys, pb.leafcap, d_stop,
                              pb.nstop, lb, s));
      }
      int *h_tab = reinterpret_cast<int *>(static_cast<char *>(ws.host_small) +
                                           3072);
      CK(hipMemcpyAsync(h_tab, d_lvl, sizeof(int) * (kTab + 1),
                        hipMemcpyDeviceToHost, s));
      CK(hipStreamSynchronize(s));
      if (h_tab[kTab] & 4) {
        set_error("non-finite particle coordinates");
        return hipErrorInvalidValue;
      }
      if (h_tab[kTab] != 0) {
        set_error("tree build: a level launch gave up waiting (internal "
                  "error)");
        return hipErrorUnknown;
      }
      for (int d = d0; d <= d1; d++) {
        const int cnt = h_tab[2 * d + 3];
        if (cnt == 0) {
          done = true;
          break;
        }
        lstart.push_back(h_tab[2 * d + 2]);
        lcount.push_back(cnt);
        nnodes += cnt;
      }
      ub = std::min(8 * ubs.back(), 8 * max_split);
    }
  }
  const int depth = (int)lstart.size() - 1;
  const int nn = (int)nnodes;
  ws.depth_hint = depth;

  // ---- leaves
  // ---- leaves and internal nodes (the latter level by level, for the
  // M2M / L2L level products); one read brings both counts
  CK(ws.leaf_list.reserve(sizeof(int) * (nn + 1)));
  CK(ws.internal_list.reserve(sizeof(int) * (nn + 1)));
  CK(ws.split_ofs.reserve(sizeof(int) * (nn + 1)));
  // Far-field tiles (all forms but the plain one on its per-pair kernel):
  // 64 targets of one (level, octant) group each.
  const bool tiled = (pb.precision != 0 || dense64) && depth >= 2;
  const int ngroups = (kMaxDepth + 1) * 8;
  const int tfirst = tiled ? lstart[2] : 0;
  const int tcount = tiled ? (int)nn - lstart[2] : 0;
  // One device table of the level sizes and the far-field tile tables:
  // lcount[K], lstart[K] (zero past the depth), then (tiled) gofs[ngroups]
  // (first chunk of each (level, octant) group), tile_of[chunks] (tile of
  // each chunk), tile_oct[tiles], and the far-field workgroups' tiles of
  // both schedules; the level part is copied now, the rest once the group
  // sizes are known.
  constexpr int K = kMaxDepth + 2;
  const int tpw_max = std::max(f64_wg_tiles(p), mfma_wg_tiles(p));
  const size_t max_chunks = tiled ? (size_t)tcount / 64 + ngroups : 0;
  CK(ws.tile_meta.reserve(
      sizeof(int) * (2 * K + (tiled ? ngroups + 2 * max_chunks +
                                          2 * (max_chunks +
                                               (size_t)ngroups * tpw_max)
                                    : 0))));
  int *d_lcount = ws.tile_meta.as<int>(), *d_lstart = d_lcount + K;
  {
    int h[2 * K] = {};
    for (int d = 0; d <= depth; d++) {
      h[d] = lcount[d];
      h[K + d] = lstart[d];
    }
    CK(stage_copy(ws, d_lcount, h, sizeof(h), s));
  }
  CK(scan_one_launch(ws, nd.leaf, ws.split_ofs.as<int>(), nn,
                     d_total, s));
  int *d_lvl = reinterpret_cast<int *>(ws.small.as<char>() + 3584);
  int *h_lvl = reinterpret_cast<int *>(static_cast<char *>(ws.host_small) +
                                       3584);
  CK(launch_level_internal(ws.split_ofs.as<int>(), d_total, d_lstart,
                           depth + 1, nn, d_lvl, s));
  CK(hipMemcpyAsync(h_lvl, d_lvl, sizeof(int) * (depth + 2),
                    hipMemcpyDeviceToHost, s));
  // The group sizes are counted here, so that they come back with the leaf
  // counts.
  int *d_meta = d_lstart;  // lstart[K], gofs, tile_of
  int *trank = nullptr, *d_tile_oct = nullptr, *d_tile_perm = nullptr;
  const int *h_cnt =
      reinterpret_cast<int *>(static_cast<char *>(ws.host_small) + 4096);
  if (tiled) {
    CK(ws.tile_cnt.reserve(sizeof(int) * ngroups));
    int *cnt = ws.tile_cnt.as<int>();
    CK(ws.tile_rank.reserve(sizeof(int) * (8 * (size_t)tcount + 1)));
    trank = ws.tile_rank.as<int>();
    CK(launch_octant_flags(nd, tfirst, tcount, trank, s));
    CK(scan_one_launch(ws, trank, trank, 8L * tcount, trank + 8L * tcount,
                       s));
    CK(launch_group_counts(tfirst, tcount, depth, d_meta, trank, cnt, s));
    CK(hipMemcpyAsync(const_cast<int *>(h_cnt), cnt, sizeof(int) * ngroups,
                      hipMemcpyDeviceToHost, s));
  }
  CK(hipStreamSynchronize(s));
  const int nleaves = h_lvl[0];
  std::vector<int> icount(depth + 1), ibase(depth + 1);
  for (int d = 0, b = 0; d <= depth; d++) {
    icount[d] = h_lvl[1 + d];
    ibase[d] = b;
    b += icount[d];
  }
  CK(launch_leaf_compact(nd, nn, ws.split_ofs.as<int>(), ws.leaf_list.as<int>(),
                         ws.internal_list.as<int>(), s));
  const int *leaves = ws.leaf_list.as<int>();
  const int ninternal = nn - nleaves;
  CK(ws.trans_rows.reserve(sizeof(TransRow) * std::max(1, ninternal)));
  CK(launch_trans_rows(nd, ws.internal_list.as<int>(), ninternal,
                       ws.trans_rows.as<TransRow>(), s));
  const TransRow *trows = ws.trans_rows.as<TransRow>();
  if (pb.timings) CK(hipEventRecord(ws.ev[1], s));
  if (pb.tree_only) {
    CK(hipStreamSynchronize(s));
    int lb_err = 0;
    CK(read_int(ws, lb_err_word(ws), &lb_err));
    if (lb_err != 0) {
      CK(launch_fill_int(lb_err_word(ws), 1, 0, s));
      set_error("a scan launch gave up waiting (internal error)");
      return hipErrorUnknown;
    }
    S.depth = depth;
    S.nnodes = nn;
    S.nleaves = nleaves;
    if (ti) {
      CK(read_root_cube(ws, ti->lo, &ti->width));
      ti->nsys = nsys;
      ti->depth = depth;
      ti->nnodes = nn;
      ti->nleaves = nleaves;
      ti->nfar = ti->nnear = 0;
      ti->level_start = lstart;
      ti->level_count = lcount;
    }
    ws.far_csr_nodes = -1;
    ws.semifar_nodes = -1;
    return hipSuccess;
  }

  // ---- colleagues, level by level
  CK(ws.colleagues.reserve(sizeof(int) * 27 * (size_t)nn));
  int *col = ws.colleagues.as<int>();
  CK(launch_colleagues_all(nd, nn, d_lstart, d_lcount,
                           col, s));
  // ---- far list (CSR; the mixed form reads its far field straight from
  // the colleagues and builds this only when the lists are exported)
  int nfar = 0;
  ws.far_csr_nodes = -1;
  if (pb.precision == 0 && !tiled) CK(build_far_lists(ws, nn, &nfar));
  // ---- near list (and, with a semifar threshold, the semifar lists)
  const bool semifar = pb.semifar >= 0;
  CK(ws.u_cnt.reserve(sizeof(int) * nleaves));
  CK(ws.u_ofs.reserve(sizeof(int) * (nleaves + 1)));
  CK(ws.u_pairs.reserve(sizeof(long long) * (nleaves + 1)));
  int nnear = 0, nsemi = 0;
  ws.semifar_nodes = -1;
  // With the ranges per leaf of an earlier solve the near-list arrays are
  // sized from it, so the second pass follows the count and its scan with
  // no read of the total on the host; a solve whose lists outgrow that size
  // is run again, sized exactly.  (Not on a fixed arena: the margin and a
  // second run would cost it memory.)
  const bool near_hinted = !semifar && ws.near_per_leaf > 0. && !g_arena;
  long near_cap = LONG_MAX;
  int *d_near_total = reinterpret_cast<int *>(ws.small.as<char>() + 1020);
  if (near_hinted)
    near_cap = (long)std::min<double>(
        (double)INT32_MAX, 1.25 * ws.near_per_leaf * nleaves + 1024.);
  if (!semifar) {
    CK(launch_ulist(nd, nleaves, leaves, col, 0, ws.u_cnt.as<int>(), nullptr,
                    nullptr, nullptr, nullptr, LONG_MAX, nullptr, s));
  } else {
    CK(ws.w_cnt.reserve(sizeof(int) * nleaves));
    CK(ws.w_ofs.reserve(sizeof(int) * (nleaves + 1)));
    CK(ws.s_part.reserve(sizeof(long long) * (nleaves + 1)));
    CK(ws.x_cnt.reserve(sizeof(int) * nn));
    CK(ws.x_ofs.reserve(sizeof(int) * (nn + 1)));
    CK(launch_ulist_semifar(nd, nleaves, leaves, col, pb.semifar, 0,
                            ws.u_cnt.as<int>(), nullptr, nullptr, nullptr,
                            nullptr, ws.w_cnt.as<int>(), nullptr, nullptr,
                            nullptr, s));
    CK(launch_xlist(nd, nn, col, pb.semifar, 0, ws.x_cnt.as<int>(), nullptr,
                    nullptr, s));
    CK(scan_one_launch(ws, ws.w_cnt.as<int>(), ws.w_ofs.as<int>(), nleaves,
                       ws.w_ofs.as<int>() + nleaves, s));
    CK(scan_one_launch(ws, ws.x_cnt.as<int>(), ws.x_ofs.as<int>(), nn,
                       ws.x_ofs.as<int>() + nn, s));
  }
  CK(scan_one_launch(ws, ws.u_cnt.as<int>(), ws.u_ofs.as<int>(), nleaves,
                     ws.u_ofs.as<int>() + nleaves, s));
  // ---- tile tables from the group counts (host, while the lists run);
  // within a level the tiles take chunk j of every octant in turn, so that
  // consecutive tiles cover the same region
  int ntiles = 0, nwg = 0;
  if (tiled) {
    std::vector<int> meta(ngroups, 0), toct, perm;  // gofs, then tile_of
    int *gofs = meta.data();
    int nchunks = 0;
    for (int g = 0; g < ngroups; g++) {
      gofs[g] = nchunks;
      nchunks += (h_cnt[g] + 63) / 64;
    }
    meta.resize(ngroups + nchunks);
    gofs = meta.data();
    int *tile_of = gofs + ngroups;
    for (int d = 2; d <= depth; d++) {
      int maxch = 0;
      for (int o = 0; o < 8; o++)
        maxch = std::max(maxch, (h_cnt[d * 8 + o] + 63) / 64);
      for (int j = 0; j < maxch; j++)
        for (int o = 0; o < 8; o++)
          if (j < (h_cnt[d * 8 + o] + 63) / 64) {
            tile_of[gofs[d * 8 + o] + j] = ntiles++;
            toct.push_back(o);
          }
    }
    CK(ws.tile_tgt.reserve(sizeof(int) * 64 * (size_t)std::max(1, ntiles)));
    CK(ws.tile_src.reserve(sizeof(int) * 64 * (size_t)kOctOffsets *
                           std::max(1, ntiles)));
    CK(ws.node_slot.reserve(sizeof(int) * nn));
    // MFMA far field: workgroups of consecutive chunks of one (level,
    // octant) group (they share every offset code).  Two orders of the same
    // workgroups: level, chunk block, octant (schedule 0: neighbouring
    // workgroups are neighbours in space) and octant, level, chunk block
    // (schedule 1: neighbouring workgroups share their operators).
    if (pb.precision >= 2 || pb.precision == 0) {
      std::vector<int> wgt, wgt1;
      const int tpw = pb.precision == 0 ? f64_wg_tiles(p) : mfma_wg_tiles(p);
      for (int d = 2; d <= depth; d++) {
        int maxch = 0;
        for (int o = 0; o < 8; o++)
          maxch = std::max(maxch, (h_cnt[d * 8 + o] + 63) / 64);
        for (int jb = 0; jb < maxch; jb += tpw)
          for (int o = 0; o < 8; o++) {
            const int nch = (h_cnt[d * 8 + o] + 63) / 64;
            if (jb >= nch) continue;
            for (int j = jb; j < jb + tpw; j++)
              wgt.push_back(j < nch ? tile_of[gofs[d * 8 + o] + j] : -1);
          }
      }
      for (int o = 0; o < 8; o++)
        for (int d = 2; d <= depth; d++) {
          const int nch = (h_cnt[d * 8 + o] + 63) / 64;
          for (int jb = 0; jb < nch; jb += tpw)
            for (int j = jb; j < jb + tpw; j++)
              wgt1.push_back(j < nch ? tile_of[gofs[d * 8 + o] + j] : -1);
        }
      nwg = (int)(wgt.size() / tpw);
      perm.swap(wgt);
      perm.insert(perm.end(), wgt1.begin(), wgt1.end());
    }
    // gofs, tile_of, tile_oct, the workgroups' tiles: one copy
    meta.insert(meta.end(), toct.begin(), toct.end());
    meta.insert(meta.end(), perm.begin(), perm.end());
    if (meta.size() > (size_t)ngroups + 2 * max_chunks +
                          2 * (max_chunks + (size_t)ngroups * tpw_max)) {
      set_error("far-field tile tables beyond their bound (internal error)");
      return hipErrorUnknown;
    }
    d_tile_oct = d_meta + K + ngroups + nchunks;
    d_tile_perm = d_tile_oct + ntiles;
    CK(stage_copy(ws, d_meta + K, meta.data(), sizeof(int) * meta.size(), s));
  }
  if (near_hinted) {
    CK(ws.u_beg.reserve(sizeof(int) * (size_t)near_cap));
    CK(ws.u_end.reserve(sizeof(int) * (size_t)near_cap));
  } else {
    CK(read_int(ws, ws.u_ofs.as<int>() + nleaves, &nnear));
    CK(ws.u_beg.reserve(sizeof(int) * std::max(1, nnear)));
    CK(ws.u_end.reserve(sizeof(int) * std::max(1, nnear)));
  }
  if (!semifar) {
    CK(launch_ulist(nd, nleaves, leaves, col, 1, ws.u_cnt.as<int>(),
                    ws.u_ofs.as<int>(), ws.u_beg.as<int>(), ws.u_end.as<int>(),
                    ws.u_pairs.as<long long>(), near_cap, d_near_total, s));
  } else {
    int nx = 0;
    CK(read_int(ws, ws.w_ofs.as<int>() + nleaves, &nsemi));
    CK(read_int(ws, ws.x_ofs.as<int>() + nn, &nx));
    if (nx != nsemi) {
      set_error("semifar lists disagree (internal error)");
      return hipErrorUnknown;
    }
    CK(ws.w_node.reserve(sizeof(int) * std::max(1, nsemi)));
    CK(ws.x_leaf.reserve(sizeof(int) * std::max(1, nsemi)));
    CK(launch_ulist_semifar(nd, nleaves, leaves, col, pb.semifar, 1,
                            ws.u_cnt.as<int>(), ws.u_ofs.as<int>(),
                            ws.u_beg.as<int>(), ws.u_end.as<int>(),
                            ws.u_pairs.as<long long>(), ws.w_cnt.as<int>(),
                            ws.w_ofs.as<int>(), ws.w_node.as<int>(),
                            ws.s_part.as<long long>(), s));
    CK(launch_xlist(nd, nn, col, pb.semifar, 1, ws.x_cnt.as<int>(),
                    ws.x_ofs.as<int>(), ws.x_leaf.as<int>(), s));
    ws.semifar_nodes = nn;
  }
  long long *d_pairs = reinterpret_cast<long long *>(ws.small.as<char>() + 1024);
  // (d_pairs and d_spart are each followed by the 64 partial sums of
  // launch_sum_ll)
  int *d_far = reinterpret_cast<int *>(ws.small.as<char>() + 2048);
  long long *d_spart =
      reinterpret_cast<long long *>(ws.small.as<char>() + 2560);
  CK(launch_sum_ll(ws.u_pairs.as<long long>(), nleaves, d_pairs, s));
  if (semifar)
    CK(launch_sum_ll(ws.s_part.as<long long>(), nleaves, d_spart, s));

  // ---- near field on the second stream: it needs the sorted particles and
  // the near lists only, so it overlaps the far-field tiles, the upward
  // pass, the far field and the downward pass (measured better than starting
  // it after the upward pass).  L2P waits for it (P2P writes the sorted
  // results, M2P and L2P add to them).
  CK(ws.pleaf.reserve(sizeof(int) * std::max<long>(1, n)));
  CK(hipEventRecord(ws.ev_fork, s));
  CK(hipStreamWaitEvent(ws.aux, ws.ev_fork, 0));
  if (pb.timings) CK(hipEventRecord(ws.ev[7], ws.aux));
  CK(launch_p2p(pb.precision, leaves, nleaves, nd, ws.u_ofs.as<int>(),
                ws.u_beg.as<int>(), ws.u_end.as<int>(), ws.xs.as<double>(),
                ws.ys.as<double>(), ws.zs.as<double>(), ws.qs.as<double>(),
                ws.pot_s.as<double>(), ws.grad_s.as<double>(), ws.aux));
  if (pb.timings) CK(hipEventRecord(ws.ev_near, ws.aux));
  CK(launch_particle_leaf(nd, leaves, nleaves, ws.pleaf.as<int>(), ws.aux));
  CK(hipEventRecord(ws.ev_join, ws.aux));
  // ---- tiles of the far field: targets and sources of every tile
  if (tiled) {
    const int *tofs = d_meta + kMaxDepth + 2;
    CK(launch_fill_int(ws.tile_tgt.as<int>(), 64L * ntiles, -1, s));
    CK(launch_tile_fill(nd, tfirst, tcount, trank, d_meta, tofs,
                        tofs + ngroups, ws.tile_tgt.as<int>(),
                        ws.node_slot.as<int>(), s));
    CK(launch_fill_int(d_far, 1, 0, s));
    CK(launch_tile_src(nd, ntiles, ws.tile_tgt.as<int>(),
                       d_tile_oct, col, nd.cmap,
                       ws.tab.oct_codes.as<short>(), ws.tile_src.as<int>(),
                       d_far, s));
  }
  if (pb.timings) CK(hipEventRecord(ws.ev[2], s));

  // ---- upward pass
  const int mc = tri(p), lc = tri(p + 1);
  CK(ws.mult.reserve(sizeof(double2) * (size_t)nn * mc));
  CK(ws.loc.reserve(sizeof(double2) * (size_t)nn * lc));
  double2 *M = ws.mult.as<double2>();
  double2 *L = ws.loc.as<double2>();
  OpTables tb = ws.tab.view();
  if (pb.timings) CK(hipEventRecord(ws.evk[0], s));
  // reduced-precision forms: FP32 P2M and L2P for the leaves below the top
  // two levels (the leaf list is in node order, so the leaves of levels 0
  // and 1 come first), FP64 for those; the plain form FP64 throughout
  int top = nleaves;
  if (pb.precision != 0) {
    top = 0;
    for (int d = 0; d < 2 && d <= depth; d++) top += lcount[d] - icount[d];
  }
  {
    if (top > 0)
      CK(launch_p2m(tb, nd, leaves, top, ws.xs.as<double>(),
                    ws.ys.as<double>(), ws.zs.as<double>(), ws.qs.as<double>(),
                    M, s, false));
    if (top < nleaves)
      CK(launch_p2m(tb, nd, leaves + top, nleaves - top, ws.xs.as<double>(),
                    ws.ys.as<double>(), ws.zs.as<double>(), ws.qs.as<double>(),
                    M, s, true));
  }
  if (pb.timings) CK(hipEventRecord(ws.evk[1], s));
  for (int d = depth - 1; d >= 0; d--) {
    if (pb.flav_m2m != kFlavRotate)
      CK(launch_flavour_level(tb, kFlavM2M, pb.flav_m2m, nd, trows + ibase[d],
                              icount[d], M, s));
    else
      CK(launch_m2m_level(tb, trows + ibase[d], icount[d], M, s));
  }
  if (pb.timings) CK(hipEventRecord(ws.evk[2], s));

  // ---- far field (the M2L timer brackets the M2L kernel alone)
  if (pb.precision == 0 && !tiled) {
    if (pb.timings) CK(hipEventRecord(ws.ev[3], s));
    if (pb.flav_m2l != kFlavRotate)
      CK(launch_flavour_m2l(tb, pb.flav_m2l, nd, nn, ws.v_ofs.as<int>(),
                            ws.v_src.as<int>(), M, L, s));
    else
      CK(launch_m2l(tb, nd, nn, ws.v_ofs.as<int>(), ws.v_src.as<int>(),
                    ws.v_code.as<short>(), M, L, s));
    if (pb.timings) CK(hipEventRecord(ws.ev[4], s));
  } else if (pb.precision >= 2) {
    const MfmaDims D = mfma_dims(p);
    const int first = depth >= 2 ? lstart[2] : nn;
    const int nsplit = m2l_mfma_split(p);
    const size_t wgt_len = (size_t)nwg * mfma_wg_tiles(p);
    auto run = [&](int sched) {
      return launch_m2l_mfma(pb.precision == 2 ? 3 : 6, p, nwg,
                             d_tile_perm + sched * wgt_len, ntiles,
                             ws.tile_src.as<int>(), d_tile_oct,
                             ws.tab.oct_codes.as<short>(), ws.tab.mfma_ops.ptr,
                             ws.mult_bf.ptr, nn, nsplit, sched,
                             ws.loc_mfma.as<float>(), s);
    };
    // the tuner times the far field with its slice reduction
    auto run_all = [&](int sched) {
      hipError_t e = run(sched);
      if (e != hipSuccess) return e;
      return launch_unscale_mfma(nd, p, first, ntiles, ws.tile_tgt.as<int>(),
                                 ws.loc_mfma.as<float>(), nsplit, L, s);
    };
    CK(ws.mult_bf.reserve(16 * (size_t)(nn + 1) * mfma_mult_units(p)));
    const size_t slice = sizeof(float) * std::max(1, ntiles) * 64 * D.ro;
    CK(launch_mfma_pack_mult(nd, nn, p, pb.precision == 2 ? 2 : 3, M,
                             ws.mult_bf.ptr, s));
    CK(ws.loc_mfma.reserve(slice * nsplit));
    int sched = mfma_default_schedule(0);
    if (pb.autotune && ntiles > 0)
      CK(tune_schedule(ws, tune_key(pb.precision, p, nwg), kMfmaSchedules,
                       run_all, &sched));
    if (pb.timings) CK(hipEventRecord(ws.ev[3], s));
    if (ntiles > 0) CK(run(sched));
    if (pb.timings) CK(hipEventRecord(ws.ev[4], s));
    CK(launch_unscale_mfma(nd, p, first, ntiles, ws.tile_tgt.as<int>(),
                           ws.loc_mfma.as<float>(), nsplit, L, s));
    S.m2l_split = nsplit;
    S.m2l_sched = sched;
    S.m2l_tiles = ntiles;
  } else if (pb.precision == 0) {
    // plain form, dense FP64 products on MFMA (tiled levels >= 2)
    const int first = lstart[2];
    const int nsplit = m2l_f64_split(p);
    const size_t wgt_len = (size_t)nwg * f64_wg_tiles(p);
    auto run = [&](int sched) {
      return launch_m2l_f64(p, nwg, d_tile_perm + sched * wgt_len,
                            ntiles, ws.tile_src.as<int>(),
                            d_tile_oct,
                            ws.tab.oct_codes.as<short>(),
                            ws.tab.f64_ops.as<double>(), ws.mult64.as<double>(),
                            nn, nsplit, sched, ws.loc64.as<double>(), s);
    };
    auto run_all = [&](int sched) {
      hipError_t e = run(sched);
      if (e != hipSuccess) return e;
      return launch_unscale_mfma(nd, p, first, ntiles, ws.tile_tgt.as<int>(),
                                 ws.loc64.as<double>(), nsplit, L, s);
    };
    CK(ws.mult64.reserve(sizeof(double) * (size_t)(nn + 1) *
                         f64_mult_doubles(p)));
    CK(launch_f64_pack_mult(nd, nn, p, M, ws.mult64.as<double>(), s));
    const size_t slice = sizeof(double) * std::max(1, ntiles) * 64 *
                         (size_t)mfma_dims(p).ro;
    CK(ws.loc64.reserve(slice * nsplit));
    int sched = mfma_default_schedule(1);
    if (pb.autotune && ntiles > 0)
      CK(tune_schedule(ws, tune_key(0, p, nwg), kMfmaSchedules, run_all,
                       &sched));
    if (pb.timings) CK(hipEventRecord(ws.ev[3], s));
    if (ntiles > 0) CK(run(sched));
    if (pb.timings) CK(hipEventRecord(ws.ev[4], s));
    CK(launch_unscale_mfma(nd, p, first, ntiles, ws.tile_tgt.as<int>(),
                           ws.loc64.as<double>(), nsplit, L, s));
    S.m2l_split = nsplit;
    S.m2l_sched = sched;
    S.m2l_tiles = ntiles;
  } else {
    CK(ws.mult32.reserve(sizeof(float2) * (size_t)nn * mc));
    const int nsplit = m2l_choose_split(p);
    const size_t slice = sizeof(float2) * std::max(1, ntiles) * 64 * lc;
    CK(launch_scale_mult(1, nd, nn, p, M, ws.mult32.ptr, s));
    M2LTileArgs ta;
    ta.tile_src = ws.tile_src.as<int>();
    ta.tile_oct = d_tile_oct;
    ta.oct_codes = ws.tab.oct_codes.as<short>();
    ta.dir = ws.tab.dir.as<DirEntry>();
    ta.ephi = ws.tab.ephi32.ptr;
    ta.rot = ws.tab.rot32.ptr;
    ta.vfac = ws.tab.vfac32.ptr;
    ta.tile_major = m2l_tile_major();
    const int first = depth >= 2 ? lstart[2] : nn;
    ta.nsplit = nsplit;
    auto run = [&](int major) {
      M2LTileArgs a = ta;
      a.tile_major = major;
      return launch_m2l_tile(1, p, ntiles, a, ws.mult32.ptr, ws.loc32.ptr, s);
    };
    auto run_all = [&](int major) {
      hipError_t e = run(major);
      if (e != hipSuccess) return e;
      return launch_unscale_local(1, nd, p, first, ntiles,
                                  ws.tile_tgt.as<int>(), ws.loc32.ptr, nsplit,
                                  L, s);
    };
    CK(ws.loc32.reserve(slice * nsplit));
    int major = ta.tile_major;
    if (pb.autotune && ntiles > 0)
      CK(tune_schedule(ws, tune_key(1, p, ntiles), 2, run_all, &major));
    if (pb.timings) CK(hipEventRecord(ws.ev[3], s));
    if (ntiles > 0) CK(run(major));
    if (pb.timings) CK(hipEventRecord(ws.ev[4], s));
    CK(launch_unscale_local(1, nd, p, first, ntiles, ws.tile_tgt.as<int>(),
                            ws.loc32.ptr, nsplit, L, s));
    S.m2l_split = nsplit;
    S.m2l_sched = major;
    S.m2l_tiles = ntiles;
  }
  // ---- semifar P2L into the locals (before they move down)
  if (semifar && nsemi > 0)
    CK(launch_p2l(tb, nd, nn, ws.x_ofs.as<int>(), ws.x_leaf.as<int>(),
                  ws.xs.as<double>(), ws.ys.as<double>(), ws.zs.as<double>(),
                  ws.qs.as<double>(), L, s));
  // ---- downward pass
  // locals of levels 0 and 1 receive nothing without semifar pairs (far
  // interactions start at level 2), so their translations would add zeros
  if (pb.timings) CK(hipEventRecord(ws.evk[3], s));
  for (int d = semifar ? 1 : 3; d <= depth; d++) {
    if (pb.flav_l2l != kFlavRotate)
      CK(launch_flavour_level(tb, kFlavL2L, pb.flav_l2l, nd,
                              trows + ibase[d - 1], icount[d - 1], L, s));
    else
      CK(launch_l2l_level(tb, trows + ibase[d - 1], icount[d - 1], L, s));
  }
  if (pb.timings) CK(hipEventRecord(ws.ev[5], s));
  // ---- join the near field, then M2P and L2P add to its results
  CK(hipStreamWaitEvent(s, ws.ev_join, 0));
  if (pb.timings) CK(hipEventRecord(ws.evk[4], s));
  if (semifar && nsemi > 0)
    CK(launch_m2p(tb, nd, leaves, nleaves, ws.w_ofs.as<int>(),
                  ws.w_node.as<int>(), ws.xs.as<double>(), ws.ys.as<double>(),
                  ws.zs.as<double>(), M, ws.pot_s.as<double>(),
                  ws.grad_s.as<double>(), s));
  // L2P adds the last term of every particle (each L2P launch takes its
  // own particles) and writes the sums in the caller's order
  const int *perm = ws.idx_a.as<int>();
  if (pb.precision == 0) {
    CK(launch_l2p(tb, nd, ws.pleaf.as<int>(), (int)n, ws.xs.as<double>(),
                  ws.ys.as<double>(), ws.zs.as<double>(), L,
                  ws.pot_s.as<double>(), ws.grad_s.as<double>(), s, 0, perm,
                  pb.pot, pb.grad));
  } else {
    CK(launch_l2p(tb, nd, ws.pleaf.as<int>(), (int)n, ws.xs.as<double>(),
                  ws.ys.as<double>(), ws.zs.as<double>(), L,
                  ws.pot_s.as<double>(), ws.grad_s.as<double>(), s, 1, perm,
                  pb.pot, pb.grad));
    if (top > 0)
      CK(launch_l2p(tb, nd, ws.pleaf.as<int>(), (int)n, ws.xs.as<double>(),
                    ws.ys.as<double>(), ws.zs.as<double>(), L,
                    ws.pot_s.as<double>(), ws.grad_s.as<double>(), s, 2, perm,
                    pb.pot, pb.grad));
  }
  if (pb.timings) CK(hipEventRecord(ws.evk[5], s));
  if (pb.timings) CK(hipEventRecord(ws.ev[6], s));

  long long *hp = reinterpret_cast<long long *>(
      static_cast<char *>(ws.host_small) + 1024);
  int *hfar = reinterpret_cast<int *>(static_cast<char *>(ws.host_small) +
                                      2048);
  // the look-back error word and the near-list total sit just before the
  // pair count
  int *herr = reinterpret_cast<int *>(static_cast<char *>(ws.host_small) +
                                      1016);
  CK(hipMemcpyAsync(herr, lb_err_word(ws), 8 + sizeof(long long),
                    hipMemcpyDeviceToHost, s));
  long long *hsp = reinterpret_cast<long long *>(
      static_cast<char *>(ws.host_small) + 2560);
  *hsp = 0;
  if (semifar)
    CK(hipMemcpyAsync(hsp, d_spart, sizeof(long long), hipMemcpyDeviceToHost,
                      s));
  if (tiled)
    CK(hipMemcpyAsync(hfar, d_far, sizeof(int), hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  if (*herr != 0) {
    CK(launch_fill_int(lb_err_word(ws), 1, 0, s));
    set_error("a scan launch gave up waiting (internal error)");
    return hipErrorUnknown;
  }
  if (near_hinted) {
    nnear = herr[1];
    if (nnear > near_cap) {  // lists outgrew their arrays: again, sized
      ws.near_per_leaf = 0.;
      return solve(ws, pb, st, ti);
    }
  }
  if (!semifar) ws.near_per_leaf = (double)nnear / std::max(1, nleaves);
  if (pb.precision != 0 || tiled) nfar = tiled ? *hfar : 0;
  S.zero = nn;
  S.p2p = (long)(*hp / 2);
  S.p2m = n;
  S.m2m = nn - nsys;
  S.m2l = nfar;
  S.l2l = nn - nsys;
  S.l2p = n;
  S.p2l = (long)*hsp;
  S.m2p = (long)*hsp;
  S.semifar = nsemi;
  S.depth = depth;
  S.nnodes = nn;
  S.nleaves = nleaves;
  if (pb.timings) {
    CK(hipEventElapsedTime(&S.t_tree, ws.ev[0], ws.ev[1]));
    CK(hipEventElapsedTime(&S.t_lists, ws.ev[1], ws.ev[2]));
    // up: P2M, M2M (and the FP32 scaling); down: L2L (and unscaling)
    CK(hipEventElapsedTime(&S.t_up, ws.ev[2], ws.ev[3]));
    CK(hipEventElapsedTime(&S.t_m2l, ws.ev[3], ws.ev[4]));
    CK(hipEventElapsedTime(&S.t_down, ws.ev[4], ws.ev[5]));
    // near: the P2P on the second stream (it overlaps the tiles, up, m2l
    // and down, so the phases do not add up to the total)
    CK(hipEventElapsedTime(&S.t_near, ws.ev[7], ws.ev_near));
    CK(hipEventElapsedTime(&S.t_total, ws.ev[0], ws.ev[6]));
    CK(hipEventElapsedTime(&S.t_p2m, ws.evk[0], ws.evk[1]));
    CK(hipEventElapsedTime(&S.t_m2m, ws.evk[1], ws.evk[2]));
    CK(hipEventElapsedTime(&S.t_l2l, ws.evk[3], ws.ev[5]));
    CK(hipEventElapsedTime(&S.t_l2p, ws.evk[4], ws.evk[5]));
  }
  if (ti) {
    CK(read_root_cube(ws, ti->lo, &ti->width));
    ti->nsys = nsys;
    ti->depth = depth;
    ti->nnodes = nn;
    ti->nleaves = nleaves;
    ti->nfar = nfar;
    ti->nnear = nnear;
    ti->level_start = lstart;
    ti->level_count = lcount;
  }
  return hipSuccess;
}

}  // namespace aran
