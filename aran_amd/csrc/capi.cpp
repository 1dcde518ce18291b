// capi.cpp -- the C-ABI of the engine: LibAran's AranSolver3d /
// AranDevelopment3d API (aran/aransolver3d.h, aran/arandevelopment3d.h), the
// LibVsg names it exposes, and the device-resident aran_fmm3d_* entry points.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

extern "C" {
#include <aran/aran.h>
#include <aran/aranfmm3d.h>
#include <aran/arandevelopment3d.h>
#include <aran/aransolver3d.h>
#include <vsg/vsgd.h>
}

#include "engine.h"

using aran::DevBuf;

namespace {

void critical(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  aran::set_error(buf);
  fprintf(stderr, "** (aran) CRITICAL **: %s\n", buf);
}

int code_of_err(hipError_t e) {
  if (e == hipSuccess) return 0;
  if (e == hipErrorInvalidValue) return ARAN_FMM3D_EINVAL;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return ARAN_FMM3D_ENOMEM;
  return ARAN_FMM3D_EHIP;
}

// Pinned host staging buffer.
struct HostBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  hipError_t reserve(size_t n) {
    if (n <= bytes && ptr) return hipSuccess;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
    size_t want = std::max<size_t>(n + n / 4, 256);
    hipError_t e = hipHostMalloc(&ptr, want);
    if (e == hipSuccess) bytes = want;
    return e;
  }
  ~HostBuf() {
    if (ptr) (void)hipHostFree(ptr);
  }
  template <class T> T *as() { return static_cast<T *>(ptr); }
};

constexpr GType kDevelopmentType = 0x41524e33;  // "ARN3"

// f(begin, end) over [0, n) on up to 16 host threads, at least 64K items
// each (the gather from and the scatter to the caller's particle structs:
// pointer-chasing loops that would otherwise take longer than the solve).
template <class F> void parallel_ranges(long n, F f) {
  constexpr long kMinItems = 1L << 16;
  const long hw = std::max(1u, std::thread::hardware_concurrency());
  const long nt = std::max(1L, std::min({hw, 16L, n / kMinItems}));
  if (nt <= 1) {
    f(0L, n);
    return;
  }
  const long chunk = (n + nt - 1) / nt;
  std::vector<std::thread> th;
  th.reserve(nt);
  for (long t = 1; t < nt; t++) {
    const long a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) continue;
    try {
      th.emplace_back(f, a, b);
    } catch (...) {  // no thread to be had: this range on the caller's
      f(a, b);
    }
  }
  f(0L, std::min(n, chunk));
  for (auto &x : th) x.join();
}

}  // namespace

// ================================================================ aran.h

extern "C" void aran_init(void) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    critical("aran_init: no HIP device available (%s)", hipGetErrorString(e));
    return;
  }
  (void)hipFree(nullptr);
}

extern "C" const gchar *aran_last_error(void) { return aran::last_error(); }

// ================================================================ vsgd.h

struct _VsgPRTree3d {
  VsgVector3d lb, ub;
  guint maxbox;
};

extern "C" guint vsg_vector3d_vector3d_locfunc(const VsgPoint3 candidate,
                                               const VsgVector3d *center) {
  const VsgVector3d *v = static_cast<const VsgVector3d *>(candidate);
  return (guint)(v->x >= center->x) | ((guint)(v->y >= center->y) << 1) |
         ((guint)(v->z >= center->z) << 2);
}

extern "C" gdouble vsg_vector3d_dist(const VsgPoint3 one,
                                     const VsgPoint3 other) {
  const VsgVector3d *a = static_cast<const VsgVector3d *>(one);
  const VsgVector3d *b = static_cast<const VsgVector3d *>(other);
  double dx = a->x - b->x, dy = a->y - b->y, dz = a->z - b->z;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

extern "C" VsgPRTree3d *vsg_prtree3d_new_full(const VsgVector3d *lbound,
                                              const VsgVector3d *ubound,
                                              VsgPoint3dLocFunc point_locfunc,
                                              VsgPoint3dDistFunc point_distfunc,
                                              VsgRegion3dLocFunc region_locfunc,
                                              guint maxbox) {
  if (!lbound || !ubound || maxbox < 1) {
    critical("vsg_prtree3d_new_full: bounds and maxbox >= 1 are required");
    return nullptr;
  }
  if ((point_locfunc && point_locfunc != vsg_vector3d_vector3d_locfunc) ||
      (point_distfunc && point_distfunc != vsg_vector3d_dist) ||
      region_locfunc) {
    critical("vsg_prtree3d_new_full: only vsg_vector3d_vector3d_locfunc and "
             "vsg_vector3d_dist on point structs are supported");
    return nullptr;
  }
  VsgPRTree3d *t = new VsgPRTree3d;
  t->lb = *lbound;
  t->ub = *ubound;
  t->maxbox = maxbox;
  return t;
}

extern "C" void vsg_prtree3d_free(VsgPRTree3d *prtree) { delete prtree; }

// ================================================================ device API

struct _AranFmm3d {
  aran::Workspace ws;
  aran::Stats st;
  aran::TreeInfo ti;
  bool timings = false;
  int precision = ARAN_FMM3D_PLAIN;
  long semifar = -1;
  bool autotune = false;
  long n_last = 0;
};

extern "C" AranFmm3d *aran_fmm3d_new(void *hip_stream) {
  AranFmm3d *c = new AranFmm3d;
  hipError_t e =
      aran::workspace_init(c->ws, static_cast<hipStream_t>(hip_stream));
  if (e != hipSuccess) {
    critical("aran_fmm3d_new: %s", aran::last_error());
    delete c;
    return nullptr;
  }
  return c;
}

extern "C" void aran_fmm3d_free(AranFmm3d *ctx) { delete ctx; }

extern "C" void *aran_fmm3d_stream(const AranFmm3d *ctx) {
  return ctx ? (void *)ctx->ws.stream : nullptr;
}

extern "C" int aran_fmm3d_solve(AranFmm3d *ctx, long n, const double *x,
                                const double *y, const double *z,
                                const double *q, const double *lb,
                                const double *ub, int order, int leafcap,
                                double *pot, double *grad) {
  if (!ctx || n < 0 || (n > 0 && (!x || !y || !z || !q)) || !lb || !ub) {
    aran::set_error("aran_fmm3d_solve: invalid argument");
    return ARAN_FMM3D_EINVAL;
  }
  aran::Problem pb;
  pb.n = n;
  pb.order = order;
  pb.leafcap = leafcap;
  for (int k = 0; k < 3; k++) {
    pb.lb[k] = lb[k];
    pb.ub[k] = ub[k];
  }
  pb.x = x;
  pb.y = y;
  pb.z = z;
  pb.q = q;
  pb.pot = pot;
  pb.grad = grad;
  pb.timings = ctx->timings;
  pb.precision = ctx->precision;
  pb.semifar = ctx->semifar;
  pb.autotune = ctx->autotune;
  hipError_t e = aran::solve(ctx->ws, pb, &ctx->st, &ctx->ti);
  ctx->n_last = (e == hipSuccess) ? n : 0;
  return code_of_err(e);
}

extern "C" int aran_fmm3d_solve_batch(AranFmm3d *ctx, int nsys,
                                      const long *offsets, const double *x,
                                      const double *y, const double *z,
                                      const double *q, const double *lbs,
                                      const double *ubs, int order, int leafcap,
                                      double *pot, double *grad) {
  if (!ctx || nsys < 1 || !offsets || !lbs || !ubs) {
    aran::set_error("aran_fmm3d_solve_batch: invalid argument");
    return ARAN_FMM3D_EINVAL;
  }
  const long n = offsets[nsys];
  if (n < 0 || (n > 0 && (!x || !y || !z || !q))) {
    aran::set_error("aran_fmm3d_solve_batch: invalid argument");
    return ARAN_FMM3D_EINVAL;
  }
  aran::Problem pb;
  pb.n = n;
  pb.order = order;
  pb.leafcap = leafcap;
  pb.nsys = nsys;
  pb.soff = offsets;
  pb.lbs = lbs;
  pb.ubs = ubs;
  pb.x = x;
  pb.y = y;
  pb.z = z;
  pb.q = q;
  pb.pot = pot;
  pb.grad = grad;
  pb.timings = ctx->timings;
  pb.precision = ctx->precision;
  pb.semifar = ctx->semifar;
  pb.autotune = ctx->autotune;
  hipError_t e = aran::solve(ctx->ws, pb, &ctx->st, &ctx->ti);
  ctx->n_last = (e == hipSuccess) ? n : 0;
  return code_of_err(e);
}

extern "C" int aran_fmm3d_get_stats(const AranFmm3d *ctx, long *o) {
  if (!ctx || !o) return ARAN_FMM3D_EINVAL;
  const aran::Stats &s = ctx->st;
  long v[15] = {s.zero,   s.p2p,     s.p2m,        s.m2m,         s.m2l,
                s.l2l,    s.l2p,     (long)s.depth, s.nnodes,     s.nleaves,
                ctx->ti.nfar, ctx->ti.nnear, s.p2l, s.m2p,        s.semifar};
  memcpy(o, v, sizeof v);
  return 0;
}

extern "C" int aran_fmm3d_set_autotune(AranFmm3d *ctx, int on) {
  if (!ctx) return ARAN_FMM3D_EINVAL;
  ctx->autotune = on != 0;
  return 0;
}

extern "C" int aran_fmm3d_set_semifar(AranFmm3d *ctx, long threshold) {
  if (!ctx) return ARAN_FMM3D_EINVAL;
  ctx->semifar = threshold < 0 ? -1 : threshold;
  return 0;
}

extern "C" long aran_fmm3d_semifar_threshold_auto(int order) {
  if (order < 1 || order > aran::kMaxOrder) return -1;
  // (P2L + M2P per particle) / P2P per pair, in algorithmic flops:
  // P2L 14 and M2P 30 per local coefficient, P2P 20 per pair
  const long lc = (long)(order + 1) * (order + 2) / 2;
  return (44 * lc + 19) / 20;
}


extern "C" int aran_fmm3d_set_precision(AranFmm3d *ctx, int precision) {
  if (!ctx || precision < ARAN_FMM3D_PLAIN || precision > ARAN_FMM3D_BF16X6) {
    aran::set_error("aran_fmm3d_set_precision: invalid precision");
    return ARAN_FMM3D_EINVAL;
  }
  ctx->precision = precision;
  return 0;
}

extern "C" int aran_fmm3d_set_timings(AranFmm3d *ctx, int on) {
  if (!ctx) return ARAN_FMM3D_EINVAL;
  ctx->timings = on != 0;
  return 0;
}

extern "C" int aran_fmm3d_get_timings(const AranFmm3d *ctx, float *o) {
  if (!ctx || !o) return ARAN_FMM3D_EINVAL;
  const aran::Stats &s = ctx->st;
  float v[7] = {s.t_tree, s.t_lists, s.t_up, s.t_m2l, s.t_down, s.t_near,
                s.t_total};
  memcpy(o, v, sizeof v);
  return 0;
}

template <class T>
static hipError_t d2h(T *dst, const DevBuf &src, size_t count,
                      hipStream_t s) {
  if (count == 0) return hipSuccess;
  return hipMemcpyAsync(dst, src.ptr, sizeof(T) * count, hipMemcpyDeviceToHost,
                        s);
}

extern "C" int aran_fmm3d_export_tree(const AranFmm3d *ctx, long *perm,
                                      int *nodes, double *centers, double *lo,
                                      double *width) {
  if (!ctx) return ARAN_FMM3D_EINVAL;
  const aran::Workspace &ws = ctx->ws;
  hipStream_t s = ws.stream;
  long n = ctx->n_last, nn = ctx->ti.nnodes;
  std::vector<int> tmp((size_t)std::max(n, nn));
  hipError_t e = hipSuccess;
  if (perm && n > 0) {
    e = d2h(tmp.data(), ws.idx_a, n, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (long i = 0; i < n && e == hipSuccess; i++) perm[i] = tmp[i];
  }
  if (nodes && nn > 0) {
    const DevBuf *f[10] = {&ws.n_depth, &ws.n_ix,     &ws.n_iy,     &ws.n_iz,
                           &ws.n_pbegin, &ws.n_pend,  &ws.n_parent, &ws.n_leaf,
                           &ws.n_child,  &ws.n_nchild};
    for (int k = 0; k < 10 && e == hipSuccess; k++) {
      e = d2h(tmp.data(), *f[k], nn, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      for (long i = 0; i < nn && e == hipSuccess; i++) nodes[10 * i + k] = tmp[i];
    }
  }
  if (centers && nn > 0) {
    std::vector<double> c((size_t)nn);
    const DevBuf *f[3] = {&ws.n_cx, &ws.n_cy, &ws.n_cz};
    for (int k = 0; k < 3 && e == hipSuccess; k++) {
      e = d2h(c.data(), *f[k], nn, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      for (long i = 0; i < nn && e == hipSuccess; i++) centers[3 * i + k] = c[i];
    }
  }
  if (lo)
    for (int k = 0; k < 3; k++) lo[k] = ctx->ti.lo[k];
  if (width) *width = ctx->ti.width;
  if (e != hipSuccess) aran::set_error(hipGetErrorString(e));
  return code_of_err(e);
}

extern "C" int aran_fmm3d_export_lists(const AranFmm3d *ctx, int *far_ofs,
                                       int *far_src, short *far_code,
                                       int *leaves, int *near_ofs,
                                       int *near_begin, int *near_end) {
  if (!ctx) return ARAN_FMM3D_EINVAL;
  aran::Workspace &ws = const_cast<aran::Workspace &>(ctx->ws);
  hipStream_t s = ws.stream;
  long nn = ctx->ti.nnodes, nl = ctx->ti.nleaves, nf = ctx->ti.nfar,
       nu = ctx->ti.nnear;
  hipError_t e = hipSuccess;
  if (nn > 0 && ws.far_csr_nodes != nn && (far_ofs || far_src || far_code)) {
    int nfar = 0;  // the mixed form keeps no CSR far lists: build them now
    e = aran::build_far_lists(ws, nn, &nfar);
    if (e == hipSuccess && nfar != nf) e = hipErrorUnknown;
  }
  if (far_ofs && nn) e = d2h(far_ofs, ws.v_ofs, nn + 1, s);
  if (e == hipSuccess && far_src) e = d2h(far_src, ws.v_src, nf, s);
  if (e == hipSuccess && far_code) e = d2h(far_code, ws.v_code, nf, s);
  if (e == hipSuccess && leaves) e = d2h(leaves, ws.leaf_list, nl, s);
  if (e == hipSuccess && near_ofs && nl) e = d2h(near_ofs, ws.u_ofs, nl + 1, s);
  if (e == hipSuccess && near_begin) e = d2h(near_begin, ws.u_beg, nu, s);
  if (e == hipSuccess && near_end) e = d2h(near_end, ws.u_end, nu, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) aran::set_error(hipGetErrorString(e));
  return code_of_err(e);
}

extern "C" int aran_fmm3d_export_semifar(const AranFmm3d *ctx, int *w_ofs,
                                         int *w_node, int *x_ofs,
                                         int *x_leaf) {
  if (!ctx || !w_ofs || !w_node || !x_ofs || !x_leaf) {
    aran::set_error("aran_fmm3d_export_semifar: invalid argument");
    return ARAN_FMM3D_EINVAL;
  }
  const aran::Workspace &ws = ctx->ws;
  if (ws.semifar_nodes < 0 || ws.semifar_nodes != ctx->ti.nnodes) {
    aran::set_error("aran_fmm3d_export_semifar: the last solve had no "
                    "semifar lists");
    return ARAN_FMM3D_EINVAL;
  }
  const long nl = ctx->ti.nleaves, nn = ctx->ti.nnodes, ns = ctx->st.semifar;
  hipError_t e = d2h(w_ofs, ws.w_ofs, nl + 1, ws.stream);
  if (e == hipSuccess && ns > 0) e = d2h(w_node, ws.w_node, ns, ws.stream);
  if (e == hipSuccess) e = d2h(x_ofs, ws.x_ofs, nn + 1, ws.stream);
  if (e == hipSuccess && ns > 0) e = d2h(x_leaf, ws.x_leaf, ns, ws.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ws.stream);
  return code_of_err(e);
}

extern "C" int aran_fmm3d_direct(AranFmm3d *ctx, long n, const double *x,
                                 const double *y, const double *z,
                                 const double *q, long nt,
                                 const long *targets, double *pot,
                                 double *grad) {
  if (!ctx || n < 0 || nt < 0 || (nt > 0 && (!targets || !pot || !grad))) {
    aran::set_error("aran_fmm3d_direct: invalid argument");
    return ARAN_FMM3D_EINVAL;
  }
  hipError_t e =
      aran::launch_direct(x, y, z, q, n, targets, nt, pot, grad, ctx->ws.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->ws.stream);
  if (e != hipSuccess) aran::set_error(hipGetErrorString(e));
  return code_of_err(e);
}

// ================================================================ development

namespace {

int dev_local_degree(const AranDevelopment3d *d) {
  return std::max(d->posdeg, d->negdeg);
}

// Process-wide context of the single-operator entry points.
struct OpContext {
  std::mutex mu;
  aran::Workspace ws;
  DevBuf buf;
};

OpContext &opctx() {
  static OpContext *c = new OpContext;
  return *c;
}

// Direction table for one arbitrary translation vector (dst - src), with
// the z flip of the reference (aran/aransphericalseriesd-rotate.c:186-197).
struct OneDir {
  double theta, r, c;
  double cp, sp;
};

OneDir one_dir(const VsgVector3d &a, const VsgVector3d &b) {
  OneDir d;
  double v[3] = {b.x - a.x, b.y - a.y, b.z - a.z};
  d.c = 1.;
  if (v[2] < 0.) {
    d.c = -1.;
    for (double &t : v) t = -t;
  }
  double rho = sqrt(v[0] * v[0] + v[1] * v[1]);
  d.r = sqrt(rho * rho + v[2] * v[2]);
  d.theta = atan2(rho, v[2]);
  d.cp = rho > 0. ? v[0] / rho : 1.;
  d.sp = rho > 0. ? v[1] / rho : 0.;
  return d;
}

// Runs one translation on the GPU and adds it into dst (host arrays).
int translate_one(int kind, int p, const double *src, const OneDir &d,
                  double *dst) {
  OpContext &oc = opctx();
  std::lock_guard<std::mutex> lock(oc.mu);
  aran::Workspace &ws = oc.ws;
  hipError_t e = aran::workspace_init(ws, nullptr);
  if (e == hipSuccess) e = aran::ensure_tables(ws, p);
  if (e != hipSuccess) return code_of_err(e);
  hipStream_t s = ws.stream;
  const int mc = aran::tri(p), lc = aran::tri(p + 1), rb = aran::rot_block(p);
  const int nin = (kind == 2) ? lc : mc, nout = (kind == 0) ? mc : lc;
  const int w = 2 * p + 1;
  auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
  size_t off_rot = 0, off_dscr = up(off_rot + sizeof(double) * 4 * rb);
  size_t off_eiphi = up(off_dscr + sizeof(double) * (p + 1) * w * w);
  size_t off_dir = up(off_eiphi + sizeof(double) * 2 * (p + 1));
  size_t off_theta = up(off_dir + sizeof(aran::DirEntry));
  size_t off_src = up(off_theta + sizeof(double));
  size_t off_dst = up(off_src + sizeof(double) * 2 * nin);
  size_t total = up(off_dst + sizeof(double) * 2 * nout);
  e = oc.buf.reserve(total);
  if (e != hipSuccess) return code_of_err(e);
  char *b = oc.buf.as<char>();
  std::vector<double> eiphi(2 * (p + 1));
  double er = 1., ei = 0.;
  for (int m = 0; m <= p; m++) {
    eiphi[2 * m] = er;
    eiphi[2 * m + 1] = ei;
    double nr = er * d.cp - ei * d.sp, ni = er * d.sp + ei * d.cp;
    er = nr;
    ei = ni;
  }
  aran::DirEntry de;
  de.theta = 0;
  de.c = d.c;
  de.norm = d.r;
  std::vector<double> out(2 * nout);
  e = hipMemsetAsync(b + off_dscr, 0, sizeof(double) * (p + 1) * w * w, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b + off_theta, &d.theta, sizeof(double),
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b + off_eiphi, eiphi.data(), sizeof(double) * eiphi.size(),
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b + off_dir, &de, sizeof de, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b + off_src, src, sizeof(double) * 2 * nin,
                       hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = aran::launch_wigner_tables(reinterpret_cast<double *>(b + off_theta), 1,
                                   p, reinterpret_cast<double *>(b + off_dscr),
                                   reinterpret_cast<double *>(b + off_rot), s);
  aran::OpTables tb = ws.tab.view();
  tb.rot = reinterpret_cast<double *>(b + off_rot);
  tb.eiphi = reinterpret_cast<double2 *>(b + off_eiphi);
  tb.dir = reinterpret_cast<aran::DirEntry *>(b + off_dir);
  if (e == hipSuccess)
    e = aran::launch_translate_one(kind, tb, d.r,
                                   reinterpret_cast<double2 *>(b + off_src),
                                   reinterpret_cast<double2 *>(b + off_dst), s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out.data(), b + off_dst, sizeof(double) * out.size(),
                       hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    aran::set_error(hipGetErrorString(e));
    return code_of_err(e);
  }
  for (size_t i = 0; i < out.size(); i++) dst[i] += out[i];
  return 0;
}

// One particle through the batched leaf kernels (one node that is both the
// expansion's node and the particle's leaf).  kOpP2M / kOpP2L: M / L (host)
// receive the added expansion; kOpL2P / kOpM2P: pot / grad receive the
// values of the expansion (L or M) at the particle.
enum { kOpP2M = 0, kOpL2P = 1, kOpP2L = 2, kOpM2P = 3 };

int leaf_op_one(int op, int p, const VsgVector3d &center,
                const VsgVector3d &pos, double charge, double *M,
                const double *L, double *pot, double *grad);

int leaf_op_one(bool p2m, int p, const VsgVector3d &center,
                const VsgVector3d &pos, double charge, double *M,
                const double *L, double *pot, double *grad) {
  return leaf_op_one(p2m ? kOpP2M : kOpL2P, p, center, pos, charge, M, L, pot,
                     grad);
}

int leaf_op_one(int op, int p, const VsgVector3d &center,
                const VsgVector3d &pos, double charge, double *M,
                const double *L, double *pot, double *grad) {
  OpContext &oc = opctx();
  std::lock_guard<std::mutex> lock(oc.mu);
  aran::Workspace &ws = oc.ws;
  hipError_t e = aran::workspace_init(ws, nullptr);
  if (e == hipSuccess) e = aran::ensure_tables(ws, p);
  if (e != hipSuccess) return code_of_err(e);
  hipStream_t s = ws.stream;
  const int mc = aran::tri(p), lc = aran::tri(p + 1);
  // node record (ints and doubles) and particle data in one buffer
  struct Host {
    int ints[10];
    int leaf0;
    int lofs[2];  // one semifar list entry: node 0
    uint64_t key;
    double c[3];
    double part[4];
    double res[4];
  } h;
  memset(&h, 0, sizeof h);
  h.ints[4] = 0;  // pbegin
  h.ints[5] = 1;  // pend
  h.ints[6] = -1;
  h.ints[7] = 1;  // leaf
  h.ints[8] = -1;
  h.leaf0 = 0;
  h.lofs[0] = 0;
  h.lofs[1] = 1;
  h.c[0] = center.x;
  h.c[1] = center.y;
  h.c[2] = center.z;
  h.part[0] = pos.x;
  h.part[1] = pos.y;
  h.part[2] = pos.z;
  h.part[3] = charge;
  size_t off_exp = sizeof(Host);
  off_exp = (off_exp + 255) & ~(size_t)255;
  size_t total = off_exp + sizeof(double) * 2 * std::max(mc, lc);
  e = oc.buf.reserve(total);
  if (e != hipSuccess) return code_of_err(e);
  char *b = oc.buf.as<char>();
  Host *dh = reinterpret_cast<Host *>(b);
  double2 *dexp = reinterpret_cast<double2 *>(b + off_exp);
  e = hipMemcpyAsync(b, &h, sizeof h, hipMemcpyHostToDevice, s);
  aran::NodeArrays nd;
  nd.depth = &dh->ints[0];
  nd.ix = &dh->ints[1];
  nd.iy = &dh->ints[2];
  nd.iz = &dh->ints[3];
  nd.pbegin = &dh->ints[4];
  nd.pend = &dh->ints[5];
  nd.parent = &dh->ints[6];
  nd.leaf = &dh->ints[7];
  nd.child = &dh->ints[8];
  nd.nchild = &dh->ints[9];
  nd.sys = nullptr;  // the leaf kernels use centres, not system boxes
  nd.box = nullptr;
  nd.key = &dh->key;
  nd.cx = &dh->c[0];
  nd.cy = &dh->c[1];
  nd.cz = &dh->c[2];
  aran::OpTables tb = ws.tab.view();
  if (op == kOpP2M || op == kOpP2L) {
    const int nc = op == kOpP2M ? mc : lc;
    if (e == hipSuccess && op == kOpP2M)
      e = aran::launch_p2m(tb, nd, &dh->leaf0, 1, &dh->part[0], &dh->part[1],
                           &dh->part[2], &dh->part[3], dexp, s);
    if (e == hipSuccess && op == kOpP2L)
      e = hipMemsetAsync(dexp, 0, sizeof(double2) * lc, s);
    if (e == hipSuccess && op == kOpP2L)
      e = aran::launch_p2l(tb, nd, 1, &dh->lofs[0], &dh->leaf0, &dh->part[0],
                           &dh->part[1], &dh->part[2], &dh->part[3], dexp, s);
    std::vector<double> out(2 * nc);
    if (e == hipSuccess)
      e = hipMemcpyAsync(out.data(), dexp, sizeof(double) * out.size(),
                         hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    double *dst = op == kOpP2M ? M : const_cast<double *>(L);
    if (e == hipSuccess)
      for (size_t i = 0; i < out.size(); i++) dst[i] += out[i];
  } else {
    const double *src = op == kOpL2P ? L : M;
    const int nc = op == kOpL2P ? lc : mc;
    if (e == hipSuccess)
      e = hipMemcpyAsync(dexp, src, sizeof(double) * 2 * nc,
                         hipMemcpyHostToDevice, s);
    if (e == hipSuccess && op == kOpL2P)
      e = aran::launch_l2p(tb, nd, &dh->leaf0, 1, &dh->part[0], &dh->part[1],
                           &dh->part[2], dexp, &dh->res[0], &dh->res[1], s);
    if (e == hipSuccess && op == kOpM2P)
      e = aran::launch_m2p(tb, nd, &dh->leaf0, 1, &dh->lofs[0], &dh->leaf0,
                           &dh->part[0], &dh->part[1], &dh->part[2], dexp,
                           &dh->res[0], &dh->res[1], s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(h.res, dh->res, sizeof h.res, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) {
      if (pot) *pot = h.res[0];
      if (grad) {
        grad[0] = h.res[1];
        grad[1] = h.res[2];
        grad[2] = h.res[3];
      }
    }
  }
  if (e != hipSuccess) {
    aran::set_error(hipGetErrorString(e));
    return code_of_err(e);
  }
  return 0;
}

bool dev_ok(const AranDevelopment3d *d, const char *fn) {
  if (!d || d->posdeg > d->negdeg || d->negdeg < 1 ||
      d->negdeg > aran::kMaxOrder) {
    critical("%s: development must have posdeg <= negdeg, 1 <= negdeg <= 15",
             fn);
    return false;
  }
  return true;
}

}  // namespace

extern "C" GType aran_development3d_get_type(void) { return kDevelopmentType; }

extern "C" AranDevelopment3d *aran_development3d_new(guint8 posdeg,
                                                     guint8 negdeg) {
  AranDevelopment3d *d =
      static_cast<AranDevelopment3d *>(calloc(1, sizeof(AranDevelopment3d)));
  d->posdeg = posdeg;
  d->negdeg = negdeg;
  int ld = std::max(posdeg, negdeg);
  d->multipole = static_cast<double *>(
      calloc(2 * (size_t)std::max(1, aran::tri(negdeg)), sizeof(double)));
  d->local =
      static_cast<double *>(calloc(2 * (size_t)aran::tri(ld + 1), sizeof(double)));
  return d;
}

extern "C" void aran_development3d_free(AranDevelopment3d *ad) {
  if (!ad) return;
  free(ad->multipole);
  free(ad->local);
  free(ad);
}

extern "C" void aran_development3d_copy(const AranDevelopment3d *src,
                                        AranDevelopment3d *dst) {
  if (!src || !dst || src->posdeg != dst->posdeg || src->negdeg != dst->negdeg) {
    critical("aran_development3d_copy: degrees must coincide");
    return;
  }
  memcpy(dst->multipole, src->multipole,
         sizeof(double) * 2 * aran::tri(src->negdeg));
  memcpy(dst->local, src->local,
         sizeof(double) * 2 * aran::tri(dev_local_degree(src) + 1));
}

extern "C" AranDevelopment3d *aran_development3d_clone(AranDevelopment3d *src) {
  if (!src) return nullptr;
  AranDevelopment3d *d = aran_development3d_new(src->posdeg, src->negdeg);
  aran_development3d_copy(src, d);
  return d;
}

extern "C" void aran_development3d_set_zero(AranDevelopment3d *ad) {
  if (!ad) return;
  memset(ad->multipole, 0, sizeof(double) * 2 * aran::tri(ad->negdeg));
  memset(ad->local, 0, sizeof(double) * 2 * aran::tri(dev_local_degree(ad) + 1));
}

// Text dump in the layout of aran_spherical_seriesd_write
// (aran/aransphericalseriesd.c:272-298).
extern "C" void aran_development3d_write(AranDevelopment3d *ad, FILE *file) {
  if (!ad || !file) return;
  fprintf(file, "{m:[");
  for (int l = 0; l < ad->negdeg; l++)
    for (int m = 0; m <= l; m++) {
      const double *c = ad->multipole + 2 * (aran::tri(l) + m);
      fprintf(file, "[%d,%d]:(%e,%e), ", -l - 1, m, c[0], c[1]);
    }
  fprintf(file, "], l:[");
  for (int l = 0; l <= dev_local_degree(ad); l++)
    for (int m = 0; m <= l; m++) {
      const double *c = ad->local + 2 * (aran::tri(l) + m);
      fprintf(file, "[%d,%d]:(%e,%e), ", l, m, c[0], c[1]);
    }
  fprintf(file, "]}");
}

extern "C" void aran_development3d_p2m(const VsgVector3d *position,
                                       const gdouble charge,
                                       const VsgPRTree3dNodeInfo *dst_node,
                                       AranDevelopment3d *dst) {
  if (!position || !dst_node || !dev_ok(dst, "aran_development3d_p2m")) return;
  if (leaf_op_one(true, dst->negdeg, dst_node->center, *position, charge,
                  dst->multipole, nullptr, nullptr, nullptr) != 0)
    critical("aran_development3d_p2m: %s", aran::last_error());
}

static void translate_api(int kind, const char *fn,
                          const VsgPRTree3dNodeInfo *src_node,
                          AranDevelopment3d *src,
                          const VsgPRTree3dNodeInfo *dst_node,
                          AranDevelopment3d *dst) {
  if (!src_node || !dst_node || !dev_ok(src, fn) || !dev_ok(dst, fn)) return;
  if (src->negdeg != dst->negdeg) {
    critical("%s: source and destination orders differ", fn);
    return;
  }
  OneDir d = one_dir(src_node->center, dst_node->center);
  const double *in = (kind == 2) ? src->local : src->multipole;
  double *out = (kind == 0) ? dst->multipole : dst->local;
  if (translate_one(kind, src->negdeg, in, d, out) != 0)
    critical("%s: %s", fn, aran::last_error());
}

extern "C" void aran_development3d_m2m_rotate(const VsgPRTree3dNodeInfo *src_node,
                                              AranDevelopment3d *src,
                                              const VsgPRTree3dNodeInfo *dst_node,
                                              AranDevelopment3d *dst) {
  translate_api(0, "aran_development3d_m2m_rotate", src_node, src, dst_node,
                dst);
}

extern "C" void aran_development3d_m2l_rotate(const VsgPRTree3dNodeInfo *src_node,
                                              AranDevelopment3d *src,
                                              const VsgPRTree3dNodeInfo *dst_node,
                                              AranDevelopment3d *dst) {
  translate_api(1, "aran_development3d_m2l_rotate", src_node, src, dst_node,
                dst);
}

extern "C" void aran_development3d_l2l_rotate(const VsgPRTree3dNodeInfo *src_node,
                                              AranDevelopment3d *src,
                                              const VsgPRTree3dNodeInfo *dst_node,
                                              AranDevelopment3d *dst) {
  translate_api(2, "aran_development3d_l2l_rotate", src_node, src, dst_node,
                dst);
}

extern "C" gcomplex128 aran_development3d_local_evaluate(
    const VsgPRTree3dNodeInfo *devel_node, AranDevelopment3d *devel,
    const VsgVector3d *pos) {
  double pot = 0.;
  if (devel_node && pos && dev_ok(devel, "aran_development3d_local_evaluate"))
    if (leaf_op_one(false, devel->negdeg, devel_node->center, *pos, 0., nullptr,
                    devel->local, &pot, nullptr) != 0)
      critical("aran_development3d_local_evaluate: %s", aran::last_error());
  return pot;
}

extern "C" void aran_development3d_local_gradient_evaluate(
    const VsgPRTree3dNodeInfo *devel_node, AranDevelopment3d *devel,
    const VsgVector3d *pos, VsgVector3d *grad) {
  double g[3] = {0., 0., 0.};
  if (devel_node && pos && grad &&
      dev_ok(devel, "aran_development3d_local_gradient_evaluate"))
    if (leaf_op_one(false, devel->negdeg, devel_node->center, *pos, 0., nullptr,
                    devel->local, nullptr, g) != 0)
      critical("aran_development3d_local_gradient_evaluate: %s",
               aran::last_error());
  if (grad) {
    grad->x = g[0];
    grad->y = g[1];
    grad->z = g[2];
  }
}

extern "C" void aran_development3d_p2l(const VsgVector3d *position,
                                       const gdouble charge,
                                       const VsgPRTree3dNodeInfo *dst_node,
                                       AranDevelopment3d *dst) {
  if (!position || !dst_node || !dev_ok(dst, "aran_development3d_p2l")) return;
  if (leaf_op_one(kOpP2L, dev_local_degree(dst), dst_node->center, *position,
                  charge, nullptr, dst->local, nullptr, nullptr) != 0)
    critical("aran_development3d_p2l: %s", aran::last_error());
}

extern "C" gcomplex128 aran_development3d_multipole_evaluate(
    const VsgPRTree3dNodeInfo *devel_node, AranDevelopment3d *devel,
    const VsgVector3d *pos) {
  double pot = 0.;
  if (devel_node && pos &&
      dev_ok(devel, "aran_development3d_multipole_evaluate"))
    if (leaf_op_one(kOpM2P, devel->negdeg, devel_node->center, *pos, 0.,
                    devel->multipole, nullptr, &pot, nullptr) != 0)
      critical("aran_development3d_multipole_evaluate: %s",
               aran::last_error());
  return pot;
}

extern "C" void aran_development3d_multipole_gradient_evaluate(
    const VsgPRTree3dNodeInfo *devel_node, AranDevelopment3d *devel,
    const VsgVector3d *pos, VsgVector3d *grad) {
  double g[3] = {0., 0., 0.};
  if (devel_node && pos && grad &&
      dev_ok(devel, "aran_development3d_multipole_gradient_evaluate"))
    if (leaf_op_one(kOpM2P, devel->negdeg, devel_node->center, *pos, 0.,
                    devel->multipole, nullptr, nullptr, g) != 0)
      critical("aran_development3d_multipole_gradient_evaluate: %s",
               aran::last_error());
  if (grad) {
    grad->x = g[0];
    grad->y = g[1];
    grad->z = g[2];
  }
}

extern "C" gcomplex128 aran_development3d_m2p(const VsgPRTree3dNodeInfo *n,
                                              AranDevelopment3d *d,
                                              const VsgVector3d *pos) {
  return aran_development3d_multipole_evaluate(n, d, pos);
}

extern "C" void aran_development3d_m2pv(const VsgPRTree3dNodeInfo *n,
                                        AranDevelopment3d *d,
                                        const VsgVector3d *pos,
                                        VsgVector3d *grad) {
  aran_development3d_multipole_gradient_evaluate(n, d, pos, grad);
}

extern "C" gcomplex128 aran_development3d_l2p(const VsgPRTree3dNodeInfo *n,
                                              AranDevelopment3d *d,
                                              const VsgVector3d *pos) {
  return aran_development3d_local_evaluate(n, d, pos);
}

extern "C" void aran_development3d_l2pv(const VsgPRTree3dNodeInfo *n,
                                        AranDevelopment3d *d,
                                        const VsgVector3d *pos,
                                        VsgVector3d *grad) {
  aran_development3d_local_gradient_evaluate(n, d, pos, grad);
}

// ================================================================ solver

namespace {

struct Layout {
  gssize charge = 24;
  gssize pot = -1;
  bool pot_complex = false;
  gssize field = -1;
};

Layout g_last_layout;  // used by the library callbacks called directly

}  // namespace

struct _AranSolver3d {
  VsgPRTree3d *prtree = nullptr;
  GType devel_type = 0;
  AranDevelopment3d *devel = nullptr;
  AranZeroFunc zero = nullptr;
  AranParticle2ParticleFunc3d p2p = nullptr;
  AranParticle2MultipoleFunc3d p2m = nullptr;
  AranMultipole2MultipoleFunc3d m2m = nullptr;
  AranMultipole2LocalFunc3d m2l = nullptr;
  AranLocal2LocalFunc3d l2l = nullptr;
  AranLocal2ParticleFunc3d l2p = nullptr;
  AranParticle2LocalFunc3d p2l = nullptr;
  AranMultipole2ParticleFunc3d m2p = nullptr;
  guint semifar_threshold = 0xffffffffu;
  gdouble tolerance = 0.;
  Layout layout;
  std::vector<VsgPoint3> points;
  std::vector<double> soa_x, soa_y, soa_z, soa_q, res_pot, res_field;
  bool soa = false;
  long stats[11] = {0};
  guint depth = 0;
  gint status = 0;
  int precision = ARAN_FMM3D_PLAIN;
  AranFmm3d *ctx = nullptr;
  DevBuf dx, dy, dz, dq, dpot, dgrad;
  HostBuf hx, hy, hz, hq, hpot, hgrad;
};

extern "C" void aran_solver3d_set_development(AranSolver3d *solver,
                                              GType devel_type, gpointer devel,
                                              AranZeroFunc zero) {
  if (!solver) return;
  if (devel_type != kDevelopmentType) {
    critical("aran_solver3d_set_development: only ARAN_TYPE_DEVELOPMENT3D is "
             "supported");
    solver->status = ARAN_FMM3D_EINVAL;
    return;
  }
  if (solver->devel && solver->devel != devel)
    aran_development3d_free(solver->devel);
  solver->devel_type = devel_type;
  solver->devel = static_cast<AranDevelopment3d *>(devel);
  solver->zero = zero;
}

extern "C" AranSolver3d *aran_solver3d_new(VsgPRTree3d *prtree,
                                           GType devel_type, gpointer devel,
                                           AranZeroFunc zero) {
  if (!prtree) {
    critical("aran_solver3d_new: prtree is NULL");
    return nullptr;
  }
  AranSolver3d *s = new AranSolver3d;
  s->prtree = prtree;
  aran_solver3d_set_development(s, devel_type, devel, zero);
  g_last_layout = s->layout;
  return s;
}

extern "C" void aran_solver3d_free(AranSolver3d *solver) {
  if (!solver) return;
  vsg_prtree3d_free(solver->prtree);
  if (solver->devel) aran_development3d_free(solver->devel);
  if (solver->ctx) aran_fmm3d_free(solver->ctx);
  DevBuf *b[] = {&solver->dx, &solver->dy, &solver->dz, &solver->dq,
                 &solver->dpot, &solver->dgrad};
  for (DevBuf *d : b) d->release();
  delete solver;
}

extern "C" void aran_solver3d_set_functions_full(
    AranSolver3d *solver, AranParticle2ParticleFunc3d p2p,
    AranParticle2MultipoleFunc3d p2m, AranMultipole2MultipoleFunc3d m2m,
    AranMultipole2LocalFunc3d m2l, AranLocal2LocalFunc3d l2l,
    AranLocal2ParticleFunc3d l2p, AranParticle2LocalFunc3d p2l,
    AranMultipole2ParticleFunc3d m2p, guint semifar_threshold) {
  if (!solver) return;
  solver->p2p = p2p;
  solver->p2m = p2m;
  solver->m2m = m2m;
  solver->m2l = m2l;
  solver->l2l = l2l;
  solver->l2p = l2p;
  solver->p2l = p2l;
  solver->m2p = m2p;
  solver->semifar_threshold = semifar_threshold;
}

extern "C" void aran_solver3d_set_functions(AranSolver3d *solver,
                                            AranParticle2ParticleFunc3d p2p,
                                            AranParticle2MultipoleFunc3d p2m,
                                            AranMultipole2MultipoleFunc3d m2m,
                                            AranMultipole2LocalFunc3d m2l,
                                            AranLocal2LocalFunc3d l2l,
                                            AranLocal2ParticleFunc3d l2p) {
  aran_solver3d_set_functions_full(solver, p2p, p2m, m2m, m2l, l2l, l2p,
                                   nullptr, nullptr, 0xffffffffu);
}

extern "C" void aran_solver3d_get_functions_full(
    AranSolver3d *solver, AranParticle2ParticleFunc3d *p2p,
    AranParticle2MultipoleFunc3d *p2m, AranMultipole2MultipoleFunc3d *m2m,
    AranMultipole2LocalFunc3d *m2l, AranLocal2LocalFunc3d *l2l,
    AranLocal2ParticleFunc3d *l2p, AranParticle2LocalFunc3d *p2l,
    AranMultipole2ParticleFunc3d *m2p, guint *semifar_threshold) {
  if (!solver) return;
  if (p2p) *p2p = solver->p2p;
  if (p2m) *p2m = solver->p2m;
  if (m2m) *m2m = solver->m2m;
  if (m2l) *m2l = solver->m2l;
  if (l2l) *l2l = solver->l2l;
  if (l2p) *l2p = solver->l2p;
  if (p2l) *p2l = solver->p2l;
  if (m2p) *m2p = solver->m2p;
  if (semifar_threshold) *semifar_threshold = solver->semifar_threshold;
}

extern "C" void aran_solver3d_reinit_stats(AranSolver3d *solver) {
  if (solver) memset(solver->stats, 0, sizeof solver->stats);
}

extern "C" void aran_solver3d_get_stats(AranSolver3d *solver, glong *zero_count,
                                        glong *p2p_count, glong *p2m_count,
                                        glong *m2m_count, glong *m2l_count,
                                        glong *l2l_count, glong *l2p_count,
                                        glong *p2l_count, glong *m2p_count,
                                        glong *p2p_remote_count,
                                        glong *m2l_remote_count) {
  if (!solver) return;
  glong *o[11] = {zero_count, p2p_count, p2m_count, m2m_count,
                  m2l_count,  l2l_count, l2p_count, p2l_count,
                  m2p_count,  p2p_remote_count, m2l_remote_count};
  for (int k = 0; k < 11; k++)
    if (o[k]) *o[k] = solver->stats[k];
}

extern "C" gdouble aran_solver3d_get_tolerance(AranSolver3d *solver) {
  return solver ? solver->tolerance : -1.;
}

extern "C" void aran_solver3d_set_tolerance(AranSolver3d *solver,
                                            gdouble tolerance) {
  if (solver) solver->tolerance = tolerance;
}

extern "C" void aran_solver3d_get_bounds(AranSolver3d *solver,
                                         VsgVector3d *lbound,
                                         VsgVector3d *ubound) {
  if (!solver) return;
  if (lbound) *lbound = solver->prtree->lb;
  if (ubound) *ubound = solver->prtree->ub;
}

extern "C" guint aran_solver3d_depth(const AranSolver3d *solver) {
  return solver ? solver->depth : 0;
}

extern "C" guint aran_solver3d_point_count(const AranSolver3d *solver) {
  if (!solver) return 0;
  return (guint)(solver->soa ? solver->soa_x.size() : solver->points.size());
}

extern "C" void aran_solver3d_insert_point(AranSolver3d *solver,
                                           VsgPoint3 point) {
  if (!solver || !point) return;
  if (solver->soa) {
    critical("aran_solver3d_insert_point: solver holds SoA particles");
    return;
  }
  solver->points.push_back(point);
}

extern "C" gboolean aran_solver3d_remove_point(AranSolver3d *solver,
                                               VsgPoint3 point) {
  if (!solver) return FALSE;
  auto it = std::find(solver->points.begin(), solver->points.end(), point);
  if (it == solver->points.end()) return FALSE;
  solver->points.erase(it);
  return TRUE;
}

// First inserted particle within the tolerance distance of selector.
extern "C" VsgPoint3 aran_solver3d_find_point(AranSolver3d *solver,
                                              VsgPoint3 selector) {
  if (!solver || !selector) return nullptr;
  for (VsgPoint3 p : solver->points)
    if (vsg_vector3d_dist(p, selector) <= solver->tolerance) return p;
  return nullptr;
}

extern "C" void aran_solver3d_foreach_point(AranSolver3d *solver, GFunc func,
                                            gpointer user_data) {
  if (!solver || !func) return;
  for (VsgPoint3 p : solver->points) func(p, user_data);
}

extern "C" gint aran_solver3d_status(const AranSolver3d *solver) {
  return solver ? solver->status : ARAN_FMM3D_EINVAL;
}

extern "C" gint aran_solver3d_set_precision(AranSolver3d *solver,
                                            gint precision) {
  if (!solver || precision < ARAN_FMM3D_PLAIN ||
      precision > ARAN_FMM3D_BF16X6) {
    aran::set_error("aran_solver3d_set_precision: invalid argument");
    return ARAN_FMM3D_EINVAL;
  }
  solver->precision = precision;
  return 0;
}

extern "C" void aran_solver3d_set_particle_layout(AranSolver3d *solver,
                                                  gssize off_charge,
                                                  gssize off_pot,
                                                  gboolean pot_is_complex,
                                                  gssize off_field) {
  if (!solver) return;
  if (off_charge < (gssize)sizeof(VsgVector3d)) {
    critical("aran_solver3d_set_particle_layout: the charge must follow the "
             "leading VsgVector3d");
    solver->status = ARAN_FMM3D_EINVAL;
    return;
  }
  solver->layout.charge = off_charge;
  solver->layout.pot = off_pot;
  solver->layout.pot_complex = pot_is_complex != 0;
  solver->layout.field = off_field;
  g_last_layout = solver->layout;
}

extern "C" gint aran_solver3d_insert_points_soa(AranSolver3d *solver,
                                                glong n, const gdouble *x,
                                                const gdouble *y,
                                                const gdouble *z,
                                                const gdouble *charge) {
  if (!solver || n < 0 || (n > 0 && (!x || !y || !z || !charge)) ||
      !solver->points.empty()) {
    critical("aran_solver3d_insert_points_soa: invalid argument or solver "
             "holds pointer particles");
    return ARAN_FMM3D_EINVAL;
  }
  solver->soa = true;
  solver->soa_x.insert(solver->soa_x.end(), x, x + n);
  solver->soa_y.insert(solver->soa_y.end(), y, y + n);
  solver->soa_z.insert(solver->soa_z.end(), z, z + n);
  solver->soa_q.insert(solver->soa_q.end(), charge, charge + n);
  return 0;
}

// ---- FMA particle files (the formats of tests/newtonfield3parallel.c)

extern "C" gint aran_solver3d_write_fma(AranSolver3d *solver, FILE *file) {
  if (!solver || !file) {
    critical("aran_solver3d_write_fma: invalid argument");
    return ARAN_FMM3D_EINVAL;
  }
  // header: count, half-width of the bounds' box and its centre
  // (newtonfield3parallel.c:1193-1212)
  const VsgVector3d lb = solver->prtree->lb, ub = solver->prtree->ub;
  const double cx = 0.5 * (lb.x + ub.x), cy = 0.5 * (lb.y + ub.y),
               cz = 0.5 * (lb.z + ub.z);
  double box = ub.x - cx;
  box = std::max(box, ub.y - cy);
  box = std::max(box, ub.z - cz);
  const unsigned long np = aran_solver3d_point_count(solver);
  if (fprintf(file, "%lu %g %g %g %g\n", np, box, cx, cy, cz) < 0)
    return ARAN_FMM3D_EINVAL;
  // one line per particle: position and charge (newtonfield3parallel.c:
  // 1186-1191), in insertion order
  for (unsigned long i = 0; i < np; i++) {
    double x, y, z, q;
    if (solver->soa) {
      x = solver->soa_x[i];
      y = solver->soa_y[i];
      z = solver->soa_z[i];
      q = solver->soa_q[i];
    } else {
      const char *pt = static_cast<const char *>(solver->points[i]);
      const VsgVector3d *v = reinterpret_cast<const VsgVector3d *>(pt);
      x = v->x;
      y = v->y;
      z = v->z;
      q = *reinterpret_cast<const double *>(pt + solver->layout.charge);
    }
    if (fprintf(file, "%g %g %g %g\n", x, y, z, q) < 0)
      return ARAN_FMM3D_EINVAL;
  }
  return 0;
}

extern "C" glong aran_fma_read_points(FILE *file, glong max_points,
                                      gdouble *x, gdouble *y, gdouble *z,
                                      gdouble *charge) {
  if (!file || max_points < 0 ||
      (max_points > 0 && (!x || !y || !z || !charge))) {
    critical("aran_fma_read_points: invalid argument");
    return ARAN_FMM3D_EINVAL;
  }
  // lines "index x y z charge" (the -dist load fill,
  // newtonfield3parallel.c:1020-1086); the index is read and not used
  glong n = 0;
  unsigned long idx;
  double a, b, c, d;
  for (;;) {
    const int got = fscanf(file, "%lu %lf %lf %lf %lf", &idx, &a, &b, &c, &d);
    if (got == EOF) break;
    if (got != 5) {
      critical("aran_fma_read_points: malformed line after %ld points", n);
      return ARAN_FMM3D_EINVAL;
    }
    if (n < max_points) {
      x[n] = a;
      y[n] = b;
      z[n] = c;
      charge[n] = d;
    }
    n++;
  }
  return n;
}

extern "C" gint aran_solver3d_get_results_soa(AranSolver3d *solver,
                                              gdouble *pot, gdouble *field) {
  if (!solver || !solver->soa) return ARAN_FMM3D_EINVAL;
  size_t n = solver->soa_x.size();
  if (solver->res_pot.size() != n) return ARAN_FMM3D_EINVAL;
  if (pot) memcpy(pot, solver->res_pot.data(), sizeof(double) * n);
  if (field) memcpy(field, solver->res_field.data(), sizeof(double) * 3 * n);
  return 0;
}

// ---- library Newton callbacks (host semantics of the reference harness)

static inline const VsgVector3d *vec_of(VsgPoint3 p) {
  return static_cast<const VsgVector3d *>(p);
}
static inline double charge_of(VsgPoint3 p) {
  return *reinterpret_cast<const double *>(static_cast<char *>(p) +
                                           g_last_layout.charge);
}
static inline void add_pot(VsgPoint3 p, double v) {
  if (g_last_layout.pot < 0) return;
  *reinterpret_cast<double *>(static_cast<char *>(p) + g_last_layout.pot) += v;
}
static inline void add_field(VsgPoint3 p, double gx, double gy, double gz) {
  if (g_last_layout.field < 0) return;
  double *f =
      reinterpret_cast<double *>(static_cast<char *>(p) + g_last_layout.field);
  f[0] += gx;
  f[1] += gy;
  f[2] += gz;
}

static void p2p_host(VsgPoint3 one, VsgPoint3 other, bool pot, bool field) {
  if (one == other) return;
  const VsgVector3d *a = vec_of(one), *b = vec_of(other);
  double dx = a->x - b->x, dy = a->y - b->y, dz = a->z - b->z;
  double ir = 1. / sqrt(dx * dx + dy * dy + dz * dz), ir3 = ir * ir * ir;
  double qa = charge_of(one), qb = charge_of(other);
  if (pot) {
    add_pot(one, ir * qb);
    add_pot(other, ir * qa);
  }
  if (field) {
    add_field(one, -ir3 * qb * dx, -ir3 * qb * dy, -ir3 * qb * dz);
    add_field(other, ir3 * qa * dx, ir3 * qa * dy, ir3 * qa * dz);
  }
}

extern "C" void aran_newton3d_p2p_pot(VsgPoint3 one, VsgPoint3 other) {
  p2p_host(one, other, true, false);
}
extern "C" void aran_newton3d_p2p_field(VsgPoint3 one, VsgPoint3 other) {
  p2p_host(one, other, false, true);
}
extern "C" void aran_newton3d_p2p_potfield(VsgPoint3 one, VsgPoint3 other) {
  p2p_host(one, other, true, true);
}

extern "C" void aran_newton3d_p2m(VsgPoint3 particle,
                                  const VsgPRTree3dNodeInfo *dst_node,
                                  gpointer dst) {
  aran_development3d_p2m(vec_of(particle), charge_of(particle), dst_node,
                         static_cast<AranDevelopment3d *>(dst));
}

extern "C" void aran_newton3d_l2p_pot(const VsgPRTree3dNodeInfo *src_node,
                                      gpointer src, VsgPoint3 particle) {
  gcomplex128 v = aran_development3d_local_evaluate(
      src_node, static_cast<AranDevelopment3d *>(src), vec_of(particle));
  add_pot(particle, __real__ v);
}

extern "C" void aran_newton3d_l2p_field(const VsgPRTree3dNodeInfo *src_node,
                                        gpointer src, VsgPoint3 particle) {
  VsgVector3d g;
  aran_development3d_local_gradient_evaluate(
      src_node, static_cast<AranDevelopment3d *>(src), vec_of(particle), &g);
  add_field(particle, g.x, g.y, g.z);
}

extern "C" void aran_newton3d_l2p_potfield(const VsgPRTree3dNodeInfo *src_node,
                                           gpointer src, VsgPoint3 particle) {
  aran_newton3d_l2p_pot(src_node, src, particle);
  aran_newton3d_l2p_field(src_node, src, particle);
}

extern "C" void aran_newton3d_p2l(VsgPoint3 particle,
                                  const VsgPRTree3dNodeInfo *dst_node,
                                  gpointer dst) {
  aran_development3d_p2l(vec_of(particle), charge_of(particle), dst_node,
                         static_cast<AranDevelopment3d *>(dst));
}

extern "C" void aran_newton3d_m2p_pot(const VsgPRTree3dNodeInfo *src_node,
                                      gpointer src, VsgPoint3 particle) {
  gcomplex128 v = aran_development3d_multipole_evaluate(
      src_node, static_cast<AranDevelopment3d *>(src), vec_of(particle));
  add_pot(particle, __real__ v);
}

extern "C" void aran_newton3d_m2p_field(const VsgPRTree3dNodeInfo *src_node,
                                        gpointer src, VsgPoint3 particle) {
  VsgVector3d g;
  aran_development3d_multipole_gradient_evaluate(
      src_node, static_cast<AranDevelopment3d *>(src), vec_of(particle), &g);
  add_field(particle, g.x, g.y, g.z);
}

extern "C" void aran_newton3d_m2p_potfield(const VsgPRTree3dNodeInfo *src_node,
                                           gpointer src, VsgPoint3 particle) {
  aran_newton3d_m2p_pot(src_node, src, particle);
  aran_newton3d_m2p_field(src_node, src, particle);
}

// ---- the solve

template <class A, class B> static bool same_fn(A a, B b) {
  return reinterpret_cast<void *>(a) == reinterpret_cast<void *>(b);
}

extern "C" void aran_solver3d_solve(AranSolver3d *solver) {
  if (!solver) return;
  solver->status = 0;
  // recognised operator set
  int pot_p2p = -1, pot_l2p = -1;
  if (solver->p2p == aran_newton3d_p2p_pot) pot_p2p = 1;
  if (solver->p2p == aran_newton3d_p2p_field) pot_p2p = 2;
  if (solver->p2p == aran_newton3d_p2p_potfield) pot_p2p = 3;
  if (solver->l2p == aran_newton3d_l2p_pot) pot_l2p = 1;
  if (solver->l2p == aran_newton3d_l2p_field) pot_l2p = 2;
  if (solver->l2p == aran_newton3d_l2p_potfield) pot_l2p = 3;
  bool ops_ok = pot_p2p > 0 && pot_p2p == pot_l2p &&
                solver->p2m == aran_newton3d_p2m &&
                same_fn(solver->m2m, aran_development3d_m2m_rotate) &&
                same_fn(solver->m2l, aran_development3d_m2l_rotate) &&
                same_fn(solver->l2l, aran_development3d_l2l_rotate);
  if (!ops_ok) {
    critical("aran_solver3d_solve: unsupported callbacks; the GPU engine runs "
             "the library's Newton callbacks (aran_newton3d_*) with the "
             "aran_development3d_*_rotate operators");
    solver->status = ARAN_FMM3D_EINVAL;
    return;
  }
  if (!dev_ok(solver->devel, "aran_solver3d_solve")) {
    solver->status = ARAN_FMM3D_EINVAL;
    return;
  }
  const int order = solver->devel->negdeg;
  // semifar pairs when both P2L and M2P are set (aran/aransolver3d.c:
  // 1152-1161): the library's P2L with the M2P of the same outputs as P2P
  long semifar = -1;
  if (solver->p2l && solver->m2p) {
    int kind = -1;
    if (solver->m2p == aran_newton3d_m2p_pot) kind = 1;
    if (solver->m2p == aran_newton3d_m2p_field) kind = 2;
    if (solver->m2p == aran_newton3d_m2p_potfield) kind = 3;
    if (solver->p2l != aran_newton3d_p2l || kind != pot_p2p) {
      critical("aran_solver3d_solve: unsupported semifar callbacks; the GPU "
               "engine runs aran_newton3d_p2l with the aran_newton3d_m2p_* "
               "callback of the same outputs as the P2P callback");
      solver->status = ARAN_FMM3D_EINVAL;
      return;
    }
    if (solver->semifar_threshold == 0) {
      // automatic threshold (the reference derives it from profiled
      // operator times, aran/aransolver3d.c:1090-1127; here from the
      // operators' algorithmic costs)
      solver->semifar_threshold =
          (guint)aran_fmm3d_semifar_threshold_auto(order);
    }
    if (solver->semifar_threshold != 0xffffffffu)
      semifar = (long)solver->semifar_threshold;
  }
  const bool want_pot = (pot_p2p & 1) != 0, want_field = (pot_p2p & 2) != 0;
  const long n = (long)(solver->soa ? solver->soa_x.size() : solver->points.size());
  if (!solver->ctx) {
    solver->ctx = aran_fmm3d_new(nullptr);
    if (!solver->ctx) {
      solver->status = ARAN_FMM3D_EHIP;
      return;
    }
  }
  aran_fmm3d_set_precision(solver->ctx, solver->precision);
  hipStream_t s = solver->ctx->ws.stream;
  hipError_t e = hipSuccess;
  auto fail = [&](hipError_t err) {
    critical("aran_solver3d_solve: %s", hipGetErrorString(err));
    solver->status = code_of_err(err);
  };
  size_t nb = sizeof(double) * std::max<long>(n, 1);
  HostBuf *hb[] = {&solver->hx, &solver->hy, &solver->hz, &solver->hq,
                   &solver->hpot};
  for (HostBuf *b : hb)
    if ((e = b->reserve(nb)) != hipSuccess) return fail(e);
  if ((e = solver->hgrad.reserve(3 * nb)) != hipSuccess) return fail(e);
  DevBuf *db[] = {&solver->dx, &solver->dy, &solver->dz, &solver->dq,
                  &solver->dpot};
  for (DevBuf *b : db)
    if ((e = b->reserve(nb)) != hipSuccess) return fail(e);
  if ((e = solver->dgrad.reserve(3 * nb)) != hipSuccess) return fail(e);
  double *hx = solver->hx.as<double>(), *hy = solver->hy.as<double>(),
         *hz = solver->hz.as<double>(), *hq = solver->hq.as<double>();
  if (solver->soa) {
    memcpy(hx, solver->soa_x.data(), sizeof(double) * n);
    memcpy(hy, solver->soa_y.data(), sizeof(double) * n);
    memcpy(hz, solver->soa_z.data(), sizeof(double) * n);
    memcpy(hq, solver->soa_q.data(), sizeof(double) * n);
  } else {
    const gssize oc = solver->layout.charge;
    void *const *pts = solver->points.data();
    parallel_ranges(n, [=](long a, long b) {
      for (long i = a; i < b; i++) {
        const char *p = static_cast<const char *>(pts[i]);
        const VsgVector3d *v = reinterpret_cast<const VsgVector3d *>(p);
        hx[i] = v->x;
        hy[i] = v->y;
        hz[i] = v->z;
        hq[i] = *reinterpret_cast<const double *>(p + oc);
      }
    });
  }
  const DevBuf *dsts[4] = {&solver->dx, &solver->dy, &solver->dz, &solver->dq};
  double *srcs[4] = {hx, hy, hz, hq};
  for (int k = 0; k < 4 && n > 0; k++)
    if ((e = hipMemcpyAsync(dsts[k]->ptr, srcs[k], sizeof(double) * n,
                            hipMemcpyHostToDevice, s)) != hipSuccess)
      return fail(e);
  const double lb[3] = {solver->prtree->lb.x, solver->prtree->lb.y,
                        solver->prtree->lb.z};
  const double ub[3] = {solver->prtree->ub.x, solver->prtree->ub.y,
                        solver->prtree->ub.z};
  aran_fmm3d_set_semifar(solver->ctx, semifar);
  int rc = aran_fmm3d_solve(solver->ctx, n, solver->dx.as<double>(),
                            solver->dy.as<double>(), solver->dz.as<double>(),
                            solver->dq.as<double>(), lb, ub, order,
                            (int)solver->prtree->maxbox,
                            solver->dpot.as<double>(), solver->dgrad.as<double>());
  if (rc != 0) {
    critical("aran_solver3d_solve: %s", aran::last_error());
    solver->status = rc;
    return;
  }
  double *hp = solver->hpot.as<double>(), *hg = solver->hgrad.as<double>();
  if (n > 0) {
    if ((e = hipMemcpyAsync(hp, solver->dpot.ptr, sizeof(double) * n,
                            hipMemcpyDeviceToHost, s)) != hipSuccess)
      return fail(e);
    if ((e = hipMemcpyAsync(hg, solver->dgrad.ptr, sizeof(double) * 3 * n,
                            hipMemcpyDeviceToHost, s)) != hipSuccess)
      return fail(e);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(e);
  }
  if (solver->soa) {
    solver->res_pot.assign(hp, hp + n);
    solver->res_field.assign(hg, hg + 3 * n);
  } else {
    // each inserted point is its own struct (a point inserted twice would
    // interact with itself at distance 0), so the ranges write disjoint
    // memory
    const Layout L = solver->layout;
    void *const *pts = solver->points.data();
    parallel_ranges(n, [=](long a, long b) {
      for (long i = a; i < b; i++) {
        char *p = static_cast<char *>(pts[i]);
        if (want_pot && L.pot >= 0)
          *reinterpret_cast<double *>(p + L.pot) += hp[i];
        if (want_field && L.field >= 0) {
          double *f = reinterpret_cast<double *>(p + L.field);
          f[0] += hg[3 * i + 0];
          f[1] += hg[3 * i + 1];
          f[2] += hg[3 * i + 2];
        }
      }
    });
  }
  long st[15];
  aran_fmm3d_get_stats(solver->ctx, st);
  for (int k = 0; k < 7; k++) solver->stats[k] += st[k];
  solver->stats[7] += st[12];  // p2l
  solver->stats[8] += st[13];  // m2p
  solver->depth = (guint)st[7];
}
