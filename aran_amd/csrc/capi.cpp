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
#include <aran/aransphericalseriesd.h>
#include <aran/aransolver3d.h>
#include <vsg/vsgd.h>
}

#include "engine.h"
#include "flavours.h"

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

extern "C" void vsg_prtree3d_write(VsgPRTree3d *prtree, FILE *file) {
  if (!prtree || !file) return;
  fprintf(file, "VsgPRTree3d { lbound: [%g, %g, %g], ubound: [%g, %g, %g], "
                "maxbox: %u }\n",
          prtree->lb.x, prtree->lb.y, prtree->lb.z, prtree->ub.x, prtree->ub.y,
          prtree->ub.z, prtree->maxbox);
}

extern "C" void vsg_vector3d_set(VsgVector3d *vec, gdouble x, gdouble y,
                                 gdouble z) {
  vec->x = x;
  vec->y = y;
  vec->z = z;
}

extern "C" void vsg_vector3d_copy(const VsgVector3d *src, VsgVector3d *dst) {
  *dst = *src;
}

extern "C" void vsg_vector3d_add(const VsgVector3d *src,
                                 const VsgVector3d *other, VsgVector3d *dst) {
  const VsgVector3d r = {src->x + other->x, src->y + other->y,
                         src->z + other->z};
  *dst = r;
}

extern "C" void vsg_vector3d_sub(const VsgVector3d *src,
                                 const VsgVector3d *other, VsgVector3d *dst) {
  const VsgVector3d r = {src->x - other->x, src->y - other->y,
                         src->z - other->z};
  *dst = r;
}

extern "C" void vsg_vector3d_scalp(const VsgVector3d *src, gdouble scal,
                                   VsgVector3d *dst) {
  const VsgVector3d r = {scal * src->x, scal * src->y, scal * src->z};
  *dst = r;
}

extern "C" gdouble vsg_vector3d_dotp(const VsgVector3d *one,
                                     const VsgVector3d *other) {
  return one->x * other->x + one->y * other->y + one->z * other->z;
}

extern "C" gdouble vsg_vector3d_norm(const VsgVector3d *vec) {
  return sqrt(vsg_vector3d_dotp(vec, vec));
}

extern "C" void vsg_vector3d_write(const VsgVector3d *vec, FILE *file) {
  fprintf(file, "[%g, %g, %g]", vec->x, vec->y, vec->z);
}

extern "C" void vsg_vector3d_to_spherical_internal(const VsgVector3d *vec,
                                                   gdouble *r, gdouble *cost,
                                                   gdouble *sint,
                                                   gdouble *cosp,
                                                   gdouble *sinp) {
  const double rho = sqrt(vec->x * vec->x + vec->y * vec->y);
  const double rr = sqrt(rho * rho + vec->z * vec->z);
  *r = rr;
  *cost = rr > 0. ? vec->z / rr : 1.;
  *sint = rr > 0. ? rho / rr : 0.;
  *cosp = rho > 0. ? vec->x / rho : 1.;
  *sinp = rho > 0. ? vec->y / rho : 0.;
}

extern "C" void vsg_vector3d_from_spherical_internal(VsgVector3d *vec,
                                                     gdouble r, gdouble cost,
                                                     gdouble sint,
                                                     gdouble cosp,
                                                     gdouble sinp) {
  vec->x = r * sint * cosp;
  vec->y = r * sint * sinp;
  vec->z = r * cost;
}

extern "C" void vsg_packed_msg_trace_set_active(gboolean active) {
  (void)active;
}

// ================================================================ device API

struct _AranFmm3d {
  aran::Workspace ws;
  aran::Stats st;
  aran::TreeInfo ti;
  bool timings = false;
  int precision = ARAN_FMM3D_PLAIN;
  long semifar = -1;
  bool autotune = false;
  int flav[3] = {0, 0, 0};  // M2M, M2L, L2L translation flavours
  long n_last = 0;
  int order_last = 0;
  // workspace of aran_fmm3d_solve_noalloc: its device buffers come from the
  // device scratch below, its pinned staging is one fixed block
  aran::Workspace na;
};

// ---- the no-allocation solve's device scratch: one array per device (the
// loader gives every device its own copy), used through the address
// hipGetSymbolAddress returns on that device (the array's name in host code
// is not a device address).  Every aran_fmm3d_solve_noalloc call on a device
// shares it: calls that overlap on one device race.
static __device__ __align__(256) char scratch[ARAN_FMM3D_NOALLOC_SCRATCH];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};
constexpr size_t kNoAllocStage = 4u << 20;  // pinned staging per context
}  // namespace

extern "C" AranFmm3d *aran_fmm3d_new(void *hip_stream) {
  AranFmm3d *c = new AranFmm3d;
  hipError_t e =
      aran::workspace_init(c->ws, static_cast<hipStream_t>(hip_stream));
  // set-up of the no-allocation solve: the scratch address on this device
  // (once per device), a workspace on the same stream and its staging
  if (e == hipSuccess) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    const int dev = c->ws.device;
    if (dev < 0 || dev >= kMaxDevices) {
      aran::set_error("device index beyond the scratch table");
      e = hipErrorInvalidDevice;
    } else if (!g_scratch[dev]) {
      void *ptr = nullptr;
      e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (e == hipSuccess) g_scratch[dev] = ptr;
      else aran::set_error(std::string("hipGetSymbolAddress: ") +
                           hipGetErrorString(e));
    }
  }
  if (e == hipSuccess) e = aran::workspace_init(c->na, c->ws.stream);
  if (e == hipSuccess) {
    e = hipHostMalloc(reinterpret_cast<void **>(&c->na.host_stage),
                      kNoAllocStage);
    if (e == hipSuccess) {
      c->na.host_stage_bytes = kNoAllocStage;
      c->na.fixed_stage = true;
    }
  }
  if (e != hipSuccess) {
    critical("aran_fmm3d_new: %s", aran::last_error());
    delete c;
    return nullptr;
  }
  return c;
}

// aran_fmm3d_solve on the device scratch: no hipMalloc, hipFree or other
// device or pinned allocation (the tables of the order are rebuilt in the
// scratch every call).
extern "C" int aran_fmm3d_solve_noalloc(AranFmm3d *ctx, long n,
                                        const double *x, const double *y,
                                        const double *z, const double *q,
                                        const double *lb, const double *ub,
                                        int order, int leafcap, double *pot,
                                        double *grad) {
  if (!ctx || n < 0 || (n > 0 && (!x || !y || !z || !q)) || !lb || !ub) {
    aran::set_error("aran_fmm3d_solve_noalloc: invalid argument");
    return ARAN_FMM3D_EINVAL;
  }
  if (n > ARAN_FMM3D_NOALLOC_MAX_N || order < 1 ||
      order > ARAN_FMM3D_NOALLOC_MAX_ORDER ||
      (ctx->precision != ARAN_FMM3D_PLAIN &&
       ctx->precision != ARAN_FMM3D_MIXED) ||
      ctx->flav[0] || ctx->flav[1] || ctx->flav[2]) {
    aran::set_error("aran_fmm3d_solve_noalloc: beyond its limits (n <= "
                    "65536, order <= 10, plain or mixed form, rotate "
                    "translations)");
    return ARAN_FMM3D_EINVAL;
  }
  void *base = g_scratch[ctx->ws.device];
  if (!base) {
    aran::set_error("aran_fmm3d_solve_noalloc: no scratch on this device");
    return ARAN_FMM3D_EINVAL;
  }
  aran::DevArena arena;
  arena.base = static_cast<char *>(base);
  arena.size = ARAN_FMM3D_NOALLOC_SCRATCH;
  ctx->na.forget_arena_buffers();
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
  pb.precision = ctx->precision;
  pb.semifar = ctx->semifar;
  // the plain form's dense FP64 operator tables would not fit the scratch:
  // its far field runs on the per-pair rotate kernel here
  pb.dense_far = false;
  aran::Stats st;
  aran::set_arena(&arena);
  hipError_t e = aran::solve(ctx->na, pb, &st, nullptr);
  aran::set_arena(nullptr);
  return code_of_err(e);
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
  pb.flav_m2m = ctx->flav[0];
  pb.flav_m2l = ctx->flav[1];
  pb.flav_l2l = ctx->flav[2];
  hipError_t e = aran::solve(ctx->ws, pb, &ctx->st, &ctx->ti);
  ctx->n_last = (e == hipSuccess) ? n : 0;
  ctx->order_last = (e == hipSuccess) ? order : 0;
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
  pb.flav_m2m = ctx->flav[0];
  pb.flav_m2l = ctx->flav[1];
  pb.flav_l2l = ctx->flav[2];
  hipError_t e = aran::solve(ctx->ws, pb, &ctx->st, &ctx->ti);
  ctx->n_last = (e == hipSuccess) ? n : 0;
  ctx->order_last = (e == hipSuccess) ? order : 0;
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

extern "C" int aran_fmm3d_get_kernel_timings(const AranFmm3d *ctx,
                                             float *o) {
  if (!ctx || !o) return ARAN_FMM3D_EINVAL;
  const aran::Stats &s = ctx->st;
  float v[4] = {s.t_p2m, s.t_m2m, s.t_l2l, s.t_l2p};
  memcpy(o, v, sizeof v);
  return 0;
}

extern "C" int aran_fmm3d_get_far_plan(const AranFmm3d *ctx, int *o) {
  if (!ctx || !o) return ARAN_FMM3D_EINVAL;
  o[0] = (int)ctx->st.m2l_tiles;
  o[1] = (int)ctx->st.m2l_split;
  o[2] = (int)ctx->st.m2l_sched;
  return 0;
}

extern "C" int aran_fmm3d_set_autotune(AranFmm3d *ctx, int on) {
  if (!ctx) return ARAN_FMM3D_EINVAL;
  ctx->autotune = on != 0;
  return 0;
}

extern "C" int aran_fmm3d_set_translations(AranFmm3d *ctx, int m2m, int m2l,
                                           int l2l) {
  const int f[3] = {m2m, m2l, l2l};
  if (!ctx) return ARAN_FMM3D_EINVAL;
  for (int v : f)
    if (v < ARAN_FMM3D_ROTATE || v > ARAN_FMM3D_KYLIN) {
      aran::set_error("aran_fmm3d_set_translations: invalid flavour");
      return ARAN_FMM3D_EINVAL;
    }
  for (int k = 0; k < 3; k++) ctx->flav[k] = f[k];
  return 0;
}

extern "C" int aran_fmm3d_export_expansions(const AranFmm3d *ctx,
                                            double *multipoles,
                                            double *locals) {
  if (!ctx || ctx->order_last < 1 || ctx->n_last <= 0) {
    aran::set_error("aran_fmm3d_export_expansions: no solve to export");
    return ARAN_FMM3D_EINVAL;
  }
  const aran::Workspace &ws = ctx->ws;
  const long nn = ctx->ti.nnodes;
  const int p = ctx->order_last;
  hipError_t e = hipSuccess;
  if (multipoles)
    e = hipMemcpyAsync(multipoles, ws.mult.ptr,
                       sizeof(double) * 2 * nn * aran::tri(p),
                       hipMemcpyDeviceToHost, ws.stream);
  if (e == hipSuccess && locals)
    e = hipMemcpyAsync(locals, ws.loc.ptr,
                       sizeof(double) * 2 * nn * aran::tri(p + 1),
                       hipMemcpyDeviceToHost, ws.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ws.stream);
  if (e != hipSuccess) aran::set_error(hipGetErrorString(e));
  return code_of_err(e);
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
  nd.anc = nullptr;
  nd.cmap = nullptr;
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

// ---- spherical series (storage of the reference's AranSphericalSeriesd:
// local terms l <= posdeg, then multipole terms l < negdeg, m >= 0, here
// 16-byte aligned after the header)

}  // namespace

struct _AranSphericalSeriesd {
  guint8 posdeg, negdeg;
  guint8 pad_[14];
};

namespace {

static_assert(sizeof(AranSphericalSeriesd) == 16, "series header");

double *ser_pos(const AranSphericalSeriesd *a) {
  return reinterpret_cast<double *>(
      reinterpret_cast<char *>(const_cast<AranSphericalSeriesd *>(a)) + 16);
}
double *ser_neg(const AranSphericalSeriesd *a) {
  return ser_pos(a) + 2 * aran::tri(a->posdeg + 1);
}
size_t ser_count(int posdeg, int negdeg) {
  return (size_t)aran::tri(posdeg + 1) + aran::tri(negdeg);
}

bool all_zero(const double *v, size_t n) {
  for (size_t i = 0; i < n; i++)
    if (v[i] != 0.) return false;
  return true;
}

// The translation engine's order for source / destination degrees: M2M
// takes multipole degree counts, L2L local degrees, M2L a multipole count
// and a local degree (kinds as aran::kFlav*).
int xlate_order(int kind, int sdeg, int ddeg) {
  return std::max(1, std::max(sdeg, ddeg));
}

// dst += T src on the GPU: one translation of the given kind and flavour
// (0 rotate, 1 normal, 2 Kylin) from centre xs to centre xd; src has sdeg,
// dst ddeg degrees (see xlate_order).  Inputs are zero-padded and outputs
// truncated to the engine's order (the terms the reference's loops skip).
int xlate(int kind, int flavour, int sdeg, const double *src, int ddeg,
          double *dst, const VsgVector3d &xs, const VsgVector3d &xd) {
  const int nin_s = kind == aran::kFlavL2L ? aran::tri(sdeg + 1)
                                           : aran::tri(sdeg);
  const int nout_d = kind == aran::kFlavM2M ? aran::tri(ddeg)
                                            : aran::tri(ddeg + 1);
  if (nin_s == 0 || nout_d == 0 || all_zero(src, 2 * (size_t)nin_s)) return 0;
  const int p = xlate_order(kind, sdeg, ddeg);
  if (p > aran::kMaxOrder) {
    aran::set_error("expansion degrees above 30");
    return ARAN_FMM3D_EINVAL;
  }
  const int nin = kind == aran::kFlavL2L ? aran::tri(p + 1) : aran::tri(p);
  const int nout = kind == aran::kFlavM2M ? aran::tri(p) : aran::tri(p + 1);
  std::vector<double> in(2 * (size_t)nin, 0.), out(2 * (size_t)nout, 0.);
  std::copy(src, src + 2 * nin_s, in.begin());
  if (flavour == aran::kFlavRotate) {
    const int rc = translate_one(kind, p, in.data(), one_dir(xs, xd),
                                 out.data());
    if (rc != 0) return rc;
  } else {
    OpContext &oc = opctx();
    std::lock_guard<std::mutex> lock(oc.mu);
    aran::Workspace &ws = oc.ws;
    hipError_t e = aran::workspace_init(ws, nullptr);
    if (e == hipSuccess) e = aran::ensure_tables(ws, p);
    if (e != hipSuccess) return code_of_err(e);
    hipStream_t s = ws.stream;
    e = oc.buf.reserve(sizeof(double) * 2 * ((size_t)nin + nout));
    if (e != hipSuccess) return code_of_err(e);
    double2 *din = oc.buf.as<double2>(), *dout = din + nin;
    aran::FlavVec v;
    const double d[3] = {xd.x - xs.x, xd.y - xs.y, xd.z - xs.z};
    const double rho = sqrt(d[0] * d[0] + d[1] * d[1]);
    v.r = sqrt(rho * rho + d[2] * d[2]);
    v.cost = v.r > 0. ? d[2] / v.r : 1.;
    v.sint = v.r > 0. ? rho / v.r : 0.;
    v.cosp = rho > 0. ? d[0] / rho : 1.;
    v.sinp = rho > 0. ? d[1] / rho : 0.;
    // the engine's degrees: M2M counts and L2L degrees are p on both sides
    // (Kylin's M2M needs equal counts); M2L reads a p-term multipole into a
    // degree-p local
    e = hipMemcpyAsync(din, in.data(), sizeof(double) * in.size(),
                       hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemsetAsync(dout, 0, sizeof(double) * out.size(), s);
    if (e == hipSuccess)
      e = aran::launch_flavour_one(ws.tab.view(), kind, flavour, p, din, p,
                                   dout, v, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(out.data(), dout, sizeof(double) * out.size(),
                         hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      aran::set_error(hipGetErrorString(e));
      return code_of_err(e);
    }
  }
  for (int i = 0; i < 2 * nout_d; i++) dst[i] += out[i];
  return 0;
}

// Series translations: local terms to local terms, and multipole terms to
// multipole terms (to_local: to local terms).  The reference's Kylin
// operators refuse sources of higher degree than the destination
// (aran/aransphericalseriesd-kkylin.c:53,226,387).
void ser_translate(int flavour, bool to_local, const char *fn,
                   const AranSphericalSeriesd *src, const VsgVector3d *xs,
                   AranSphericalSeriesd *dst, const VsgVector3d *xd) {
  if (!src || !dst || !xs || !xd) {
    critical("%s: invalid argument", fn);
    return;
  }
  if (flavour == aran::kFlavKylin &&
      (src->posdeg > dst->posdeg ||
       (!to_local && src->negdeg > dst->negdeg) ||
       (to_local && src->negdeg > 0 && src->negdeg - 1 > dst->posdeg))) {
    critical("%s: source degrees above the destination's", fn);
    return;
  }
  int rc = xlate(aran::kFlavL2L, flavour, src->posdeg, ser_pos(src),
                 dst->posdeg, ser_pos(dst), *xs, *xd);
  if (rc == 0 && src->negdeg > 0) {
    if (to_local)
      rc = xlate(aran::kFlavM2L, flavour, src->negdeg, ser_neg(src),
                 dst->posdeg, ser_pos(dst), *xs, *xd);
    else if (dst->negdeg > 0)
      rc = xlate(aran::kFlavM2M, flavour, src->negdeg, ser_neg(src),
                 dst->negdeg, ser_neg(dst), *xs, *xd);
  }
  if (rc != 0) critical("%s: %s", fn, aran::last_error());
}

// Value and gradient of a series at x relative to its centre: the local
// terms by the L2P kernel, the multipole terms by the M2P kernel.
int ser_eval(const AranSphericalSeriesd *a, const VsgVector3d &x, double *pot,
             double *grad) {
  const VsgVector3d zero = {0., 0., 0.};
  double pl = 0., gl[3] = {0., 0., 0.}, pm = 0., gm[3] = {0., 0., 0.};
  const int lp = std::max(1, (int)a->posdeg);
  if (lp > aran::kMaxOrder || a->negdeg > aran::kMaxOrder) {
    aran::set_error("expansion degrees above 30");
    return ARAN_FMM3D_EINVAL;
  }
  std::vector<double> L(2 * (size_t)aran::tri(lp + 1), 0.);
  std::copy(ser_pos(a), ser_pos(a) + 2 * aran::tri(a->posdeg + 1), L.begin());
  int rc = 0;
  if (!all_zero(L.data(), L.size()))
    rc = leaf_op_one(kOpL2P, lp, zero, x, 0., nullptr, L.data(), &pl, gl);
  if (rc == 0 && a->negdeg > 0)
    rc = leaf_op_one(kOpM2P, a->negdeg, zero, x, 0., ser_neg(a), nullptr, &pm,
                     gm);
  if (pot) *pot = pl + pm;
  if (grad)
    for (int k = 0; k < 3; k++) grad[k] = gl[k] + gm[k];
  return rc;
}

// The engine's FMM order of a development (the solver's nodes): its
// multipole degree count; its local degree must be the same number.
bool dev_ok(const AranDevelopment3d *d, const char *fn) {
  if (!d || !d->multipole || !d->local || d->multipole->negdeg < 1 ||
      d->multipole->negdeg > aran::kMaxOrder ||
      d->local->posdeg != d->multipole->negdeg) {
    critical("%s: the solver's development must be aran_development3d_new "
             "(posdeg, p) with posdeg <= p and 1 <= p <= 30", fn);
    return false;
  }
  return true;
}

bool dev_any(const AranDevelopment3d *d, const char *fn) {
  if (!d || !d->multipole || !d->local) {
    critical("%s: invalid development", fn);
    return false;
  }
  return true;
}

}  // namespace

// ================================================================ series

extern "C" AranSphericalSeriesd *aran_spherical_seriesd_new(guint8 posdeg,
                                                            guint8 negdeg) {
  void *mem = calloc(1, 16 + 16 * ser_count(posdeg, negdeg));
  AranSphericalSeriesd *a = static_cast<AranSphericalSeriesd *>(mem);
  a->posdeg = posdeg;
  a->negdeg = negdeg;
  return a;
}

extern "C" void aran_spherical_seriesd_free(AranSphericalSeriesd *ass) {
  free(ass);
}

// Copies the common degrees (aran/aransphericalseriesd.c:158-182).
extern "C" void aran_spherical_seriesd_copy(const AranSphericalSeriesd *src,
                                            AranSphericalSeriesd *dst) {
  if (!src || !dst) return;
  if (src->posdeg > dst->posdeg || src->negdeg > dst->negdeg)
    critical("aran_spherical_seriesd_copy: could lose precision");
  aran_spherical_seriesd_set_zero(dst);
  const int pd = std::min(src->posdeg, dst->posdeg);
  const int nd = std::min(src->negdeg, dst->negdeg);
  memcpy(ser_pos(dst), ser_pos(src), sizeof(double) * 2 * aran::tri(pd + 1));
  memcpy(ser_neg(dst), ser_neg(src), sizeof(double) * 2 * aran::tri(nd));
}

extern "C" AranSphericalSeriesd *
aran_spherical_seriesd_clone(const AranSphericalSeriesd *src) {
  if (!src) return nullptr;
  AranSphericalSeriesd *a =
      aran_spherical_seriesd_new(src->posdeg, src->negdeg);
  memcpy(ser_pos(a), ser_pos(src),
         16 * ser_count(src->posdeg, src->negdeg));
  return a;
}

extern "C" gcomplex128 *aran_spherical_seriesd_get_term(
    AranSphericalSeriesd *ass, gint i, gint j) {
  if (!ass || j < 0 || j > (i < 0 ? -i - 1 : i)) return nullptr;
  if (i >= 0) {
    if (i > ass->posdeg) return nullptr;
    return reinterpret_cast<gcomplex128 *>(ser_pos(ass)) + aran::tri(i) + j;
  }
  const int l = -i - 1;
  if (l >= ass->negdeg) return nullptr;
  return reinterpret_cast<gcomplex128 *>(ser_neg(ass)) + aran::tri(l) + j;
}

extern "C" guint8 aran_spherical_seriesd_get_posdeg(
    const AranSphericalSeriesd *ass) {
  return ass ? ass->posdeg : 0;
}

extern "C" guint8 aran_spherical_seriesd_get_negdeg(
    const AranSphericalSeriesd *ass) {
  return ass ? ass->negdeg : 0;
}

extern "C" void aran_spherical_seriesd_set_zero(AranSphericalSeriesd *ass) {
  if (ass) memset(ser_pos(ass), 0, 16 * ser_count(ass->posdeg, ass->negdeg));
}

// "[i,j]:(re,im), " for i = -negdeg..posdeg (aran/aransphericalseriesd.c:
// 272-298)
extern "C" void aran_spherical_seriesd_write(const AranSphericalSeriesd *ass,
                                             FILE *file) {
  if (!ass || !file) return;
  fprintf(file, "[");
  for (int i = -ass->negdeg; i <= ass->posdeg; i++) {
    const int maxj = i < 0 ? -i - 1 : i;
    for (int j = 0; j <= maxj; j++) {
      const double *c =
          i < 0 ? ser_neg(ass) + 2 * (aran::tri(-i - 1) + j)
                : ser_pos(ass) + 2 * (aran::tri(i) + j);
      fprintf(file, "[%d,%d]:(%e,%e), ", i, j, c[0], c[1]);
    }
  }
  fprintf(file, "]");
}

extern "C" gcomplex128 aran_spherical_seriesd_evaluate(
    const AranSphericalSeriesd *ass, const VsgVector3d *x) {
  double v = 0.;
  if (ass && x && ser_eval(ass, *x, &v, nullptr) != 0)
    critical("aran_spherical_seriesd_evaluate: %s", aran::last_error());
  return v;
}

extern "C" void aran_spherical_seriesd_gradient_evaluate(
    const AranSphericalSeriesd *ass, const VsgVector3d *x, VsgVector3d *grad) {
  double g[3] = {0., 0., 0.};
  if (ass && x && ser_eval(ass, *x, nullptr, g) != 0)
    critical("aran_spherical_seriesd_gradient_evaluate: %s",
             aran::last_error());
  if (grad) {
    grad->x = g[0];
    grad->y = g[1];
    grad->z = g[2];
  }
}

#define ARAN_SERIES_XLATE(name, flavour, to_local)                           \
  extern "C" void name(const AranSphericalSeriesd *src,                      \
                       const VsgVector3d *xsrc, AranSphericalSeriesd *dst,   \
                       const VsgVector3d *xdst) {                            \
    ser_translate(flavour, to_local, #name, src, xsrc, dst, xdst);           \
  }
ARAN_SERIES_XLATE(aran_spherical_seriesd_translate, aran::kFlavNormal, false)
ARAN_SERIES_XLATE(aran_spherical_seriesd_to_local, aran::kFlavNormal, true)
ARAN_SERIES_XLATE(aran_spherical_seriesd_translate_kkylin, aran::kFlavKylin,
                  false)
ARAN_SERIES_XLATE(aran_spherical_seriesd_to_local_kkylin, aran::kFlavKylin,
                  true)
ARAN_SERIES_XLATE(aran_spherical_seriesd_translate_rotate, aran::kFlavRotate,
                  false)
ARAN_SERIES_XLATE(aran_spherical_seriesd_to_local_rotate, aran::kFlavRotate,
                  true)
#undef ARAN_SERIES_XLATE

// ================================================================ development

extern "C" GType aran_development3d_get_type(void) { return kDevelopmentType; }

// (aran/arandevelopment3d.c:64-73)
extern "C" AranDevelopment3d *aran_development3d_new(guint8 posdeg,
                                                     guint8 negdeg) {
  AranDevelopment3d *d =
      static_cast<AranDevelopment3d *>(calloc(1, sizeof(AranDevelopment3d)));
  d->multipole = aran_spherical_seriesd_new(posdeg, negdeg);
  d->local = aran_spherical_seriesd_new(std::max(posdeg, negdeg), 0);
  return d;
}

extern "C" void aran_development3d_free(AranDevelopment3d *ad) {
  if (!ad) return;
  aran_spherical_seriesd_free(ad->multipole);
  aran_spherical_seriesd_free(ad->local);
  free(ad);
}

extern "C" void aran_development3d_copy(const AranDevelopment3d *src,
                                        AranDevelopment3d *dst) {
  if (!src || !dst) {
    critical("aran_development3d_copy: invalid argument");
    return;
  }
  aran_spherical_seriesd_copy(src->multipole, dst->multipole);
  aran_spherical_seriesd_copy(src->local, dst->local);
}

extern "C" AranDevelopment3d *aran_development3d_clone(AranDevelopment3d *src) {
  if (!src) return nullptr;
  AranDevelopment3d *d = aran_development3d_new(src->multipole->posdeg,
                                                src->multipole->negdeg);
  aran_development3d_copy(src, d);
  return d;
}

extern "C" void aran_development3d_set_zero(AranDevelopment3d *ad) {
  if (!ad) return;
  aran_spherical_seriesd_set_zero(ad->multipole);
  aran_spherical_seriesd_set_zero(ad->local);
}

extern "C" void aran_development3d_write(AranDevelopment3d *ad, FILE *file) {
  if (!ad || !file) return;
  fprintf(file, "{m:");
  aran_spherical_seriesd_write(ad->multipole, file);
  fprintf(file, ", l:");
  aran_spherical_seriesd_write(ad->local, file);
  fprintf(file, "}");
}

// q (4 pi/(2l+1)) r^l conj(Y_l^m) into the multipole terms
// (aran/arandevelopment3d.c:169-216)
extern "C" void aran_development3d_p2m(const VsgVector3d *position,
                                       const gdouble charge,
                                       const VsgPRTree3dNodeInfo *dst_node,
                                       AranDevelopment3d *dst) {
  if (!position || !dst_node || !dev_any(dst, "aran_development3d_p2m")) return;
  const int nm = dst->multipole->negdeg;
  if (nm == 0) return;
  if (leaf_op_one(kOpP2M, nm, dst_node->center, *position, charge,
                  ser_neg(dst->multipole), nullptr, nullptr, nullptr) != 0)
    critical("aran_development3d_p2m: %s", aran::last_error());
}

// q (4 pi/(2l+1)) conj(Y_l^m) / r^(l+1) into the local terms
// (aran/arandevelopment3d.c:227-277)
extern "C" void aran_development3d_p2l(const VsgVector3d *position,
                                       const gdouble charge,
                                       const VsgPRTree3dNodeInfo *dst_node,
                                       AranDevelopment3d *dst) {
  if (!position || !dst_node || !dev_any(dst, "aran_development3d_p2l")) return;
  const int pd = dst->local->posdeg, lp = std::max(1, pd);
  std::vector<double> L(2 * (size_t)aran::tri(lp + 1), 0.);
  if (leaf_op_one(kOpP2L, lp, dst_node->center, *position, charge, nullptr,
                  L.data(), nullptr, nullptr) != 0) {
    critical("aran_development3d_p2l: %s", aran::last_error());
    return;
  }
  double *out = ser_pos(dst->local);
  for (int i = 0; i < 2 * aran::tri(pd + 1); i++) out[i] += L[i];
}

#define ARAN_DEV_XLATE(name, flavour, part_src, part_dst, to_local)          \
  extern "C" void name(const VsgPRTree3dNodeInfo *src_node,                  \
                       AranDevelopment3d *src,                               \
                       const VsgPRTree3dNodeInfo *dst_node,                  \
                       AranDevelopment3d *dst) {                             \
    if (!src_node || !dst_node || !dev_any(src, #name) || !dev_any(dst, #name)) \
      return;                                                                \
    ser_translate(flavour, to_local, #name, src->part_src, &src_node->center, \
                  dst->part_dst, &dst_node->center);                         \
  }
// (aran/arandevelopment3d.c:288-331, 343-390, 402-451)
ARAN_DEV_XLATE(aran_development3d_m2m, aran::kFlavNormal, multipole,
               multipole, false)
ARAN_DEV_XLATE(aran_development3d_m2l, aran::kFlavNormal, multipole, local,
               true)
ARAN_DEV_XLATE(aran_development3d_l2l, aran::kFlavNormal, local, local, false)
ARAN_DEV_XLATE(aran_development3d_m2m_kkylin, aran::kFlavKylin, multipole,
               multipole, false)
ARAN_DEV_XLATE(aran_development3d_m2l_kkylin, aran::kFlavKylin, multipole,
               local, true)
ARAN_DEV_XLATE(aran_development3d_l2l_kkylin, aran::kFlavKylin, local, local,
               false)
ARAN_DEV_XLATE(aran_development3d_m2m_rotate, aran::kFlavRotate, multipole,
               multipole, false)
ARAN_DEV_XLATE(aran_development3d_m2l_rotate, aran::kFlavRotate, multipole,
               local, true)
ARAN_DEV_XLATE(aran_development3d_l2l_rotate, aran::kFlavRotate, local, local,
               false)
#undef ARAN_DEV_XLATE

static VsgVector3d rel(const VsgVector3d &p, const VsgVector3d &c) {
  VsgVector3d r = {p.x - c.x, p.y - c.y, p.z - c.z};
  return r;
}

// (aran/arandevelopment3d.c:463-568): the series at pos - centre
extern "C" gcomplex128 aran_development3d_local_evaluate(
    const VsgPRTree3dNodeInfo *devel_node, AranDevelopment3d *devel,
    const VsgVector3d *pos) {
  if (!devel_node || !pos || !dev_any(devel, "aran_development3d_local_evaluate"))
    return 0.;
  const VsgVector3d x = rel(*pos, devel_node->center);
  return aran_spherical_seriesd_evaluate(devel->local, &x);
}

extern "C" void aran_development3d_local_gradient_evaluate(
    const VsgPRTree3dNodeInfo *devel_node, AranDevelopment3d *devel,
    const VsgVector3d *pos, VsgVector3d *grad) {
  if (!devel_node || !pos || !grad ||
      !dev_any(devel, "aran_development3d_local_gradient_evaluate"))
    return;
  const VsgVector3d x = rel(*pos, devel_node->center);
  aran_spherical_seriesd_gradient_evaluate(devel->local, &x, grad);
}

extern "C" gcomplex128 aran_development3d_multipole_evaluate(
    const VsgPRTree3dNodeInfo *devel_node, AranDevelopment3d *devel,
    const VsgVector3d *pos) {
  if (!devel_node || !pos ||
      !dev_any(devel, "aran_development3d_multipole_evaluate"))
    return 0.;
  const VsgVector3d x = rel(*pos, devel_node->center);
  return aran_spherical_seriesd_evaluate(devel->multipole, &x);
}

extern "C" void aran_development3d_multipole_gradient_evaluate(
    const VsgPRTree3dNodeInfo *devel_node, AranDevelopment3d *devel,
    const VsgVector3d *pos, VsgVector3d *grad) {
  if (!devel_node || !pos || !grad ||
      !dev_any(devel, "aran_development3d_multipole_gradient_evaluate"))
    return;
  const VsgVector3d x = rel(*pos, devel_node->center);
  aran_spherical_seriesd_gradient_evaluate(devel->multipole, &x, grad);
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
  // near/far leaf predicate (aran_solver3d_set_nf_isleaf)
  VsgPRTree3dNFIsleafFunc nf_isleaf = nullptr;
  gpointer nf_isleaf_data = nullptr;
  bool hilbert = false;  // children order of aran_solver3d_traverse
  bool solved = false;   // expansions of the current particles on the device
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

// The tree's bounds grow to take a particle that falls outside them and
// keep that size when it is removed again, as the reference's tree does
// (the Plummer fill inserts and removes corner points for that,
// tests/newtonpot3.c:528-542): the root cube doubles about its centre until
// it holds the particle (the growth rule of the engine's root box).
static void grow_bounds(VsgPRTree3d *t, double x, double y, double z) {
  double c[3], h = 0.;
  const double lo[3] = {t->lb.x, t->lb.y, t->lb.z},
               hi[3] = {t->ub.x, t->ub.y, t->ub.z}, v[3] = {x, y, z};
  bool inside = true;
  for (int k = 0; k < 3; k++)
    if (v[k] < lo[k] || v[k] > hi[k]) inside = false;
  if (inside || !(std::isfinite(x) && std::isfinite(y) && std::isfinite(z)))
    return;
  for (int k = 0; k < 3; k++) {
    c[k] = 0.5 * (lo[k] + hi[k]);
    h = std::max(h, 0.5 * (hi[k] - lo[k]));
  }
  if (!(h > 0.)) h = 1.;
  for (;;) {
    bool in = true;
    for (int k = 0; k < 3; k++)
      if (v[k] < c[k] - h || v[k] > c[k] + h) in = false;
    if (in) break;
    h *= 2.;
  }
  t->lb.x = c[0] - h;
  t->lb.y = c[1] - h;
  t->lb.z = c[2] - h;
  t->ub.x = c[0] + h;
  t->ub.y = c[1] + h;
  t->ub.z = c[2] + h;
}

extern "C" void aran_solver3d_insert_point(AranSolver3d *solver,
                                           VsgPoint3 point) {
  if (!solver || !point) return;
  if (solver->soa) {
    critical("aran_solver3d_insert_point: solver holds SoA particles");
    return;
  }
  const VsgVector3d *v = static_cast<const VsgVector3d *>(point);
  grow_bounds(solver->prtree, v->x, v->y, v->z);
  solver->points.push_back(point);
  solver->solved = false;
}

// Without a distributed tree every insertion is local
// (aran/aransolver3d.c:980-1000).
extern "C" void aran_solver3d_insert_point_local(AranSolver3d *solver,
                                                 VsgPoint3 point) {
  aran_solver3d_insert_point(solver, point);
}

extern "C" void aran_solver3d_set_nf_isleaf(AranSolver3d *solver,
                                            VsgPRTree3dNFIsleafFunc isleaf,
                                            gpointer user_data) {
  if (!solver) return;
  solver->nf_isleaf = isleaf;
  solver->nf_isleaf_data = user_data;
  solver->solved = false;
}

extern "C" void aran_solver3d_set_children_order_hilbert(AranSolver3d *solver) {
  if (solver) solver->hilbert = true;
}

extern "C" void aran_solver3d_set_children_order_default(AranSolver3d *solver) {
  if (solver) solver->hilbert = false;
}

extern "C" gboolean aran_solver3d_remove_point(AranSolver3d *solver,
                                               VsgPoint3 point) {
  if (!solver) return FALSE;
  auto it = std::find(solver->points.begin(), solver->points.end(), point);
  if (it == solver->points.end()) return FALSE;
  solver->points.erase(it);
  solver->solved = false;
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
  solver->solved = false;
  for (long i = 0; i < n; i++) grow_bounds(solver->prtree, x[i], y[i], z[i]);
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

namespace {

// Flavour of a registered translation: 0 rotate, 1 normal, 2 Kylin, -1 not a
// library operator.
template <class F, class A, class B, class C>
int flavour_of(F f, A rotate, B normal, C kylin) {
  if (same_fn(f, rotate)) return aran::kFlavRotate;
  if (same_fn(f, normal)) return aran::kFlavNormal;
  if (same_fn(f, kylin)) return aran::kFlavKylin;
  return -1;
}

struct SolveOps {
  int outputs = 0;  // 1 potential, 2 field, 3 both
  int flav[3] = {0, 0, 0};
  long semifar = -1;
  int order = 0;
};

// The recognised operator set (see aransolver3d.h); false with a critical
// message otherwise.
bool solver_ops(AranSolver3d *solver, SolveOps *o) {
  int pot_p2p = -1, pot_l2p = -1;
  if (solver->p2p == aran_newton3d_p2p_pot) pot_p2p = 1;
  if (solver->p2p == aran_newton3d_p2p_field) pot_p2p = 2;
  if (solver->p2p == aran_newton3d_p2p_potfield) pot_p2p = 3;
  if (solver->l2p == aran_newton3d_l2p_pot) pot_l2p = 1;
  if (solver->l2p == aran_newton3d_l2p_field) pot_l2p = 2;
  if (solver->l2p == aran_newton3d_l2p_potfield) pot_l2p = 3;
  o->flav[0] = flavour_of(
      solver->m2m, aran_development3d_m2m_rotate, aran_development3d_m2m,
      aran_development3d_m2m_kkylin);
  o->flav[1] = flavour_of(
      solver->m2l, aran_development3d_m2l_rotate, aran_development3d_m2l,
      aran_development3d_m2l_kkylin);
  o->flav[2] = flavour_of(
      solver->l2l, aran_development3d_l2l_rotate, aran_development3d_l2l,
      aran_development3d_l2l_kkylin);
  const bool ops_ok = pot_p2p > 0 && pot_p2p == pot_l2p &&
                      solver->p2m == aran_newton3d_p2m && o->flav[0] >= 0 &&
                      o->flav[1] >= 0 && o->flav[2] >= 0;
  if (!ops_ok) {
    critical("aran_solver3d_solve: unsupported callbacks; the GPU engine runs "
             "the library's Newton callbacks (aran_newton3d_*) with the "
             "aran_development3d_{m2m,m2l,l2l}[_rotate|_kkylin] operators");
    return false;
  }
  if ((o->flav[0] | o->flav[1] | o->flav[2]) != 0 &&
      solver->precision != ARAN_FMM3D_PLAIN) {
    critical("aran_solver3d_solve: the normal and Kylin operators run in the "
             "plain FP64 form only");
    return false;
  }
  if (!dev_ok(solver->devel, "aran_solver3d_solve")) return false;
  o->order = solver->devel->multipole->negdeg;
  o->outputs = pot_p2p;
  // semifar pairs when both P2L and M2P are set (aran/aransolver3d.c:
  // 1152-1161): the library's P2L with the M2P of the same outputs as P2P
  if (solver->p2l && solver->m2p) {
    int kind = -1;
    if (solver->m2p == aran_newton3d_m2p_pot) kind = 1;
    if (solver->m2p == aran_newton3d_m2p_field) kind = 2;
    if (solver->m2p == aran_newton3d_m2p_potfield) kind = 3;
    if (solver->p2l != aran_newton3d_p2l || kind != pot_p2p) {
      critical("aran_solver3d_solve: unsupported semifar callbacks; the GPU "
               "engine runs aran_newton3d_p2l with the aran_newton3d_m2p_* "
               "callback of the same outputs as the P2P callback");
      return false;
    }
    if (solver->semifar_threshold == 0) {
      // automatic threshold (the reference derives it from profiled
      // operator times, aran/aransolver3d.c:1090-1127; here from the
      // operators' algorithmic costs)
      solver->semifar_threshold =
          (guint)aran_fmm3d_semifar_threshold_auto(o->order);
    }
    if (solver->semifar_threshold != 0xffffffffu)
      o->semifar = (long)solver->semifar_threshold;
  }
  return true;
}

// Particle positions and charges on the device (gathered from the caller's
// structs through the declared layout, or the SoA arrays); *n their count.
int solver_upload(AranSolver3d *solver, long *n_out) {
  const long n =
      (long)(solver->soa ? solver->soa_x.size() : solver->points.size());
  *n_out = n;
  if (!solver->ctx) {
    solver->ctx = aran_fmm3d_new(nullptr);
    if (!solver->ctx) return ARAN_FMM3D_EHIP;
  }
  hipStream_t s = solver->ctx->ws.stream;
  hipError_t e = hipSuccess;
  const size_t nb = sizeof(double) * std::max<long>(n, 1);
  HostBuf *hb[] = {&solver->hx, &solver->hy, &solver->hz, &solver->hq,
                   &solver->hpot};
  for (HostBuf *b : hb)
    if ((e = b->reserve(nb)) != hipSuccess) return code_of_err(e);
  if ((e = solver->hgrad.reserve(3 * nb)) != hipSuccess) return code_of_err(e);
  DevBuf *db[] = {&solver->dx, &solver->dy, &solver->dz, &solver->dq,
                  &solver->dpot};
  for (DevBuf *b : db)
    if ((e = b->reserve(nb)) != hipSuccess) return code_of_err(e);
  if ((e = solver->dgrad.reserve(3 * nb)) != hipSuccess) return code_of_err(e);
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
  const DevBuf *dsts[4] = {&solver->dx, &solver->dy, &solver->dz,
                           &solver->dq};
  double *srcs[4] = {hx, hy, hz, hq};
  for (int k = 0; k < 4 && n > 0; k++)
    if ((e = hipMemcpyAsync(dsts[k]->ptr, srcs[k], sizeof(double) * n,
                            hipMemcpyHostToDevice, s)) != hipSuccess)
      return code_of_err(e);
  return 0;
}

// The engine's problem for the solver's uploaded particles.
aran::Problem solver_problem(AranSolver3d *solver, long n, int order) {
  aran::Problem pb;
  pb.n = n;
  pb.order = order;
  pb.leafcap = (int)solver->prtree->maxbox;
  pb.lb[0] = solver->prtree->lb.x;
  pb.lb[1] = solver->prtree->lb.y;
  pb.lb[2] = solver->prtree->lb.z;
  pb.ub[0] = solver->prtree->ub.x;
  pb.ub[1] = solver->prtree->ub.y;
  pb.ub[2] = solver->prtree->ub.z;
  pb.x = solver->dx.as<double>();
  pb.y = solver->dy.as<double>();
  pb.z = solver->dz.as<double>();
  pb.q = solver->dq.as<double>();
  return pb;
}

// Host view of the engine's last tree: node records and LibVsg node infos
// (bounds, centre, particle count, leaf flag, depth, father, key); the
// particle lists are not kept on the host (point_list is NULL).
struct HostTree {
  long nn = 0;
  std::vector<int> nodes;  // 10 ints per node (aran_fmm3d_export_tree)
  std::vector<double> centers;
  std::vector<VsgPRTree3dNodeInfo> info;
};

int host_tree(AranFmm3d *ctx, HostTree *t) {
  long st[15];
  aran_fmm3d_get_stats(ctx, st);
  t->nn = st[8];
  t->nodes.assign(10 * (size_t)t->nn, 0);
  t->centers.assign(3 * (size_t)t->nn, 0.);
  double lo[3], width = 1.;
  int rc = aran_fmm3d_export_tree(ctx, nullptr, t->nodes.data(),
                                  t->centers.data(), lo, &width);
  if (rc != 0) return rc;
  t->info.assign(t->nn, VsgPRTree3dNodeInfo());
  for (long i = 0; i < t->nn; i++) {
    const int *r = &t->nodes[10 * i];
    VsgPRTree3dNodeInfo &f = t->info[i];
    memset(&f, 0, sizeof f);
    const double h = 0.5 * ldexp(width, -r[0]);
    f.center.x = t->centers[3 * i];
    f.center.y = t->centers[3 * i + 1];
    f.center.z = t->centers[3 * i + 2];
    f.lbound.x = f.center.x - h;
    f.lbound.y = f.center.y - h;
    f.lbound.z = f.center.z - h;
    f.ubound.x = f.center.x + h;
    f.ubound.y = f.center.y + h;
    f.ubound.z = f.center.z + h;
    f.point_count = (guint)(r[5] - r[4]);
    f.point_list = nullptr;
    f.isleaf = r[7] ? TRUE : FALSE;
    f.depth = (guint8)r[0];
    f.father_info = r[6] >= 0 ? &t->info[r[6]] : nullptr;
    f.user_data = nullptr;
    f.id.x = (guint64)r[1];
    f.id.y = (guint64)r[2];
    f.id.z = (guint64)r[3];
    f.id.depth = (guint8)r[0];
  }
  return 0;
}

// Nodes the near/far leaf predicate turns into leaves: the first node on
// each root-to-leaf path (breadth-first node order, parents first) for
// which it returns TRUE, as (depth, Morton key), sorted.
int solver_stop_list(AranSolver3d *solver, long n, int order,
                     std::vector<unsigned long long> *stop) {
  aran::Problem pb = solver_problem(solver, n, order);
  pb.tree_only = true;
  AranFmm3d *ctx = solver->ctx;
  hipError_t e = aran::solve(ctx->ws, pb, &ctx->st, &ctx->ti);
  ctx->n_last = (e == hipSuccess) ? n : 0;
  ctx->order_last = 0;
  if (e != hipSuccess) return code_of_err(e);
  HostTree t;
  int rc = host_tree(ctx, &t);
  if (rc != 0) return rc;
  std::vector<char> cut(t.nn, 0);
  std::vector<std::pair<unsigned long long, unsigned long long>> v;
  for (long i = 0; i < t.nn; i++) {
    const int *r = &t.nodes[10 * i];
    if (r[6] >= 0 && cut[r[6]]) {
      cut[i] = 1;
      continue;
    }
    if (r[7]) continue;  // a leaf already
    if (solver->nf_isleaf(&t.info[i], solver->nf_isleaf_data)) {
      cut[i] = 1;
      v.emplace_back((unsigned long long)r[0],
                     (unsigned long long)aran::morton_key(r[1], r[2], r[3]));
    }
  }
  std::sort(v.begin(), v.end());
  stop->clear();
  for (auto &x : v) {
    stop->push_back(x.first);
    stop->push_back(x.second);
  }
  return 0;
}

// A Hilbert index of integer coordinates (b bits each) by J. Skilling's
// transpose algorithm ("Programming the Hilbert curve", AIP Conf. Proc. 707,
// 2004): the order in which aran_solver3d_traverse visits siblings after
// aran_solver3d_set_children_order_hilbert.
uint64_t hilbert_index(unsigned x, unsigned y, unsigned z, int b) {
  if (b <= 0) return 0;
  unsigned X[3] = {x, y, z};
  const unsigned M = 1u << (b - 1);
  for (unsigned Q = M; Q > 1; Q >>= 1) {
    const unsigned P = Q - 1;
    for (int i = 0; i < 3; i++) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const unsigned t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; i++) X[i] ^= X[i - 1];
  unsigned t = 0;
  for (unsigned Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; i++) X[i] ^= t;
  uint64_t h = 0;
  for (int bit = b - 1; bit >= 0; bit--)
    for (int i = 0; i < 3; i++) h = (h << 1) | ((X[i] >> bit) & 1u);
  return h;
}

}  // namespace

// Visits the nodes of the octree of the inserted particles (the last
// solve's, or one built now) with func (aran/aransolver3d.c:1204-1217).
// After a solve of the same particles each node's user_data is an
// AranDevelopment3d holding its multipole and local; otherwise NULL.
extern "C" void aran_solver3d_traverse(AranSolver3d *solver,
                                       GTraverseType order,
                                       VsgPRTree3dFunc func,
                                       gpointer user_data) {
  if (!solver || !func) return;
  long n = 0;
  if (!solver->solved) {
    if (!dev_ok(solver->devel, "aran_solver3d_traverse")) return;
    int rc = solver_upload(solver, &n);
    if (rc == 0) {
      aran::Problem pb =
          solver_problem(solver, n, solver->devel->multipole->negdeg);
      pb.tree_only = true;
      AranFmm3d *ctx = solver->ctx;
      hipError_t e = aran::solve(ctx->ws, pb, &ctx->st, &ctx->ti);
      ctx->n_last = (e == hipSuccess) ? n : 0;
      ctx->order_last = 0;
      rc = code_of_err(e);
    }
    if (rc != 0) {
      critical("aran_solver3d_traverse: %s", aran::last_error());
      return;
    }
  }
  HostTree t;
  if (solver->ctx->n_last <= 0 || host_tree(solver->ctx, &t) != 0) {
    if (solver->ctx->n_last > 0)
      critical("aran_solver3d_traverse: %s", aran::last_error());
    return;
  }
  // node developments from the device expansions of the last solve
  std::vector<AranDevelopment3d *> devs;
  if (solver->solved) {
    const int p = solver->ctx->order_last;
    const int mc = aran::tri(p), lc = aran::tri(p + 1);
    std::vector<double> M(2 * (size_t)t.nn * mc), L(2 * (size_t)t.nn * lc);
    if (aran_fmm3d_export_expansions(solver->ctx, M.data(), L.data()) != 0) {
      critical("aran_solver3d_traverse: %s", aran::last_error());
      return;
    }
    devs.resize(t.nn);
    for (long i = 0; i < t.nn; i++) {
      devs[i] = aran_development3d_clone(solver->devel);
      aran_development3d_set_zero(devs[i]);
      memcpy(ser_neg(devs[i]->multipole), &M[2 * (size_t)i * mc],
             sizeof(double) * 2 * mc);
      memcpy(ser_pos(devs[i]->local), &L[2 * (size_t)i * lc],
             sizeof(double) * 2 * lc);
      t.info[i].user_data = devs[i];
    }
  }
  // children of each node, in octant order or along the Hilbert curve
  auto children = [&](long v) {
    std::vector<long> c;
    const int *r = &t.nodes[10 * v];
    for (int k = 0; k < r[9]; k++) c.push_back(r[8] + k);
    if (solver->hilbert)
      std::stable_sort(c.begin(), c.end(), [&](long a, long b) {
        const int *ra = &t.nodes[10 * a], *rb = &t.nodes[10 * b];
        return hilbert_index(ra[1], ra[2], ra[3], ra[0]) <
               hilbert_index(rb[1], rb[2], rb[3], rb[0]);
      });
    return c;
  };
  if (order == G_LEVEL_ORDER) {
    for (long i = 0; i < t.nn; i++) func(&t.info[i], user_data);
  } else {
    // depth-first from the root (node 0) without recursion
    struct Frame {
      long node;
      std::vector<long> kids;
      size_t next;
    };
    std::vector<Frame> stack;
    auto push = [&](long v) {
      if (order == G_PRE_ORDER) func(&t.info[v], user_data);
      stack.push_back(Frame{v, children(v), 0});
    };
    if (t.nn > 0) push(0);
    while (!stack.empty()) {
      Frame &f = stack.back();
      if (order == G_IN_ORDER && f.next == 1)  // after the first child
        func(&t.info[f.node], user_data);
      if (f.next < f.kids.size()) {
        const long c = f.kids[f.next++];
        push(c);
        continue;
      }
      const long v = f.node;
      const bool leafish = f.kids.empty();
      stack.pop_back();
      if (order == G_POST_ORDER || (order == G_IN_ORDER && leafish))
        func(&t.info[v], user_data);
    }
  }
  for (AranDevelopment3d *d : devs) aran_development3d_free(d);
}

extern "C" void aran_solver3d_solve(AranSolver3d *solver) {
  if (!solver) return;
  solver->status = 0;
  solver->solved = false;
  SolveOps ops;
  if (!solver_ops(solver, &ops)) {
    solver->status = ARAN_FMM3D_EINVAL;
    return;
  }
  const bool want_pot = (ops.outputs & 1) != 0,
             want_field = (ops.outputs & 2) != 0;
  long n = 0;
  int rc = solver_upload(solver, &n);
  if (rc != 0) {
    critical("aran_solver3d_solve: %s", aran::last_error());
    solver->status = rc;
    return;
  }
  std::vector<unsigned long long> stop;
  if (solver->nf_isleaf && n > 0) {
    rc = solver_stop_list(solver, n, ops.order, &stop);
    if (rc != 0) {
      critical("aran_solver3d_solve: %s", aran::last_error());
      solver->status = rc;
      return;
    }
  }
  AranFmm3d *ctx = solver->ctx;
  aran_fmm3d_set_precision(ctx, solver->precision);
  aran_fmm3d_set_semifar(ctx, ops.semifar);
  aran_fmm3d_set_translations(ctx, ops.flav[0], ops.flav[1], ops.flav[2]);
  hipStream_t s = ctx->ws.stream;
  aran::Problem pb = solver_problem(solver, n, ops.order);
  pb.pot = solver->dpot.as<double>();
  pb.grad = solver->dgrad.as<double>();
  pb.timings = ctx->timings;
  pb.precision = ctx->precision;
  pb.semifar = ctx->semifar;
  pb.autotune = ctx->autotune;
  pb.flav_m2m = ctx->flav[0];
  pb.flav_m2l = ctx->flav[1];
  pb.flav_l2l = ctx->flav[2];
  pb.stop = stop.empty() ? nullptr : stop.data();
  pb.nstop = (int)(stop.size() / 2);
  hipError_t e = aran::solve(ctx->ws, pb, &ctx->st, &ctx->ti);
  ctx->n_last = (e == hipSuccess) ? n : 0;
  ctx->order_last = (e == hipSuccess) ? ops.order : 0;
  if (e != hipSuccess) {
    critical("aran_solver3d_solve: %s", aran::last_error());
    solver->status = code_of_err(e);
    return;
  }
  auto fail = [&](hipError_t err) {
    critical("aran_solver3d_solve: %s", hipGetErrorString(err));
    solver->status = code_of_err(err);
  };
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
        if (want_pot && L.pot >= 0) {
          // a gcomplex128 potential takes the value in its real part
          *reinterpret_cast<double *>(p + L.pot) += hp[i];
        }
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
  aran_fmm3d_get_stats(ctx, st);
  for (int k = 0; k < 7; k++) solver->stats[k] += st[k];
  solver->stats[7] += st[12];  // p2l
  solver->stats[8] += st[13];  // m2p
  solver->depth = (guint)st[7];
  solver->solved = true;
}
