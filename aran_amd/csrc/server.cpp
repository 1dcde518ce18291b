// server.cpp -- continuous-batching request loop over the batched engine.
//
// No counterpart in the reference (LibAran solves one tree per call,
// aran/aransolver3d.c:1135-1202).  Requests are copied in at submission and
// queued.  Each worker owns an engine context on its own HIP stream; when a
// worker is free it admits the oldest waiting requests that share the
// oldest one's order and leaf capacity, while the estimated device
// footprint of everything in flight stays under the budget, and solves them
// together as independent systems.  With two workers one batch's copies
// and kernels overlap another's.  A system's result does not depend on the
// batch it ran in: the engine gives every system its own octree, and every
// form sums a target's far field over offset ranges fixed per order, each
// in a fixed order, then adds the ranges in range order.  Served results
// therefore equal standalone solves bit for bit in every form.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

extern "C" {
#include <aran/aranfmm3d.h>
#include <aran/aranserver3d.h>
}

#include "engine.h"

namespace {

using Clock = std::chrono::steady_clock;

// A request's host array (left uninitialised on allocation: every element
// is written by a copy before it is read).
struct HostArr {
  std::unique_ptr<double[]> p;
  void alloc(long k) { p.reset(k > 0 ? new double[k] : nullptr); }
  void release() { p.reset(); }
  double *data() const { return p.get(); }
};

struct Request {
  long id = 0;
  long n = 0;
  int order = 8, leafcap = 32, want_grad = 1;
  double lb[3], ub[3];
  // inputs x, y, z, q and outputs pot, grad: in a block of the server's
  // pinned pool ([x | y | z | q | pot | grad], 8 n doubles) when it has
  // room, else in pageable arrays staged through the worker's buffers
  double *pin = nullptr;
  size_t pin_off = 0, pin_bytes = 0;
  HostArr x, y, z, q, pot, grad;
  double bytes = 0.;
  Clock::time_point t_submit, t_start, t_done;
  bool done = false;
  bool claimed = false;  // a waiter holds this request
  int status = 0;
};

// Host copies of request data run on several threads when they are large
// (one thread moves well under the host's memory bandwidth, and the copies
// in and out of the pinned staging otherwise bound the served rate).
static void par_copy(void *dst, const void *src, size_t bytes) {
  const size_t kChunk = size_t(4) << 20;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>(std::min(8u, hw), bytes / kChunk);
  if (nt < 2) {
    if (bytes) memcpy(dst, src, bytes);
    return;
  }
  const size_t part = (bytes / nt + 63) / 64 * 64;
  std::vector<std::thread> th;
  for (size_t t = 1; t < nt; t++) {
    const size_t b = t * part;
    if (b >= bytes) break;
    const size_t e = std::min(bytes, b + part);
    th.emplace_back([=] {
      memcpy(static_cast<char *>(dst) + b, static_cast<const char *>(src) + b,
             e - b);
    });
  }
  memcpy(dst, src, std::min(bytes, part));
  for (auto &t : th) t.join();
}

struct PinnedBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  bool reserve(size_t n) {
    if (n <= bytes && ptr) return true;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
    size_t want = std::max<size_t>(n + n / 4, 4096);
    if (hipHostMalloc(&ptr, want) != hipSuccess) return false;
    bytes = want;
    return true;
  }
  ~PinnedBuf() {
    if (ptr) (void)hipHostFree(ptr);
  }
};

// First-fit allocator over one pinned host region (offsets and sizes in
// 256-byte units), under the server's lock.  Request data copied once into
// it at submission goes to the GPU and comes back without staging copies.
struct PinnedPool {
  char *base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes, coalesced
  bool init(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (hipHostMalloc(reinterpret_cast<void **>(&base), bytes) != hipSuccess) {
      base = nullptr;
      return false;
    }
    size = bytes;
    free_.clear();
    free_[0] = bytes;
    return true;
  }
  char *alloc(size_t bytes, size_t *off) {
    bytes = (bytes + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes, at = it->first + bytes;
        free_.erase(it);
        if (rest) free_[at] = rest;
        return base + *off;
      }
    return nullptr;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  ~PinnedPool() {
    if (base) (void)hipHostFree(base);
  }
};

struct Worker {
  AranFmm3d *ctx = nullptr;
  aran::DevBuf din, dout;
  PinnedBuf hin, hout;
  std::thread th;
};

}  // namespace

struct _AranServer3d {
  double budget = 0., inflight = 0., peak_inflight = 0.;
  int precision = ARAN_FMM3D_MIXED;
  int max_batch = 64;
  int device = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> queue;
  std::map<long, std::unique_ptr<Request>> all;
  long next_id = 0;
  bool stop = false, paused = false, preparing = false;
  long batches = 0, served = 0, max_batch_seen = 0;
  int running = 0;  // batches taken and not yet finished
  std::vector<std::unique_ptr<Worker>> workers;
  PinnedPool pool;  // request blocks (set up by aran_server3d_prepare)
};

extern "C" double aran_server3d_footprint(long n, int order, int leafcap) {
  if (n < 0 || order < 1 || leafcap < 1) return -1.;
  const double mc = order * (order + 1) / 2., lc = (order + 1) * (order + 2) / 2.;
  const double nodes = 64. + 4. * ((double)n / leafcap + 1.);
  // particle arrays (inputs, outputs, sort keys, sorted copies) and, per
  // node, tree arrays, colleagues, far list, tile table, FP64 and FP32
  // expansions (8 partial FP32 locals)
  const double per_particle = 200.;
  const double per_node = 2600. + 24. * (mc + lc) + 64. * lc;
  return 1.25 * (n * per_particle + nodes * per_node) + 1e6;
}

// Tables (of up to four orders) and fixed buffers of one worker context,
// outside the requests.
static const double kWorkerBytes = ARAN_SERVER3D_WORKER_BYTES;

// Host and device staging of a batch of n particles (inputs x, y, z, q;
// outputs potential and gradient).
static bool reserve_io(Worker *W, long n) {
  const size_t nb = sizeof(double) * std::max<long>(n, 1);
  return W->hin.reserve(4 * nb) && W->hout.reserve(4 * nb) &&
         W->din.reserve(4 * nb) == hipSuccess &&
         W->dout.reserve(4 * nb) == hipSuccess;
}

static void worker_loop(AranServer3d *S, Worker *W) {
  (void)hipSetDevice(S->device);
  for (;;) {
    std::vector<Request *> batch;
    double bytes = 0.;
    {
      std::unique_lock<std::mutex> lock(S->mu);
      for (;;) {
        if (S->stop && S->queue.empty()) return;
        if (!S->queue.empty() && !S->preparing && (!S->paused || S->stop)) {
          Request *r0 = S->queue.front();
          const double cap = S->budget / (double)S->workers.size();
          if (S->inflight + r0->bytes <= S->budget || S->inflight == 0.) {
            batch.push_back(r0);
            bytes = r0->bytes;
            for (auto it = S->queue.begin() + 1; it != S->queue.end(); ++it) {
              Request *r = *it;
              if ((int)batch.size() >= S->max_batch) break;
              if (r->order != r0->order || r->leafcap != r0->leafcap) continue;
              if (bytes + r->bytes > cap) continue;
              if (S->inflight + bytes + r->bytes > S->budget) continue;
              batch.push_back(r);
              bytes += r->bytes;
            }
            for (Request *r : batch)
              S->queue.erase(std::find(S->queue.begin(), S->queue.end(), r));
            S->inflight += bytes;
            S->running++;
            S->peak_inflight = std::max(S->peak_inflight, S->inflight);
            break;
          }
        }
        S->cv_work.wait(lock);
      }
    }
    const Clock::time_point t0 = Clock::now();
    const int nsys = (int)batch.size();
    std::vector<long> off(nsys + 1, 0);
    std::vector<double> lbs(3 * nsys), ubs(3 * nsys);
    for (int k = 0; k < nsys; k++) {
      off[k + 1] = off[k] + batch[k]->n;
      for (int a = 0; a < 3; a++) {
        lbs[3 * k + a] = batch[k]->lb[a];
        ubs[3 * k + a] = batch[k]->ub[a];
      }
    }
    const long n = off[nsys];
    int rc = 0;
    hipStream_t st = static_cast<hipStream_t>(aran_fmm3d_stream(W->ctx));
    if (!reserve_io(W, n)) rc = ARAN_FMM3D_ENOMEM;
    double *hin = static_cast<double *>(W->hin.ptr);
    double *din = W->din.as<double>();
    double *hout = static_cast<double *>(W->hout.ptr);
    double *dout = W->dout.as<double>();
    if (rc == 0 && n > 0) {
      // pinned requests go to the device from their own block; the others
      // through the worker's pinned staging
      bool staged = false;
      for (int k = 0; k < nsys && rc == 0; k++) {
        const Request *r = batch[k];
        const size_t nb = sizeof(double) * r->n;
        for (int a = 0; a < 4 && rc == 0; a++) {
          double *d = din + a * n + off[k];
          if (r->pin) {
            if (r->n > 0 &&
                hipMemcpyAsync(d, r->pin + a * r->n, nb, hipMemcpyHostToDevice,
                               st) != hipSuccess)
              rc = ARAN_FMM3D_EHIP;
          } else {
            const HostArr *src[4] = {&r->x, &r->y, &r->z, &r->q};
            par_copy(hin + a * n + off[k], src[a]->data(), nb);
            staged = true;
          }
        }
      }
      if (rc == 0 && staged)
        for (int k = 0; k < nsys && rc == 0; k++) {
          const Request *r = batch[k];
          if (r->pin || r->n == 0) continue;
          for (int a = 0; a < 4 && rc == 0; a++)
            if (hipMemcpyAsync(din + a * n + off[k], hin + a * n + off[k],
                               sizeof(double) * r->n, hipMemcpyHostToDevice,
                               st) != hipSuccess)
              rc = ARAN_FMM3D_EHIP;
        }
    }
    if (rc == 0)
      rc = aran_fmm3d_solve_batch(W->ctx, nsys, off.data(), din, din + n,
                                  din + 2 * n, din + 3 * n, lbs.data(),
                                  ubs.data(), batch[0]->order,
                                  batch[0]->leafcap, dout, dout + n);
    if (rc == 0 && n > 0) {
      // results of pinned requests straight into their blocks, the rest
      // into the worker's staging
      for (int k = 0; k < nsys && rc == 0; k++) {
        const Request *r = batch[k];
        if (r->n == 0) continue;
        double *hp = r->pin ? r->pin + 4 * r->n : hout + off[k];
        double *hg = r->pin ? r->pin + 5 * r->n : hout + n + 3 * off[k];
        if (hipMemcpyAsync(hp, dout + off[k], sizeof(double) * r->n,
                           hipMemcpyDeviceToHost, st) != hipSuccess ||
            (r->want_grad &&
             hipMemcpyAsync(hg, dout + n + 3 * off[k],
                            sizeof(double) * 3 * r->n, hipMemcpyDeviceToHost,
                            st) != hipSuccess))
          rc = ARAN_FMM3D_EHIP;
      }
      if (rc == 0 && hipStreamSynchronize(st) != hipSuccess)
        rc = ARAN_FMM3D_EHIP;
    }
    {
      std::lock_guard<std::mutex> lock(S->mu);
      const Clock::time_point t1 = Clock::now();
      for (int k = 0; k < nsys; k++) {
        Request *r = batch[k];
        r->status = rc;
        if (rc == 0 && !r->pin) {
          r->pot.alloc(r->n);
          par_copy(r->pot.data(), hout + off[k], sizeof(double) * r->n);
          if (r->want_grad) {
            r->grad.alloc(3 * r->n);
            par_copy(r->grad.data(), hout + n + 3 * off[k],
                     sizeof(double) * 3 * r->n);
          }
        }
        r->x.release();
        r->y.release();
        r->z.release();
        r->q.release();
        r->t_start = t0;
        r->t_done = t1;
        r->done = true;
      }
      S->inflight -= bytes;
      S->running--;
      S->batches++;
      S->served += nsys;
      S->max_batch_seen = std::max<long>(S->max_batch_seen, nsys);
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();
  }
}

extern "C" AranServer3d *aran_server3d_new(double hbm_budget_bytes,
                                           int nworkers, int precision,
                                           int max_batch) {
  if (!(hbm_budget_bytes > 0.) || nworkers < 1 || nworkers > 8 ||
      max_batch < 1 ||
      precision < ARAN_FMM3D_PLAIN || precision > ARAN_FMM3D_BF16X6) {
    aran::set_error("aran_server3d_new: invalid argument");
    return nullptr;
  }
  if (hbm_budget_bytes <= kWorkerBytes * nworkers) {
    aran::set_error("aran_server3d_new: budget below the workers' tables");
    return nullptr;
  }
  AranServer3d *S = new AranServer3d;
  S->budget = hbm_budget_bytes - kWorkerBytes * nworkers;
  S->precision = precision;
  S->max_batch = max_batch;
  (void)hipGetDevice(&S->device);
  for (int k = 0; k < nworkers; k++) {
    std::unique_ptr<Worker> w(new Worker);
    w->ctx = aran_fmm3d_new(nullptr);
    if (!w->ctx) {
      for (auto &o : S->workers) aran_fmm3d_free(o->ctx);
      delete S;
      return nullptr;
    }
    aran_fmm3d_set_precision(w->ctx, precision);
    S->workers.push_back(std::move(w));
  }
  for (auto &w : S->workers) w->th = std::thread(worker_loop, S, w.get());
  return S;
}

extern "C" void aran_server3d_free(AranServer3d *S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lock(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  for (auto &w : S->workers) {
    if (w->th.joinable()) w->th.join();
    aran_fmm3d_free(w->ctx);
    w->din.release();
    w->dout.release();
  }
  delete S;
}

extern "C" long aran_server3d_submit(AranServer3d *S, long n, const double *x,
                                     const double *y, const double *z,
                                     const double *q, const double *lb,
                                     const double *ub, int order, int leafcap,
                                     int want_grad) {
  if (!S || n < 0 || (n > 0 && (!x || !y || !z || !q)) || !lb || !ub ||
      order < 1 || order > aran::kMaxOrderFast || leafcap < 1) {
    aran::set_error("aran_server3d_submit: invalid argument");
    return ARAN_FMM3D_EINVAL;
  }
  const double bytes = aran_server3d_footprint(n, order, leafcap);
  if (bytes > S->budget / (double)S->workers.size()) {
    aran::set_error("aran_server3d_submit: request exceeds the HBM budget");
    return ARAN_SERVER3D_EBUDGET;
  }
  std::unique_ptr<Request> r(new Request);
  r->n = n;
  r->order = order;
  r->leafcap = leafcap;
  r->want_grad = want_grad != 0;
  for (int a = 0; a < 3; a++) {
    r->lb[a] = lb[a];
    r->ub[a] = ub[a];
  }
  const size_t nb = sizeof(double) * n;
  if (n > 0 && S->pool.base) {
    std::lock_guard<std::mutex> lock(S->mu);
    r->pin_bytes = 8 * nb;
    r->pin = reinterpret_cast<double *>(S->pool.alloc(r->pin_bytes,
                                                      &r->pin_off));
  }
  if (r->pin) {
    par_copy(r->pin, x, nb);
    par_copy(r->pin + n, y, nb);
    par_copy(r->pin + 2 * n, z, nb);
    par_copy(r->pin + 3 * n, q, nb);
  } else {
    r->x.alloc(n);
    r->y.alloc(n);
    r->z.alloc(n);
    r->q.alloc(n);
    par_copy(r->x.data(), x, nb);
    par_copy(r->y.data(), y, nb);
    par_copy(r->z.data(), z, nb);
    par_copy(r->q.data(), q, nb);
  }
  r->bytes = bytes;
  r->t_submit = Clock::now();
  long id;
  {
    std::lock_guard<std::mutex> lock(S->mu);
    if (S->stop) {
      if (r->pin) S->pool.release(r->pin_off, r->pin_bytes);
      aran::set_error("aran_server3d_submit: server is stopping");
      return ARAN_FMM3D_EINVAL;
    }
    id = S->next_id++;
    r->id = id;
    S->queue.push_back(r.get());
    S->all[id] = std::move(r);
  }
  S->cv_work.notify_one();
  return id;
}

extern "C" int aran_server3d_wait(AranServer3d *S, long id, double *pot,
                                  double *grad, double *times3) {
  if (!S) return ARAN_FMM3D_EINVAL;
  std::unique_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lock(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) {
      aran::set_error("aran_server3d_wait: unknown request");
      return ARAN_FMM3D_EINVAL;
    }
    Request *rp = it->second.get();
    if (rp->claimed) {
      aran::set_error("aran_server3d_wait: request already being waited for");
      return ARAN_FMM3D_EINVAL;
    }
    // claimed under the lock: a second waiter is refused above, so the
    // entry stays in the map until this waiter removes it
    rp->claimed = true;
    S->cv_done.wait(lock, [&] { return rp->done; });
    auto again = S->all.find(id);
    r = std::move(again->second);
    S->all.erase(again);
  }
  const int status = r->status;
  if (status == 0) {
    const double *rp = r->pin ? r->pin + 4 * r->n : r->pot.data();
    const double *rg = r->pin ? r->pin + 5 * r->n : r->grad.data();
    if (pot) par_copy(pot, rp, sizeof(double) * r->n);
    if (grad && r->want_grad) par_copy(grad, rg, sizeof(double) * 3 * r->n);
  }
  if (r->pin) {
    std::lock_guard<std::mutex> lock(S->mu);
    S->pool.release(r->pin_off, r->pin_bytes);
    r->pin = nullptr;
  }
  if (status != 0) return status;
  if (times3) {
    using sec = std::chrono::duration<double>;
    times3[0] = sec(r->t_done - r->t_submit).count();       // latency
    times3[1] = sec(r->t_start - r->t_submit).count();      // queueing
    times3[2] = sec(r->t_done - r->t_start).count();        // batch service
  }
  return 0;
}

extern "C" int aran_server3d_prepare(AranServer3d *S, long n_max,
                                     const int *orders, int norders,
                                     int leafcap) {
  if (!S || n_max < 1 || norders < 0 || (norders > 0 && !orders) ||
      leafcap < 1) {
    aran::set_error("aran_server3d_prepare: invalid argument");
    return ARAN_FMM3D_EINVAL;
  }
  for (int k = 0; k < norders; k++)
    if (orders[k] < 1 || orders[k] > aran::kMaxOrderFast) {
      aran::set_error("aran_server3d_prepare: invalid order");
      return ARAN_FMM3D_EINVAL;
    }
  {
    std::lock_guard<std::mutex> lock(S->mu);
    if (S->preparing || !S->queue.empty() || S->running != 0) {
      aran::set_error("aran_server3d_prepare: requests queued or in flight");
      return ARAN_FMM3D_EINVAL;
    }
    S->preparing = true;  // workers take no batch until it is cleared
  }
  (void)hipSetDevice(S->device);
  int rc = 0;
  // the pinned pool for request data: room for 16 requests of n_max
  // particles (8 doubles each: x, y, z, q in, potential and gradient out)
  if (!S->pool.base && !S->pool.init(16 * 64 * (size_t)n_max)) {
    aran::set_error("aran_server3d_prepare: pinned pool allocation failed");
    rc = ARAN_FMM3D_ENOMEM;
  }
  for (auto &w : S->workers) {
    if (rc != 0) break;
    Worker *W = w.get();
    if (!reserve_io(W, n_max)) {
      aran::set_error("aran_server3d_prepare: staging allocation failed");
      rc = ARAN_FMM3D_ENOMEM;
      break;
    }
    // a uniform system of n_max particles (fixed LCG) sizes the context's
    // workspace; one solve per order builds that order's tables
    double *h = static_cast<double *>(W->hin.ptr);
    unsigned long long st = 0x9E3779B97F4A7C15ull;
    for (long i = 0; i < 3 * n_max; i++) {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      h[i] = (double)(st >> 11) * (1.0 / 9007199254740992.0);
    }
    for (long i = 0; i < n_max; i++) h[3 * n_max + i] = 1.0 / (double)n_max;
    hipStream_t hs = static_cast<hipStream_t>(aran_fmm3d_stream(W->ctx));
    double *din = W->din.as<double>(), *dout = W->dout.as<double>();
    if (hipMemcpyAsync(din, h, 4 * sizeof(double) * n_max,
                       hipMemcpyHostToDevice, hs) != hipSuccess) {
      rc = ARAN_FMM3D_EHIP;
      break;
    }
    const long off[2] = {0, n_max};
    const double lb[3] = {0., 0., 0.}, ub[3] = {1., 1., 1.};
    for (int k = 0; k < norders && rc == 0; k++)
      rc = aran_fmm3d_solve_batch(W->ctx, 1, off, din, din + n_max,
                                  din + 2 * n_max, din + 3 * n_max, lb, ub,
                                  orders[k], leafcap, dout, dout + n_max);
    if (rc == 0 && hipStreamSynchronize(hs) != hipSuccess) rc = ARAN_FMM3D_EHIP;
    if (rc != 0) break;
  }
  {
    std::lock_guard<std::mutex> lock(S->mu);
    S->preparing = false;
  }
  S->cv_work.notify_all();
  return rc;
}

extern "C" int aran_server3d_pause(AranServer3d *S, int paused) {
  if (!S) return ARAN_FMM3D_EINVAL;
  {
    std::lock_guard<std::mutex> lock(S->mu);
    S->paused = paused != 0;
  }
  S->cv_work.notify_all();
  return 0;
}

extern "C" int aran_server3d_stats(AranServer3d *S, double *out4) {
  if (!S || !out4) return ARAN_FMM3D_EINVAL;
  std::lock_guard<std::mutex> lock(S->mu);
  out4[0] = (double)S->batches;
  out4[1] = (double)S->served;
  out4[2] = (double)S->max_batch_seen;
  out4[3] = S->peak_inflight;
  return 0;
}
