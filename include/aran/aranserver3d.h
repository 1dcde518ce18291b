/*
 * aranserver3d.h -- serving FMM solves on one GPU with continuous batching.
 *
 * No counterpart in the reference: LibAran solves one particle system per
 * aran_solver3d_solve() call (aran/aransolver3d.c:1135-1202).  Requests
 * (host arrays, copied at submission) queue in arrival order; each worker
 * (an engine context on its own HIP stream) admits the oldest waiting
 * requests that share the oldest one's order and leaf capacity while the
 * estimated device footprint in flight stays under the budget, and solves
 * them together as independent systems (aran_fmm3d_solve_batch).  In the
 * plain form a served result is bitwise the result of solving that request
 * alone.  In the other forms each local expansion is accumulated in a number
 * of partial sums (the far-field split) chosen from the size of the whole
 * batch: a served result is bitwise the standalone one when both choose the
 * same split (as they do while the batch's far-field work fits one round of
 * workgroups on the GPU), and otherwise equal to within that form's rounding.
 */
#ifndef ARAN_SERVER3D_H
#define ARAN_SERVER3D_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _AranServer3d AranServer3d;

#define ARAN_SERVER3D_EBUDGET (-4)
/* Device bytes set aside per worker for its engine context: operator tables
 * of up to four orders (the BF16 MFMA operators alone take about 57 MB at
 * order 12 and 126 MB at order 15) and fixed buffers. */
#define ARAN_SERVER3D_WORKER_BYTES 512e6

/* Estimated device bytes of one request (upper bound used for admission). */
double aran_server3d_footprint (long n, int order, int leafcap);

/* hbm_budget_bytes: admission budget (ARAN_SERVER3D_WORKER_BYTES per worker
 * are set aside for its context; the rest bounds the requests in flight); nworkers: concurrent batches (1..8);
 * precision: ARAN_FMM3D_PLAIN, _MIXED, _BF16X3 or _BF16X6; max_batch: requests per
 * batch.  Binds the current HIP device. */
AranServer3d *aran_server3d_new (double hbm_budget_bytes, int nworkers,
                                 int precision, int max_batch);
/* Serves what is queued, then stops the workers. */
void aran_server3d_free (AranServer3d *srv);

/* Queues one solve; returns a request id >= 0, or ARAN_SERVER3D_EBUDGET when
 * the request alone exceeds one worker's share of the budget, or another
 * negative code. */
long aran_server3d_submit (AranServer3d *srv, long n, const double *x,
                           const double *y, const double *z, const double *q,
                           const double *lb, const double *ub, int order,
                           int leafcap, int want_grad);

/* Blocks until request id is solved; copies the potential (n) and, when it
 * was asked for, the gradient (n x 3) to host arrays; times3 (may be NULL)
 * receives latency (submission to results on the host), queueing and batch
 * service time in seconds.  Each id is waited for once. */
int aran_server3d_wait (AranServer3d *srv, long id, double *pot, double *grad,
                        double *times3);

/* Readies the server for traffic before the first request: every worker
 * reserves host (pinned) and device staging for batches of up to n_max
 * particles, sizes its solver workspace with a synthetic uniform system of
 * n_max particles, and builds its tables for each of the norders orders
 * (one solve per order at leaf capacity leafcap).  It also sets up a pool
 * of pinned host memory for request data (room for 16 requests of n_max
 * particles): a request that fits is copied once into it at submission,
 * goes to the GPU and comes back from it without staging copies, and its
 * results are copied once into the caller's buffers at wait.  Without it
 * the first requests of each size class and order pay for those
 * allocations and table builds, and every request is staged.  Returns
 * ARAN_FMM3D_EINVAL when requests are queued or in flight; blocks until
 * done. */
int aran_server3d_prepare (AranServer3d *srv, long n_max, const int *orders,
                           int norders, int leafcap);

/* paused != 0: workers finish their batch and take no new one until
 * resumed (queued requests wait and can then be batched together). */
int aran_server3d_pause (AranServer3d *srv, int paused);

/* batches run, requests served, largest batch, peak estimated bytes in
 * flight. */
int aran_server3d_stats (AranServer3d *srv, double *out4);

#ifdef __cplusplus
}
#endif

#endif /* ARAN_SERVER3D_H */
