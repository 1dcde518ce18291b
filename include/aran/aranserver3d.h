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

/* Estimated device bytes of one request (upper bound used for admission). */
double aran_server3d_footprint (long n, int order, int leafcap);

/* hbm_budget_bytes: admission budget (64 MB per worker is set aside for its
 * tables; the rest bounds the requests in flight); nworkers: concurrent batches (1..8);
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
