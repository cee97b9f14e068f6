/*
 * parallax/px_serve.h -- the served form: a request loop over primitive calls.
 *
 * NO COUNTERPART IN THE REFERENCE (its primitives are synchronous calls only).  For
 * callers that issue many primitive calls of mixed op, element type and size as they
 * arrive: buffers stay resident in a device pool the server owns, requests are admitted
 * against an HBM budget, requests that are waiting when a slot (one batch in flight, with
 * its own scratch) frees up are fused into ONE segmented launch per (op, dtype,
 * predicate) group, and batches are queued on the device while earlier ones run and new
 * requests keep arriving (continuous batching).  Slots share `lanes` streams (default
 * two): batches of one lane run back to back in launch order, so the device never waits
 * for the host between them; scans, compactions and sorts (whole-device look-back
 * kernels, which must not share the CUs with each other) take the first lane and
 * reduces the last.
 * A request's result does not depend on what it was batched with: each request is one
 * segment with the same tiling it would get alone.
 *
 * Requests:
 *   PX_OP_REDUCE  result = sum of the input (one element in the request's output)
 *   PX_OP_SCAN    output = inclusive scan of the input
 *   PX_OP_SORT    the input buffer is sorted in place (output = input)
 *   PX_OP_FLAGS   output = stable copy_if of the input by `flags` (count reported)
 * Inputs must live in the server's pool (pxs_alloc); outputs are allocated by the
 * server at admission and stay valid until pxs_release.  Thread-safe: any thread may
 * submit, wait and release.
 */
#ifndef PARALLAX_PX_SERVE_H
#define PARALLAX_PX_SERVE_H

#include <stddef.h>
#include <stdint.h>

#include "parallax/descriptor.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pxs_server* pxs_server_t;

typedef struct pxs_config {
  int slots;            /* batches in flight at once; default 8 */
  size_t hbm_budget;    /* bytes of the device pool (inputs + outputs); default 64 GiB */
  int max_batch;        /* most requests fused into one launch; default 4096 */
  double max_wait_us;   /* how long a lone request may wait for batch-mates; default 0 */
  int lanes;            /* streams the slots share, 1 or 2 (more count as 2); default 2:
                           reduces and sorts' counting on the second, the rest on the first */
} pxs_config;

typedef struct pxs_request {
  int op;               /* PX_OP_REDUCE, PX_OP_SCAN, PX_OP_SORT or PX_OP_FLAGS */
  int dtype;            /* px_dtype */
  size_t n;             /* elements */
  void* d_in;           /* input inside the server's pool (pxs_alloc) */
  px_desc flags;        /* predicate of PX_OP_FLAGS (dtype must match) */
} pxs_request;

typedef struct pxs_stats {
  uint64_t submitted, completed, batches, rejected;  /* rejected: refused by pxs_submit */
  double p50_ms, p99_ms, mean_ms, max_ms;  /* arrival -> completion latency (successes) */
  double first_arrival_s, last_completion_s;
  double req_per_s;                        /* completed / (last completion - first arrival) */
  size_t pool_bytes, pool_peak_bytes;
  uint64_t failed;                         /* submitted requests that completed with an error */
} pxs_stats;

void pxs_config_default(pxs_config* cfg);
pxs_server_t pxs_create(const pxs_config* cfg);
void pxs_destroy(pxs_server_t s);

/* Device memory in the server's pool (counts against the budget); NULL when full. */
void* pxs_alloc(pxs_server_t s, size_t bytes);
void pxs_free(pxs_server_t s, void* d_ptr);

/* Queue a request; returns its id (>= 0) or a negative PX_E* code: PX_EARG for a bad
 * request (input outside the pool or not aligned to its element size from the pool
 * base; bytes shared between an in-place sort and another request not yet complete),
 * PX_ENOMEM when its output (a sort: its scratch) could never fit in the pool beside the
 * client's inputs.  The arrival time is the time of this call. */
int64_t pxs_submit(pxs_server_t s, const pxs_request* req);
/* Queues n requests that arrive together (one lock, one wake-up of the server): ids[i]
 * receives what pxs_submit would return for reqs[i].  Returns the number refused. */
int pxs_submit_batch(pxs_server_t s, const pxs_request* reqs, size_t n, int64_t* ids);
/* Wait for completion (timeout_ms < 0: no limit).  Returns 0 when the request completed,
 * 1 on time-out, the request's negative PX_E* code when it failed, and PX_ENOMEM when it
 * waits for pool memory that only the client can free (nothing is running and outputs
 * that were not released fill the pool: release some and wait again). */
int pxs_wait(pxs_server_t s, int64_t id, double timeout_ms);
/* Device pointer of the output (reduce: one element; scan/copy_if: n elements; sort: the
 * input), the kept count (copy_if) and the latency in ms.  Returns the request's status:
 * 0, or its negative PX_E* code when it failed (the output is then undefined). */
int pxs_result(pxs_server_t s, int64_t id, void** d_out, uint64_t* count, double* latency_ms);
/* Frees the request's output (and its bookkeeping). */
int pxs_release(pxs_server_t s, int64_t id);
/* Blocks until every submitted request has completed (0), or returns PX_ENOMEM when the
 * remaining ones wait for pool memory that only the client can free. */
int pxs_drain(pxs_server_t s);
int pxs_get_stats(pxs_server_t s, pxs_stats* out);
/* Clears the latency record (keeps buffers and requests). */
int pxs_reset_stats(pxs_server_t s);

#ifdef __cplusplus
}
#endif

#endif /* PARALLAX_PX_SERVE_H */
