/* scip_amd_serve.h -- request serving on top of the propagation engine (scip_amd_prop.h).
 *
 * A server keeps propagation problems (linear systems) resident on one device and answers
 * node requests: "propagate problem p from its root fixpoint with these fixings".  It has no
 * counterpart in SCIP, whose node propagation runs one node at a time inside one solver
 * (SCIPpropagateProbing, src/scip/scip_probing.c); the server is what a pool of solver
 * workers, a strong-branching or diving driver, or a remote propagation service would call
 * instead of one sap_propagate per node.
 *
 * Continuous batching: requests are admitted into the slots of their problem's slot pool
 * (sap_slots_*) as they arrive and as slots free up; every pool with running slots is stepped
 * (a few rounds per step) on one of several HIP streams, so that pools of different problems
 * overlap on the device, and is collected and stepped again as soon as its own step has ended
 * (sap_slots_query), without waiting for the other pools.
 * Admission by HBM footprint: a problem occupies device memory for its matrix, its root
 * vector and its slot pool; when a request needs a problem that is not resident, least
 * recently used problems without running or queued requests are evicted until it fits the
 * budget (a request whose problem cannot fit waits).  A request's result does not depend on
 * what else is in flight.
 */
#ifndef SCIP_AMD_SERVE_H
#define SCIP_AMD_SERVE_H

#include <stdint.h>

#include "scip_amd_prop.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sap_server sap_server;

/* scheduling modes of sap_server_replay */
#define SAP_SERVE_CONTINUOUS 0   /* continuous batching over all problems and streams */
#define SAP_SERVE_PLAIN      1   /* one request at a time, in arrival order, one launch sequence each */

typedef struct sap_serve_stats
{
   double  wall_s;          /* first arrival to last result */
   int64_t requests;
   int64_t nnz_swept;       /* over all steps */
   int64_t rows_swept;
   int64_t steps;           /* slot-pool steps launched */
   int64_t admissions;      /* admission calls (a call admits one or more requests) */
   int64_t uploads;         /* problem uploads (first use or after eviction) */
   int64_t evictions;
   int64_t max_running;     /* most requests in flight at once */
   int64_t peak_bytes;      /* most device memory held by resident problems */
   double  p50_ms;          /* request latency: arrival to result on the host */
   double  p99_ms;
   double  max_ms;
} sap_serve_stats;

/* nstreams HIP streams on device; hbm_budget bytes for resident problems; slots per pool
 * (the default of every problem, see sap_server_set_slots) */
int sap_server_create(int32_t device, int32_t nstreams, int64_t hbm_budget, int32_t slots_per_pool,
   sap_server** srv);
int sap_server_free(sap_server* srv);

/* register a problem (the server copies the arrays); it is uploaded on first use */
int sap_server_add_problem(sap_server* srv, int64_t nrows, int64_t ncols, const int64_t* rowptr,
   const int32_t* colidx, const double* vals, const double* lhs, const double* rhs, const uint8_t* integral,
   const double* lb, const double* ub, int32_t* id);

/* slots of the problem's pool (default: the server's slots_per_pool), e.g. many for a problem
 * that receives most requests; the pool is sized when the problem is uploaded, so a resident
 * problem is released and uploaded again at its next use.  SAP_INVALIDCALL while the problem
 * has requests running or queued. */
int sap_server_set_slots(sap_server* srv, int32_t id, int32_t slots);

/* device bytes the problem occupies while resident (matrix, root vector, slot pool): what
 * admission counts against the budget */
int sap_server_footprint(sap_server* srv, int32_t id, int64_t* bytes);

/* change the HBM budget; resident problems without running or queued requests are evicted,
 * least recently used first, until the resident ones fit (SAP_NOMEMORY if they cannot) */
int sap_server_set_budget(sap_server* srv, int64_t hbm_budget);

/* upload a problem now and compute its root fixpoint (otherwise done at its first request);
 * counts as a use of the problem for the least-recently-used order of eviction */
int sap_server_warm(sap_server* srv, int32_t id);

/* Replay a request trace in real time.  Request i arrives arrival_s[i] seconds after the start
 * (ascending) and asks for problem prob[i] with the fixings ovptr[i]..ovptr[i+1]-1 (column
 * ovcol[q] gets bounds [ovlb[q], ovub[q]]) applied to the problem's root fixpoint.  Per
 * request: result (SAP_CUTOFF / SAP_REDUCEDDOM / SAP_DIDNOTFIND), rounds, number of changed
 * columns and latency; any of these may be NULL.  With chgptr != NULL the changed columns of
 * request i are returned at chgptr[i]..chgptr[i+1]-1 (sorted by column) of chgcol/chglb/chgub
 * when they fit chgcap (else SAP_NOMEMORY after the replay, with chgptr filled).  steprounds is
 * SAP_STEP_SHORT or SAP_STEP_LONG.  An infeasible root makes every request of the problem a
 * cutoff. */
int sap_server_replay(sap_server* srv, const sap_params* params, int32_t mode, int32_t steprounds, int32_t nreq,
   const double* arrival_s, const int32_t* prob, const int64_t* ovptr, const int32_t* ovcol, const double* ovlb,
   const double* ovub, int32_t* result, int32_t* rounds, int32_t* nchg, double* latency_ms, int64_t chgcap,
   int64_t* chgptr, int32_t* chgcol, double* chglb, double* chgub, sap_serve_stats* stats);

/* per request of the last replay (nreq entries): 1 if its problem was uploaded while it
 * waited (a cold request), else 0 */
int sap_server_cold(sap_server* srv, int32_t nreq, uint8_t* cold);

#ifdef __cplusplus
}
#endif

#endif
