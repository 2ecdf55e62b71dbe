/* scip_amd_prop.h -- C ABI of the MI355X linear-constraint propagation engine.
 *
 * The engine replaces the activity and bound-tightening path of SCIP's linear constraint
 * handler (src/scip/cons_linear.c: consdataCalcActivities :2283, getMinActivity/
 * getMaxActivity :2345/:2440, consdataGetActivityResiduals :2661, tightenVarBoundsEasy
 * :5380, tightenVarBounds :6700, tightenBounds :6980, propagateCons :7621, consPropLinear
 * :16126) and the round loop of src/scip/solve.c propagateDomains (:723-795) with a
 * device-resident, round-synchronous fixpoint.  It is called from the SCIP propagator
 * plugin scip_amd/plugin/prop_gpulinear.c through SCIPincludePropBasic's PROPEXEC and
 * SCIPsetPropPresol's PROPPRESOL callbacks (src/scip/scip_prop.c:118, :283).
 *
 * Conventions
 *   - Return codes have the values of SCIP_RETCODE (src/scip/type_retcode.h:40-60):
 *     SAP_OKAY = 1, errors <= 0.  No function aborts or throws across this ABI.
 *   - All array arguments are plain host pointers owned by the caller unless the
 *     function name says "device".  Nothing here allocates memory the caller must free,
 *     except the opaque handles, which have matching *_free functions.
 *   - One sap_context per SCIP instance (SCIP* is single-threaded); a context owns one
 *     HIP stream.  Calls on different contexts may run concurrently.
 *   - Arithmetic is IEEE FP64 throughout.
 */
#ifndef SCIP_AMD_PROP_H
#define SCIP_AMD_PROP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes: same values as SCIP_RETCODE (type_retcode.h:42-60) */
#define SAP_OKAY            1
#define SAP_ERROR           0
#define SAP_NOMEMORY       -1
#define SAP_INVALIDCALL    -8
#define SAP_INVALIDDATA    -9
#define SAP_NOTIMPLEMENTED -18

/* propagation verdicts: same values as SCIP_RESULT (type_result.h:42-59) */
#define SAP_DIDNOTFIND      3
#define SAP_CUTOFF          7
#define SAP_REDUCEDDOM     10

/* numerics and propagation parameters.  Defaults (sap_params_default) are SCIP's:
 * numerics/{infinity,epsilon,sumepsilon,feastol,boundstreps,hugeval,recomputefac}
 * (def.h:172-184), constraints/linear/maxeasyactivitydelta (cons_linear.c), and
 * propagating/maxrounds = -1 (unlimited). */
typedef struct sap_params
{
   double  infinity;
   double  epsilon;
   double  sumepsilon;
   double  feastol;
   double  boundstreps;
   double  hugeval;
   double  recompfac;
   double  maxeasyactivitydelta;
   int32_t maxrounds;             /* -1: until fixpoint */
   int32_t flags;                 /* SAP_FLAG_* */
} sap_params;

/* the input bound vector(s) are already a propagation fixpoint except for the columns an
 * overlay changes: only rows containing such columns are swept in the first round */
#define SAP_FLAG_BASE_IS_FIXPOINT 0x1
/* run the round loop from the host with one synchronisation per round instead of launching
 * hipGraph-captured groups of rounds (same results; for debugging and comparison) */
#define SAP_FLAG_HOSTLOOP 0x2
/* also run ranged-row propagation in every round (rangedRowPropagation, cons_linear.c:5715-6697;
 * SCIP's constraints/linear/rangedrowpropagation, on by default outside presolving): ranged
 * rows and equations of at least 3 nonzeros with both sides below 1e20 in magnitude get the
 * gcd-based fixings, bounds and infeasibility test, as propagation applies it after the root
 * node (no artificial constraints). */
#define SAP_FLAG_RANGEDROWS 0x4
/* consPropLinear with tightenbounds FALSE (cons_linear.c:16147-16161, the
 * constraints/linear/tightenboundsfreq gating): every row is only tested for infeasibility and
 * local redundancy (propagateCons, :7717-7753); no bound is tightened, and ranged-row
 * propagation, which runs only together with the tightening (:7694), is off */
#define SAP_FLAG_NOTIGHTEN 0x8

/* per-call statistics */
typedef struct sap_stats
{
   int32_t result;       /* SAP_CUTOFF, SAP_REDUCEDDOM or SAP_DIDNOTFIND (single vector) */
   int32_t rounds;       /* propagation rounds executed (max over vectors for batches) */
   int64_t nchgbds;      /* accepted bound changes */
   int64_t rows_swept;   /* rows swept over all rounds (and vectors) */
   int64_t nnz_swept;    /* nonzeros swept over all rounds (and vectors) */
   double  device_ms;    /* device time of the propagation kernels (HIP events) */
   double  sweep_ms;     /* device time of the row-sweep kernels alone */
   int64_t sweep_launches;
   int64_t nrecompute;   /* residual activities recomputed in plain double because the
                            incremental value was unreliable (the path of
                            consdataGetReliableResidualActivity, cons_linear.c:2580-2657) */
} sap_stats;

typedef struct sap_context sap_context;
typedef struct sap_matrix  sap_matrix;
typedef struct sap_bounds  sap_bounds;

void sap_params_default(sap_params* params);
const char* sap_error_string(int rc);

/* number of visible HIP devices (0 if none); never fails */
int sap_device_count(int32_t* ndevices);

/* context = device + stream; one per SCIP instance (created in PROPINITSOL/INITPRE) */
int sap_context_create(int32_t device, sap_context** ctx);
int sap_context_free(sap_context* ctx);

/* Row binning and launch shape of the matrices a context creates.  Rows are grouped in bins
 * by length (at most 4, 8, 16, 32, 64, 128 nonzeros for bins 1..6, then longer).  The rows of
 * bins 1..rowbins are swept one row per lane; in the work modes that sweep every row (dense)
 * the rows of bins rowbins+1..tilebins are packed into tiles of 64 lanes holding tilewidth
 * nonzeros each (bin 6 only with tilewidth 2 or 4); groups of lanes take the other short
 * bins, one wavefront takes a row of up to longrow nonzeros and a whole workgroup a longer
 * one; capscale (percent) scales the workgroup count of each kind of sweep work.
 * sap_autotune measures candidates on a row sample of a matrix.  No counterpart in SCIP
 * (whose propagation has no launch shape). */
typedef struct sap_tuning
{
   int32_t tilebins;    /* 0..6, default 6 (6 requires tilewidth 2 or 4) */
   int32_t longrow;     /* 64..2^24, default 1024 */
   int32_t capscale;    /* 10..1000, default 50 */
   int32_t rowbins;     /* 0..3, default 0 */
   int32_t tilewidth;   /* nonzeros per lane in a tile: 1, 2 or 4, default 4 */
} sap_tuning;
void sap_tuning_default(sap_tuning* tuning);
int sap_context_set_tuning(sap_context* ctx, const sap_tuning* tuning);
int sap_context_get_tuning(const sap_context* ctx, sap_tuning* tuning);

/* Upload a linear system lhs <= A x <= rhs in CSR form (rows = linear constraints of
 * SCIPgetVarsLinear/SCIPgetValsLinear/SCIPgetLhsLinear/SCIPgetRhsLinear,
 * cons_linear.c:18326-18436; columns = SCIPvarGetProbindex).  Column indices must be
 * distinct within a row and coefficients nonzero.  integral[j] = SCIPvarIsIntegral;
 * glb/gub are the global bounds, which decide SCIPvarIsBinary (var.c:23801) and the
 * binary maxactdelta rule.  The matrix stays resident on the device until freed. */
int sap_matrix_create(sap_context* ctx, int64_t nrows, int64_t ncols, const int64_t* rowptr,
   const int32_t* colidx, const double* vals, const double* lhs, const double* rhs,
   const uint8_t* integral, const double* glb, const double* gub, sap_matrix** mat);
int sap_matrix_free(sap_matrix* mat);
/* Hand a matrix, with its bound sets, over to another context of the same device: its later
 * work runs on that context's stream.  For a loader thread that builds matrices on a context
 * of its own while another thread propagates on others (the request server).  No work of the
 * matrix may be in flight (a launched slots step must have been waited for).  No counterpart
 * in SCIP. */
int sap_matrix_set_context(sap_matrix* mat, sap_context* ctx);
/* Autotuner: builds the system restricted to a sample of samplerows rows (all rows if fewer;
 * chosen by seed) once per candidate tuning ((rowbins, tilebins) in (0,3) (0,4) (0,5) (0,6)
 * (2,6) (3,6) x tilewidth 2/4 x longrow 1024/4096/16384 x capscale 25/50/100), propagates lb/ub
 * with params on it (rounds driven from the host) and returns the candidate with the least
 * row-sweep time in best (times in ms: best_ms, default_ms for the default tuning).
 * Results are identical under every tuning; only the launch shape differs.  The context's
 * own tuning is left unchanged: apply the result with sap_context_set_tuning. */
int sap_autotune(sap_context* ctx, int64_t nrows, int64_t ncols, const int64_t* rowptr, const int32_t* colidx,
   const double* vals, const double* lhs, const double* rhs, const uint8_t* integral, const double* lb,
   const double* ub, const sap_params* params, int64_t samplerows, uint64_t seed, sap_tuning* best,
   double* best_ms, double* default_ms);
int sap_matrix_info(const sap_matrix* mat, int64_t* nrows, int64_t* ncols, int64_t* nnz,
   int64_t* device_bytes);

/* Propagate one bound vector to fixpoint (PROPEXEC path): lb/ub are read and overwritten
 * with the propagated local bounds.  stats->result is SAP_CUTOFF when infeasibility was
 * detected (bounds are then unspecified), else SAP_REDUCEDDOM or SAP_DIDNOTFIND. */
int sap_propagate(sap_matrix* mat, const sap_params* params, double* lb, double* ub, sap_stats* stats);

/* Propagation with reasons, for conflict analysis (PROPRESPROP, prop.c:739-788; the role of
 * cons_linear.c's inferinfo, :344-369, resolvePropagation :5097-5189 and analyzeConflict).
 * Same propagation and results as sap_propagate; in addition every accepted bound change is
 * returned as a history entry, sorted by round, column and bound (lower first):
 *   hround   the round (1, 2, ...) whose column step accepted it; the row sweep of that round
 *            read the bounds after all changes of the earlier rounds;
 *   hcol, hbound (0 lower, 1 upper), hval  the column and its new bound;
 *   hrow     the reason: the smallest row that gave the winning candidate, from the bounds
 *            at the start of that round;
 *   hrule    (may be NULL) how: 0, the row's residual activity -- for a lower bound of a
 *            column with coefficient a > 0 the row's lhs against its maximal residual
 *            activity, with a < 0 the rhs against the minimal one, symmetric for upper bounds,
 *            so the row's bounds of the other columns explain it (cons_linear.c
 *            addConflictBounds); 1, the row's ranged-row gcd test (SAP_FLAG_RANGEDROWS), which
 *            reads both bounds of all the row's columns, the changed one included (its bounds
 *            before the change; cons_linear.c addConflictFixedVars).
 *            Applied in history order, each change is implied by its row and the bounds
 *            already applied.
 * *nhist is always set; if it exceeds histcap the history is not written and SAP_NOMEMORY is
 * returned (bounds and conflict are).  On a cutoff, conflict describes the infeasibility found
 * in the last round (the changes of the earlier rounds are in the history):
 *   SAP_CONFLICT_ROW     the smallest row that proves infeasibility: side 0, its minimal
 *                        activity (lower bounds where a > 0, upper where a < 0) exceeds rhs;
 *                        side 1, its maximal activity is below lhs; side 2, its ranged-row gcd
 *                        test fails (all bounds of its columns);
 *   SAP_CONFLICT_COLUMN  the smallest column whose candidate bounds crossed: rowlb implied its
 *                        lower bound (-1: the column's own lower bound), rowub its upper one
 *                        (-1: its own upper bound), by the rules in bits 0 (rowlb) and 1 (rowub)
 *                        of rules, as for history entries.
 * One instance (not a packed matrix); rows < 2^28, columns < 2^30. */
#define SAP_CONFLICT_NONE   0
#define SAP_CONFLICT_ROW    1
#define SAP_CONFLICT_COLUMN 2
typedef struct sap_conflict
{
   int32_t kind;
   int32_t round;
   int32_t row;
   int32_t side;
   int32_t col;
   int32_t rowlb;
   int32_t rowub;
   int32_t rules;
} sap_conflict;
int sap_propagate_reasons(sap_matrix* mat, const sap_params* params, double* lb, double* ub, int64_t histcap,
   int64_t* nhist, int32_t* hround, int32_t* hcol, int32_t* hbound, double* hval, int32_t* hrow, int32_t* hrule,
   sap_conflict* conflict, sap_stats* stats);

/* Device-resident bound vector(s), for callers that keep bounds in HBM across calls
 * (1 <= nvec < 2^30). */
int sap_bounds_create(sap_matrix* mat, int32_t nvec, sap_bounds** bnd);
/* as sap_bounds_create with an explicit work-tracking mode: SAP_MODE_DENSE scans round-tagged
 * row/column marks (no list atomics; best for one or few vectors), SAP_MODE_SPARSE keeps
 * per-round worklists (best for many sparse node vectors); SAP_MODE_AUTO chooses by size.
 * Results are identical in every mode. */
#define SAP_MODE_AUTO   0
#define SAP_MODE_DENSE  1
#define SAP_MODE_SPARSE 2
int sap_bounds_create_mode(sap_matrix* mat, int32_t nvec, int32_t mode, sap_bounds** bnd);
int sap_bounds_free(sap_bounds* bnd);
/* set vector v (or all vectors if v < 0) from host arrays */
int sap_bounds_upload(sap_bounds* bnd, int32_t v, const double* lb, const double* ub);
int sap_bounds_download(sap_bounds* bnd, int32_t v, double* lb, double* ub);
/* device-to-device copy of vector sv of src into vector dv of dst (same matrix) */
int sap_bounds_copy(sap_bounds* dst, int32_t dv, const sap_bounds* src, int32_t sv);
/* propagate vector 0 of bnd in place; no host transfer of bounds */
int sap_propagate_resident(sap_bounds* bnd, const sap_params* params, sap_stats* stats);

/* Batched node propagation (many B&B node bound vectors on one matrix, one launch per
 * round for all vectors).  Vector v starts from base_lb/base_ub with the overlay entries
 * ovptr[v]..ovptr[v+1]-1 applied (column ovcol[q] gets bounds [ovlb[q], ovub[q]]).
 * Per-vector verdicts go to result[v] (SAP_CUTOFF/SAP_REDUCEDDOM/SAP_DIDNOTFIND) and
 * rounds[v]; if out_lb/out_ub are non-NULL they receive nvec*ncols propagated bounds
 * (vector-major).  bnd must have been created with at least nvec vectors; it holds the
 * propagated bounds on the device afterwards. */
int sap_propagate_batch(sap_bounds* bnd, const sap_params* params, int32_t nvec, const double* base_lb,
   const double* base_ub, const int64_t* ovptr, const int32_t* ovcol, const double* ovlb,
   const double* ovub, int32_t* result, int32_t* rounds, double* out_lb, double* out_ub,
   sap_stats* stats);

/* Node propagation from a device-resident base vector (vector basevec of base, same
 * matrix), with sparse results: vector v starts from the base with the overlay entries
 * ovptr[v]..ovptr[v+1]-1 applied (replaces a loop of SCIPpropagateProbing calls,
 * scip_probing.c, one per node).  result[v] and rounds[v] as in sap_propagate_batch.  The
 * columns whose bounds differ from the node's start are returned per vector, sorted by
 * column: entries chgptr[v]..chgptr[v+1]-1 of chgcol/chglb/chgub (empty for cut-off
 * vectors).  chgptr (nvec+1) is always filled; if chgptr[nvec] > chgcap the lists are not
 * written and SAP_NOMEMORY is returned (bounds stay on the device: sap_bounds_download).
 * chgcol == NULL asks for verdicts only. */
int sap_propagate_nodes(sap_bounds* bnd, const sap_params* params, int32_t nvec, const sap_bounds* base,
   int32_t basevec, const int64_t* ovptr, const int32_t* ovcol, const double* ovlb, const double* ovub,
   int32_t* result, int32_t* rounds, int64_t chgcap, int64_t* chgptr, int32_t* chgcol, double* chglb,
   double* chgub, sap_stats* stats);

/* Probing on binary columns: for each probe column x (integral, bounds exactly [0,1] in the
 * base vector) both branches x = 0 and x = 1 are propagated to fixpoint from vector basevec of
 * base, two node vectors per probe, as many probes per launch as bnd has vector pairs.  This
 * is the work of SCIPapplyProbingVar (prop_probing.c:1203-1311) for all probe columns at once;
 * the output is what SCIPanalyzeDeductionsProbing (prop_probing.c:1314) needs from the two
 * branches:
 *   verdict[p]  SAP_PROBE_DOWN_INFEASIBLE (x = 0 is cut off) | SAP_PROBE_UP_INFEASIBLE (x = 1);
 *   rounds      (may be NULL) 2*nprobe: rounds of branch x = 0, then x = 1, per probe;
 *   the columns other than x whose bounds changed in both branches (only these can give a
 *   fixing, a tightening of the union bounds or an aggregation), sorted by column: entries
 *   dedptr[p]..dedptr[p+1]-1 of dedcol with the bounds in the x = 0 branch (dlb0, dub0) and in
 *   the x = 1 branch (dlb1, dub1); empty when a branch is cut off.
 * Implications from columns changed in one branch only are not listed (sap_propagate_nodes
 * with the same two overlays gives them).  verdict, rounds and dedptr (nprobe+1) are always
 * filled; if dedptr[nprobe] > dedcap the lists are not written and SAP_NOMEMORY is returned;
 * dedcol == NULL asks for verdicts and counts only.  A probe column that is not an unfixed
 * binary of the base is SAP_INVALIDDATA.  stats->result is SAP_CUTOFF when both branches of
 * some probe are cut off (the base itself is infeasible), SAP_REDUCEDDOM when some branch is cut
 * off or some column changed in both branches, else SAP_DIDNOTFIND.
 * SAP_FLAG_BASE_IS_FIXPOINT in params->flags as in sap_propagate_nodes. */
#define SAP_PROBE_DOWN_INFEASIBLE 1
#define SAP_PROBE_UP_INFEASIBLE   2
int sap_probe(sap_bounds* bnd, const sap_params* params, int32_t nprobe, const int32_t* probecol,
   const sap_bounds* base, int32_t basevec, int32_t* verdict, int32_t* rounds, int64_t dedcap, int64_t* dedptr,
   int32_t* dedcol, double* dlb0, double* dub0, double* dlb1, double* dub1, sap_stats* stats);
/* Split probing, the propagation part of strong branching (SCIPgetVarStrongbranchWithPropagation,
 * scip_var.c:4139-4440, for many candidates at once): probe p branches column probecol[p] into
 * x <= downub[p] (SAP_PROBE_DOWN_INFEASIBLE) and x >= uplb[p] (SAP_PROBE_UP_INFEASIBLE), with
 * base lb <= downub[p] < uplb[p] <= base ub (else SAP_INVALIDDATA); SCIP's strong branching
 * uses downub = feasCeil(solval - 1), uplb = feasFloor(solval + 1).  Outputs as sap_probe; the
 * pair lists give the valid bounds of both children (SCIP's newlbs/newubs: the union of the
 * children's bounds differs from the base only on columns changed in both), and when neither
 * branch is infeasible they include the probe column itself with its bounds in both branches. */
int sap_probe_split(sap_bounds* bnd, const sap_params* params, int32_t nprobe, const int32_t* probecol,
   const double* downub, const double* uplb, const sap_bounds* base, int32_t basevec, int32_t* verdict,
   int32_t* rounds, int64_t dedcap, int64_t* dedptr, int32_t* dedcol, double* dlb0, double* dub0, double* dlb1,
   double* dub1, sap_stats* stats);

/* Independent instances packed block-diagonally into one resident matrix (one launch per
 * round for all of them; replaces one consPropLinear fixpoint per SCIP instance).  Instance
 * s owns rows [rowoff[s], rowoff[s+1]) and columns [coloff[s], coloff[s+1]) of the
 * concatenated CSR (global column indices); a row referencing a column of another
 * instance is SAP_INVALIDDATA. */
int sap_matrix_create_packed(sap_context* ctx, int32_t ninst, const int64_t* rowoff, const int64_t* coloff,
   int64_t nrows, int64_t ncols, const int64_t* rowptr, const int32_t* colidx, const double* vals,
   const double* lhs, const double* rhs, const uint8_t* integral, const double* glb, const double* gub,
   sap_matrix** mat);
/* Propagate all instances of a packed matrix (or the one instance of a plain matrix) from
 * lb/ub (overwritten).  Per instance: result[s] (SAP_CUTOFF / SAP_REDUCEDDOM /
 * SAP_DIDNOTFIND) and rounds[s]; a cutoff of one instance does not stop the others, and every
 * instance gets exactly the result it gets when propagated alone. */
int sap_propagate_packed(sap_matrix* mat, const sap_params* params, double* lb, double* ub, int32_t* result,
   int32_t* rounds, sap_stats* stats);

/* Continuous batching (serving).  The nvec vectors of a sap_bounds are slots.
 *   sap_slots_admit   starts slots[0..n) from vector basevec of base (another sap_bounds of
 *                     the same matrix) plus the overlays; with SAP_FLAG_BASE_IS_FIXPOINT in
 *                     flags only rows of overlay columns are swept first, else every row.
 *   sap_slots_launch  enqueues SAP_STEP_SHORT or SAP_STEP_LONG rounds for every running slot
 *                     (asynchronous; admitted slots join at the next launch);
 *   sap_slots_query   *ready = 1 when the launched step has finished;
 *   sap_slots_wait    waits for the step and lists the slots that finished in it (fixpoint,
 *                     cutoff or their own maxrounds); stats cover the step;
 *   sap_slots_step    launch + wait;
 *   sap_slots_collect verdicts, rounds and changed bounds of finished slots (as in
 *                     sap_propagate_nodes), after which the slots are free.
 * Every slot gets exactly the result it gets alone: slots share launches, not state.  One
 * sap_bounds serves either as slots or for one-shot calls (batch/nodes), not both.  The
 * admission, step and collection calls of one sap_bounds must not overlap. */
#define SAP_STEP_SHORT 2
#define SAP_STEP_LONG  8
int sap_slots_admit(sap_bounds* bnd, int32_t n, const int32_t* slots, const sap_bounds* base, int32_t basevec,
   const int64_t* ovptr, const int32_t* ovcol, const double* ovlb, const double* ovub, int32_t flags);
int sap_slots_launch(sap_bounds* bnd, const sap_params* params, int32_t rounds);
int sap_slots_query(sap_bounds* bnd, int32_t* ready);
int sap_slots_wait(sap_bounds* bnd, int32_t* nfinished, int32_t* finished, sap_stats* stats);
int sap_slots_step(sap_bounds* bnd, const sap_params* params, int32_t rounds, int32_t* nfinished, int32_t* finished,
   sap_stats* stats);
int sap_slots_collect(sap_bounds* bnd, int32_t n, const int32_t* slots, int32_t* result, int32_t* rounds,
   int64_t chgcap, int64_t* chgptr, int32_t* chgcol, double* chglb, double* chgub);

/* Row activities on the given bounds with the engine's canonical double-double sums
 * (diagnostic / parity): minact/maxact finite parts (hi+lo), counts[8*i..8*i+7] =
 * {minposinf, minneginf, minposhuge, minneghuge, maxposinf, maxneginf, maxposhuge,
 * maxneghuge}, maxactdelta as in consdataRecomputeMaxActivityDelta (cons_linear.c:1542).
 * Any output pointer may be NULL. */
/* sap_propagate on bounds already in device memory, allocating nothing.  No counterpart in
 * SCIP: it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  dlb and dub are device pointers to
 * ncols doubles, the node's bounds, overwritten with the fixpoint unless the result is a
 * cutoff.  The work arrays live in one static device array per GPU (about 61 MiB, held until
 * the process ends), so the matrix may have at most 2^20 columns and 2^20 rows and one
 * instance, and ranged-row propagation is refused (SAP_INVALIDDATA); rounds run from a host
 * loop.  Calls on one GPU share that array: calls that overlap on one GPU (several threads or
 * contexts) race.  The results equal sap_propagate's. */
int sap_propagate_device_noalloc(sap_matrix* mat, const sap_params* params, double* dlb, double* dub,
   sap_stats* stats);

/* Measurement: the memory floor of the row sweep on this matrix -- its value and column-index
 * streams and the (lb, ub) gather of every nonzero from vector 0 of b, without the sweep's
 * arithmetic -- timed with HIP events over reps launches of blocks workgroups (0: one lane
 * per four nonzeros, no grid-stride loop); *ms_per_pass is the mean launch time.  No
 * counterpart in SCIP. */
int sap_bench_gather(sap_bounds* b, int32_t reps, int32_t blocks, double* ms_per_pass);

int sap_row_activities(sap_matrix* mat, const sap_params* params, const double* lb, const double* ub,
   double* minact, double* maxact, int32_t* counts, double* maxactdelta);

#ifdef __cplusplus
}
#endif

#endif
