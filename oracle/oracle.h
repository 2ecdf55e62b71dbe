/* oracle.h -- CPU restatement of SCIP's linear-constraint domain propagation.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it, as the checker.
 *
 * Two schedules are provided:
 *   orc_propagate_jacobi  - round-synchronous (Jacobi) propagation with the canonical
 *                           pairwise double-double activity sum; the bit-exact twin of
 *                           the HIP engine.
 *   orc_propagate_scip    - SCIP-order (Gauss-Seidel) propagation with incremental,
 *                           event-driven activity updates, restating
 *                           src/scip/cons_linear.c tightenBounds/propagateCons/consPropLinear
 *                           and src/scip/solve.c propagateDomains.
 */
#ifndef SCIP_AMD_ORACLE_H
#define SCIP_AMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* numerics and propagation parameters; defaults follow src/scip/def.h:172-184 and
 * cons_linear.c:17587-17709 (maxeasyactivitydelta) */
typedef struct orc_params
{
   double  infinity;              /* numerics/infinity (1e20) */
   double  epsilon;               /* numerics/epsilon (1e-9) */
   double  sumepsilon;            /* numerics/sumepsilon (1e-6) */
   double  feastol;               /* numerics/feastol (1e-6) */
   double  boundstreps;           /* numerics/boundstreps (0.05) */
   double  hugeval;               /* numerics/hugeval (1e15) */
   double  recompfac;             /* numerics/recomputefac (1e6) */
   double  maxeasyactivitydelta;  /* constraints/linear/maxeasyactivitydelta (1e6) */
   int32_t maxrounds;             /* propagating/maxrounds (-1: unlimited) */
   int32_t rangedrows;            /* constraints/linear/rangedrowpropagation (Jacobi twin only) */
   int32_t notighten;             /* consPropLinear's tightenbounds FALSE (cons_linear.c:16147-16161):
                                     rows are only tested for infeasibility and redundancy */
   int32_t pad;
} orc_params;

typedef struct orc_matrix
{
   int64_t        nrows;
   int64_t        ncols;
   const int64_t* rowptr;         /* nrows+1 */
   const int32_t* colidx;         /* nnz */
   const double*  vals;           /* nnz */
   const double*  lhs;            /* nrows */
   const double*  rhs;            /* nrows */
   const uint8_t* integral;       /* ncols: SCIPvarIsIntegral */
   const uint8_t* binary;         /* ncols: SCIPvarIsBinary (integral, global bounds within [0,1]) */
} orc_matrix;

typedef struct orc_result
{
   int32_t cutoff;                /* 1 if infeasibility was detected */
   int32_t rounds;                /* propagation rounds executed */
   int64_t nchgbds;               /* accepted bound changes */
   int64_t rows_swept;            /* rows swept over all rounds */
   int64_t nnz_swept;             /* nonzeros swept over all rounds */
   int64_t nrecompute;            /* reliable residual recomputes (cons_linear.c:2580-2657) */
} orc_result;

typedef struct orc_rowact
{
   double  minact;                /* finite part hi+lo of the min activity (canonical sum) */
   double  maxact;                /* finite part hi+lo of the max activity (canonical sum) */
   double  minact_hi, minact_lo;
   double  maxact_hi, maxact_lo;
   int32_t minposinf, minneginf, minposhuge, minneghuge;
   int32_t maxposinf, maxneginf, maxposhuge, maxneghuge;
   double  maxactdelta;
} orc_rowact;

void orc_params_default(orc_params* p);

/* SCIPvarIsBinary on global bounds (var.c:23801) */
void orc_compute_binary(int64_t ncols, const uint8_t* integral, const double* glb, const double* gub,
   uint8_t* binary);

/* per-row activities with the canonical pairwise sum (canonical=1) or SCIP's sequential
 * Sum21 order of consdataCalcActivities (canonical=0) */
void orc_row_activities(const orc_matrix* m, const orc_params* p, const double* lb, const double* ub,
   int canonical, orc_rowact* out);

/* Jacobi twin of the device engine; lb/ub are updated in place.  full_sweep=1 sweeps
 * every non-redundant row every round instead of the worklist (same result). */
int orc_propagate_jacobi(const orc_matrix* m, const orc_params* p, double* lb, double* ub,
   int full_sweep, orc_result* res);

/* the Jacobi twin started with only the rows flagged in mark0 (nrows bytes) marked: the
 * device's SAP_FLAG_BASE_IS_FIXPOINT start, where only rows of changed columns are swept in
 * the first round (same fixpoint when the rest is already one) */
int orc_propagate_jacobi_marked(const orc_matrix* m, const orc_params* p, double* lb, double* ub,
   const uint8_t* mark0, orc_result* res);

/* SCIP-order Gauss-Seidel restatement; lb/ub are updated in place */
int orc_propagate_scip(const orc_matrix* m, const orc_params* p, double* lb, double* ub,
   orc_result* res);

/* the Jacobi twin with reasons (restates the reason recording of sap_propagate_reasons):
 * history entries (round, column, bound 0 lower / 1 upper, value, the smallest row that
 * submitted the winning candidate, and its rule: 0 residual activity, 1 ranged-row test) in
 * round, column, bound order, and the certificate of a cutoff: the smallest infeasible row
 * with its side (0 minimal activity, 1 maximal activity, 2 ranged-row test), or the smallest
 * column whose bounds crossed with the rows of its two sides (-1: the column's own bound).
 * *nhist counts all entries; at most histcap are written. */
#define ORC_CONFLICT_NONE   0
#define ORC_CONFLICT_ROW    1
#define ORC_CONFLICT_COLUMN 2
typedef struct orc_conflict
{
   int32_t kind;
   int32_t round;
   int32_t row;
   int32_t side;
   int32_t col;
   int32_t rowlb;
   int32_t rowub;
   int32_t rules;   /* bit 0: rowlb's rule, bit 1: rowub's (1: ranged-row test) */
} orc_conflict;
int orc_propagate_jacobi_reasons(const orc_matrix* m, const orc_params* p, double* lb, double* ub, int64_t histcap,
   int64_t* nhist, int32_t* hround, int32_t* hcol, int32_t* hbound, double* hval, int32_t* hrow, int32_t* hrule,
   orc_conflict* conflict, orc_result* res);

/* double-double primitives restated from src/scip/dbldblarith.h, exported for the
 * known-answer tests of tests/src/misc/dbldblarith.c */
void orc_dd_sum(double a, double b, double* rhi, double* rlo);
void orc_dd_sum21(double ahi, double alo, double b, double* rhi, double* rlo);
void orc_dd_prod(double a, double b, double* rhi, double* rlo);
void orc_dd_prod21(double ahi, double alo, double b, double* rhi, double* rlo);
void orc_dd_pairadd(double ahi, double alo, double bhi, double blo, double* rhi, double* rlo);

/* canonical pairwise sum of n doubles (leaves (x_p, 0), zero padded perfect tree) */
void orc_canonical_sum(const double* x, int64_t n, double* rhi, double* rlo);

#ifdef __cplusplus
}
#endif

#endif
