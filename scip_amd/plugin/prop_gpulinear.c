/* prop_gpulinear.c -- SCIP propagator: linear-constraint domain propagation on an AMD MI355X.
 *
 * The propagator collects the rows of all linear constraints (SCIPgetVarsLinear /
 * SCIPgetValsLinear / SCIPgetLhsLinear / SCIPgetRhsLinear, cons_linear.c:18326-18436) into a
 * CSR matrix over the active problem variables (SCIPvarGetProbindex), keeps it resident on
 * the device, and in every PROPEXEC call propagates the node's local bounds to the device
 * fixpoint with sap_propagate.  The device applies the acceptance rules of
 * SCIPinferVarLbCons / SCIPinferVarUbCons (scip_var.c:6965-7174), so every returned bound is
 * handed to SCIPinferVarLbProp / SCIPinferVarUbProp (scip_var.c:7589/7705) unchanged.
 *
 * In presolving it can also probe: every unfixed binary column is fixed to 0 and to 1 and
 * both branches are propagated on the device, two node vectors per probe, many probes per
 * launch (the work of SCIPapplyProbingVar, prop_probing.c:1203, for all binaries at once).  An
 * infeasible branch fixes the column to the other value (prop_probing.c:477-556); otherwise
 * the columns changed in both branches go to SCIP's own SCIPanalyzeDeductionsProbing
 * (prop_probing.c:1314), which derives fixings, tightenings, aggregations and implications.
 *
 * With reasons (the default), PROPEXEC propagates with sap_propagate_reasons: the device
 * returns every accepted change with its round and reason row, and the changes are handed to
 * SCIP in round order with the row as inferinfo.  Each change is then implied by its row and
 * the bounds already in force, so PROPRESPROP explains it with the row's other bounds at the
 * change's time, as cons_linear's resolvePropagation does with its own inferinfo
 * (cons_linear.c:344-369, :5097-5189).  A device cutoff comes with a certificate (an
 * infeasible row, or a column whose bounds crossed with the rows of both sides), from which
 * the propagator starts conflict analysis (cons_linear.c analyzeConflict).
 *
 * Parameters mirror SCIP's (numerics/..., constraints/linear/maxeasyactivitydelta); the
 * propagator's own:
 *   propagating/gpulinear/device          HIP device ordinal
 *   propagating/gpulinear/maxrounds       device rounds per call (-1: to fixpoint)
 *   propagating/gpulinear/probing         probe the binaries in presolving
 *   propagating/gpulinear/probingvectors  node vectors of one probing launch (two per probe)
 *   propagating/gpulinear/reasons         record reasons for conflict analysis
 *
 * Error convention: every HIP or engine failure becomes a SCIP_RETCODE (the engine's return
 * codes have SCIP_RETCODE values); nothing aborts.
 */
#include <limits.h>
#include <math.h>
#include <string.h>

#include "scip/scip.h"
#include "scip/cons_linear.h"
#include "scip/prop_probing.h"
#include "scip_amd_prop.h"
#include "prop_gpulinear.h"

#define PROP_NAME              "gpulinear"
#define PROP_DESC              "linear constraint propagation on an AMD MI355X (scip_amd engine)"
#define PROP_PRIORITY          -1000          /**< after the constraint handlers */
#define PROP_FREQ                  1
#define PROP_DELAY             FALSE
#define PROP_TIMING            SCIP_PROPTIMING_BEFORELP
#define PROP_PRESOL_PRIORITY   -1000
#define PROP_PRESOL_MAXROUNDS     -1
#define PROP_PRESOLTIMING      SCIP_PRESOLTIMING_MEDIUM

#define DEFAULT_DEVICE             0
#define DEFAULT_MAXROUNDS         -1
#define DEFAULT_PROBING        FALSE
#define DEFAULT_PROBINGVECTORS  2048
#define DEFAULT_REASONS         TRUE

/** propagator data */
struct SCIP_PropData
{
   sap_context*          ctx;                /**< engine context (device + stream) */
   sap_matrix*           mat;                /**< device-resident linear rows, or NULL */
   SCIP_VAR**            vars;               /**< column -> variable (active problem variables) */
   int                   nvars;              /**< number of columns */
   int                   nrows;              /**< number of rows on the device */
   int                   device;             /**< parameter: HIP device */
   int                   maxrounds;          /**< parameter: device rounds per call */
   SCIP_Longint          ncalls;             /**< PROPEXEC calls that ran on the device */
   SCIP_Longint          ncutoffs;           /**< cutoffs found */
   SCIP_Longint          nchgbds;            /**< bound changes handed to SCIP */
   SCIP_Bool             probing;            /**< parameter: probe the binaries in presolving */
   int                   probingvectors;     /**< parameter: node vectors of one probing launch */
   sap_bounds*           probebase;          /**< probing: the presolving fixpoint (one vector) */
   sap_bounds*           probeset;           /**< probing: the branch vectors */
   SCIP_Longint          nprobed;            /**< probes run on the device */
   SCIP_Longint          nprobefixed;        /**< columns fixed by probing */
   SCIP_Longint          nprobeaggr;         /**< aggregations found by probing */
   SCIP_Longint          nprobeimpl;         /**< implications found by probing */
   SCIP_Bool             reasons;            /**< parameter: record reasons for conflict analysis */
   int64_t*              rowptr;             /**< host copy of the device rows (for PROPRESPROP) */
   int32_t*              colidx;
   double*               vals;
   int                   nnz;
   SCIP_Longint          nconflicts;         /**< conflict analyses started from device cutoffs */
   sap_bounds*           diveset;            /**< dives: one node vector per dive */
   int                   divevecs;           /**< vectors of diveset */
};

/*
 * local methods
 */

/** maps an engine return code (SCIP_RETCODE values) to a SCIP_RETCODE */
static
SCIP_RETCODE engineCall(
   int                   rc,                 /**< engine return code */
   const char*           what                /**< engine function */
   )
{
   if( rc != SAP_OKAY )
   {
      SCIPerrorMessage("%s failed: %s (%d)\n", what, sap_error_string(rc), rc);
      return (SCIP_RETCODE)rc;
   }
   return SCIP_OKAY;
}

/** engine parameters from SCIP's numerics */
static
SCIP_RETCODE getEngineParams(
   SCIP*                 scip,               /**< SCIP data structure */
   SCIP_PROPDATA*        propdata,           /**< propagator data */
   sap_params*           params              /**< parameters to fill */
   )
{
   SCIP_Real val;

   sap_params_default(params);
   params->infinity = SCIPinfinity(scip);
   params->epsilon = SCIPepsilon(scip);
   params->sumepsilon = SCIPsumepsilon(scip);
   params->feastol = SCIPfeastol(scip);
   params->hugeval = SCIPgetHugeValue(scip);
   SCIP_CALL( SCIPgetRealParam(scip, "numerics/boundstreps", &val) );
   params->boundstreps = val;
   SCIP_CALL( SCIPgetRealParam(scip, "numerics/recompfac", &val) );
   params->recompfac = val;
   SCIP_CALL( SCIPgetRealParam(scip, "constraints/linear/maxeasyactivitydelta", &val) );
   params->maxeasyactivitydelta = val;
   params->maxrounds = propdata->maxrounds;
   return SCIP_OKAY;
}

/** an int parameter of the linear constraint handler, or its default when it is not there */
static
SCIP_RETCODE linearIntParam(
   SCIP*                 scip,               /**< SCIP data structure */
   const char*           name,               /**< parameter name */
   int                   defaultvalue,       /**< cons_linear.c's default */
   int*                  value               /**< pointer to store the value */
   )
{
   if( SCIPgetIntParam(scip, name, value) != SCIP_OKAY )
      *value = defaultvalue;
   return SCIP_OKAY;
}

/** consPropLinear's gating of a propagation call (cons_linear.c:16147-16171), applied to the
 *  node propagation of PROPEXEC.  This propagator's own frequency takes the place of the linear
 *  handler's propagation frequency, which is -1 when the handler's propagation is switched off
 *  in favour of this one.
 *  - In probing the bounds are always tightened and ranged rows are never propagated.
 *  - Otherwise the bounds are tightened at the depths that constraints/linear/tightenboundsfreq
 *    selects (SAP_FLAG_NOTIGHTEN elsewhere: rows are only tested for infeasibility and
 *    redundancy), and ranged-row propagation needs constraints/linear/rangedrowpropagation,
 *    no repropagation, a depth of at most rangedrowmaxdepth, a depth that rangedrowfreq
 *    selects and a stage other than PRESOLVING. */
static
SCIP_RETCODE linearGating(
   SCIP*                 scip,               /**< SCIP data structure */
   SCIP_PROP*            prop,               /**< the propagator */
   sap_params*           params              /**< parameters to complete */
   )
{
   SCIP_Bool tightenbounds;
   SCIP_Bool ranged = FALSE;

   if( SCIPinProbing(scip) )
      tightenbounds = TRUE;
   else
   {
      int depth = SCIPgetDepth(scip);
      int propfreq = SCIPpropGetFreq(prop);
      int tbfreq;
      int rrfreq;
      int rrmaxdepth;
      int f;

      SCIP_CALL( linearIntParam(scip, "constraints/linear/tightenboundsfreq", 1, &tbfreq) );
      SCIP_CALL( linearIntParam(scip, "constraints/linear/rangedrowfreq", 1, &rrfreq) );
      SCIP_CALL( linearIntParam(scip, "constraints/linear/rangedrowmaxdepth", INT_MAX, &rrmaxdepth) );
      f = propfreq * tbfreq;
      tightenbounds = (tbfreq >= 0) && ((f == 0 && depth == 0) || (f >= 1 && (depth % f == 0)));

      if( SCIPgetBoolParam(scip, "constraints/linear/rangedrowpropagation", &ranged) != SCIP_OKAY )
         ranged = TRUE;
      ranged = ranged && !SCIPinRepropagation(scip);
      ranged = ranged && (depth <= rrmaxdepth);
      f = propfreq * rrfreq;
      ranged = ranged && (rrfreq >= 0) && ((f == 0 && depth == 0) || (f >= 1 && (depth % f == 0)));
      ranged = ranged && (SCIPgetStage(scip) != SCIP_STAGE_PRESOLVING);
   }
   if( !tightenbounds )
      params->flags |= SAP_FLAG_NOTIGHTEN;
   else if( ranged )
      params->flags |= SAP_FLAG_RANGEDROWS;
   return SCIP_OKAY;
}

/** releases the device matrix and the column map */
static
SCIP_RETCODE freeMatrix(
   SCIP*                 scip,               /**< SCIP data structure */
   SCIP_PROPDATA*        propdata            /**< propagator data */
   )
{
   /* bound sets before their matrix */
   if( propdata->diveset != NULL )
   {
      SCIP_CALL( engineCall(sap_bounds_free(propdata->diveset), "sap_bounds_free") );
      propdata->diveset = NULL;
      propdata->divevecs = 0;
   }
   if( propdata->probeset != NULL )
   {
      SCIP_CALL( engineCall(sap_bounds_free(propdata->probeset), "sap_bounds_free") );
      propdata->probeset = NULL;
   }
   if( propdata->probebase != NULL )
   {
      SCIP_CALL( engineCall(sap_bounds_free(propdata->probebase), "sap_bounds_free") );
      propdata->probebase = NULL;
   }
   if( propdata->mat != NULL )
   {
      SCIP_CALL( engineCall(sap_matrix_free(propdata->mat), "sap_matrix_free") );
      propdata->mat = NULL;
   }
   if( propdata->vars != NULL )
   {
      SCIPfreeBlockMemoryArray(scip, &propdata->vars, propdata->nvars);
      propdata->vars = NULL;
   }
   if( propdata->rowptr != NULL )
   {
      SCIPfreeBlockMemoryArray(scip, &propdata->vals, propdata->nnz);
      SCIPfreeBlockMemoryArray(scip, &propdata->colidx, propdata->nnz);
      SCIPfreeBlockMemoryArray(scip, &propdata->rowptr, propdata->nrows + 1);
      propdata->rowptr = NULL;
   }
   propdata->nnz = 0;
   propdata->nvars = 0;
   propdata->nrows = 0;
   return SCIP_OKAY;
}

/** builds the CSR matrix of all linear constraints over the active variables and uploads it */
static
SCIP_RETCODE buildMatrix(
   SCIP*                 scip,               /**< SCIP data structure */
   SCIP_PROPDATA*        propdata            /**< propagator data */
   )
{
   SCIP_CONSHDLR* conshdlr;
   SCIP_CONS** conss;
   SCIP_VAR** probvars;
   SCIP_VAR** rowvars;
   SCIP_Real* rowvals;
   int64_t* rowptr;
   int32_t* colidx;
   double* vals;
   double* lhs;
   double* rhs;
   uint8_t* integral;
   double* glb;
   double* gub;
   int nconss;
   int nvars;
   int nnzcap;
   int nnz;
   int nrows;
   int c;
   int j;

   SCIP_CALL( freeMatrix(scip, propdata) );

   conshdlr = SCIPfindConshdlr(scip, "linear");
   if( conshdlr == NULL )
      return SCIP_OKAY;
   nconss = SCIPconshdlrGetNConss(conshdlr);
   conss = SCIPconshdlrGetConss(conshdlr);
   nvars = SCIPgetNVars(scip);
   probvars = SCIPgetVars(scip);
   if( nconss == 0 || nvars == 0 )
      return SCIP_OKAY;

   nnzcap = 0;
   for( c = 0; c < nconss; ++c )
      nnzcap += SCIPgetNVarsLinear(scip, conss[c]);

   SCIP_CALL( SCIPallocBlockMemoryArray(scip, &propdata->vars, nvars) );
   propdata->nvars = nvars;
   for( j = 0; j < nvars; ++j )
      propdata->vars[j] = probvars[j];

   SCIP_CALL( SCIPallocBufferArray(scip, &rowptr, nconss + 1) );
   SCIP_CALL( SCIPallocBufferArray(scip, &colidx, nnzcap + 1) );
   SCIP_CALL( SCIPallocBufferArray(scip, &vals, nnzcap + 1) );
   SCIP_CALL( SCIPallocBufferArray(scip, &lhs, nconss) );
   SCIP_CALL( SCIPallocBufferArray(scip, &rhs, nconss) );
   SCIP_CALL( SCIPallocBufferArray(scip, &integral, nvars) );
   SCIP_CALL( SCIPallocBufferArray(scip, &glb, nvars) );
   SCIP_CALL( SCIPallocBufferArray(scip, &gub, nvars) );

   nnz = 0;
   nrows = 0;
   rowptr[0] = 0;
   for( c = 0; c < nconss; ++c )
   {
      SCIP_Real constant = 0.0;
      int nrowvars;
      int varssize;
      int requiredsize;
      int k;
      SCIP_Bool ok = TRUE;

      if( !SCIPconsIsActive(conss[c]) )
         continue;
      nrowvars = SCIPgetNVarsLinear(scip, conss[c]);
      varssize = nrowvars + 1;
      /* active representation of the row (aggregated / fixed variables resolved) */
      SCIP_CALL( SCIPallocBufferArray(scip, &rowvars, varssize) );
      SCIP_CALL( SCIPallocBufferArray(scip, &rowvals, varssize) );
      for( k = 0; k < nrowvars; ++k )
      {
         rowvars[k] = SCIPgetVarsLinear(scip, conss[c])[k];
         rowvals[k] = SCIPgetValsLinear(scip, conss[c])[k];
      }
      SCIP_CALL( SCIPgetProbvarLinearSum(scip, rowvars, rowvals, &nrowvars, varssize, &constant, &requiredsize) );
      if( requiredsize > varssize )
         ok = FALSE;   /* multi-aggregations expand the row; leave it to cons_linear */
      for( k = 0; ok && k < nrowvars; ++k )
         if( SCIPvarGetProbindex(rowvars[k]) < 0 || rowvals[k] == 0.0 )
            ok = FALSE;
      if( ok && nnz + nrowvars <= nnzcap )
      {
         SCIP_Real l = SCIPgetLhsLinear(scip, conss[c]);
         SCIP_Real r = SCIPgetRhsLinear(scip, conss[c]);
         for( k = 0; k < nrowvars; ++k )
         {
            colidx[nnz] = (int32_t)SCIPvarGetProbindex(rowvars[k]);
            vals[nnz] = rowvals[k];
            ++nnz;
         }
         lhs[nrows] = SCIPisInfinity(scip, -l) ? l : l - constant;
         rhs[nrows] = SCIPisInfinity(scip, r) ? r : r - constant;
         ++nrows;
         rowptr[nrows] = nnz;
      }
      SCIPfreeBufferArray(scip, &rowvals);
      SCIPfreeBufferArray(scip, &rowvars);
   }

   for( j = 0; j < nvars; ++j )
   {
      integral[j] = SCIPvarIsIntegral(propdata->vars[j]) ? 1 : 0;
      glb[j] = SCIPvarGetLbGlobal(propdata->vars[j]);
      gub[j] = SCIPvarGetUbGlobal(propdata->vars[j]);
   }

   if( nrows > 0 )
   {
      SCIP_CALL( engineCall(sap_matrix_create(propdata->ctx, nrows, nvars, rowptr, colidx, vals, lhs, rhs, integral,
            glb, gub, &propdata->mat), "sap_matrix_create") );
      /* the rows stay on the host too: PROPRESPROP and conflict analysis read them */
      SCIP_CALL( SCIPallocBlockMemoryArray(scip, &propdata->rowptr, nrows + 1) );
      SCIP_CALL( SCIPallocBlockMemoryArray(scip, &propdata->colidx, nnz) );
      SCIP_CALL( SCIPallocBlockMemoryArray(scip, &propdata->vals, nnz) );
      memcpy(propdata->rowptr, rowptr, sizeof(int64_t) * (size_t)(nrows + 1));
      memcpy(propdata->colidx, colidx, sizeof(int32_t) * (size_t)nnz);
      memcpy(propdata->vals, vals, sizeof(double) * (size_t)nnz);
      propdata->nnz = nnz;
   }
   propdata->nrows = nrows;

   SCIPfreeBufferArray(scip, &gub);
   SCIPfreeBufferArray(scip, &glb);
   SCIPfreeBufferArray(scip, &integral);
   SCIPfreeBufferArray(scip, &rhs);
   SCIPfreeBufferArray(scip, &lhs);
   SCIPfreeBufferArray(scip, &vals);
   SCIPfreeBufferArray(scip, &colidx);
   SCIPfreeBufferArray(scip, &rowptr);

   return SCIP_OKAY;
}

/** propagates the given (local or global) bounds on the device; returns the new bounds */
static
SCIP_RETCODE propagateOnDevice(
   SCIP*                 scip,               /**< SCIP data structure */
   SCIP_PROP*            prop,               /**< the propagator */
   SCIP_PROPDATA*        propdata,           /**< propagator data */
   SCIP_Bool             local,              /**< local (node) bounds, else global bounds */
   double*               lb,                 /**< array of nvars: new lower bounds */
   double*               ub,                 /**< array of nvars: new upper bounds */
   SCIP_Bool*            cutoff              /**< pointer to store whether infeasibility was detected */
   )
{
   sap_params params;
   sap_stats stats;
   int j;

   for( j = 0; j < propdata->nvars; ++j )
   {
      SCIP_VAR* var = propdata->vars[j];
      lb[j] = local ? SCIPvarGetLbLocal(var) : SCIPvarGetLbGlobal(var);
      ub[j] = local ? SCIPvarGetUbLocal(var) : SCIPvarGetUbGlobal(var);
   }
   SCIP_CALL( getEngineParams(scip, propdata, &params) );
   if( local )
   {
      SCIP_CALL( linearGating(scip, prop, &params) );
   }
   memset(&stats, 0, sizeof(stats));
   SCIP_CALL( engineCall(sap_propagate(propdata->mat, &params, lb, ub, &stats), "sap_propagate") );
   *cutoff = (stats.result == SAP_CUTOFF);
   ++propdata->ncalls;
   if( *cutoff )
      ++propdata->ncutoffs;
   return SCIP_OKAY;
}

/** probing on the device from the presolving fixpoint lb/ub (the bounds just handed to SCIP):
 *  both branches of every unfixed binary, then the deductions of prop_probing.c's applyProbing
 *  (:477-556 for an infeasible branch, SCIPanalyzeDeductionsProbing for the others) */
static
SCIP_RETCODE probeOnDevice(
   SCIP*                 scip,               /**< SCIP data structure */
   SCIP_PROPDATA*        propdata,           /**< propagator data */
   const double*         lb,                 /**< array of nvars: presolving fixpoint, lower bounds */
   const double*         ub,                 /**< array of nvars: presolving fixpoint, upper bounds */
   int*                  nfixedvars,         /**< pointer to count fixings */
   int*                  naggrvars,          /**< pointer to count aggregations */
   int*                  nchgbds,            /**< pointer to count tightenings */
   SCIP_Bool*            cutoff              /**< pointer to store whether infeasibility was detected */
   )
{
   sap_params params;
   sap_stats stats;
   int32_t* probecol;
   int32_t* verdict;
   int64_t* dedptr;
   int32_t* dedcol = NULL;
   double* ded[4] = { NULL, NULL, NULL, NULL };
   SCIP_VAR** pairvars = NULL;
   int64_t cap;
   int nprobe = 0;
   int rc;
   int p;
   int j;
   int k;

   *cutoff = FALSE;
   SCIP_CALL( SCIPallocBufferArray(scip, &probecol, propdata->nvars) );
   for( j = 0; j < propdata->nvars; ++j )
      if( SCIPvarIsBinary(propdata->vars[j]) && lb[j] == 0.0 && ub[j] == 1.0 )
         probecol[nprobe++] = (int32_t)j;
   if( nprobe == 0 )
   {
      SCIPfreeBufferArray(scip, &probecol);
      return SCIP_OKAY;
   }
   if( propdata->probebase == NULL )
   {
      SCIP_CALL( engineCall(sap_bounds_create(propdata->mat, 1, &propdata->probebase), "sap_bounds_create") );
   }
   if( propdata->probeset == NULL )
   {
      SCIP_CALL( engineCall(sap_bounds_create(propdata->mat, propdata->probingvectors, &propdata->probeset),
            "sap_bounds_create") );
   }
   SCIP_CALL( engineCall(sap_bounds_upload(propdata->probebase, 0, lb, ub), "sap_bounds_upload") );
   SCIP_CALL( getEngineParams(scip, propdata, &params) );
   params.flags |= SAP_FLAG_BASE_IS_FIXPOINT;

   SCIP_CALL( SCIPallocBufferArray(scip, &verdict, nprobe) );
   SCIP_CALL( SCIPallocBufferArray(scip, &dedptr, nprobe + 1) );
   cap = 16 * (int64_t)nprobe;
   for( ;; )
   {
      SCIP_CALL( SCIPallocBufferArray(scip, &dedcol, cap) );
      for( k = 0; k < 4; ++k )
      {
         SCIP_CALL( SCIPallocBufferArray(scip, &ded[k], cap) );
      }
      memset(&stats, 0, sizeof(stats));
      rc = sap_probe(propdata->probeset, &params, nprobe, probecol, propdata->probebase, 0, verdict, NULL, cap, dedptr,
         dedcol, ded[0], ded[1], ded[2], ded[3], &stats);
      if( rc != SAP_NOMEMORY || dedptr[nprobe] <= cap )
         break;
      /* too little room for the pairs: again with enough (same results) */
      cap = dedptr[nprobe];
      for( k = 3; k >= 0; --k )
         SCIPfreeBufferArray(scip, &ded[k]);
      SCIPfreeBufferArray(scip, &dedcol);
   }
   SCIP_CALL( engineCall(rc, "sap_probe") );
   propdata->nprobed += nprobe;

   SCIP_CALL( SCIPallocBufferArray(scip, &pairvars, propdata->nvars) );
   for( p = 0; p < nprobe && !*cutoff; ++p )
   {
      SCIP_VAR* var = propdata->vars[probecol[p]];
      SCIP_Bool infeasible = FALSE;
      SCIP_Bool fixed = FALSE;
      int n = (int)(dedptr[p + 1] - dedptr[p]);
      int nfix = 0;
      int naggr = 0;
      int nimpl = 0;
      int nchg = 0;

      /* fixed by the deductions of an earlier probe */
      if( SCIPvarGetLbGlobal(var) > 0.5 || SCIPvarGetUbGlobal(var) < 0.5 )
         continue;
      if( verdict[p] == (SAP_PROBE_DOWN_INFEASIBLE | SAP_PROBE_UP_INFEASIBLE) )
      {
         *cutoff = TRUE;
         break;
      }
      if( verdict[p] != 0 )
      {
         SCIP_CALL( SCIPfixVar(scip, var, (verdict[p] & SAP_PROBE_UP_INFEASIBLE) ? 0.0 : 1.0, &infeasible, &fixed) );
         *cutoff = infeasible;
         if( fixed )
         {
            ++(*nfixedvars);
            ++propdata->nprobefixed;
         }
         continue;
      }
      if( n == 0 )
         continue;
      for( k = 0; k < n; ++k )
         pairvars[k] = propdata->vars[dedcol[dedptr[p] + k]];
      SCIP_CALL( SCIPanalyzeDeductionsProbing(scip, var, 0.0, 1.0, n, pairvars, NULL, NULL, ded[0] + dedptr[p],
            ded[1] + dedptr[p], NULL, NULL, ded[2] + dedptr[p], ded[3] + dedptr[p], &nfix, &naggr, &nimpl, &nchg,
            cutoff) );
      *nfixedvars += nfix;
      *naggrvars += naggr;
      *nchgbds += nchg;
      propdata->nprobefixed += nfix;
      propdata->nprobeaggr += naggr;
      propdata->nprobeimpl += nimpl;
   }

   SCIPfreeBufferArray(scip, &pairvars);
   for( k = 3; k >= 0; --k )
      SCIPfreeBufferArray(scip, &ded[k]);
   SCIPfreeBufferArray(scip, &dedcol);
   SCIPfreeBufferArray(scip, &dedptr);
   SCIPfreeBufferArray(scip, &verdict);
   SCIPfreeBufferArray(scip, &probecol);
   return SCIP_OKAY;
}

/** adds to the conflict the bounds of row's columns other than skip that make up its minimal
 *  (maximal) activity: lower bounds where the coefficient is positive and upper bounds where it
 *  is negative (the reverse for the maximal activity), at the time of bdchgidx (NULL: now);
 *  cons_linear.c addConflictBounds (:5097) */
static
SCIP_RETCODE addRowBounds(
   SCIP*                 scip,               /**< SCIP data structure */
   SCIP_PROPDATA*        propdata,           /**< propagator data */
   int                   row,                /**< row of the device matrix */
   int                   skip,               /**< column left out, or -1 */
   SCIP_Bool             maximal,            /**< the maximal activity's bounds, else the minimal's */
   SCIP_BDCHGIDX*        bdchgidx            /**< time of the bounds, or NULL for the current ones */
   )
{
   int64_t k;

   for( k = propdata->rowptr[row]; k < propdata->rowptr[row + 1]; ++k )
   {
      SCIP_VAR* var = propdata->vars[propdata->colidx[k]];

      if( propdata->colidx[k] == skip )
         continue;
      if( (propdata->vals[k] > 0.0) != maximal )
      {
         SCIP_CALL( SCIPaddConflictLb(scip, var, bdchgidx) );
      }
      else
      {
         SCIP_CALL( SCIPaddConflictUb(scip, var, bdchgidx) );
      }
   }
   return SCIP_OKAY;
}

/** adds both bounds of every column of row at the time of bdchgidx (NULL: now): the reason of a
 *  ranged-row deduction, which reads the fixed columns' values and the other columns' ranges
 *  (cons_linear.c addConflictFixedVars :4981 and analyzeConflictRangedRow :5657 add the fixed
 *  columns and the columns whose bounds are not global; all bounds are a superset) */
static
SCIP_RETCODE addRowAllBounds(
   SCIP*                 scip,               /**< SCIP data structure */
   SCIP_PROPDATA*        propdata,           /**< propagator data */
   int                   row,                /**< row of the device matrix */
   SCIP_BDCHGIDX*        bdchgidx            /**< time of the bounds, or NULL for the current ones */
   )
{
   int64_t k;

   for( k = propdata->rowptr[row]; k < propdata->rowptr[row + 1]; ++k )
   {
      SCIP_VAR* var = propdata->vars[propdata->colidx[k]];

      SCIP_CALL( SCIPaddConflictLb(scip, var, bdchgidx) );
      SCIP_CALL( SCIPaddConflictUb(scip, var, bdchgidx) );
   }
   return SCIP_OKAY;
}

/** the reason of bound boundtype of column j derived from row by rule (0: residual activity,
 *  1: ranged-row test), at the time of bdchgidx */
static
SCIP_RETCODE addReason(
   SCIP*                 scip,               /**< SCIP data structure */
   SCIP_PROPDATA*        propdata,           /**< propagator data */
   int                   row,                /**< row of the device matrix */
   int                   rule,               /**< reason rule */
   int                   j,                  /**< the column whose bound was derived */
   SCIP_BOUNDTYPE        boundtype,          /**< the bound derived */
   SCIP_BDCHGIDX*        bdchgidx            /**< time of the bounds, or NULL for the current ones */
   );

/** coefficient of column j in row (0.0 if absent) */
static
double rowCoef(
   SCIP_PROPDATA*        propdata,           /**< propagator data */
   int                   row,                /**< row of the device matrix */
   int                   j                   /**< column */
   )
{
   int64_t k;

   for( k = propdata->rowptr[row]; k < propdata->rowptr[row + 1]; ++k )
      if( propdata->colidx[k] == j )
         return propdata->vals[k];
   return 0.0;
}

static
SCIP_RETCODE addReason(
   SCIP*                 scip,               /**< SCIP data structure */
   SCIP_PROPDATA*        propdata,           /**< propagator data */
   int                   row,                /**< row of the device matrix */
   int                   rule,               /**< reason rule */
   int                   j,                  /**< the column whose bound was derived */
   SCIP_BOUNDTYPE        boundtype,          /**< the bound derived */
   SCIP_BDCHGIDX*        bdchgidx            /**< time of the bounds, or NULL for the current ones */
   )
{
   double a;

   if( rule == 1 )
      return addRowAllBounds(scip, propdata, row, bdchgidx);
   /* a lower bound of a column with a positive coefficient came from lhs and the maximal
    * residual activity, with a negative one from rhs and the minimal residual; symmetric for
    * upper bounds */
   a = rowCoef(propdata, row, j);
   SCIP_CALL( addRowBounds(scip, propdata, row, j, (boundtype == SCIP_BOUNDTYPE_LOWER) == (a > 0.0), bdchgidx) );
   return SCIP_OKAY;
}

/** PROPEXEC with reasons: the device's changes in round order, each with its reason row and
 *  rule as inferinfo (row * 2 + rule); on a cutoff, the changes of the earlier rounds, then
 *  conflict analysis on the certificate */
static
SCIP_RETCODE execWithReasons(
   SCIP*                 scip,               /**< SCIP data structure */
   SCIP_PROP*            prop,               /**< the propagator */
   SCIP_PROPDATA*        propdata,           /**< propagator data */
   SCIP_RESULT*          result              /**< pointer to store the result */
   )
{
   sap_params params;
   sap_stats stats;
   sap_conflict conflict;
   double* lb;
   double* ub;
   double* hval = NULL;
   int32_t* hround = NULL;
   int32_t* hcol = NULL;
   int32_t* hbound = NULL;
   int32_t* hrow = NULL;
   int32_t* hrule = NULL;
   int64_t nhist = 0;
   int64_t cap = 4 * (int64_t)propdata->nvars + 64;
   SCIP_Bool cutoff;
   int nchg = 0;
   int64_t q;
   int rc;
   int j;

   SCIP_CALL( SCIPallocBufferArray(scip, &lb, propdata->nvars) );
   SCIP_CALL( SCIPallocBufferArray(scip, &ub, propdata->nvars) );
   SCIP_CALL( getEngineParams(scip, propdata, &params) );
   SCIP_CALL( linearGating(scip, prop, &params) );
   for( ;; )
   {
      for( j = 0; j < propdata->nvars; ++j )
      {
         lb[j] = SCIPvarGetLbLocal(propdata->vars[j]);
         ub[j] = SCIPvarGetUbLocal(propdata->vars[j]);
      }
      SCIP_CALL( SCIPallocBufferArray(scip, &hround, cap) );
      SCIP_CALL( SCIPallocBufferArray(scip, &hcol, cap) );
      SCIP_CALL( SCIPallocBufferArray(scip, &hbound, cap) );
      SCIP_CALL( SCIPallocBufferArray(scip, &hval, cap) );
      SCIP_CALL( SCIPallocBufferArray(scip, &hrow, cap) );
      SCIP_CALL( SCIPallocBufferArray(scip, &hrule, cap) );
      memset(&stats, 0, sizeof(stats));
      rc = sap_propagate_reasons(propdata->mat, &params, lb, ub, cap, &nhist, hround, hcol, hbound, hval, hrow,
         hrule, &conflict, &stats);
      if( rc != SAP_NOMEMORY || nhist <= cap )
         break;
      /* a longer history than the room given: again with room for all of it (same results) */
      cap = nhist;
      SCIPfreeBufferArray(scip, &hrule);
      SCIPfreeBufferArray(scip, &hrow);
      SCIPfreeBufferArray(scip, &hval);
      SCIPfreeBufferArray(scip, &hbound);
      SCIPfreeBufferArray(scip, &hcol);
      SCIPfreeBufferArray(scip, &hround);
   }
   SCIP_CALL( engineCall(rc, "sap_propagate_reasons") );
   ++propdata->ncalls;
   cutoff = FALSE;

   /* the changes in round order: each is implied by its row and the bounds applied before it */
   for( q = 0; q < nhist && !cutoff; ++q )
   {
      SCIP_VAR* var = propdata->vars[hcol[q]];
      SCIP_Bool infeasible = FALSE;
      SCIP_Bool tightened = FALSE;

      if( hbound[q] == 0 )
      {
         SCIP_CALL( SCIPinferVarLbProp(scip, var, hval[q], prop, 2 * hrow[q] + hrule[q], TRUE, &infeasible,
               &tightened) );
      }
      else
      {
         SCIP_CALL( SCIPinferVarUbProp(scip, var, hval[q], prop, 2 * hrow[q] + hrule[q], TRUE, &infeasible,
               &tightened) );
      }
      nchg += tightened ? 1 : 0;
      cutoff = infeasible;
   }
   propdata->nchgbds += nchg;

   if( !cutoff && stats.result == SAP_CUTOFF )
   {
      cutoff = TRUE;
      ++propdata->ncutoffs;
      /* conflict analysis on the certificate (cons_linear.c analyzeConflict) */
      if( (SCIPgetStage(scip) == SCIP_STAGE_SOLVING || SCIPinProbing(scip)) && SCIPisConflictAnalysisApplicable(scip)
         && conflict.kind != SAP_CONFLICT_NONE )
      {
         SCIP_CALL( SCIPinitConflictAnalysis(scip, SCIP_CONFTYPE_PROPAGATION, FALSE) );
         if( conflict.kind == SAP_CONFLICT_ROW && conflict.side == 2 )
         {
            /* the ranged-row test failed */
            SCIP_CALL( addRowAllBounds(scip, propdata, conflict.row, NULL) );
         }
         else if( conflict.kind == SAP_CONFLICT_ROW )
         {
            /* side 0: the minimal activity exceeds rhs; side 1: the maximal one misses lhs */
            SCIP_CALL( addRowBounds(scip, propdata, conflict.row, -1, conflict.side == 1, NULL) );
         }
         else
         {
            SCIP_VAR* var = propdata->vars[conflict.col];

            /* each side: the row that implied it, else the column's own bound */
            if( conflict.rowlb >= 0 )
            {
               SCIP_CALL( addReason(scip, propdata, conflict.rowlb, conflict.rules & 1, conflict.col,
                     SCIP_BOUNDTYPE_LOWER, NULL) );
            }
            else
            {
               SCIP_CALL( SCIPaddConflictLb(scip, var, NULL) );
            }
            if( conflict.rowub >= 0 )
            {
               SCIP_CALL( addReason(scip, propdata, conflict.rowub, (conflict.rules >> 1) & 1, conflict.col,
                     SCIP_BOUNDTYPE_UPPER, NULL) );
            }
            else
            {
               SCIP_CALL( SCIPaddConflictUb(scip, var, NULL) );
            }
         }
         SCIP_CALL( SCIPanalyzeConflict(scip, 0, NULL) );
         ++propdata->nconflicts;
      }
   }
   *result = cutoff ? SCIP_CUTOFF : (nchg > 0 ? SCIP_REDUCEDDOM : SCIP_DIDNOTFIND);

   SCIPfreeBufferArray(scip, &hrule);
   SCIPfreeBufferArray(scip, &hrow);
   SCIPfreeBufferArray(scip, &hval);
   SCIPfreeBufferArray(scip, &hbound);
   SCIPfreeBufferArray(scip, &hcol);
   SCIPfreeBufferArray(scip, &hround);
   SCIPfreeBufferArray(scip, &ub);
   SCIPfreeBufferArray(scip, &lb);
   return SCIP_OKAY;
}

/*
 * callback methods of propagator
 */

/** destructor of propagator to free user data (called when SCIP is exiting) */
static
SCIP_DECL_PROPFREE(propFreeGpulinear)
{  /*lint --e{715}*/
   SCIP_PROPDATA* propdata = SCIPpropGetData(prop);

   SCIP_CALL( freeMatrix(scip, propdata) );
   if( propdata->ctx != NULL )
   {
      SCIP_CALL( engineCall(sap_context_free(propdata->ctx), "sap_context_free") );
      propdata->ctx = NULL;
   }
   SCIPfreeBlockMemory(scip, &propdata);
   SCIPpropSetData(prop, NULL);
   return SCIP_OKAY;
}

/** creates the device context on first use */
static
SCIP_RETCODE ensureContext(
   SCIP_PROPDATA*        propdata            /**< propagator data */
   )
{
   if( propdata->ctx == NULL )
   {
      SCIP_CALL( engineCall(sap_context_create(propdata->device, &propdata->ctx), "sap_context_create") );
   }
   return SCIP_OKAY;
}

/** presolving initialization: matrix over the current (presolving) problem */
static
SCIP_DECL_PROPINITPRE(propInitpreGpulinear)
{  /*lint --e{715}*/
   SCIP_PROPDATA* propdata = SCIPpropGetData(prop);

   SCIP_CALL( ensureContext(propdata) );
   SCIP_CALL( buildMatrix(scip, propdata) );
   return SCIP_OKAY;
}

/** presolving deinitialization: the presolved problem gets a new matrix in INITSOL */
static
SCIP_DECL_PROPEXITPRE(propExitpreGpulinear)
{  /*lint --e{715}*/
   SCIP_CALL( freeMatrix(scip, SCIPpropGetData(prop)) );
   return SCIP_OKAY;
}

/** solving process initialization: matrix over the presolved problem */
static
SCIP_DECL_PROPINITSOL(propInitsolGpulinear)
{  /*lint --e{715}*/
   SCIP_PROPDATA* propdata = SCIPpropGetData(prop);

   SCIP_CALL( ensureContext(propdata) );
   SCIP_CALL( buildMatrix(scip, propdata) );
   return SCIP_OKAY;
}

/** solving process deinitialization */
static
SCIP_DECL_PROPEXITSOL(propExitsolGpulinear)
{  /*lint --e{715}*/
   SCIP_CALL( freeMatrix(scip, SCIPpropGetData(prop)) );
   return SCIP_OKAY;
}

/** presolving: global bounds to the device fixpoint (the bound tightening of consPresolLinear) */
static
SCIP_DECL_PROPPRESOL(propPresolGpulinear)
{  /*lint --e{715}*/
   SCIP_PROPDATA* propdata = SCIPpropGetData(prop);
   double* lb;
   double* ub;
   SCIP_Bool cutoff;
   int nchg = 0;
   int j;

   *result = SCIP_DIDNOTRUN;
   if( propdata->mat == NULL || propdata->nvars != SCIPgetNVars(scip) )
   {
      /* presolving changed the variable set: rebuild over the current problem */
      SCIP_CALL( ensureContext(propdata) );
      SCIP_CALL( buildMatrix(scip, propdata) );
      if( propdata->mat == NULL )
         return SCIP_OKAY;
   }

   SCIP_CALL( SCIPallocBufferArray(scip, &lb, propdata->nvars) );
   SCIP_CALL( SCIPallocBufferArray(scip, &ub, propdata->nvars) );
   SCIP_CALL( propagateOnDevice(scip, prop, propdata, FALSE, lb, ub, &cutoff) );
   if( cutoff )
      *result = SCIP_CUTOFF;
   else
   {
      for( j = 0; j < propdata->nvars && !cutoff; ++j )
      {
         SCIP_VAR* var = propdata->vars[j];
         SCIP_Bool infeasible = FALSE;
         SCIP_Bool tightened = FALSE;

         if( lb[j] > SCIPvarGetLbGlobal(var) )
         {
            SCIP_CALL( SCIPtightenVarLb(scip, var, lb[j], TRUE, &infeasible, &tightened) );
            nchg += tightened ? 1 : 0;
            cutoff = cutoff || infeasible;
         }
         if( !cutoff && ub[j] < SCIPvarGetUbGlobal(var) )
         {
            SCIP_CALL( SCIPtightenVarUb(scip, var, ub[j], TRUE, &infeasible, &tightened) );
            nchg += tightened ? 1 : 0;
            cutoff = cutoff || infeasible;
         }
      }
      *nchgbds += nchg;
      propdata->nchgbds += nchg;
      if( !cutoff && propdata->probing )
      {
         int oldfixed = *nfixedvars;
         int oldaggr = *naggrvars;
         int oldchg = *nchgbds;

         SCIP_CALL( probeOnDevice(scip, propdata, lb, ub, nfixedvars, naggrvars, nchgbds, &cutoff) );
         nchg += (*nfixedvars - oldfixed) + (*naggrvars - oldaggr) + (*nchgbds - oldchg);
      }
      *result = cutoff ? SCIP_CUTOFF : (nchg > 0 ? SCIP_SUCCESS : SCIP_DIDNOTFIND);
   }
   SCIPfreeBufferArray(scip, &ub);
   SCIPfreeBufferArray(scip, &lb);
   return SCIP_OKAY;
}

/** execution method: local bounds of the node to the device fixpoint */
static
SCIP_DECL_PROPEXEC(propExecGpulinear)
{  /*lint --e{715}*/
   SCIP_PROPDATA* propdata = SCIPpropGetData(prop);
   double* lb;
   double* ub;
   SCIP_Bool cutoff;
   int nchg = 0;
   int j;

   *result = SCIP_DIDNOTRUN;
   if( propdata->mat == NULL )
      return SCIP_OKAY;
   if( propdata->reasons )
      return execWithReasons(scip, prop, propdata, result);

   SCIP_CALL( SCIPallocBufferArray(scip, &lb, propdata->nvars) );
   SCIP_CALL( SCIPallocBufferArray(scip, &ub, propdata->nvars) );
   SCIP_CALL( propagateOnDevice(scip, prop, propdata, TRUE, lb, ub, &cutoff) );
   if( cutoff )
      *result = SCIP_CUTOFF;
   else
   {
      for( j = 0; j < propdata->nvars && !cutoff; ++j )
      {
         SCIP_VAR* var = propdata->vars[j];
         SCIP_Bool infeasible = FALSE;
         SCIP_Bool tightened = FALSE;

         /* the device accepted these changes under SCIP's rules (including forced changes of
          * single-variable rows), so they are applied as forced */
         if( lb[j] > SCIPvarGetLbLocal(var) )
         {
            SCIP_CALL( SCIPinferVarLbProp(scip, var, lb[j], prop, 0, TRUE, &infeasible, &tightened) );
            nchg += tightened ? 1 : 0;
            cutoff = cutoff || infeasible;
         }
         if( !cutoff && ub[j] < SCIPvarGetUbLocal(var) )
         {
            SCIP_CALL( SCIPinferVarUbProp(scip, var, ub[j], prop, 0, TRUE, &infeasible, &tightened) );
            nchg += tightened ? 1 : 0;
            cutoff = cutoff || infeasible;
         }
      }
      propdata->nchgbds += nchg;
      *result = cutoff ? SCIP_CUTOFF : (nchg > 0 ? SCIP_REDUCEDDOM : SCIP_DIDNOTFIND);
   }
   SCIPfreeBufferArray(scip, &ub);
   SCIPfreeBufferArray(scip, &lb);
   return SCIP_OKAY;
}

/** conflict resolution (prop.c:739-788): inferinfo is the reason row * 2 + rule of the change
 *  (cons_linear.c keeps the position and PROPRULE there, :344-369).  Rule 0: the bounds making
 *  up the row's residual activity at the change's time explain it (resolvePropagation /
 *  addConflictBounds); rule 1 (ranged-row test): the bounds of all the row's columns.  Without
 *  a recorded row (reasons off, or a change from before the current matrix) the change is
 *  reported as unexplained, which conflict analysis accepts (the change itself stays in the
 *  conflict). */
static
SCIP_DECL_PROPRESPROP(propRespropGpulinear)
{  /*lint --e{715}*/
   SCIP_PROPDATA* propdata = SCIPpropGetData(prop);
   int j = SCIPvarGetProbindex(infervar);
   int row = inferinfo >> 1;

   *result = SCIP_DIDNOTFIND;
   if( !propdata->reasons || propdata->rowptr == NULL || inferinfo < 0 || row >= propdata->nrows || j < 0
      || j >= propdata->nvars || propdata->vars[j] != infervar || rowCoef(propdata, row, j) == 0.0 )
      return SCIP_OKAY;
   SCIP_CALL( addReason(scip, propdata, row, inferinfo & 1, j, boundtype, bdchgidx) );
   *result = SCIP_SUCCESS;
   return SCIP_OKAY;
}

/*
 * propagator specific interface methods
 */

/** the propagation part of strong branching for many candidates at once on the device
 *  (SCIPgetVarStrongbranchWithPropagation, scip_var.c:4139-4440, without the LP solve): the
 *  children x <= feasCeil(solval - 1) and x >= feasFloor(solval + 1) of every candidate are
 *  propagated from the node's local bounds in one batched device call (sap_probe_split), and
 *  the deductions a strong-branching rule draws from them (branch_relpscost.c:1992-2070) are
 *  collected against the node:
 *  - a child whose branching bound leaves the local domain, or whose propagation ends in a
 *    cutoff, is infeasible (downinf / upinf);
 *  - both children infeasible: the node is infeasible (*cutoff);
 *  - one child infeasible: the candidate's bound of the other child holds at the node;
 *  - both feasible: the union of the children's bounds holds at the node (SCIP's newlbs /
 *    newubs), which differs from the node's bounds only on columns changed in both children.
 *  newlbs / newubs (indexed by SCIPvarGetProbindex, or NULL) receive the node's local bounds
 *  tightened by all of these; the caller applies them as a branching rule does. */
SCIP_RETCODE SCIPgpulinearStrongbranch(
   SCIP*                 scip,               /**< SCIP data structure */
   int                   ncands,             /**< number of candidates */
   SCIP_VAR**            cands,              /**< integral candidate variables (columns of the matrix) */
   const SCIP_Real*      solvals,            /**< LP values of the candidates */
   SCIP_Bool*            downinf,            /**< array of ncands: the down child is infeasible */
   SCIP_Bool*            upinf,              /**< array of ncands: the up child is infeasible */
   SCIP_Real*            newlbs,             /**< array of the matrix's columns: valid lower bounds, or NULL */
   SCIP_Real*            newubs,             /**< array of the matrix's columns: valid upper bounds, or NULL */
   SCIP_Bool*            cutoff              /**< pointer to store whether the node is infeasible */
   )
{
   SCIP_PROP* prop = SCIPfindProp(scip, PROP_NAME);
   SCIP_PROPDATA* propdata;
   sap_params params;
   sap_stats stats;
   double* lb;
   double* ub;
   int32_t* pcol;
   int32_t* pidx;
   double* downub;
   double* uplb;
   int32_t* verdict;
   int64_t* dedptr;
   int32_t* dedcol = NULL;
   double* ded[4] = { NULL, NULL, NULL, NULL };
   int64_t cap;
   int np = 0;
   int rc;
   int c;
   int j;
   int k;

   if( prop == NULL || ncands < 0 || (ncands > 0 && (cands == NULL || solvals == NULL || downinf == NULL
            || upinf == NULL)) || cutoff == NULL || ((newlbs == NULL) != (newubs == NULL)) )
      return SCIP_INVALIDDATA;
   propdata = SCIPpropGetData(prop);
   if( propdata->mat == NULL )
      return SCIP_INVALIDCALL;   /* no device matrix (before INITSOL/presolving) */
   *cutoff = FALSE;
   SCIP_CALL( SCIPallocBufferArray(scip, &lb, propdata->nvars) );
   SCIP_CALL( SCIPallocBufferArray(scip, &ub, propdata->nvars) );
   for( j = 0; j < propdata->nvars; ++j )
   {
      lb[j] = SCIPvarGetLbLocal(propdata->vars[j]);
      ub[j] = SCIPvarGetUbLocal(propdata->vars[j]);
   }
   if( newlbs != NULL )
   {
      memcpy(newlbs, lb, sizeof(SCIP_Real) * (size_t)propdata->nvars);
      memcpy(newubs, ub, sizeof(SCIP_Real) * (size_t)propdata->nvars);
   }
   SCIP_CALL( SCIPallocBufferArray(scip, &pcol, ncands + 1) );
   SCIP_CALL( SCIPallocBufferArray(scip, &pidx, ncands + 1) );
   SCIP_CALL( SCIPallocBufferArray(scip, &downub, ncands + 1) );
   SCIP_CALL( SCIPallocBufferArray(scip, &uplb, ncands + 1) );
   for( c = 0; c < ncands; ++c )
   {
      int col = SCIPvarGetProbindex(cands[c]);
      double dub = SCIPfeasCeil(scip, solvals[c] - 1.0);
      double ulb = SCIPfeasFloor(scip, solvals[c] + 1.0);

      if( col < 0 || col >= propdata->nvars || propdata->vars[col] != cands[c] || !SCIPvarIsIntegral(cands[c]) )
      {
         SCIPfreeBufferArray(scip, &uplb);
         SCIPfreeBufferArray(scip, &downub);
         SCIPfreeBufferArray(scip, &pidx);
         SCIPfreeBufferArray(scip, &pcol);
         SCIPfreeBufferArray(scip, &ub);
         SCIPfreeBufferArray(scip, &lb);
         return SCIP_INVALIDDATA;
      }
      /* a branching bound outside the local domain makes that child infeasible without any
       * propagation (scip_var.c:4266-4313) */
      upinf[c] = ulb > ub[col] + 0.5;
      downinf[c] = dub < lb[col] - 0.5;
      if( !upinf[c] && !downinf[c] )
      {
         pcol[np] = col;
         pidx[np] = c;
         downub[np] = dub;
         uplb[np] = ulb;
         ++np;
      }
   }
   if( np > 0 )
   {
      if( propdata->probebase == NULL )
      {
         SCIP_CALL( engineCall(sap_bounds_create(propdata->mat, 1, &propdata->probebase), "sap_bounds_create") );
      }
      if( propdata->probeset == NULL )
      {
         SCIP_CALL( engineCall(sap_bounds_create(propdata->mat, propdata->probingvectors, &propdata->probeset),
               "sap_bounds_create") );
      }
      SCIP_CALL( engineCall(sap_bounds_upload(propdata->probebase, 0, lb, ub), "sap_bounds_upload") );
      SCIP_CALL( getEngineParams(scip, propdata, &params) );   /* node propagation: every row in round 1 */
      SCIP_CALL( SCIPallocBufferArray(scip, &verdict, np) );
      SCIP_CALL( SCIPallocBufferArray(scip, &dedptr, np + 1) );
      cap = 16 * (int64_t)np + 64;
      for( ;; )
      {
         SCIP_CALL( SCIPallocBufferArray(scip, &dedcol, cap) );
         for( k = 0; k < 4; ++k )
         {
            SCIP_CALL( SCIPallocBufferArray(scip, &ded[k], cap) );
         }
         memset(&stats, 0, sizeof(stats));
         rc = sap_probe_split(propdata->probeset, &params, np, pcol, downub, uplb, propdata->probebase, 0, verdict,
            NULL, cap, dedptr, dedcol, ded[0], ded[1], ded[2], ded[3], &stats);
         if( rc != SAP_NOMEMORY || dedptr[np] <= cap )
            break;
         cap = dedptr[np];   /* too little room for the pair lists: again with enough (same results) */
         for( k = 3; k >= 0; --k )
            SCIPfreeBufferArray(scip, &ded[k]);
         SCIPfreeBufferArray(scip, &dedcol);
      }
      if( rc == SAP_OKAY )
      {
         for( k = 0; k < np; ++k )
         {
            c = pidx[k];
            downinf[c] = (verdict[k] & SAP_PROBE_DOWN_INFEASIBLE) != 0;
            upinf[c] = (verdict[k] & SAP_PROBE_UP_INFEASIBLE) != 0;
            if( newlbs != NULL && !downinf[c] && !upinf[c] )
            {
               int64_t q;
               /* columns changed in both children: the union of their bounds */
               for( q = dedptr[k]; q < dedptr[k + 1]; ++q )
               {
                  int v = dedcol[q];
                  double ulbv = ded[0][q] < ded[2][q] ? ded[0][q] : ded[2][q];
                  double uubv = ded[1][q] > ded[3][q] ? ded[1][q] : ded[3][q];
                  newlbs[v] = ulbv > newlbs[v] ? ulbv : newlbs[v];
                  newubs[v] = uubv < newubs[v] ? uubv : newubs[v];
               }
            }
         }
      }
      for( k = 3; k >= 0; --k )
         SCIPfreeBufferArray(scip, &ded[k]);
      SCIPfreeBufferArray(scip, &dedcol);
      SCIPfreeBufferArray(scip, &dedptr);
      SCIPfreeBufferArray(scip, &verdict);
      if( rc != SAP_OKAY )
      {
         SCIPfreeBufferArray(scip, &uplb);
         SCIPfreeBufferArray(scip, &downub);
         SCIPfreeBufferArray(scip, &pidx);
         SCIPfreeBufferArray(scip, &pcol);
         SCIPfreeBufferArray(scip, &ub);
         SCIPfreeBufferArray(scip, &lb);
         return engineCall(rc, "sap_probe_split");
      }
   }
   /* the branching bound of the feasible child of a candidate with one infeasible child */
   for( c = 0; c < ncands && !*cutoff; ++c )
   {
      int col = SCIPvarGetProbindex(cands[c]);
      if( downinf[c] && upinf[c] )
         *cutoff = TRUE;
      else if( newlbs != NULL && downinf[c] )
      {
         double ulb = SCIPfeasFloor(scip, solvals[c] + 1.0);
         newlbs[col] = ulb > newlbs[col] ? ulb : newlbs[col];
      }
      else if( newlbs != NULL && upinf[c] )
      {
         double dub = SCIPfeasCeil(scip, solvals[c] - 1.0);
         newubs[col] = dub < newubs[col] ? dub : newubs[col];
      }
   }
   /* deductions of different candidates may contradict each other only if the node is infeasible */
   for( j = 0; newlbs != NULL && j < propdata->nvars && !*cutoff; ++j )
      if( newlbs[j] > newubs[j] + SCIPfeastol(scip) )
         *cutoff = TRUE;
   SCIPfreeBufferArray(scip, &uplb);
   SCIPfreeBufferArray(scip, &downub);
   SCIPfreeBufferArray(scip, &pidx);
   SCIPfreeBufferArray(scip, &pcol);
   SCIPfreeBufferArray(scip, &ub);
   SCIPfreeBufferArray(scip, &lb);
   return SCIP_OKAY;
}

/** batched propagation dives (the fix-and-propagate loop of heur_fixandinfer.c:255-280 and of
 *  SCIP's diving heuristics, SCIPfixVarProbing / SCIPpropagateProbing per step), many dives at
 *  once on the device.  Dive k visits the candidates in its order order[k*ncands ...]; at each
 *  step it fixes the first candidate still unfixed in its current bounds to fixvals[k*ncands +
 *  c] (rounded to the nearest integer for an integral variable, then moved into the current
 *  domain) and propagates; it ends on a cutoff, when no candidate is left unfixed or after
 *  maxdepth fixings.  Every step of all live dives is one batched device call: each dive is a
 *  node vector started from the node's propagated bounds with all of its fixings so far
 *  (sap_propagate_nodes).  Outputs per dive: depth[k] fixings made, cutoff[k], and with
 *  divelbs / diveubs (ndives x the matrix's columns, indexed by SCIPvarGetProbindex) its
 *  final bounds. */
SCIP_RETCODE SCIPgpulinearDive(
   SCIP*                 scip,               /**< SCIP data structure */
   int                   ndives,             /**< number of dives */
   int                   ncands,             /**< number of candidates */
   SCIP_VAR**            cands,              /**< candidate variables (columns of the matrix) */
   const int*            order,              /**< ndives x ncands: the order in which dive k visits them */
   const SCIP_Real*      fixvals,            /**< ndives x ncands: the value dive k fixes candidate c to */
   int                   maxdepth,           /**< most fixings per dive (-1: no limit) */
   int*                  depth,              /**< array of ndives: fixings made */
   SCIP_Bool*            cutoff,             /**< array of ndives: the dive ended in a cutoff */
   SCIP_Real*            divelbs,            /**< ndives x columns: final lower bounds, or NULL */
   SCIP_Real*            diveubs             /**< ndives x columns: final upper bounds, or NULL */
   )
{
   SCIP_PROP* prop = SCIPfindProp(scip, PROP_NAME);
   SCIP_PROPDATA* propdata;
   sap_params params;
   sap_stats stats;
   double* lb0;
   double* ub0;
   double* cur;        /* per dive: current lower bounds, then upper bounds (2 x nvars) */
   int* pos;           /* per dive: next index into its order */
   int* fixcol;        /* per dive: its fixings so far (ncands slots each) */
   double* fixval;
   SCIP_Bool* live;
   int* vmap;          /* vector of the batched call -> dive */
   int32_t* result;
   int64_t* ovptr;
   int32_t* ovcol;
   double* ovval;
   int64_t* chgptr;
   int32_t* chgcol;
   double* chglb;
   double* chgub;
   int64_t cap;
   int nvars;
   int nlive;
   int rc = SAP_OKAY;
   int j;
   int k;
   int q;

   if( prop == NULL || ndives < 0 || ncands < 0 || (ndives > 0 && ncands > 0 && (cands == NULL || order == NULL
            || fixvals == NULL)) || (ndives > 0 && (depth == NULL || cutoff == NULL))
      || ((divelbs == NULL) != (diveubs == NULL)) )
      return SCIP_INVALIDDATA;
   propdata = SCIPpropGetData(prop);
   if( propdata->mat == NULL )
      return SCIP_INVALIDCALL;
   nvars = propdata->nvars;
   for( k = 0; k < ndives * ncands; ++k )
      if( order[k] < 0 || order[k] >= ncands )
         return SCIP_INVALIDDATA;
   for( k = 0; k < ncands; ++k )
   {
      int col = SCIPvarGetProbindex(cands[k]);
      if( col < 0 || col >= nvars || propdata->vars[col] != cands[k] )
         return SCIP_INVALIDDATA;
   }
   for( k = 0; k < ndives; ++k )
   {
      depth[k] = 0;
      cutoff[k] = FALSE;
   }
   if( ndives == 0 )
      return SCIP_OKAY;

   SCIP_CALL( SCIPallocBufferArray(scip, &lb0, nvars) );
   SCIP_CALL( SCIPallocBufferArray(scip, &ub0, nvars) );
   for( j = 0; j < nvars; ++j )
   {
      lb0[j] = SCIPvarGetLbLocal(propdata->vars[j]);
      ub0[j] = SCIPvarGetUbLocal(propdata->vars[j]);
   }
   SCIP_CALL( getEngineParams(scip, propdata, &params) );
   if( propdata->probebase == NULL )
   {
      SCIP_CALL( engineCall(sap_bounds_create(propdata->mat, 1, &propdata->probebase), "sap_bounds_create") );
   }
   /* the node's bounds, propagated to their fixpoint: the base of every dive */
   SCIP_CALL( engineCall(sap_bounds_upload(propdata->probebase, 0, lb0, ub0), "sap_bounds_upload") );
   memset(&stats, 0, sizeof(stats));
   SCIP_CALL( engineCall(sap_propagate_resident(propdata->probebase, &params, &stats), "sap_propagate_resident") );
   if( stats.result == SAP_CUTOFF )
   {
      for( k = 0; k < ndives; ++k )
         cutoff[k] = TRUE;
      SCIPfreeBufferArray(scip, &ub0);
      SCIPfreeBufferArray(scip, &lb0);
      return SCIP_OKAY;
   }
   SCIP_CALL( engineCall(sap_bounds_download(propdata->probebase, 0, lb0, ub0), "sap_bounds_download") );
   if( propdata->diveset != NULL && propdata->divevecs < ndives )
   {
      SCIP_CALL( engineCall(sap_bounds_free(propdata->diveset), "sap_bounds_free") );
      propdata->diveset = NULL;
   }
   if( propdata->diveset == NULL )
   {
      SCIP_CALL( engineCall(sap_bounds_create(propdata->mat, ndives, &propdata->diveset), "sap_bounds_create") );
      propdata->divevecs = ndives;
   }

   SCIP_CALL( SCIPallocBufferArray(scip, &cur, 2 * (size_t)ndives * (size_t)nvars) );
   SCIP_CALL( SCIPallocBufferArray(scip, &pos, ndives) );
   SCIP_CALL( SCIPallocBufferArray(scip, &fixcol, (size_t)ndives * (size_t)(ncands > 0 ? ncands : 1)) );
   SCIP_CALL( SCIPallocBufferArray(scip, &fixval, (size_t)ndives * (size_t)(ncands > 0 ? ncands : 1)) );
   SCIP_CALL( SCIPallocBufferArray(scip, &live, ndives) );
   SCIP_CALL( SCIPallocBufferArray(scip, &vmap, ndives) );
   SCIP_CALL( SCIPallocBufferArray(scip, &result, ndives) );
   SCIP_CALL( SCIPallocBufferArray(scip, &ovptr, ndives + 1) );
   SCIP_CALL( SCIPallocBufferArray(scip, &ovcol, (size_t)ndives * (size_t)(ncands > 0 ? ncands : 1)) );
   SCIP_CALL( SCIPallocBufferArray(scip, &ovval, (size_t)ndives * (size_t)(ncands > 0 ? ncands : 1)) );
   SCIP_CALL( SCIPallocBufferArray(scip, &chgptr, ndives + 1) );
   cap = 64 * (int64_t)ndives + 1024;
   SCIP_CALL( SCIPallocBufferArray(scip, &chgcol, cap) );
   SCIP_CALL( SCIPallocBufferArray(scip, &chglb, cap) );
   SCIP_CALL( SCIPallocBufferArray(scip, &chgub, cap) );
   for( k = 0; k < ndives; ++k )
   {
      memcpy(cur + 2 * (size_t)k * (size_t)nvars, lb0, sizeof(double) * (size_t)nvars);
      memcpy(cur + (2 * (size_t)k + 1) * (size_t)nvars, ub0, sizeof(double) * (size_t)nvars);
      pos[k] = 0;
      live[k] = TRUE;
   }
   params.flags |= SAP_FLAG_BASE_IS_FIXPOINT;

   for( ;; )
   {
      /* every live dive fixes its next unfixed candidate; its node is all of its fixings */
      nlive = 0;
      ovptr[0] = 0;
      for( k = 0; k < ndives; ++k )
      {
         const double* klb = cur + 2 * (size_t)k * (size_t)nvars;
         const double* kub = klb + nvars;
         int c = -1;

         if( !live[k] )
            continue;
         if( maxdepth >= 0 && depth[k] >= maxdepth )
         {
            live[k] = FALSE;
            continue;
         }
         while( pos[k] < ncands && c < 0 )
         {
            int cand = order[(size_t)k * (size_t)ncands + (size_t)pos[k]];
            int col = SCIPvarGetProbindex(cands[cand]);
            ++pos[k];
            if( klb[col] < kub[col] )
               c = cand;
         }
         if( c < 0 )
         {
            live[k] = FALSE;
            continue;
         }
         {
            int col = SCIPvarGetProbindex(cands[c]);
            double v = fixvals[(size_t)k * (size_t)ncands + (size_t)c];
            if( SCIPvarIsIntegral(cands[c]) )
               v = floor(v + 0.5);
            v = v < klb[col] ? klb[col] : (v > kub[col] ? kub[col] : v);
            fixcol[(size_t)k * (size_t)ncands + (size_t)depth[k]] = col;
            fixval[(size_t)k * (size_t)ncands + (size_t)depth[k]] = v;
            ++depth[k];
         }
         for( q = 0; q < depth[k]; ++q )
         {
            ovcol[ovptr[nlive] + q] = (int32_t)fixcol[(size_t)k * (size_t)ncands + (size_t)q];
            ovval[ovptr[nlive] + q] = fixval[(size_t)k * (size_t)ncands + (size_t)q];
         }
         ovptr[nlive + 1] = ovptr[nlive] + depth[k];
         vmap[nlive] = k;
         ++nlive;
      }
      if( nlive == 0 )
         break;
      for( ;; )
      {
         memset(&stats, 0, sizeof(stats));
         rc = sap_propagate_nodes(propdata->diveset, &params, nlive, propdata->probebase, 0, ovptr, ovcol, ovval,
            ovval, result, NULL, cap, chgptr, chgcol, chglb, chgub, &stats);
         if( rc != SAP_NOMEMORY || chgptr[nlive] <= cap )
            break;
         cap = chgptr[nlive];   /* room for every change: again (same results) */
         SCIPfreeBufferArray(scip, &chgub);
         SCIPfreeBufferArray(scip, &chglb);
         SCIPfreeBufferArray(scip, &chgcol);
         SCIP_CALL( SCIPallocBufferArray(scip, &chgcol, cap) );
         SCIP_CALL( SCIPallocBufferArray(scip, &chglb, cap) );
         SCIP_CALL( SCIPallocBufferArray(scip, &chgub, cap) );
      }
      if( rc != SAP_OKAY )
         break;
      for( j = 0; j < nlive; ++j )
      {
         double* klb;
         double* kub;
         int64_t e;
         k = vmap[j];
         if( result[j] == SAP_CUTOFF )
         {
            cutoff[k] = TRUE;
            live[k] = FALSE;
            continue;
         }
         /* the dive's bounds: the base, its fixings, then the propagated changes */
         klb = cur + 2 * (size_t)k * (size_t)nvars;
         kub = klb + nvars;
         memcpy(klb, lb0, sizeof(double) * (size_t)nvars);
         memcpy(kub, ub0, sizeof(double) * (size_t)nvars);
         for( e = ovptr[j]; e < ovptr[j + 1]; ++e )
            klb[ovcol[e]] = kub[ovcol[e]] = ovval[e];
         for( e = chgptr[j]; e < chgptr[j + 1]; ++e )
         {
            klb[chgcol[e]] = chglb[e];
            kub[chgcol[e]] = chgub[e];
         }
      }
   }
   if( rc == SAP_OKAY && divelbs != NULL )
      for( k = 0; k < ndives; ++k )
      {
         memcpy(divelbs + (size_t)k * (size_t)nvars, cur + 2 * (size_t)k * (size_t)nvars,
            sizeof(double) * (size_t)nvars);
         memcpy(diveubs + (size_t)k * (size_t)nvars, cur + (2 * (size_t)k + 1) * (size_t)nvars,
            sizeof(double) * (size_t)nvars);
      }
   SCIPfreeBufferArray(scip, &chgub);
   SCIPfreeBufferArray(scip, &chglb);
   SCIPfreeBufferArray(scip, &chgcol);
   SCIPfreeBufferArray(scip, &chgptr);
   SCIPfreeBufferArray(scip, &ovval);
   SCIPfreeBufferArray(scip, &ovcol);
   SCIPfreeBufferArray(scip, &ovptr);
   SCIPfreeBufferArray(scip, &result);
   SCIPfreeBufferArray(scip, &vmap);
   SCIPfreeBufferArray(scip, &live);
   SCIPfreeBufferArray(scip, &fixval);
   SCIPfreeBufferArray(scip, &fixcol);
   SCIPfreeBufferArray(scip, &pos);
   SCIPfreeBufferArray(scip, &cur);
   SCIPfreeBufferArray(scip, &ub0);
   SCIPfreeBufferArray(scip, &lb0);
   return engineCall(rc, "sap_propagate_nodes");
}

/** creates the gpulinear propagator and includes it in SCIP */
SCIP_RETCODE SCIPincludePropGpulinear(
   SCIP*                 scip                /**< SCIP data structure */
   )
{
   SCIP_PROPDATA* propdata;
   SCIP_PROP* prop;

   SCIP_CALL( SCIPallocBlockMemory(scip, &propdata) );
   memset(propdata, 0, sizeof(*propdata));

   SCIP_CALL( SCIPincludePropBasic(scip, &prop, PROP_NAME, PROP_DESC, PROP_PRIORITY, PROP_FREQ, PROP_DELAY,
         PROP_TIMING, propExecGpulinear, propdata) );
   SCIP_CALL( SCIPsetPropFree(scip, prop, propFreeGpulinear) );
   SCIP_CALL( SCIPsetPropInitpre(scip, prop, propInitpreGpulinear) );
   SCIP_CALL( SCIPsetPropExitpre(scip, prop, propExitpreGpulinear) );
   SCIP_CALL( SCIPsetPropInitsol(scip, prop, propInitsolGpulinear) );
   SCIP_CALL( SCIPsetPropExitsol(scip, prop, propExitsolGpulinear) );
   SCIP_CALL( SCIPsetPropPresol(scip, prop, propPresolGpulinear, PROP_PRESOL_PRIORITY, PROP_PRESOL_MAXROUNDS,
         PROP_PRESOLTIMING) );
   SCIP_CALL( SCIPsetPropResprop(scip, prop, propRespropGpulinear) );

   SCIP_CALL( SCIPaddIntParam(scip, "propagating/" PROP_NAME "/device", "HIP device ordinal of the engine",
         &propdata->device, TRUE, DEFAULT_DEVICE, 0, 1024, NULL, NULL) );
   SCIP_CALL( SCIPaddIntParam(scip, "propagating/" PROP_NAME "/maxrounds",
         "maximal number of device propagation rounds per call (-1: until fixpoint)",
         &propdata->maxrounds, FALSE, DEFAULT_MAXROUNDS, -1, INT_MAX, NULL, NULL) );
   SCIP_CALL( SCIPaddBoolParam(scip, "propagating/" PROP_NAME "/probing",
         "probe the unfixed binaries on the device in presolving", &propdata->probing, FALSE, DEFAULT_PROBING,
         NULL, NULL) );
   SCIP_CALL( SCIPaddBoolParam(scip, "propagating/" PROP_NAME "/reasons",
         "record a reason row for every device bound change (conflict analysis)", &propdata->reasons, FALSE,
         DEFAULT_REASONS, NULL, NULL) );
   SCIP_CALL( SCIPaddIntParam(scip, "propagating/" PROP_NAME "/probingvectors",
         "node vectors of one probing launch (two per probed binary)", &propdata->probingvectors, TRUE,
         DEFAULT_PROBINGVECTORS, 2, 1 << 20, NULL, NULL) );

   return SCIP_OKAY;
}
