/* testhost.c -- test double of the SCIP API subset used by prop_gpulinear.c.
 *
 * A problem is a set of variables with global and local bounds and a set of linear rows;
 * the propagator registered through SCIPincludePropBasic is driven by the th_* entry points
 * (the order of SCIP's solving stages: include, INITPRE/PRESOL/EXITPRE, INITSOL, PROPEXEC,
 * EXITSOL, FREE).  Bound changes go through the acceptance logic of SCIPinferVarLbCons /
 * SCIPinferVarUbCons (src/scip/scip_var.c:6965-7174, var.c:1957-2021, set.c:7711-7753), so a
 * change the plugin hands over is accepted or rejected as SCIP would.
 *
 * Probing deductions handed to SCIPanalyzeDeductionsProbing and fixings made with SCIPfixVar
 * are recorded for the tests (th_get_probe_records).  Local bound changes made through
 * SCIPinferVarLbProp / SCIPinferVarUbProp are kept in order with their inferinfo; a bound
 * change index is a position in that list, and SCIPaddConflictLb / SCIPaddConflictUb record
 * the bound in force before that position (the bound "at the time" of SCIP's bdchgidx), so a
 * test can check that a conflict set explains a change.
 *
 * Test infrastructure: built into scip_amd/plugin/libgpulinear_testhost.so for tests/ only.
 */
#include <math.h>
#include <string.h>

#include "scip/scip.h"
#include "scip/cons_linear.h"
#include "scip/prop_probing.h"
#include "prop_gpulinear.h"

struct SCIP_Var
{
   int probindex;
   int integral;
   double glb, gub, llb, lub;
   double llb0, lub0;   /* local bounds at the start of the node (before the recorded changes) */
};

struct SCIP_BdChgIdx
{
   int pos;             /* position in the change list */
};

struct SCIP_Cons
{
   int nvars;
   SCIP_VAR** vars;
   double* vals;
   double lhs, rhs;
};

struct SCIP_Conshdlr
{
   int nconss;
   int consssize;
   SCIP_CONS** conss;
};

struct SCIP_Prop
{
   SCIP_PROPDATA* data;
   SCIP_DECL_PROPEXEC((*exec));
   SCIP_DECL_PROPFREE((*pfree));
   SCIP_DECL_PROPINITPRE((*initpre));
   SCIP_DECL_PROPEXITPRE((*exitpre));
   SCIP_DECL_PROPINITSOL((*initsol));
   SCIP_DECL_PROPEXITSOL((*exitsol));
   SCIP_DECL_PROPPRESOL((*presol));
   SCIP_DECL_PROPRESPROP((*resprop));
   int freq;
   char name[64];
};

#define TH_MAXPARAMS 16

struct Scip
{
   int nvars;
   int varssize;
   SCIP_VAR** vars;
   SCIP_CONSHDLR linear;
   SCIP_PROP* prop;
   double infinity, epsilon, sumepsilon, feastol, hugeval, boundstreps, recompfac, maxeasy;
   int nintparams;
   char intname[TH_MAXPARAMS][96];
   int* intptr[TH_MAXPARAMS];
   /* records: SCIPfixVar calls (var, value) and SCIPanalyzeDeductionsProbing entries (probing
    * var, var, bounds in the left / right branch) */
   int nfix, fixcap;
   struct FixRec { int var; double val; }* fix;
   int nded, dedcap;
   struct DedRec { int probe, var; double bnd[4]; }* ded;
   int naggr, nimpl;
   /* local bound changes of the node, in order, and the conflict candidate set */
   int nchg, chgcap;
   struct ChgRec { int var, bound, inferinfo; double oldval, newval; }* chg;
   int nconf, confcap;
   struct ConfRec { int var, bound; double val; }* conf;
   int nanalyzed;
   int stage;
   int rangedrowprop;   /* constraints/linear/rangedrowpropagation (SCIP's default TRUE) */
   /* constraints/linear/{tightenboundsfreq,rangedrowfreq,rangedrowmaxdepth} (cons_linear.c's
    * defaults 1, 1, INT_MAX) */
   int tightenboundsfreq, rangedrowfreq, rangedrowmaxdepth;
   /* the node: in probing, depth, in repropagation */
   int inprobing, depth, inreprop;
};

/* ---- numerics (def.h, set.c) ---- */
SCIP_Real SCIPinfinity(SCIP* scip) { return scip->infinity; }
SCIP_Real SCIPepsilon(SCIP* scip) { return scip->epsilon; }
SCIP_Real SCIPsumepsilon(SCIP* scip) { return scip->sumepsilon; }
SCIP_Real SCIPfeastol(SCIP* scip) { return scip->feastol; }
SCIP_Real SCIPgetHugeValue(SCIP* scip) { return scip->hugeval; }
SCIP_Bool SCIPisInfinity(SCIP* scip, SCIP_Real val) { return val >= scip->infinity; }
/* set.c SCIPsetFeasFloor / SCIPsetFeasCeil: floor(x + feastol), ceil(x - feastol) */
SCIP_Real SCIPfeasFloor(SCIP* scip, SCIP_Real val) { return floor(val + scip->feastol); }
SCIP_Real SCIPfeasCeil(SCIP* scip, SCIP_Real val) { return ceil(val - scip->feastol); }

static double relDiff(double a, double b)
{
   double q = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
   q = q > 1.0 ? q : 1.0;
   return (a - b) / q;
}

SCIP_RETCODE SCIPgetRealParam(SCIP* scip, const char* name, SCIP_Real* value)
{
   if( strcmp(name, "numerics/boundstreps") == 0 )
      *value = scip->boundstreps;
   else if( strcmp(name, "numerics/recompfac") == 0 )
      *value = scip->recompfac;
   else if( strcmp(name, "constraints/linear/maxeasyactivitydelta") == 0 )
      *value = scip->maxeasy;
   else
      return SCIP_PARAMETERUNKNOWN;
   return SCIP_OKAY;
}

SCIP_RETCODE SCIPgetIntParam(SCIP* scip, const char* name, int* value)
{
   if( strcmp(name, "constraints/linear/tightenboundsfreq") == 0 )
      *value = scip->tightenboundsfreq;
   else if( strcmp(name, "constraints/linear/rangedrowfreq") == 0 )
      *value = scip->rangedrowfreq;
   else if( strcmp(name, "constraints/linear/rangedrowmaxdepth") == 0 )
      *value = scip->rangedrowmaxdepth;
   else
      return SCIP_PARAMETERUNKNOWN;
   return SCIP_OKAY;
}

SCIP_RETCODE SCIPgetBoolParam(SCIP* scip, const char* name, SCIP_Bool* value)
{
   if( strcmp(name, "constraints/linear/rangedrowpropagation") != 0 )
      return SCIP_PARAMETERUNKNOWN;
   *value = scip->rangedrowprop ? TRUE : FALSE;
   return SCIP_OKAY;
}

SCIP_RETCODE SCIPaddBoolParam(SCIP* scip, const char* name, const char* desc, SCIP_Bool* valueptr,
   SCIP_Bool isadvanced, SCIP_Bool defaultvalue, SCIP_DECL_PARAMCHGD((*paramchgd)), SCIP_PARAMDATA* paramdata)
{
   (void)desc; (void)isadvanced; (void)paramchgd; (void)paramdata;
   if( scip->nintparams >= TH_MAXPARAMS )
      return SCIP_NOMEMORY;
   /* kept with the int parameters (SCIP_Bool is an unsigned int) */
   strncpy(scip->intname[scip->nintparams], name, 95);
   scip->intptr[scip->nintparams] = (int*)valueptr;
   scip->nintparams++;
   if( valueptr != NULL )
      *valueptr = defaultvalue;
   return SCIP_OKAY;
}

SCIP_RETCODE SCIPaddIntParam(SCIP* scip, const char* name, const char* desc, int* valueptr, SCIP_Bool isadvanced,
   int defaultvalue, int minvalue, int maxvalue, SCIP_DECL_PARAMCHGD((*paramchgd)), SCIP_PARAMDATA* paramdata)
{
   (void)desc; (void)isadvanced; (void)minvalue; (void)maxvalue; (void)paramchgd; (void)paramdata;
   if( scip->nintparams >= TH_MAXPARAMS )
      return SCIP_NOMEMORY;
   strncpy(scip->intname[scip->nintparams], name, 95);
   scip->intptr[scip->nintparams] = valueptr;
   scip->nintparams++;
   if( valueptr != NULL )
      *valueptr = defaultvalue;
   return SCIP_OKAY;
}

/* ---- propagator registration ---- */
SCIP_RETCODE SCIPincludePropBasic(SCIP* scip, SCIP_PROP** propptr, const char* name, const char* desc, int priority,
   int freq, SCIP_Bool delay, SCIP_PROPTIMING timingmask, SCIP_DECL_PROPEXEC((*propexec)), SCIP_PROPDATA* propdata)
{
   (void)desc; (void)priority; (void)delay; (void)timingmask;
   if( scip->prop != NULL )
      return SCIP_INVALIDDATA;   /* one propagator per test host */
   scip->prop = (SCIP_PROP*)calloc(1, sizeof(SCIP_PROP));
   if( scip->prop == NULL )
      return SCIP_NOMEMORY;
   strncpy(scip->prop->name, name, 63);
   scip->prop->exec = propexec;
   scip->prop->data = propdata;
   scip->prop->freq = freq;
   if( propptr != NULL )
      *propptr = scip->prop;
   return SCIP_OKAY;
}
SCIP_RETCODE SCIPsetPropFree(SCIP* scip, SCIP_PROP* prop, SCIP_DECL_PROPFREE((*f))) { (void)scip; prop->pfree = f; return SCIP_OKAY; }
SCIP_RETCODE SCIPsetPropInitpre(SCIP* scip, SCIP_PROP* prop, SCIP_DECL_PROPINITPRE((*f))) { (void)scip; prop->initpre = f; return SCIP_OKAY; }
SCIP_RETCODE SCIPsetPropExitpre(SCIP* scip, SCIP_PROP* prop, SCIP_DECL_PROPEXITPRE((*f))) { (void)scip; prop->exitpre = f; return SCIP_OKAY; }
SCIP_RETCODE SCIPsetPropInitsol(SCIP* scip, SCIP_PROP* prop, SCIP_DECL_PROPINITSOL((*f))) { (void)scip; prop->initsol = f; return SCIP_OKAY; }
SCIP_RETCODE SCIPsetPropExitsol(SCIP* scip, SCIP_PROP* prop, SCIP_DECL_PROPEXITSOL((*f))) { (void)scip; prop->exitsol = f; return SCIP_OKAY; }
SCIP_RETCODE SCIPsetPropPresol(SCIP* scip, SCIP_PROP* prop, SCIP_DECL_PROPPRESOL((*f)), int presolpriority,
   int presolmaxrounds, SCIP_PRESOLTIMING presoltiming)
{
   (void)scip; (void)presolpriority; (void)presolmaxrounds; (void)presoltiming;
   prop->presol = f;
   return SCIP_OKAY;
}
SCIP_RETCODE SCIPsetPropResprop(SCIP* scip, SCIP_PROP* prop, SCIP_DECL_PROPRESPROP((*f))) { (void)scip; prop->resprop = f; return SCIP_OKAY; }
SCIP_PROPDATA* SCIPpropGetData(SCIP_PROP* prop) { return prop->data; }
int SCIPpropGetFreq(SCIP_PROP* prop) { return prop->freq; }
SCIP_PROP* SCIPfindProp(SCIP* scip, const char* name)
{
   return (scip->prop != NULL && strcmp(scip->prop->name, name) == 0) ? scip->prop : NULL;
}
void SCIPpropSetData(SCIP_PROP* prop, SCIP_PROPDATA* propdata) { prop->data = propdata; }

/* ---- problem access ---- */
SCIP_CONSHDLR* SCIPfindConshdlr(SCIP* scip, const char* name) { return strcmp(name, "linear") == 0 ? &scip->linear : NULL; }
int SCIPconshdlrGetNConss(SCIP_CONSHDLR* h) { return h->nconss; }
SCIP_CONS** SCIPconshdlrGetConss(SCIP_CONSHDLR* h) { return h->conss; }
SCIP_Bool SCIPconsIsActive(SCIP_CONS* cons) { (void)cons; return TRUE; }
int SCIPgetNVars(SCIP* scip) { return scip->nvars; }
SCIP_VAR** SCIPgetVars(SCIP* scip) { return scip->vars; }
int SCIPvarGetProbindex(SCIP_VAR* var) { return var->probindex; }
SCIP_Bool SCIPvarIsIntegral(SCIP_VAR* var) { return var->integral != 0; }
/* pub_var.h: an integral variable with global bounds within [0,1] */
SCIP_Bool SCIPvarIsBinary(SCIP_VAR* var) { return var->integral != 0 && var->glb >= 0.0 && var->gub <= 1.0; }
SCIP_Real SCIPvarGetLbGlobal(SCIP_VAR* var) { return var->glb; }
SCIP_Real SCIPvarGetUbGlobal(SCIP_VAR* var) { return var->gub; }
SCIP_Real SCIPvarGetLbLocal(SCIP_VAR* var) { return var->llb; }
SCIP_Real SCIPvarGetUbLocal(SCIP_VAR* var) { return var->lub; }
SCIP_Real SCIPgetLhsLinear(SCIP* scip, SCIP_CONS* cons) { (void)scip; return cons->lhs; }
SCIP_Real SCIPgetRhsLinear(SCIP* scip, SCIP_CONS* cons) { (void)scip; return cons->rhs; }
int SCIPgetNVarsLinear(SCIP* scip, SCIP_CONS* cons) { (void)scip; return cons->nvars; }
SCIP_VAR** SCIPgetVarsLinear(SCIP* scip, SCIP_CONS* cons) { (void)scip; return cons->vars; }
SCIP_Real* SCIPgetValsLinear(SCIP* scip, SCIP_CONS* cons) { (void)scip; return cons->vals; }

/* all variables of the test host are active */
SCIP_RETCODE SCIPgetProbvarLinearSum(SCIP* scip, SCIP_VAR** vars, SCIP_Real* scalars, int* nvars, int varssize,
   SCIP_Real* constant, int* requiredsize)
{
   (void)scip; (void)vars; (void)scalars; (void)varssize; (void)constant;
   *requiredsize = *nvars;
   return SCIP_OKAY;
}

/* ---- bound changes: SCIPinferVarLbCons / SCIPinferVarUbCons semantics ---- */
static double adjLb(SCIP* scip, int integral, double lb)
{
   if( lb < 0.0 && -lb >= scip->infinity )
      return -scip->infinity;
   if( lb > 0.0 && lb >= scip->infinity )
      return scip->infinity;
   if( integral )
      return ceil(lb - scip->feastol);
   if( lb > 0.0 && lb < scip->epsilon )
      return 0.0;
   return lb;
}

static double adjUb(SCIP* scip, int integral, double ub)
{
   if( ub > 0.0 && ub >= scip->infinity )
      return scip->infinity;
   if( ub < 0.0 && -ub >= scip->infinity )
      return -scip->infinity;
   if( integral )
      return floor(ub + scip->feastol);
   if( ub < 0.0 && ub > -scip->epsilon )
      return 0.0;
   return ub;
}

static int lbBetter(SCIP* scip, double newlb, double oldlb, double oldub)
{
   double t;
   if( oldlb < 0.0 && newlb >= 0.0 )
      return 1;
   t = (oldub - oldlb) <= fabs(oldlb) ? (oldub - oldlb) : fabs(oldlb);
   t = t >= 1e-3 ? t : 1e-3;
   return newlb - oldlb > scip->boundstreps * t;
}

static int ubBetter(SCIP* scip, double newub, double oldlb, double oldub)
{
   double t;
   if( oldub > 0.0 && newub <= 0.0 )
      return 1;
   t = (oldub - oldlb) <= fabs(oldub) ? (oldub - oldlb) : fabs(oldub);
   t = t >= 1e-3 ? t : 1e-3;
   return newub - oldub < -(scip->boundstreps * t);
}

static SCIP_RETCODE chgLb(SCIP* scip, SCIP_VAR* var, double nb, SCIP_Bool force, SCIP_Bool global,
   SCIP_Bool* infeasible, SCIP_Bool* tightened)
{
   double lb = global ? var->glb : var->llb, ub = global ? var->gub : var->lub;
   *infeasible = FALSE;
   if( tightened )
      *tightened = FALSE;
   nb = adjLb(scip, var->integral, nb);
   if( nb >= scip->infinity || relDiff(nb, ub) > scip->feastol )
   {
      *infeasible = TRUE;
      return SCIP_OKAY;
   }
   nb = nb <= ub ? nb : ub;
   if( (force && nb - lb <= scip->epsilon) || (!force && !lbBetter(scip, nb, lb, ub)) )
      return SCIP_OKAY;
   if( nb > lb )
   {
      if( global )
         var->glb = nb;
      var->llb = nb > var->llb ? nb : var->llb;
      if( tightened )
         *tightened = TRUE;
   }
   return SCIP_OKAY;
}

static SCIP_RETCODE chgUb(SCIP* scip, SCIP_VAR* var, double nb, SCIP_Bool force, SCIP_Bool global,
   SCIP_Bool* infeasible, SCIP_Bool* tightened)
{
   double lb = global ? var->glb : var->llb, ub = global ? var->gub : var->lub;
   *infeasible = FALSE;
   if( tightened )
      *tightened = FALSE;
   nb = adjUb(scip, var->integral, nb);
   if( -nb >= scip->infinity || relDiff(nb, lb) < -scip->feastol )
   {
      *infeasible = TRUE;
      return SCIP_OKAY;
   }
   nb = nb >= lb ? nb : lb;
   if( (force && nb - ub >= -scip->epsilon) || (!force && !ubBetter(scip, nb, lb, ub)) )
      return SCIP_OKAY;
   if( nb < ub )
   {
      if( global )
         var->gub = nb;
      var->lub = nb < var->lub ? nb : var->lub;
      if( tightened )
         *tightened = TRUE;
   }
   return SCIP_OKAY;
}

static int grow(void** p, int* cap, int need, size_t elem);

static SCIP_RETCODE noteChange(SCIP* scip, SCIP_VAR* var, int bound, int inferinfo, double oldval, double newval)
{
   struct ChgRec* r;
   if( !grow((void**)&scip->chg, &scip->chgcap, scip->nchg + 1, sizeof(*scip->chg)) )
      return SCIP_NOMEMORY;
   r = &scip->chg[scip->nchg++];
   r->var = var->probindex;
   r->bound = bound;
   r->inferinfo = inferinfo;
   r->oldval = oldval;
   r->newval = newval;
   return SCIP_OKAY;
}

SCIP_RETCODE SCIPinferVarLbProp(SCIP* scip, SCIP_VAR* var, SCIP_Real newbound, SCIP_PROP* inferprop, int inferinfo,
   SCIP_Bool force, SCIP_Bool* infeasible, SCIP_Bool* tightened)
{
   double old = var->llb;
   SCIP_Bool t = FALSE;
   (void)inferprop;
   SCIP_CALL( chgLb(scip, var, newbound, force, FALSE, infeasible, &t) );
   if( tightened )
      *tightened = t;
   return t ? noteChange(scip, var, 0, inferinfo, old, var->llb) : SCIP_OKAY;
}
SCIP_RETCODE SCIPinferVarUbProp(SCIP* scip, SCIP_VAR* var, SCIP_Real newbound, SCIP_PROP* inferprop, int inferinfo,
   SCIP_Bool force, SCIP_Bool* infeasible, SCIP_Bool* tightened)
{
   double old = var->lub;
   SCIP_Bool t = FALSE;
   (void)inferprop;
   SCIP_CALL( chgUb(scip, var, newbound, force, FALSE, infeasible, &t) );
   if( tightened )
      *tightened = t;
   return t ? noteChange(scip, var, 1, inferinfo, old, var->lub) : SCIP_OKAY;
}

/* ---- conflict analysis ---- */
SCIP_STAGE SCIPgetStage(SCIP* scip) { return (SCIP_STAGE)scip->stage; }
SCIP_Bool SCIPinProbing(SCIP* scip) { return scip->inprobing ? TRUE : FALSE; }
int SCIPgetDepth(SCIP* scip) { return scip->depth; }
SCIP_Bool SCIPinRepropagation(SCIP* scip) { return scip->inreprop ? TRUE : FALSE; }
SCIP_Bool SCIPisConflictAnalysisApplicable(SCIP* scip) { (void)scip; return TRUE; }
SCIP_RETCODE SCIPinitConflictAnalysis(SCIP* scip, SCIP_CONFTYPE conftype, SCIP_Bool iscutoffinvolved)
{
   (void)conftype; (void)iscutoffinvolved;
   scip->nconf = 0;
   return SCIP_OKAY;
}

/* the local bound of var in force just before change position pos (all recorded changes:
 * pos = nchg) */
static double boundAt(SCIP* scip, SCIP_VAR* var, int bound, int pos)
{
   double v = bound == 0 ? var->llb0 : var->lub0;
   int i;
   for( i = 0; i < pos && i < scip->nchg; ++i )
      if( scip->chg[i].var == var->probindex && scip->chg[i].bound == bound )
         v = scip->chg[i].newval;
   return v;
}

static SCIP_RETCODE addConflict(SCIP* scip, SCIP_VAR* var, int bound, SCIP_BDCHGIDX* bdchgidx)
{
   struct ConfRec* r;
   if( !grow((void**)&scip->conf, &scip->confcap, scip->nconf + 1, sizeof(*scip->conf)) )
      return SCIP_NOMEMORY;
   r = &scip->conf[scip->nconf++];
   r->var = var->probindex;
   r->bound = bound;
   r->val = boundAt(scip, var, bound, bdchgidx != NULL ? bdchgidx->pos : scip->nchg);
   return SCIP_OKAY;
}
SCIP_RETCODE SCIPaddConflictLb(SCIP* scip, SCIP_VAR* var, SCIP_BDCHGIDX* bdchgidx) { return addConflict(scip, var, 0, bdchgidx); }
SCIP_RETCODE SCIPaddConflictUb(SCIP* scip, SCIP_VAR* var, SCIP_BDCHGIDX* bdchgidx) { return addConflict(scip, var, 1, bdchgidx); }
SCIP_RETCODE SCIPanalyzeConflict(SCIP* scip, int validdepth, SCIP_Bool* success)
{
   (void)validdepth;
   scip->nanalyzed++;
   if( success != NULL )
      *success = TRUE;
   return SCIP_OKAY;
}
SCIP_RETCODE SCIPtightenVarLb(SCIP* scip, SCIP_VAR* var, SCIP_Real newbound, SCIP_Bool force, SCIP_Bool* infeasible,
   SCIP_Bool* tightened)
{
   return chgLb(scip, var, newbound, force, TRUE, infeasible, tightened);
}
SCIP_RETCODE SCIPtightenVarUb(SCIP* scip, SCIP_VAR* var, SCIP_Real newbound, SCIP_Bool force, SCIP_Bool* infeasible,
   SCIP_Bool* tightened)
{
   return chgUb(scip, var, newbound, force, TRUE, infeasible, tightened);
}

/* ---- presolving fixings and probing deductions ---- */
static int grow(void** p, int* cap, int need, size_t elem)
{
   void* q;
   int c;
   if( need <= *cap )
      return 1;
   c = *cap > 0 ? 2 * *cap : 256;
   while( c < need )
      c *= 2;
   q = realloc(*p, elem * (size_t)c);
   if( q == NULL )
      return 0;
   *p = q;
   *cap = c;
   return 1;
}

/* SCIPfixVar in presolving (scip_var.c): infeasible if the value lies outside the global
 * bounds (feasibility tolerance) or is fractional for an integral variable; else both bounds
 * are set to the value */
SCIP_RETCODE SCIPfixVar(SCIP* scip, SCIP_VAR* var, SCIP_Real fixedval, SCIP_Bool* infeasible, SCIP_Bool* fixed)
{
   *fixed = FALSE;
   *infeasible = (relDiff(fixedval, var->glb) < -scip->feastol || relDiff(fixedval, var->gub) > scip->feastol
      || (var->integral && fabs(fixedval - floor(fixedval + 0.5)) > scip->feastol));
   if( *infeasible )
      return SCIP_OKAY;
   if( !grow((void**)&scip->fix, &scip->fixcap, scip->nfix + 1, sizeof(*scip->fix)) )
      return SCIP_NOMEMORY;
   scip->fix[scip->nfix].var = var->probindex;
   scip->fix[scip->nfix].val = fixedval;
   scip->nfix++;
   *fixed = !(var->glb == fixedval && var->gub == fixedval);
   var->glb = var->gub = var->llb = var->lub = fixedval;
   return SCIP_OKAY;
}

/* test double of SCIPanalyzeDeductionsProbing (prop_probing.c:1314-1640), for a binary probing
 * variable with left branch x <= 0 and right branch x >= 1: records every entry, and applies
 * what SCIP derives from the union of the two branches -- a fixing when the union is one value,
 * else tightenings (integral: by more than 0.5; continuous: SCIPisLb/UbBetter) -- and counts
 * the aggregations (fixed in both branches) and the implications (fixed at a union end in one
 * branch) that SCIP would add */
SCIP_RETCODE SCIPanalyzeDeductionsProbing(SCIP* scip, SCIP_VAR* probingvar, SCIP_Real leftub, SCIP_Real rightlb,
   int nvars, SCIP_VAR** vars, SCIP_Real* leftimpllbs, SCIP_Real* leftimplubs, SCIP_Real* leftproplbs,
   SCIP_Real* leftpropubs, SCIP_Real* rightimpllbs, SCIP_Real* rightimplubs, SCIP_Real* rightproplbs,
   SCIP_Real* rightpropubs, int* nfixedvars, int* naggrvars, int* nimplications, int* nchgbds, SCIP_Bool* cutoff)
{
   int j;
   (void)leftub; (void)rightlb; (void)leftimpllbs; (void)leftimplubs; (void)rightimpllbs; (void)rightimplubs;
   *cutoff = FALSE;
   for( j = 0; j < nvars && !*cutoff; ++j )
   {
      SCIP_VAR* var = vars[j];
      double newlb = leftproplbs[j] < rightproplbs[j] ? leftproplbs[j] : rightproplbs[j];
      double newub = leftpropubs[j] > rightpropubs[j] ? leftpropubs[j] : rightpropubs[j];
      SCIP_Bool tightened;
      struct DedRec* r;

      if( var == probingvar )
         continue;
      if( !grow((void**)&scip->ded, &scip->dedcap, scip->nded + 1, sizeof(*scip->ded)) )
         return SCIP_NOMEMORY;
      r = &scip->ded[scip->nded++];
      r->probe = probingvar->probindex;
      r->var = var->probindex;
      r->bnd[0] = leftproplbs[j];
      r->bnd[1] = leftpropubs[j];
      r->bnd[2] = rightproplbs[j];
      r->bnd[3] = rightpropubs[j];

      if( fabs(newlb - newub) <= scip->epsilon )
      {
         SCIP_Bool fixed;
         SCIP_CALL( SCIPfixVar(scip, var, newlb, cutoff, &fixed) );
         if( fixed )
            (*nfixedvars)++;
         continue;
      }
      if( var->integral ? newlb > var->glb + 0.5 : lbBetter(scip, newlb, var->glb, var->gub) )
      {
         SCIP_CALL( SCIPtightenVarLb(scip, var, newlb, TRUE, cutoff, &tightened) );
         *nchgbds += tightened ? 1 : 0;
      }
      if( !*cutoff && (var->integral ? newub < var->gub - 0.5 : ubBetter(scip, newub, var->glb, var->gub)) )
      {
         SCIP_CALL( SCIPtightenVarUb(scip, var, newub, TRUE, cutoff, &tightened) );
         *nchgbds += tightened ? 1 : 0;
      }
      if( fabs(leftproplbs[j] - leftpropubs[j]) <= scip->epsilon
         && fabs(rightproplbs[j] - rightpropubs[j]) <= scip->epsilon )
      {
         (*naggrvars)++;
         scip->naggr++;
      }
      else if( fabs(newlb - leftpropubs[j]) <= scip->epsilon || fabs(newub - leftproplbs[j]) <= scip->epsilon
         || fabs(newlb - rightpropubs[j]) <= scip->epsilon || fabs(newub - rightproplbs[j]) <= scip->epsilon )
      {
         (*nimplications)++;
         scip->nimpl++;
      }
   }
   return SCIP_OKAY;
}

/* ---- driver (ctypes entry points for tests) ---- */
SCIP_EXPORT SCIP* th_create(void)
{
   SCIP* s = (SCIP*)calloc(1, sizeof(SCIP));
   if( s == NULL )
      return NULL;
   s->infinity = 1e20;
   s->epsilon = 1e-9;
   s->sumepsilon = 1e-6;
   s->feastol = 1e-6;
   s->hugeval = 1e15;
   s->boundstreps = 0.05;
   s->recompfac = 1e6;
   s->maxeasy = 1e6;
   s->stage = SCIP_STAGE_SOLVING;
   s->rangedrowprop = 1;
   s->tightenboundsfreq = 1;
   s->rangedrowfreq = 1;
   s->rangedrowmaxdepth = INT_MAX;
   return s;
}

SCIP_EXPORT int th_add_vars(SCIP* s, int n, const double* lb, const double* ub, const unsigned char* integral)
{
   int i;
   s->vars = (SCIP_VAR**)malloc(sizeof(SCIP_VAR*) * (size_t)(n > 0 ? n : 1));
   if( s->vars == NULL )
      return SCIP_NOMEMORY;
   for( i = 0; i < n; ++i )
   {
      SCIP_VAR* v = (SCIP_VAR*)calloc(1, sizeof(SCIP_VAR));
      if( v == NULL )
         return SCIP_NOMEMORY;
      v->probindex = i;
      v->integral = integral[i];
      v->glb = v->llb = v->llb0 = lb[i];
      v->gub = v->lub = v->lub0 = ub[i];
      s->vars[i] = v;
   }
   s->nvars = n;
   return SCIP_OKAY;
}

/* rows in CSR form */
SCIP_EXPORT int th_add_rows(SCIP* s, int nrows, const long long* rowptr, const int* colidx, const double* vals,
   const double* lhs, const double* rhs)
{
   int i;
   s->linear.conss = (SCIP_CONS**)malloc(sizeof(SCIP_CONS*) * (size_t)(nrows > 0 ? nrows : 1));
   if( s->linear.conss == NULL )
      return SCIP_NOMEMORY;
   for( i = 0; i < nrows; ++i )
   {
      long long k, b = rowptr[i], e = rowptr[i + 1];
      SCIP_CONS* c = (SCIP_CONS*)calloc(1, sizeof(SCIP_CONS));
      if( c == NULL )
         return SCIP_NOMEMORY;
      c->nvars = (int)(e - b);
      c->vars = (SCIP_VAR**)malloc(sizeof(SCIP_VAR*) * (size_t)(c->nvars > 0 ? c->nvars : 1));
      c->vals = (double*)malloc(sizeof(double) * (size_t)(c->nvars > 0 ? c->nvars : 1));
      for( k = b; k < e; ++k )
      {
         c->vars[k - b] = s->vars[colidx[k]];
         c->vals[k - b] = vals[k];
      }
      c->lhs = lhs[i];
      c->rhs = rhs[i];
      s->linear.conss[i] = c;
   }
   s->linear.nconss = nrows;
   return SCIP_OKAY;
}

SCIP_EXPORT int th_include_gpulinear(SCIP* s) { return SCIPincludePropGpulinear(s); }

/* SCIP's own parameter constraints/linear/rangedrowpropagation */
SCIP_EXPORT void th_set_rangedrowprop(SCIP* s, int value) { s->rangedrowprop = value; }

/* SCIP's int parameters constraints/linear/{tightenboundsfreq,rangedrowfreq,rangedrowmaxdepth} */
SCIP_EXPORT int th_set_linear_int(SCIP* s, const char* name, int value)
{
   if( strcmp(name, "tightenboundsfreq") == 0 )
      s->tightenboundsfreq = value;
   else if( strcmp(name, "rangedrowfreq") == 0 )
      s->rangedrowfreq = value;
   else if( strcmp(name, "rangedrowmaxdepth") == 0 )
      s->rangedrowmaxdepth = value;
   else
      return SCIP_PARAMETERUNKNOWN;
   return SCIP_OKAY;
}

/* the node the next PROPEXEC runs at: SCIP stage (type_set.h values), probing, depth,
 * repropagation */
SCIP_EXPORT void th_set_node(SCIP* s, int stage, int inprobing, int depth, int inreprop)
{
   s->stage = stage;
   s->inprobing = inprobing;
   s->depth = depth;
   s->inreprop = inreprop;
}

SCIP_EXPORT int th_set_int_param(SCIP* s, const char* name, int value)
{
   int i;
   for( i = 0; i < s->nintparams; ++i )
      if( strcmp(s->intname[i], name) == 0 && s->intptr[i] != NULL )
      {
         *s->intptr[i] = value;
         return SCIP_OKAY;
      }
   return SCIP_PARAMETERUNKNOWN;
}

SCIP_EXPORT int th_initpre(SCIP* s) { return s->prop->initpre ? s->prop->initpre(s, s->prop) : SCIP_OKAY; }
SCIP_EXPORT int th_exitpre(SCIP* s) { return s->prop->exitpre ? s->prop->exitpre(s, s->prop) : SCIP_OKAY; }
SCIP_EXPORT int th_initsol(SCIP* s) { return s->prop->initsol ? s->prop->initsol(s, s->prop) : SCIP_OKAY; }
SCIP_EXPORT int th_exitsol(SCIP* s) { return s->prop->exitsol ? s->prop->exitsol(s, s->prop, FALSE) : SCIP_OKAY; }

SCIP_EXPORT int th_propexec(SCIP* s, int* result)
{
   SCIP_RESULT r = SCIP_DIDNOTRUN;
   SCIP_RETCODE rc = s->prop->exec(s, s->prop, SCIP_PROPTIMING_BEFORELP, &r);
   *result = (int)r;
   return rc;
}

SCIP_EXPORT int th_presol(SCIP* s, int* result, int* nchgbds)
{
   int z[10] = { 0 };
   SCIP_RESULT r = SCIP_DIDNOTRUN;
   SCIP_RETCODE rc = s->prop->presol(s, s->prop, 0, SCIP_PRESOLTIMING_MEDIUM, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
      &z[0], &z[1], &z[2], nchgbds, &z[4], &z[5], &z[6], &z[7], &z[8], &z[9], &r);
   *result = (int)r;
   return rc;
}

/* conflict analysis asking for the reason of the bound change of column j (prop.c:739-788) */
SCIP_EXPORT int th_resprop(SCIP* s, int j, int upper, int* result)
{
   SCIP_RESULT r = SCIP_DIDNOTRUN;
   SCIP_RETCODE rc;
   if( s->prop->resprop == NULL )
      return SCIP_PLUGINNOTFOUND;   /* what SCIPpropResolvePropagation reports */
   rc = s->prop->resprop(s, s->prop, s->vars[j], 0, upper ? SCIP_BOUNDTYPE_UPPER : SCIP_BOUNDTYPE_LOWER, NULL,
      upper ? s->vars[j]->lub : s->vars[j]->llb, &r);
   *result = (int)r;
   return rc;
}

SCIP_EXPORT void th_set_local_bounds(SCIP* s, const double* lb, const double* ub)
{
   int i;
   for( i = 0; i < s->nvars; ++i )
   {
      s->vars[i]->llb = s->vars[i]->llb0 = lb[i];
      s->vars[i]->lub = s->vars[i]->lub0 = ub[i];
   }
   s->nchg = 0;   /* a new node: no local changes yet */
}

/* conflict analysis resolving recorded change pos (prop.c:739-788): the propagator's
 * PROPRESPROP gets the changed variable, its inferinfo, the bound type and the change's index;
 * the conflict set it adds is then read with th_get_conflict */
SCIP_EXPORT int th_resprop_change(SCIP* s, int pos, int* result)
{
   SCIP_RESULT r = SCIP_DIDNOTRUN;
   SCIP_BDCHGIDX idx;
   SCIP_RETCODE rc;
   const struct ChgRec* c;
   if( pos < 0 || pos >= s->nchg )
      return SCIP_INVALIDDATA;
   if( s->prop->resprop == NULL )
      return SCIP_PLUGINNOTFOUND;
   c = &s->chg[pos];
   idx.pos = pos;
   s->nconf = 0;
   rc = s->prop->resprop(s, s->prop, s->vars[c->var], c->inferinfo,
      c->bound ? SCIP_BOUNDTYPE_UPPER : SCIP_BOUNDTYPE_LOWER, &idx, c->newval, &r);
   *result = (int)r;
   return rc;
}

/* strong branching on the device (SCIPgpulinearStrongbranch) for candidate columns cand with
 * LP values solval, at the current local bounds */
SCIP_EXPORT int th_strongbranch(SCIP* s, int ncands, const int* cand, const double* solval, int* downinf, int* upinf,
   double* newlbs, double* newubs, int* cutoff)
{
   SCIP_VAR** vars = (SCIP_VAR**)malloc(sizeof(SCIP_VAR*) * (size_t)(ncands > 0 ? ncands : 1));
   SCIP_Bool* di = (SCIP_Bool*)malloc(sizeof(SCIP_Bool) * (size_t)(ncands > 0 ? ncands : 1));
   SCIP_Bool* ui = (SCIP_Bool*)malloc(sizeof(SCIP_Bool) * (size_t)(ncands > 0 ? ncands : 1));
   SCIP_Bool cut = FALSE;
   SCIP_RETCODE rc;
   int c;
   if( vars == NULL || di == NULL || ui == NULL )
      return SCIP_NOMEMORY;
   for( c = 0; c < ncands; ++c )
      vars[c] = s->vars[cand[c]];
   rc = SCIPgpulinearStrongbranch(s, ncands, vars, solval, di, ui, newlbs, newubs, &cut);
   for( c = 0; c < ncands; ++c )
   {
      downinf[c] = (int)di[c];
      upinf[c] = (int)ui[c];
   }
   *cutoff = (int)cut;
   free(vars);
   free(di);
   free(ui);
   return rc;
}

/* batched dives on the device (SCIPgpulinearDive) over candidate columns cand, at the current
 * local bounds */
SCIP_EXPORT int th_dive(SCIP* s, int ndives, int ncands, const int* cand, const int* order, const double* fixvals,
   int maxdepth, int* depth, int* cutoff, double* divelbs, double* diveubs)
{
   SCIP_VAR** vars = (SCIP_VAR**)malloc(sizeof(SCIP_VAR*) * (size_t)(ncands > 0 ? ncands : 1));
   SCIP_Bool* cut = (SCIP_Bool*)malloc(sizeof(SCIP_Bool) * (size_t)(ndives > 0 ? ndives : 1));
   SCIP_RETCODE rc;
   int c;
   if( vars == NULL || cut == NULL )
      return SCIP_NOMEMORY;
   for( c = 0; c < ncands; ++c )
      vars[c] = s->vars[cand[c]];
   rc = SCIPgpulinearDive(s, ndives, ncands, vars, order, fixvals, maxdepth, depth, cut, divelbs, diveubs);
   for( c = 0; c < ndives; ++c )
      cutoff[c] = (int)cut[c];
   free(vars);
   free(cut);
   return rc;
}

/* the recorded local changes; with non-NULL arrays they are copied out */
SCIP_EXPORT int th_get_changes(SCIP* s, int* var, int* bound, int* inferinfo, double* oldval, double* newval)
{
   int i;
   for( i = 0; var != NULL && i < s->nchg; ++i )
   {
      var[i] = s->chg[i].var;
      bound[i] = s->chg[i].bound;
      inferinfo[i] = s->chg[i].inferinfo;
      oldval[i] = s->chg[i].oldval;
      newval[i] = s->chg[i].newval;
   }
   return s->nchg;
}

/* the conflict candidate set and the number of SCIPanalyzeConflict calls */
SCIP_EXPORT int th_get_conflict(SCIP* s, int* var, int* bound, double* val, int* nanalyzed)
{
   int i;
   for( i = 0; var != NULL && i < s->nconf; ++i )
   {
      var[i] = s->conf[i].var;
      bound[i] = s->conf[i].bound;
      val[i] = s->conf[i].val;
   }
   *nanalyzed = s->nanalyzed;
   return s->nconf;
}

SCIP_EXPORT void th_get_bounds(SCIP* s, int global, double* lb, double* ub)
{
   int i;
   for( i = 0; i < s->nvars; ++i )
   {
      lb[i] = global ? s->vars[i]->glb : s->vars[i]->llb;
      ub[i] = global ? s->vars[i]->gub : s->vars[i]->lub;
   }
}

/* counts of the records; with non-NULL arrays they are copied out as well */
SCIP_EXPORT int th_get_probe_records(SCIP* s, int* nfix, int* fixvar, double* fixval, int* nded, int* dedprobe,
   int* dedvar, double* dedbnd, int* naggr, int* nimpl)
{
   int i;
   *nfix = s->nfix;
   *nded = s->nded;
   *naggr = s->naggr;
   *nimpl = s->nimpl;
   for( i = 0; fixvar != NULL && i < s->nfix; ++i )
   {
      fixvar[i] = s->fix[i].var;
      fixval[i] = s->fix[i].val;
   }
   for( i = 0; dedprobe != NULL && i < s->nded; ++i )
   {
      dedprobe[i] = s->ded[i].probe;
      dedvar[i] = s->ded[i].var;
      memcpy(dedbnd + 4 * (size_t)i, s->ded[i].bnd, 4 * sizeof(double));
   }
   return SCIP_OKAY;
}

SCIP_EXPORT int th_free(SCIP* s)
{
   int i, rc = SCIP_OKAY;
   if( s->prop != NULL && s->prop->pfree != NULL )
      rc = s->prop->pfree(s, s->prop);
   free(s->prop);
   for( i = 0; i < s->linear.nconss; ++i )
   {
      free(s->linear.conss[i]->vars);
      free(s->linear.conss[i]->vals);
      free(s->linear.conss[i]);
   }
   free(s->linear.conss);
   for( i = 0; i < s->nvars; ++i )
      free(s->vars[i]);
   free(s->vars);
   free(s->fix);
   free(s->ded);
   free(s->chg);
   free(s->conf);
   free(s);
   return rc;
}
