/* Test double of SCIP's public API (scip/scip.h) for prop_gpulinear.c.
 *
 * It declares only the subset of SCIP's API that the gpulinear propagator calls, with the
 * signatures of SCIP's own declarations (src/scip/scip_prop.h, scip_var.h, scip_param.h,
 * prop_probing.h (in scip/prop_probing.h here), scip_conflict.h, scip_general.h,
 * scip_numerics.h, scip_mem.h, pub_cons.h, pub_prop.h, pub_var.h), so that the plugin
 * compiles and runs under tests/ without a SCIP library.  Built against an installed SCIP,
 * the plugin uses SCIP's own headers instead and this directory is not on the include path.
 * The implementation (../testhost.c) is this repository's own; it mirrors the role of SCIP's
 * fake-stage test fixture (tests/include/scip_test.c).
 */
#ifndef SCIP_AMD_TESTHOST_SCIP_H
#define SCIP_AMD_TESTHOST_SCIP_H

#include <limits.h>
#include <stdio.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCIP_EXPORT __attribute__((visibility("default")))

typedef enum SCIP_Retcode
{
   SCIP_OKAY = 1, SCIP_ERROR = 0, SCIP_NOMEMORY = -1, SCIP_READERROR = -2, SCIP_WRITEERROR = -3,
   SCIP_NOFILE = -4, SCIP_FILECREATEERROR = -5, SCIP_LPERROR = -6, SCIP_NOPROBLEM = -7,
   SCIP_INVALIDCALL = -8, SCIP_INVALIDDATA = -9, SCIP_INVALIDRESULT = -10, SCIP_PLUGINNOTFOUND = -11,
   SCIP_PARAMETERUNKNOWN = -12, SCIP_PARAMETERWRONGTYPE = -13, SCIP_PARAMETERWRONGVAL = -14,
   SCIP_KEYALREADYEXISTING = -15, SCIP_MAXDEPTHLEVEL = -16, SCIP_BRANCHERROR = -17, SCIP_NOTIMPLEMENTED = -18
} SCIP_RETCODE;

typedef enum SCIP_Result
{
   SCIP_DIDNOTRUN = 1, SCIP_DELAYED = 2, SCIP_DIDNOTFIND = 3, SCIP_FEASIBLE = 4, SCIP_INFEASIBLE = 5,
   SCIP_UNBOUNDED = 6, SCIP_CUTOFF = 7, SCIP_SEPARATED = 8, SCIP_NEWROUND = 9, SCIP_REDUCEDDOM = 10,
   SCIP_CONSADDED = 11, SCIP_CONSCHANGED = 12, SCIP_BRANCHED = 13, SCIP_SOLVELP = 14, SCIP_FOUNDSOL = 15,
   SCIP_SUSPENDED = 16, SCIP_SUCCESS = 17, SCIP_DELAYNODE = 18
} SCIP_RESULT;

typedef unsigned int SCIP_Bool;
typedef double SCIP_Real;
typedef long long SCIP_Longint;
typedef unsigned int SCIP_PROPTIMING;
typedef unsigned int SCIP_PRESOLTIMING;
#define TRUE 1
#define FALSE 0

#define SCIP_PROPTIMING_BEFORELP     0x001u
#define SCIP_PRESOLTIMING_FAST       0x004u
#define SCIP_PRESOLTIMING_MEDIUM     0x008u

typedef struct Scip SCIP;
typedef struct SCIP_Var SCIP_VAR;
typedef struct SCIP_Cons SCIP_CONS;
typedef struct SCIP_Conshdlr SCIP_CONSHDLR;
typedef struct SCIP_Prop SCIP_PROP;
typedef struct SCIP_PropData SCIP_PROPDATA;
typedef struct SCIP_ParamData SCIP_PARAMDATA;
typedef struct SCIP_Param SCIP_PARAM;
typedef struct SCIP_BdChgIdx SCIP_BDCHGIDX;
typedef enum SCIP_BoundType { SCIP_BOUNDTYPE_LOWER = 0, SCIP_BOUNDTYPE_UPPER = 1 } SCIP_BOUNDTYPE;
typedef enum SCIP_Stage
{
   SCIP_STAGE_INIT = 0, SCIP_STAGE_PROBLEM = 1, SCIP_STAGE_TRANSFORMING = 2, SCIP_STAGE_TRANSFORMED = 3,
   SCIP_STAGE_INITPRESOLVE = 4, SCIP_STAGE_PRESOLVING = 5, SCIP_STAGE_EXITPRESOLVE = 6, SCIP_STAGE_PRESOLVED = 7,
   SCIP_STAGE_INITSOLVE = 8, SCIP_STAGE_SOLVING = 9, SCIP_STAGE_SOLVED = 10, SCIP_STAGE_EXITSOLVE = 11,
   SCIP_STAGE_FREETRANS = 12, SCIP_STAGE_FREE = 13
} SCIP_STAGE;   /* type_set.h:41-56 values */
typedef enum SCIP_ConflictType { SCIP_CONFTYPE_UNKNOWN = 0, SCIP_CONFTYPE_PROPAGATION = 1 } SCIP_CONFTYPE;

#define SCIP_DECL_PARAMCHGD(x) SCIP_RETCODE x (SCIP* scip, SCIP_PARAM* param)
#define SCIP_DECL_PROPFREE(x) SCIP_RETCODE x (SCIP* scip, SCIP_PROP* prop)
#define SCIP_DECL_PROPINITPRE(x) SCIP_RETCODE x (SCIP* scip, SCIP_PROP* prop)
#define SCIP_DECL_PROPEXITPRE(x) SCIP_RETCODE x (SCIP* scip, SCIP_PROP* prop)
#define SCIP_DECL_PROPINITSOL(x) SCIP_RETCODE x (SCIP* scip, SCIP_PROP* prop)
#define SCIP_DECL_PROPEXITSOL(x) SCIP_RETCODE x (SCIP* scip, SCIP_PROP* prop, SCIP_Bool restart)
#define SCIP_DECL_PROPEXEC(x) SCIP_RETCODE x (SCIP* scip, SCIP_PROP* prop, SCIP_PROPTIMING proptiming, SCIP_RESULT* result)
#define SCIP_DECL_PROPRESPROP(x) SCIP_RETCODE x (SCIP* scip, SCIP_PROP* prop, SCIP_VAR* infervar, int inferinfo, \
      SCIP_BOUNDTYPE boundtype, SCIP_BDCHGIDX* bdchgidx, SCIP_Real relaxedbd, SCIP_RESULT* result)
#define SCIP_DECL_PROPPRESOL(x) SCIP_RETCODE x (SCIP* scip, SCIP_PROP* prop, int nrounds, SCIP_PRESOLTIMING presoltiming, \
      int nnewfixedvars, int nnewaggrvars, int nnewchgvartypes, int nnewchgbds, int nnewholes, \
      int nnewdelconss, int nnewaddconss, int nnewupgdconss, int nnewchgcoefs, int nnewchgsides, \
      int* nfixedvars, int* naggrvars, int* nchgvartypes, int* nchgbds, int* naddholes, \
      int* ndelconss, int* naddconss, int* nupgdconss, int* nchgcoefs, int* nchgsides, SCIP_RESULT* result)

#define SCIP_CALL(x) do { SCIP_RETCODE _restat_ = (x); if( _restat_ != SCIP_OKAY ) return _restat_; } while( 0 )
#define SCIPerrorMessage(...) fprintf(stderr, __VA_ARGS__)

/* memory: plain heap allocation */
#define SCIPallocBlockMemory(scip, ptr) ((*(ptr) = malloc(sizeof(**(ptr)))) == NULL ? SCIP_NOMEMORY : SCIP_OKAY)
#define SCIPallocBlockMemoryArray(scip, ptr, num) \
   ((*(ptr) = malloc(sizeof(**(ptr)) * (size_t)((num) > 0 ? (num) : 1))) == NULL ? SCIP_NOMEMORY : SCIP_OKAY)
#define SCIPallocBufferArray(scip, ptr, num) SCIPallocBlockMemoryArray(scip, ptr, num)
#define SCIPfreeBlockMemory(scip, ptr) do { free(*(ptr)); *(ptr) = NULL; } while( 0 )
#define SCIPfreeBlockMemoryArray(scip, ptr, num) do { free(*(ptr)); *(ptr) = NULL; } while( 0 )
#define SCIPfreeBufferArray(scip, ptr) do { free(*(ptr)); *(ptr) = NULL; } while( 0 )

/* scip_prop.h */
SCIP_RETCODE SCIPincludePropBasic(SCIP* scip, SCIP_PROP** propptr, const char* name, const char* desc, int priority,
   int freq, SCIP_Bool delay, SCIP_PROPTIMING timingmask, SCIP_DECL_PROPEXEC((*propexec)), SCIP_PROPDATA* propdata);
SCIP_RETCODE SCIPsetPropFree(SCIP* scip, SCIP_PROP* prop, SCIP_DECL_PROPFREE((*propfree)));
SCIP_RETCODE SCIPsetPropInitpre(SCIP* scip, SCIP_PROP* prop, SCIP_DECL_PROPINITPRE((*propinitpre)));
SCIP_RETCODE SCIPsetPropExitpre(SCIP* scip, SCIP_PROP* prop, SCIP_DECL_PROPEXITPRE((*propexitpre)));
SCIP_RETCODE SCIPsetPropInitsol(SCIP* scip, SCIP_PROP* prop, SCIP_DECL_PROPINITSOL((*propinitsol)));
SCIP_RETCODE SCIPsetPropExitsol(SCIP* scip, SCIP_PROP* prop, SCIP_DECL_PROPEXITSOL((*propexitsol)));
SCIP_RETCODE SCIPsetPropPresol(SCIP* scip, SCIP_PROP* prop, SCIP_DECL_PROPPRESOL((*proppresol)), int presolpriority,
   int presolmaxrounds, SCIP_PRESOLTIMING presoltiming);
SCIP_RETCODE SCIPsetPropResprop(SCIP* scip, SCIP_PROP* prop, SCIP_DECL_PROPRESPROP((*propresprop)));
/* pub_prop.h */
SCIP_PROPDATA* SCIPpropGetData(SCIP_PROP* prop);
SCIP_PROP* SCIPfindProp(SCIP* scip, const char* name);
int SCIPpropGetFreq(SCIP_PROP* prop);
void SCIPpropSetData(SCIP_PROP* prop, SCIP_PROPDATA* propdata);
/* scip_param.h */
SCIP_RETCODE SCIPaddIntParam(SCIP* scip, const char* name, const char* desc, int* valueptr, SCIP_Bool isadvanced,
   int defaultvalue, int minvalue, int maxvalue, SCIP_DECL_PARAMCHGD((*paramchgd)), SCIP_PARAMDATA* paramdata);
SCIP_RETCODE SCIPaddBoolParam(SCIP* scip, const char* name, const char* desc, SCIP_Bool* valueptr,
   SCIP_Bool isadvanced, SCIP_Bool defaultvalue, SCIP_DECL_PARAMCHGD((*paramchgd)), SCIP_PARAMDATA* paramdata);
SCIP_RETCODE SCIPgetRealParam(SCIP* scip, const char* name, SCIP_Real* value);
SCIP_RETCODE SCIPgetBoolParam(SCIP* scip, const char* name, SCIP_Bool* value);
SCIP_RETCODE SCIPgetIntParam(SCIP* scip, const char* name, int* value);
/* scip_general.h / scip_probing.h / scip_tree.h */
SCIP_STAGE SCIPgetStage(SCIP* scip);
SCIP_Bool SCIPinProbing(SCIP* scip);
int SCIPgetDepth(SCIP* scip);
SCIP_Bool SCIPinRepropagation(SCIP* scip);
/* scip_conflict.h */
SCIP_Bool SCIPisConflictAnalysisApplicable(SCIP* scip);
SCIP_RETCODE SCIPinitConflictAnalysis(SCIP* scip, SCIP_CONFTYPE conftype, SCIP_Bool iscutoffinvolved);
SCIP_RETCODE SCIPaddConflictLb(SCIP* scip, SCIP_VAR* var, SCIP_BDCHGIDX* bdchgidx);
SCIP_RETCODE SCIPaddConflictUb(SCIP* scip, SCIP_VAR* var, SCIP_BDCHGIDX* bdchgidx);
SCIP_RETCODE SCIPanalyzeConflict(SCIP* scip, int validdepth, SCIP_Bool* success);
/* scip_numerics.h */
SCIP_Real SCIPinfinity(SCIP* scip);
SCIP_Real SCIPepsilon(SCIP* scip);
SCIP_Real SCIPsumepsilon(SCIP* scip);
SCIP_Real SCIPfeastol(SCIP* scip);
SCIP_Real SCIPgetHugeValue(SCIP* scip);
SCIP_Bool SCIPisInfinity(SCIP* scip, SCIP_Real val);
SCIP_Real SCIPfeasFloor(SCIP* scip, SCIP_Real val);
SCIP_Real SCIPfeasCeil(SCIP* scip, SCIP_Real val);
/* scip_cons.h / pub_cons.h */
SCIP_CONSHDLR* SCIPfindConshdlr(SCIP* scip, const char* name);
int SCIPconshdlrGetNConss(SCIP_CONSHDLR* conshdlr);
SCIP_CONS** SCIPconshdlrGetConss(SCIP_CONSHDLR* conshdlr);
SCIP_Bool SCIPconsIsActive(SCIP_CONS* cons);
/* scip_prob.h */
int SCIPgetNVars(SCIP* scip);
SCIP_VAR** SCIPgetVars(SCIP* scip);
/* pub_var.h / scip_var.h */
int SCIPvarGetProbindex(SCIP_VAR* var);
SCIP_Bool SCIPvarIsIntegral(SCIP_VAR* var);
SCIP_Bool SCIPvarIsBinary(SCIP_VAR* var);
SCIP_Real SCIPvarGetLbGlobal(SCIP_VAR* var);
SCIP_Real SCIPvarGetUbGlobal(SCIP_VAR* var);
SCIP_Real SCIPvarGetLbLocal(SCIP_VAR* var);
SCIP_Real SCIPvarGetUbLocal(SCIP_VAR* var);
SCIP_RETCODE SCIPgetProbvarLinearSum(SCIP* scip, SCIP_VAR** vars, SCIP_Real* scalars, int* nvars, int varssize,
   SCIP_Real* constant, int* requiredsize);
SCIP_RETCODE SCIPinferVarLbProp(SCIP* scip, SCIP_VAR* var, SCIP_Real newbound, SCIP_PROP* inferprop, int inferinfo,
   SCIP_Bool force, SCIP_Bool* infeasible, SCIP_Bool* tightened);
SCIP_RETCODE SCIPinferVarUbProp(SCIP* scip, SCIP_VAR* var, SCIP_Real newbound, SCIP_PROP* inferprop, int inferinfo,
   SCIP_Bool force, SCIP_Bool* infeasible, SCIP_Bool* tightened);
SCIP_RETCODE SCIPtightenVarLb(SCIP* scip, SCIP_VAR* var, SCIP_Real newbound, SCIP_Bool force, SCIP_Bool* infeasible,
   SCIP_Bool* tightened);
SCIP_RETCODE SCIPtightenVarUb(SCIP* scip, SCIP_VAR* var, SCIP_Real newbound, SCIP_Bool force, SCIP_Bool* infeasible,
   SCIP_Bool* tightened);
SCIP_RETCODE SCIPfixVar(SCIP* scip, SCIP_VAR* var, SCIP_Real fixedval, SCIP_Bool* infeasible, SCIP_Bool* fixed);

#ifdef __cplusplus
}
#endif

#endif
