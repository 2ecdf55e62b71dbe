/* Test double of scip/prop_probing.h: the one function of SCIP's probing propagator that
 * prop_gpulinear.c calls, with the signature of src/scip/prop_probing.h. */
#ifndef SCIP_AMD_TESTHOST_PROP_PROBING_H
#define SCIP_AMD_TESTHOST_PROP_PROBING_H

#include "scip/scip.h"

#ifdef __cplusplus
extern "C" {
#endif

SCIP_RETCODE SCIPanalyzeDeductionsProbing(SCIP* scip, SCIP_VAR* probingvar, SCIP_Real leftub, SCIP_Real rightlb,
   int nvars, SCIP_VAR** vars, SCIP_Real* leftimpllbs, SCIP_Real* leftimplubs, SCIP_Real* leftproplbs,
   SCIP_Real* leftpropubs, SCIP_Real* rightimpllbs, SCIP_Real* rightimplubs, SCIP_Real* rightproplbs,
   SCIP_Real* rightpropubs, int* nfixedvars, int* naggrvars, int* nimplications, int* nchgbds, SCIP_Bool* cutoff);

#ifdef __cplusplus
}
#endif

#endif
