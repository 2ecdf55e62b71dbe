/* prop_gpulinear.h -- SCIP propagator running linear-constraint domain propagation on an
 * AMD MI355X through the scip_amd engine (include/scip_amd_prop.h).
 *
 * Drop-in for the activity and bound-tightening path of the linear constraint handler
 * (src/scip/cons_linear.c consPropLinear :16126 and the presolving bound tightening of
 * consPresolLinear :16356): with constraints/linear/propfreq = -1 the linear handler stops
 * propagating and this propagator takes over, registered through SCIPincludePropBasic
 * (src/scip/scip_prop.c:118) and SCIPsetPropPresol (:283).
 */
#ifndef __SCIP_PROP_GPULINEAR_H__
#define __SCIP_PROP_GPULINEAR_H__

#include "scip/scip.h"

#ifdef __cplusplus
extern "C" {
#endif

/** creates the gpulinear propagator and includes it in SCIP */
SCIP_EXPORT
SCIP_RETCODE SCIPincludePropGpulinear(
   SCIP*                 scip                /**< SCIP data structure */
   );

/** the propagation part of strong branching (SCIPgetVarStrongbranchWithPropagation,
 *  src/scip/scip_var.c:4139-4440, without the LP) for many candidates at once on the device:
 *  per candidate whether the down child x <= feasCeil(solval - 1) and the up child
 *  x >= feasFloor(solval + 1) are infeasible, and the bounds valid at the node that follow
 *  (the union of both children's bounds, the feasible child's branching bound when the other
 *  child is infeasible, branch_relpscost.c:1992-2070), indexed by SCIPvarGetProbindex;
 *  *cutoff when some candidate has two infeasible children */
SCIP_EXPORT
SCIP_RETCODE SCIPgpulinearStrongbranch(
   SCIP*                 scip,               /**< SCIP data structure */
   int                   ncands,             /**< number of candidates */
   SCIP_VAR**            cands,              /**< integral candidate variables */
   const SCIP_Real*      solvals,            /**< LP values of the candidates */
   SCIP_Bool*            downinf,            /**< array of ncands: the down child is infeasible */
   SCIP_Bool*            upinf,              /**< array of ncands: the up child is infeasible */
   SCIP_Real*            newlbs,             /**< array of the active variables: valid lower bounds, or NULL */
   SCIP_Real*            newubs,             /**< array of the active variables: valid upper bounds, or NULL */
   SCIP_Bool*            cutoff              /**< pointer to store whether the node is infeasible */
   );

#ifdef __cplusplus
}
#endif

#endif
