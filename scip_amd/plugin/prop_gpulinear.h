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

/** batched propagation dives (the fix-and-propagate loop of heur_fixandinfer.c:255-280 and of
 *  SCIP's diving heuristics, SCIPfixVarProbing / SCIPpropagateProbing per step): ndives dives
 *  from the node's propagated bounds advance together, one batched device call per step; dive k
 *  fixes, in its order order[k*ncands ...], the next candidate still unfixed to
 *  fixvals[k*ncands + c] (nearest integer for an integral variable, moved into the current
 *  domain) and propagates, until a cutoff, no unfixed candidate or maxdepth fixings (-1: no
 *  limit); per dive the fixings made, whether it ended in a cutoff and (divelbs / diveubs, or
 *  NULL) its final bounds, indexed by SCIPvarGetProbindex */
SCIP_EXPORT
SCIP_RETCODE SCIPgpulinearDive(
   SCIP*                 scip,               /**< SCIP data structure */
   int                   ndives,             /**< number of dives */
   int                   ncands,             /**< number of candidates */
   SCIP_VAR**            cands,              /**< candidate variables */
   const int*            order,              /**< ndives x ncands: the order in which dive k visits them */
   const SCIP_Real*      fixvals,            /**< ndives x ncands: the value dive k fixes candidate c to */
   int                   maxdepth,           /**< most fixings per dive (-1: no limit) */
   int*                  depth,              /**< array of ndives: fixings made */
   SCIP_Bool*            cutoff,             /**< array of ndives: the dive ended in a cutoff */
   SCIP_Real*            divelbs,            /**< ndives x active variables: final lower bounds, or NULL */
   SCIP_Real*            diveubs             /**< ndives x active variables: final upper bounds, or NULL */
   );

#ifdef __cplusplus
}
#endif

#endif
