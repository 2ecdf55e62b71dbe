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

#ifdef __cplusplus
}
#endif

#endif
