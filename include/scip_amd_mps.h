/* scip_amd_mps.h -- MPS reader producing the engine's CSR input.
 *
 * Reads a linear (mixed-integer) program in free or fixed MPS format into a sap_instance
 * (include/scip_amd_gen.h): the linear rows with their sides, the column bounds and
 * integrality, as SCIP's MPS reader builds them (src/scip/reader_mps.c):
 *   - ROWS: the first N row is the objective and is not part of the system; further N rows are
 *     ignored; E/L/G rows start at [0,0] / [-inf,0] / [0,+inf] (readRows);
 *   - COLUMNS: columns between 'MARKER' 'INTORG' and 'INTEND' are integer with default bounds
 *     [0,1], the others continuous with [0,+inf) (readCols :1144-1155); coefficients with
 *     |a| <= 1e-9 are dropped (SCIPisZero), repeated entries of a row are summed;
 *   - RHS, RANGES, BOUNDS: only the first vector name of each section is used; RHS of the
 *     objective row is an offset and ignored; ranges follow the table of readRanges
 *     (:1697-1704); bound types UP LO FX FR MI PL BV LI UI SC with readBounds' conversion of
 *     a default-binary column into a general integer with upper bound +inf (:2089-2111);
 *   - semi-continuous columns (SC) get bounds [0, ub]; their bound disjunction is not a linear
 *     row and is not represented; sections other than the above (SOS, QMATRIX, ...) are
 *     skipped;
 *   - bounds of integer columns are rounded inward with the feasibility tolerance 1e-6.
 * Rows are returned with ascending column indices.
 */
#ifndef SCIP_AMD_MPS_H
#define SCIP_AMD_MPS_H

#include <stdint.h>

#include "scip_amd_gen.h"

#ifdef __cplusplus
extern "C" {
#endif

/* return codes besides those of scip_amd_prop.h: SCIP_READERROR and SCIP_NOFILE values */
#define SAP_READERROR -2
#define SAP_NOFILE    -4

/* read `path` into `out` (release with sap_instance_free); *nsemicont (may be NULL) receives
 * the number of semi-continuous columns */
int sap_read_mps(const char* path, sap_instance* out, int32_t* nsemicont);

#ifdef __cplusplus
}
#endif

#endif
