/* scip_amd_gen.h -- seeded synthetic instance generators for the propagation engine.
 *
 * These produce the benchmark workloads as plain CSR arrays (host memory owned by the
 * library; release with sap_instance_free).  They have no counterpart in SCIP, whose
 * instances come from readers (src/scip/reader_mps.c) or the C API
 * (SCIPcreateConsBasicLinear, src/scip/cons_linear.c); the generated rows are what such a
 * caller would hand to SCIPcreateConsBasicLinear.
 *
 * Random numbers: xoshiro256** seeded through splitmix64; identical seeds give identical
 * instances on every x86-64 host with the same libm.
 */
#ifndef SCIP_AMD_GEN_H
#define SCIP_AMD_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sap_instance
{
   int64_t  nrows;
   int64_t  ncols;
   int64_t  nnz;
   int64_t* rowptr;      /* nrows+1 */
   int32_t* colidx;      /* nnz, ascending within a row, distinct */
   double*  vals;        /* nnz, nonzero */
   double*  lhs;         /* nrows, -1e20 for none */
   double*  rhs;         /* nrows, +1e20 for none */
   uint8_t* integral;    /* ncols */
   double*  lb;          /* ncols, global lower bounds */
   double*  ub;          /* ncols, global upper bounds */
} sap_instance;

/* Pure-binary covering/packing rows: nnz_per_row distinct uniform columns per row,
 * coefficients uniform in {1..100}, side = 0.4 * row coefficient sum, each row packing
 * (a x <= side) or covering (a x >= side) with probability 1/2, bounds [0,1]. */
int sap_gen_covering_packing(int64_t nrows, int64_t ncols, int32_t nnz_per_row, uint64_t seed,
   sap_instance* out);

/* Mixed integer/continuous rows.
 *   lenmode 0: row length round(exp(ln(median) + sigma*N(0,1))) clamped to [minlen, maxlen];
 *   lenmode 1: row length uniform in {minlen..maxlen}.
 * 40% integer columns with bounds [0, U{1..1000}]; 60% continuous with bounds
 * [-U[0,1e4], U[0,1e4]], 5% of those with one side infinite (half each side).
 * Coefficients +-10^U[-3,3].  Ranged rows: with [m, M] the finite part of the activity
 * range and S = M-m, lhs = c - 0.3 S, rhs = c + 0.3 S, where
 *   sidemode 0: c = (m+M)/2 (jointly infeasible with high probability beyond ~10^4 rows);
 *   sidemode 1: c = a.x* for a planted point x* drawn uniformly within the bounds (finite
 *               sides of one-sided columns +- U[0,1e4]), so the system is feasible. */
int sap_gen_mixed(int64_t nrows, int64_t ncols, int32_t lenmode, double median, double sigma, int32_t minlen,
   int32_t maxlen, int32_t sidemode, uint64_t seed, sap_instance* out);

/* Node bound vectors: node k fixes F_k in U{minfix..maxfix} distinct random integral columns
 * to 0 or 1 (p = 1/2), as lb = ub = value.  fixptr has nnodes+1 entries; fixcol/fixval
 * must hold nnodes*maxfix entries. */
int sap_gen_node_fixings(const sap_instance* inst, int32_t nnodes, int32_t minfix, int32_t maxfix, uint64_t seed,
   int64_t* fixptr, int32_t* fixcol, double* fixval);

void sap_instance_free(sap_instance* inst);

#ifdef __cplusplus
}
#endif

#endif
