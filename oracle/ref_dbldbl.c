/* ref_dbldbl.c -- exports the reference's own double-double macros for the oracle checks.
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled by oracle/build.py against the unmodified header
 * src/scip/dbldblarith.h where it lies in the SCIP source tree (the header is
 * self-contained: it includes only <math.h>).  The result, oracle/_ref/libdbldbl_ref.so,
 * pins the oracle's TwoSum / Sum21 / TwoProd restatement bit for bit.
 */
#include "scip/dbldblarith.h"

void ref_dd_sum(double a, double b, double* rhi, double* rlo)
{
   double hi, lo;
   SCIPdbldblSum(hi, lo, a, b);
   *rhi = hi;
   *rlo = lo;
}

void ref_dd_sum21(double ahi, double alo, double b, double* rhi, double* rlo)
{
   double hi, lo;
   SCIPdbldblSum21(hi, lo, ahi, alo, b);
   *rhi = hi;
   *rlo = lo;
}

void ref_dd_sum22(double ahi, double alo, double bhi, double blo, double* rhi, double* rlo)
{
   double hi, lo;
   SCIPdbldblSum22(hi, lo, ahi, alo, bhi, blo);
   *rhi = hi;
   *rlo = lo;
}

void ref_dd_prod(double a, double b, double* rhi, double* rlo)
{
   double hi, lo;
   SCIPdbldblProd(hi, lo, a, b);
   *rhi = hi;
   *rlo = lo;
}

void ref_dd_prod21(double ahi, double alo, double b, double* rhi, double* rlo)
{
   double hi, lo;
   SCIPdbldblProd21(hi, lo, ahi, alo, b);
   *rhi = hi;
   *rlo = lo;
}

/* sequential activity sum exactly as consdataCalcActivities accumulates it:
 * QUAD_ASSIGN(act, 0.0); for each nonzero contribution: SCIPquadprecSumQD(act, act, c) */
void ref_activity_sum(const double* x, long n, double* rhi, double* rlo)
{
   double acthi = 0.0, actlo = 0.0;
   long i;
   for( i = 0; i < n; ++i )
      if( x[i] != 0.0 )
         SCIPdbldblSum21(acthi, actlo, acthi, actlo, x[i]);
   *rhi = acthi;
   *rlo = actlo;
}
