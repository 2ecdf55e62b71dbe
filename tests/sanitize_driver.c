/* sanitize_driver.c -- host code under AddressSanitizer and UndefinedBehaviorSanitizer.
 *
 * Built by tests/test_host_sanitizers.py with gcc -fsanitize=address,undefined together with
 * the plain-C host sources (scip_amd/csrc/mps.c, scip_amd/csrc/gen.c, oracle/oracle.c); no GPU
 * code is involved.  It reads every MPS file named on the command line (and the malformed
 * inputs the test writes), generates instances with every generator, propagates them with
 * both oracle schedules, and exits 0 when everything ran; a sanitizer report aborts it.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "scip_amd_gen.h"
#include "scip_amd_mps.h"

static int propagate_both(const sap_instance* inst)
{
   orc_matrix m;
   orc_params p;
   orc_result r;
   uint8_t* binary = (uint8_t*)malloc((size_t)(inst->ncols > 0 ? inst->ncols : 1));
   double* lb = (double*)malloc(sizeof(double) * (size_t)(inst->ncols > 0 ? inst->ncols : 1));
   double* ub = (double*)malloc(sizeof(double) * (size_t)(inst->ncols > 0 ? inst->ncols : 1));
   int k;
   if( binary == NULL || lb == NULL || ub == NULL )
      return 1;
   orc_compute_binary(inst->ncols, inst->integral, inst->lb, inst->ub, binary);
   m.nrows = inst->nrows;
   m.ncols = inst->ncols;
   m.rowptr = inst->rowptr;
   m.colidx = inst->colidx;
   m.vals = inst->vals;
   m.lhs = inst->lhs;
   m.rhs = inst->rhs;
   m.integral = inst->integral;
   m.binary = binary;
   for( k = 0; k < 3; ++k )
   {
      orc_params_default(&p);
      p.rangedrows = (k == 1);
      p.notighten = (k == 2);
      memcpy(lb, inst->lb, sizeof(double) * (size_t)inst->ncols);
      memcpy(ub, inst->ub, sizeof(double) * (size_t)inst->ncols);
      orc_propagate_jacobi(&m, &p, lb, ub, 0, &r);
      memcpy(lb, inst->lb, sizeof(double) * (size_t)inst->ncols);
      memcpy(ub, inst->ub, sizeof(double) * (size_t)inst->ncols);
      orc_propagate_scip(&m, &p, lb, ub, &r);
   }
   free(binary);
   free(lb);
   free(ub);
   return 0;
}

int main(int argc, char** argv)
{
   sap_instance inst;
   int64_t* fixptr;
   int32_t* fixcol;
   double* fixval;
   int i;
   int bad = 0;

   /* every MPS file given: well-formed ones must read, malformed ones must fail cleanly */
   for( i = 1; i < argc; ++i )
   {
      int32_t nsc = 0;
      int rc;
      memset(&inst, 0, sizeof(inst));
      rc = sap_read_mps(argv[i], &inst, &nsc);
      printf("%s: rc %d rows %lld cols %lld nnz %lld\n", argv[i], rc, (long long)inst.nrows, (long long)inst.ncols,
         (long long)inst.nnz);
      if( rc == 1 )
      {
         bad |= propagate_both(&inst);
         sap_instance_free(&inst);
      }
   }

   /* generators, and the oracle on what they produce */
   memset(&inst, 0, sizeof(inst));
   if( sap_gen_covering_packing(400, 900, 10, 1, &inst) != 1 )
      return 2;
   fixptr = (int64_t*)malloc(sizeof(int64_t) * 17);
   fixcol = (int32_t*)malloc(sizeof(int32_t) * 16 * 50);
   fixval = (double*)malloc(sizeof(double) * 16 * 50);
   if( sap_gen_node_fixings(&inst, 16, 1, 50, 3, fixptr, fixcol, fixval) != 1 )
      return 3;
   bad |= propagate_both(&inst);
   sap_instance_free(&inst);
   free(fixptr);
   free(fixcol);
   free(fixval);
   for( i = 0; i < 2; ++i )
   {
      memset(&inst, 0, sizeof(inst));
      if( sap_gen_mixed(3000, 3000, 0, 8.0, 1.0, 1, 10000, i, 2, &inst) != 1 )
         return 4;
      bad |= propagate_both(&inst);
      sap_instance_free(&inst);
      memset(&inst, 0, sizeof(inst));
      if( sap_gen_mixed(800, 1200, 1, 8.0, 1.0, 4, 64, 1, 4000 + i, &inst) != 1 )
         return 5;
      bad |= propagate_both(&inst);
      sap_instance_free(&inst);
   }
   printf("sanitize driver done\n");
   return bad;
}
