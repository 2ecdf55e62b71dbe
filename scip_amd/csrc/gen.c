/* gen.c -- seeded synthetic instances for the propagation engine (host, C). */
#include "scip_amd_gen.h"
#include "scip_amd_prop.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t* x)
{
   uint64_t z = (*x += UINT64_C(0x9E3779B97F4A7C15));
   z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
   z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
   return z ^ (z >> 31);
}

static void rng_seed(rng_t* r, uint64_t seed)
{
   uint64_t x = seed;
   int i;
   for( i = 0; i < 4; ++i )
      r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* xoshiro256** */
static inline uint64_t rng_next(rng_t* r)
{
   uint64_t* s = r->s;
   uint64_t result = rotl(s[1] * 5, 7) * 9;
   uint64_t t = s[1] << 17;
   s[2] ^= s[0];
   s[3] ^= s[1];
   s[1] ^= s[2];
   s[0] ^= s[3];
   s[2] ^= t;
   s[3] = rotl(s[3], 45);
   return result;
}

/* uniform in [0,1) */
static inline double rng_unif(rng_t* r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

/* uniform integer in {lo..hi} */
static inline int64_t rng_int(rng_t* r, int64_t lo, int64_t hi)
{
   uint64_t range = (uint64_t)(hi - lo) + 1;
   return lo + (int64_t)(rng_next(r) % range);
}

static double rng_normal(rng_t* r)
{
   double u1 = 1.0 - rng_unif(r);   /* (0,1] */
   double u2 = rng_unif(r);
   return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static int cmp_i32(const void* a, const void* b)
{
   int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
   return (x > y) - (x < y);
}

static int alloc_instance(sap_instance* o, int64_t nrows, int64_t ncols, int64_t nnzcap)
{
   memset(o, 0, sizeof(*o));
   o->nrows = nrows;
   o->ncols = ncols;
   o->rowptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nrows + 1));
   o->colidx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnzcap > 0 ? nnzcap : 1));
   o->vals = (double*)malloc(sizeof(double) * (size_t)(nnzcap > 0 ? nnzcap : 1));
   o->lhs = (double*)malloc(sizeof(double) * (size_t)(nrows > 0 ? nrows : 1));
   o->rhs = (double*)malloc(sizeof(double) * (size_t)(nrows > 0 ? nrows : 1));
   o->integral = (uint8_t*)malloc((size_t)(ncols > 0 ? ncols : 1));
   o->lb = (double*)malloc(sizeof(double) * (size_t)(ncols > 0 ? ncols : 1));
   o->ub = (double*)malloc(sizeof(double) * (size_t)(ncols > 0 ? ncols : 1));
   if( !o->rowptr || !o->colidx || !o->vals || !o->lhs || !o->rhs || !o->integral || !o->lb || !o->ub )
   {
      sap_instance_free(o);
      return SAP_NOMEMORY;
   }
   return SAP_OKAY;
}

void sap_instance_free(sap_instance* o)
{
   if( o == NULL )
      return;
   free(o->rowptr);
   free(o->colidx);
   free(o->vals);
   free(o->lhs);
   free(o->rhs);
   free(o->integral);
   free(o->lb);
   free(o->ub);
   memset(o, 0, sizeof(*o));
}

/* draw n distinct columns in [0, ncols) into dst (sorted ascending) */
static void draw_columns(rng_t* r, int64_t ncols, int64_t n, int64_t stampval, int64_t* stamp, int32_t* dst)
{
   int64_t k = 0;
   while( k < n )
   {
      int64_t c = rng_int(r, 0, ncols - 1);
      if( stamp[c] == stampval )
         continue;
      stamp[c] = stampval;
      dst[k++] = (int32_t)c;
   }
   qsort(dst, (size_t)n, sizeof(int32_t), cmp_i32);
}

int sap_gen_covering_packing(int64_t nrows, int64_t ncols, int32_t nnz_per_row, uint64_t seed, sap_instance* out)
{
   rng_t r;
   int64_t i, k, *stamp;
   int rc;
   if( out == NULL || nrows < 0 || ncols <= 0 || nnz_per_row <= 0 || nnz_per_row > ncols )
      return SAP_INVALIDDATA;
   rc = alloc_instance(out, nrows, ncols, nrows * nnz_per_row);
   if( rc != SAP_OKAY )
      return rc;
   stamp = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncols);
   if( stamp == NULL )
   {
      sap_instance_free(out);
      return SAP_NOMEMORY;
   }
   for( k = 0; k < ncols; ++k )
      stamp[k] = -1;
   rng_seed(&r, seed);
   out->rowptr[0] = 0;
   for( i = 0; i < nrows; ++i )
   {
      int64_t s = i * nnz_per_row;
      double sum = 0.0, side;
      draw_columns(&r, ncols, nnz_per_row, i, stamp, out->colidx + s);
      for( k = 0; k < nnz_per_row; ++k )
      {
         out->vals[s + k] = (double)rng_int(&r, 1, 100);
         sum += out->vals[s + k];
      }
      side = 0.4 * sum;
      if( rng_next(&r) & 1 )
      {
         out->lhs[i] = -1e20;
         out->rhs[i] = side;
      }
      else
      {
         out->lhs[i] = side;
         out->rhs[i] = 1e20;
      }
      out->rowptr[i + 1] = s + nnz_per_row;
   }
   out->nnz = nrows * nnz_per_row;
   for( k = 0; k < ncols; ++k )
   {
      out->integral[k] = 1;
      out->lb[k] = 0.0;
      out->ub[k] = 1.0;
   }
   free(stamp);
   return SAP_OKAY;
}

int sap_gen_mixed(int64_t nrows, int64_t ncols, int32_t lenmode, double median, double sigma, int32_t minlen,
   int32_t maxlen, int32_t sidemode, uint64_t seed, sap_instance* out)
{
   rng_t r;
   int64_t i, k, nnz, *stamp, *len;
   double* xstar = NULL;
   int rc;
   if( out == NULL || nrows < 0 || ncols <= 0 || minlen < 1 || maxlen < minlen || sidemode < 0 || sidemode > 1 )
      return SAP_INVALIDDATA;
   rng_seed(&r, seed);
   len = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nrows > 0 ? nrows : 1));
   if( len == NULL )
      return SAP_NOMEMORY;
   nnz = 0;
   for( i = 0; i < nrows; ++i )
   {
      int64_t L;
      if( lenmode == 0 )
      {
         double x = exp(log(median) + sigma * rng_normal(&r));
         L = (int64_t)floor(x + 0.5);
      }
      else
         L = rng_int(&r, minlen, maxlen);
      if( L < minlen )
         L = minlen;
      if( L > maxlen )
         L = maxlen;
      if( L > ncols )
         L = ncols;
      len[i] = L;
      nnz += L;
   }
   rc = alloc_instance(out, nrows, ncols, nnz);
   if( rc != SAP_OKAY )
   {
      free(len);
      return rc;
   }
   stamp = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncols);
   if( stamp == NULL )
   {
      free(len);
      sap_instance_free(out);
      return SAP_NOMEMORY;
   }
   for( k = 0; k < ncols; ++k )
      stamp[k] = -1;

   /* columns first, so row sides can use their bounds */
   for( k = 0; k < ncols; ++k )
   {
      if( rng_unif(&r) < 0.4 )
      {
         out->integral[k] = 1;
         out->lb[k] = 0.0;
         out->ub[k] = (double)rng_int(&r, 1, 1000);
      }
      else
      {
         double u = rng_unif(&r);
         out->integral[k] = 0;
         out->lb[k] = -1e4 * rng_unif(&r);
         out->ub[k] = 1e4 * rng_unif(&r);
         if( u < 0.025 )
            out->lb[k] = -1e20;
         else if( u < 0.05 )
            out->ub[k] = 1e20;
      }
   }

   if( sidemode == 1 )
   {
      xstar = (double*)malloc(sizeof(double) * (size_t)ncols);
      if( xstar == NULL )
      {
         free(stamp);
         free(len);
         sap_instance_free(out);
         return SAP_NOMEMORY;
      }
      for( k = 0; k < ncols; ++k )
      {
         double lo = out->lb[k], hi = out->ub[k];
         if( out->integral[k] )
            xstar[k] = (double)rng_int(&r, (int64_t)lo, (int64_t)hi);
         else if( lo <= -1e20 )
            xstar[k] = hi - 1e4 * rng_unif(&r);
         else if( hi >= 1e20 )
            xstar[k] = lo + 1e4 * rng_unif(&r);
         else
            xstar[k] = lo + (hi - lo) * rng_unif(&r);
      }
   }

   out->rowptr[0] = 0;
   for( i = 0; i < nrows; ++i )
   {
      int64_t s = out->rowptr[i];
      double mn = 0.0, mx = 0.0, ax = 0.0, c, S;
      draw_columns(&r, ncols, len[i], i, stamp, out->colidx + s);
      for( k = 0; k < len[i]; ++k )
      {
         int32_t j = out->colidx[s + k];
         double mag = pow(10.0, -3.0 + 6.0 * rng_unif(&r));
         double a = (rng_next(&r) & 1) ? mag : -mag;
         double bmin = a > 0.0 ? out->lb[j] : out->ub[j];
         double bmax = a > 0.0 ? out->ub[j] : out->lb[j];
         out->vals[s + k] = a;
         if( fabs(bmin) < 1e20 )
            mn += a * bmin;
         if( fabs(bmax) < 1e20 )
            mx += a * bmax;
         if( xstar != NULL )
            ax += a * xstar[j];
      }
      c = (xstar != NULL) ? ax : 0.5 * (mn + mx);
      S = mx - mn;
      out->lhs[i] = c - 0.3 * S;
      out->rhs[i] = c + 0.3 * S;
      out->rowptr[i + 1] = s + len[i];
   }
   out->nnz = nnz;
   free(xstar);
   free(stamp);
   free(len);
   return SAP_OKAY;
}

int sap_gen_node_fixings(const sap_instance* inst, int32_t nnodes, int32_t minfix, int32_t maxfix, uint64_t seed,
   int64_t* fixptr, int32_t* fixcol, double* fixval)
{
   rng_t r;
   int64_t k, nint = 0, *stamp;
   int32_t* intcols;
   int32_t n;
   if( inst == NULL || fixptr == NULL || fixcol == NULL || fixval == NULL || nnodes < 0 || minfix < 0
      || maxfix < minfix )
      return SAP_INVALIDDATA;
   intcols = (int32_t*)malloc(sizeof(int32_t) * (size_t)(inst->ncols > 0 ? inst->ncols : 1));
   stamp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(inst->ncols > 0 ? inst->ncols : 1));
   if( intcols == NULL || stamp == NULL )
   {
      free(intcols);
      free(stamp);
      return SAP_NOMEMORY;
   }
   for( k = 0; k < inst->ncols; ++k )
   {
      stamp[k] = -1;
      if( inst->integral[k] )
         intcols[nint++] = (int32_t)k;
   }
   if( nint < maxfix )
   {
      free(intcols);
      free(stamp);
      return SAP_INVALIDDATA;
   }
   rng_seed(&r, seed);
   fixptr[0] = 0;
   for( n = 0; n < nnodes; ++n )
   {
      int64_t F = rng_int(&r, minfix, maxfix), q = 0, base = fixptr[n];
      while( q < F )
      {
         int32_t c = intcols[rng_int(&r, 0, nint - 1)];
         if( stamp[c] == n )
            continue;
         stamp[c] = n;
         fixcol[base + q] = c;
         fixval[base + q] = (rng_next(&r) & 1) ? 1.0 : 0.0;
         ++q;
      }
      fixptr[n + 1] = base + F;
   }
   free(intcols);
   free(stamp);
   return SAP_OKAY;
}
