/* mps.c -- MPS reader (see include/scip_amd_mps.h for the semantics, which are those of
 * src/scip/reader_mps.c).  Host code; fields are whitespace-separated (free MPS; fixed MPS
 * files without blanks inside names read the same way). */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "scip_amd_mps.h"

#define INF 1e20
#define MAXLINE 4096
#define MAXFIELDS 8

/* ---- string -> index hash map (open addressing) ---- */
typedef struct
{
   char** keys;
   int32_t* vals;
   int64_t cap;
   int64_t n;
} strmap;

static uint64_t hashStr(const char* s)
{
   uint64_t h = 1469598103934665603ULL;   /* FNV-1a */
   while( *s )
      h = (h ^ (unsigned char)*s++) * 1099511628211ULL;
   return h;
}

static int mapInit(strmap* m, int64_t cap)
{
   m->cap = cap;
   m->n = 0;
   m->keys = (char**)calloc((size_t)cap, sizeof(char*));
   m->vals = (int32_t*)calloc((size_t)cap, sizeof(int32_t));
   return (m->keys && m->vals) ? 0 : -1;
}

static void mapFree(strmap* m)
{
   for( int64_t i = 0; m->keys && i < m->cap; ++i )
      free(m->keys[i]);
   free(m->keys);
   free(m->vals);
}

static int32_t mapGet(const strmap* m, const char* k)
{
   uint64_t i = hashStr(k) & (uint64_t)(m->cap - 1);
   while( m->keys[i] != NULL )
   {
      if( strcmp(m->keys[i], k) == 0 )
         return m->vals[i];
      i = (i + 1) & (uint64_t)(m->cap - 1);
   }
   return -1;
}

static int mapPut(strmap* m, const char* k, int32_t v)
{
   if( 2 * (m->n + 1) > m->cap )
   {
      strmap g;
      if( mapInit(&g, 2 * m->cap) )
         return -1;
      for( int64_t i = 0; i < m->cap; ++i )
         if( m->keys[i] )
         {
            uint64_t j = hashStr(m->keys[i]) & (uint64_t)(g.cap - 1);
            while( g.keys[j] )
               j = (j + 1) & (uint64_t)(g.cap - 1);
            g.keys[j] = m->keys[i];
            g.vals[j] = m->vals[i];
         }
      g.n = m->n;
      free(m->keys);
      free(m->vals);
      *m = g;
   }
   uint64_t i = hashStr(k) & (uint64_t)(m->cap - 1);
   while( m->keys[i] != NULL )
      i = (i + 1) & (uint64_t)(m->cap - 1);
   m->keys[i] = strdup(k);
   if( m->keys[i] == NULL )
      return -1;
   m->vals[i] = v;
   m->n++;
   return 0;
}

/* ---- growable arrays ---- */
#define GROW(ptr, cnt, cap, type)                                           \
   do                                                                      \
   {                                                                       \
      if( (cnt) >= (cap) )                                                 \
      {                                                                    \
         int64_t _nc = (cap) > 0 ? 2 * (cap) : 1024;                       \
         type* _p = (type*)realloc((ptr), sizeof(type) * (size_t)_nc);     \
         if( _p == NULL )                                                  \
            goto nomem;                                                    \
         (ptr) = _p;                                                       \
         (cap) = _nc;                                                      \
      }                                                                    \
   } while( 0 )

typedef struct
{
   int32_t row, col;
   double val;
} triplet;

static int cmpTriplet(const void* a, const void* b)
{
   const triplet* x = (const triplet*)a;
   const triplet* y = (const triplet*)b;
   if( x->row != y->row )
      return x->row < y->row ? -1 : 1;
   if( x->col != y->col )
      return x->col < y->col ? -1 : 1;
   return 0;
}

enum { SEC_NONE, SEC_NAME, SEC_ROWS, SEC_COLUMNS, SEC_RHS, SEC_RANGES, SEC_BOUNDS, SEC_SKIP, SEC_END };

/* column type while reading: 0 continuous, 1 binary (integer with default bounds), 2 integer */
int sap_read_mps(const char* path, sap_instance* out, int32_t* nsemicont)
{
   FILE* f;
   char line[MAXLINE];
   char* fld[MAXFIELDS];
   strmap rows, cols;
   char objname[MAXLINE] = "";
   char rhsname[MAXLINE] = "", rngname[MAXLINE] = "", bndname[MAXLINE] = "";
   int sec = SEC_NONE, intmode = 0, rc = 1;
   int64_t nrows = 0, rowcap = 0, ncols = 0, colcap = 0, ntrip = 0, tripcap = 0;
   char* rtype = NULL;          /* 'E', 'L', 'G' per row */
   double *lhs = NULL, *rhs = NULL, *lb = NULL, *ub = NULL;
   int8_t* ctype = NULL;
   uint8_t* semic = NULL;
   triplet* trip = NULL;
   int32_t lastcol = -1;
   int32_t nsc = 0;

   if( path == NULL || out == NULL )
      return -9;
   memset(out, 0, sizeof(*out));
   f = fopen(path, "r");
   if( f == NULL )
      return SAP_NOFILE;
   if( mapInit(&rows, 1024) || mapInit(&cols, 1024) )
   {
      fclose(f);
      return -1;
   }
   while( fgets(line, sizeof(line), f) != NULL )
   {
      int nf = 0;
      char* p = line;
      if( line[0] == '*' || line[0] == '\n' || line[0] == '\r' )
         continue;
      const int header = !isspace((unsigned char)line[0]);
      while( nf < MAXFIELDS )
      {
         while( *p && isspace((unsigned char)*p) )
            ++p;
         if( !*p )
            break;
         fld[nf++] = p;
         while( *p && !isspace((unsigned char)*p) )
            ++p;
         if( *p )
            *p++ = '\0';
      }
      if( nf == 0 )
         continue;
      if( header )
      {
         if( !strcmp(fld[0], "NAME") )
            sec = SEC_NAME;
         else if( !strcmp(fld[0], "ROWS") )
            sec = SEC_ROWS;
         else if( !strcmp(fld[0], "COLUMNS") )
            sec = SEC_COLUMNS;
         else if( !strcmp(fld[0], "RHS") )
            sec = SEC_RHS;
         else if( !strcmp(fld[0], "RANGES") )
            sec = SEC_RANGES;
         else if( !strcmp(fld[0], "BOUNDS") )
            sec = SEC_BOUNDS;
         else if( !strcmp(fld[0], "ENDATA") )
         {
            sec = SEC_END;
            break;
         }
         else
            sec = SEC_SKIP;   /* OBJSENSE, SOS, QMATRIX, ... */
         continue;
      }
      switch( sec )
      {
      case SEC_ROWS:
      {
         if( nf < 2 )
            goto readerr;
         const char t = (char)toupper((unsigned char)fld[0][0]);
         if( t == 'N' )
         {
            if( objname[0] == '\0' )
               strcpy(objname, fld[1]);
            break;   /* further free rows are ignored */
         }
         if( t != 'E' && t != 'L' && t != 'G' )
            goto readerr;
         GROW(rtype, nrows, rowcap, char);
         rtype[nrows] = t;
         if( mapPut(&rows, fld[1], (int32_t)nrows) )
            goto nomem;
         ++nrows;
         break;
      }
      case SEC_COLUMNS:
      {
         if( nf >= 3 && !strcmp(fld[1], "'MARKER'") )
         {
            if( !strcmp(fld[2], "'INTORG'") )
               intmode = 1;
            else if( !strcmp(fld[2], "'INTEND'") )
               intmode = 0;
            else
               goto readerr;
            break;
         }
         if( nf != 3 && nf != 5 )
            goto readerr;
         int32_t c = mapGet(&cols, fld[0]);
         if( c < 0 )
         {
            const int64_t oldcap = colcap;
            GROW(lb, ncols, colcap, double);
            if( colcap != oldcap )
            {
               double* nub = (double*)realloc(ub, sizeof(double) * (size_t)colcap);
               int8_t* nct = (int8_t*)realloc(ctype, (size_t)colcap);
               uint8_t* nsm = (uint8_t*)realloc(semic, (size_t)colcap);
               if( !nub || !nct || !nsm )
               {
                  if( nub ) ub = nub;
                  if( nct ) ctype = nct;
                  if( nsm ) semic = nsm;
                  goto nomem;
               }
               ub = nub;
               ctype = nct;
               semic = nsm;
            }
            c = (int32_t)ncols;
            lb[c] = 0.0;
            ub[c] = intmode ? 1.0 : INF;
            ctype[c] = intmode ? 1 : 0;
            semic[c] = 0;
            if( mapPut(&cols, fld[0], c) )
               goto nomem;
            ++ncols;
         }
         else if( c != lastcol )
            goto readerr;   /* the entries of a column must be consecutive */
         lastcol = c;
         for( int q = 1; q + 1 < nf; q += 2 )
         {
            const double v = atof(fld[q + 1]);
            if( !strcmp(fld[q], objname) )
               continue;
            const int32_t r = mapGet(&rows, fld[q]);
            if( r < 0 || fabs(v) <= 1e-9 )
               continue;   /* unknown (free) row or a zero coefficient */
            GROW(trip, ntrip, tripcap, triplet);
            trip[ntrip].row = r;
            trip[ntrip].col = c;
            trip[ntrip].val = v;
            ++ntrip;
         }
         break;
      }
      case SEC_RHS:
      case SEC_RANGES:
      {
         /* "name row value [row value]"; a missing vector name is tolerated */
         char* fv[MAXFIELDS + 1];
         int nv = nf;
         if( nf == 2 || nf == 4 )
         {
            fv[0] = NULL;   /* fv[1..] are the row/value pairs */
            for( int q = 0; q < nf; ++q )
               fv[q + 1] = fld[q];
            nv = nf + 1;
         }
         else if( nf == 3 || nf == 5 )
         {
            for( int q = 0; q < nf; ++q )
               fv[q] = fld[q];
         }
         else
            goto readerr;
         char* vecname = (nf == 2 || nf == 4) ? (char*)"" : fld[0];
         char* first = sec == SEC_RHS ? rhsname : rngname;
         if( first[0] == '\0' )
            strcpy(first, vecname[0] ? vecname : " ");
         if( strcmp(first, vecname[0] ? vecname : " ") != 0 )
            break;   /* only the first vector is read */
         for( int q = 1; q + 1 < nv; q += 2 )
         {
            const int32_t r = mapGet(&rows, fv[q]);
            const double v = atof(fv[q + 1]);
            if( r < 0 )
               continue;   /* objective row (offset) or unknown row */
            if( sec == SEC_RHS )
            {
               if( lhs == NULL )
               {
                  lhs = (double*)malloc(sizeof(double) * (size_t)(nrows > 0 ? nrows : 1));
                  rhs = (double*)malloc(sizeof(double) * (size_t)(nrows > 0 ? nrows : 1));
                  if( !lhs || !rhs )
                     goto nomem;
                  for( int64_t i = 0; i < nrows; ++i )
                  {
                     lhs[i] = rtype[i] == 'L' ? -INF : 0.0;
                     rhs[i] = rtype[i] == 'G' ? INF : 0.0;
                  }
               }
               if( rtype[r] == 'L' )
                  rhs[r] = v;
               else if( rtype[r] == 'G' )
                  lhs[r] = v;
               else
                  lhs[r] = rhs[r] = v;
            }
            else
            {
               if( lhs == NULL )
               {
                  lhs = (double*)malloc(sizeof(double) * (size_t)(nrows > 0 ? nrows : 1));
                  rhs = (double*)malloc(sizeof(double) * (size_t)(nrows > 0 ? nrows : 1));
                  if( !lhs || !rhs )
                     goto nomem;
                  for( int64_t i = 0; i < nrows; ++i )
                  {
                     lhs[i] = rtype[i] == 'L' ? -INF : 0.0;
                     rhs[i] = rtype[i] == 'G' ? INF : 0.0;
                  }
               }
               /* readRanges: the row sense is read from the current sides */
               if( lhs[r] <= -INF )
                  lhs[r] = rhs[r] - fabs(v);
               else if( rhs[r] >= INF )
                  rhs[r] = lhs[r] + fabs(v);
               else if( v >= 0.0 )
                  rhs[r] = rhs[r] + v;
               else
                  lhs[r] = lhs[r] + v;
            }
         }
         break;
      }
      case SEC_BOUNDS:
      {
         /* "type name column [value]" (a missing vector name is tolerated) */
         if( nf < 2 )
            goto readerr;
         char type[3] = { (char)toupper((unsigned char)fld[0][0]),
            fld[0][1] ? (char)toupper((unsigned char)fld[0][1]) : '\0', '\0' };
         const int noval = !strcmp(type, "FR") || !strcmp(type, "MI") || !strcmp(type, "PL") || !strcmp(type, "BV");
         const char* vec;
         const char* colname;
         const char* valstr;
         if( noval )
         {
            vec = nf >= 3 ? fld[1] : " ";
            colname = nf >= 3 ? fld[2] : fld[1];
            valstr = (nf >= 4) ? fld[3] : NULL;
         }
         else
         {
            if( nf == 3 )
            {
               vec = " ";
               colname = fld[1];
               valstr = fld[2];
            }
            else if( nf >= 4 )
            {
               vec = fld[1];
               colname = fld[2];
               valstr = fld[3];
            }
            else
               goto readerr;
         }
         if( bndname[0] == '\0' )
            strcpy(bndname, vec);
         if( strcmp(bndname, vec) != 0 )
            break;   /* only the first bound vector is read */
         int32_t c = mapGet(&cols, colname);
         if( c < 0 )
            break;   /* a column without entries: not part of the system */
         const double val = valstr ? atof(valstr) : 0.0;
         /* a default-binary column becomes a general integer with ub = +inf unless the bound
          * keeps it binary (BV, UP <= 1, FX in [0,1]) (readBounds :2089-2111) */
         if( ctype[c] == 1 && strcmp(type, "BV") != 0 && (type[0] != 'U' || val > 1.0 + 1e-6)
            && (strcmp(type, "FX") != 0 || val < -1e-6 || val > 1.0 + 1e-6) )
         {
            ctype[c] = 2;
            ub[c] = INF;
         }
         if( type[0] == 'L' )
         {
            lb[c] = val;
            if( type[1] == 'I' )
               ctype[c] = ctype[c] == 1 ? 1 : 2;
         }
         else if( type[0] == 'U' )
         {
            ub[c] = val;
            if( type[1] == 'I' )
               ctype[c] = ctype[c] == 1 ? 1 : 2;
         }
         else if( !strcmp(type, "FX") )
            lb[c] = ub[c] = val;
         else if( !strcmp(type, "FR") )
         {
            lb[c] = -INF;
            ub[c] = INF;
         }
         else if( !strcmp(type, "MI") )
            lb[c] = -INF;
         else if( !strcmp(type, "PL") )
            ub[c] = INF;
         else if( !strcmp(type, "BV") )
         {
            if( ctype[c] != 1 )
            {
               lb[c] = lb[c] > 0.0 ? lb[c] : 0.0;
               ub[c] = ub[c] < 1.0 ? ub[c] : 1.0;
               ctype[c] = 1;
            }
         }
         else if( type[0] == 'S' )
         {
            if( type[1] == 'I' )
               ctype[c] = ctype[c] == 0 ? 2 : ctype[c];
            if( val < INF || ub[c] < INF )
               ub[c] = val;
            if( !semic[c] )
               ++nsc;
            semic[c] = 1;
         }
         else
            goto readerr;
         break;
      }
      default:
         break;
      }
   }
   fclose(f);
   f = NULL;
   if( sec != SEC_END )
      goto readerr_noclose;
   if( lhs == NULL )
   {
      lhs = (double*)malloc(sizeof(double) * (size_t)(nrows > 0 ? nrows : 1));
      rhs = (double*)malloc(sizeof(double) * (size_t)(nrows > 0 ? nrows : 1));
      if( !lhs || !rhs )
         goto nomem;
      for( int64_t i = 0; i < nrows; ++i )
      {
         lhs[i] = rtype[i] == 'L' ? -INF : 0.0;
         rhs[i] = rtype[i] == 'G' ? INF : 0.0;
      }
   }
   /* semi-continuous columns: lower bound 0 (the disjunction is not a linear row) */
   for( int64_t j = 0; j < ncols; ++j )
      if( semic[j] )
         lb[j] = 0.0;
   /* CSR with ascending columns; repeated entries of a row are summed, zero sums dropped */
   if( ntrip > 1 )   /* a file without entries has no triplet array at all */
      qsort(trip, (size_t)ntrip, sizeof(triplet), cmpTriplet);
   out->nrows = nrows;
   out->ncols = ncols;
   out->rowptr = (int64_t*)calloc((size_t)nrows + 1, sizeof(int64_t));
   out->colidx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ntrip > 0 ? ntrip : 1));
   out->vals = (double*)malloc(sizeof(double) * (size_t)(ntrip > 0 ? ntrip : 1));
   out->lhs = lhs;
   out->rhs = rhs;
   out->integral = (uint8_t*)malloc((size_t)(ncols > 0 ? ncols : 1));
   out->lb = lb ? lb : (double*)malloc(sizeof(double));
   out->ub = ub ? ub : (double*)malloc(sizeof(double));
   lhs = rhs = lb = ub = NULL;
   if( !out->rowptr || !out->colidx || !out->vals || !out->integral || !out->lb || !out->ub )
      goto nomem;
   {
      int64_t k = 0;
      for( int64_t q = 0; q < ntrip; )
      {
         int64_t e = q;
         double v = 0.0;
         while( e < ntrip && trip[e].row == trip[q].row && trip[e].col == trip[q].col )
            v += trip[e++].val;
         if( fabs(v) > 1e-9 )
         {
            out->colidx[k] = trip[q].col;
            out->vals[k] = v;
            out->rowptr[trip[q].row + 1]++;
            ++k;
         }
         q = e;
      }
      for( int64_t i = 0; i < nrows; ++i )
         out->rowptr[i + 1] += out->rowptr[i];
      out->nnz = k;
   }
   for( int64_t j = 0; j < ncols; ++j )
   {
      out->integral[j] = ctype[j] != 0;
      if( ctype[j] != 0 )
      {
         if( out->lb[j] > -INF )
            out->lb[j] = ceil(out->lb[j] - 1e-6);
         if( out->ub[j] < INF )
            out->ub[j] = floor(out->ub[j] + 1e-6);
      }
   }
   if( nsemicont )
      *nsemicont = nsc;
   rc = 1;
   goto done;
nomem:
   rc = -1;
   goto fail;
readerr:
   rc = SAP_READERROR;
   goto fail;
readerr_noclose:
   rc = SAP_READERROR;
fail:
   sap_instance_free(out);
done:
   if( f )
      fclose(f);
   mapFree(&rows);
   mapFree(&cols);
   free(rtype);
   free(lhs);
   free(rhs);
   free(lb);
   free(ub);
   free(ctype);
   free(semic);
   free(trip);
   return rc;
}
