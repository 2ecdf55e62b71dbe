/* oracle.c -- CPU restatement of SCIP's linear-constraint domain propagation.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  The product library never
 * links, loads or calls this file.
 *
 * Built with -O2 -ffp-contract=off (no FMA contraction, no fast-math) so that every
 * double operation is the IEEE operation written in the source, as in SCIP's own build.
 *
 * Sources restated (paths relative to the SCIP tree):
 *   src/scip/def.h:172-204            numerics defaults and EPS* macros
 *   src/scip/set.c:6824-7753, 8032    SCIPsetIs* predicates, IsLb/UbBetter, IsUpdateUnreliable
 *   src/scip/misc.c:11162             SCIPrelDiff
 *   src/scip/dbldblarith.h:118-187    TwoSum / Sum21 / TwoProd
 *   src/scip/var.c:1957-2021          adjustedLb / adjustedUb
 *   src/scip/scip_var.c:6965-7174     SCIPinferVarLbCons / SCIPinferVarUbCons
 *   src/scip/cons_linear.c:1307-2806  activities, residuals, maxactdelta, incremental updates
 *   src/scip/cons_linear.c:3191-3395  propagation sort order (consdataCompVarProp, consdataSort)
 *   src/scip/cons_linear.c:5214-5653  canTightenBounds, tightenVarUb/Lb, tightenVarBoundsEasy
 *   src/scip/cons_linear.c:6700-7154  tightenVarBounds, tightenBounds
 *   src/scip/cons_linear.c:7621-7758  propagateCons
 *   src/scip/cons_linear.c:16126-16195 consPropLinear
 *   src/scip/cons_linear.c:17162-17279 eventExecLinear
 *   src/scip/solve.c:723-795           propagateDomains round loop
 *   src/scip/cons.c:7309-7366          delayed (un)marking of constraints for propagation
 *   src/scip/cons_linear.c:5715-6697  rangedRowPropagation (jRanged, params.rangedrows)
 *   src/scip/cons_linear.c:344-369, 4981-5189, 5657  inferinfo, conflict bounds (the reason
 *                                    recording of orc_propagate_jacobi_reasons)
 *   src/scip/prop_probing.c:1203-1640 probing branches and deductions (oracle.py probe())
 *
 * What pins each part (tests/test_oracle_arith.py, tests/test_oracle_prop.py):
 *   numerics, TwoSum/Sum21/TwoProd   bit-exact against SCIP's own dbldblarith.h (oracle/_ref) and
 *                                    the known-answer vectors of tests/src/misc/dbldblarith.c
 *   canonical tree                   exact on integers, tiling/padding invariance, fsum accuracy
 *   single-row tightening, acceptance, infeasibility, redundancy, maxactdelta, force rule:
 *                                    closed-form cases computed by hand
 *   propagation fixpoints            brute-force soundness on tiny integer systems, completeness
 *                                    on infeasible binary systems, Jacobi = SCIP order on binary
 *                                    systems (unique fixpoint), mutual fixpoints otherwise;
 *                                    both schedules = the greatest supported box, found by
 *                                    enumeration, on integer systems with boundstreps 0
 *                                    (tests/test_oracle_fixpoint.py)
 *   reliable residual recompute      hand-computed cancellation cases whose bound only the
 *                                    plain-double recompute gives (RELIABLE in
 *                                    tests/test_oracle_prop.py)
 *   tightenbounds FALSE (notighten)  closed form: only rows whose activity range misses a side
 *                                    cut off, nothing moves
 *   ranged rows                      the two examples of the reference's own documentation of
 *                                    rangedRowPropagation (cons_linear.c:5702-5711), a hand-
 *                                    computed bound, soundness against enumeration of every
 *                                    integer point (tests/test_oracle_ranged.py)
 *   reasons                          every history entry implied by its row and the bounds of the
 *                                    earlier rounds (exact rational arithmetic; enumeration for
 *                                    ranged-rule entries), every certificate proving
 *                                    infeasibility, hand-computed histories and certificates
 *                                    (tests/test_oracle_reasons.py)
 *   probing                          hand-computed pair lists and soundness against enumeration
 *                                    of every 0/1 point (tests/test_oracle_probe.py)
 * Parity unpinned (no reference vector or closed form decides it; SCIP itself cannot be
 * built here):
 *   - orc_propagate_scip's exact general-integer / continuous bounds where the fixpoint is
 *     order dependent (boundstreps > 0) and the staleness of SCIP's incremental updates
 *     matters; pinned only by soundness, the mutual-fixpoint property and exact agreement
 *     wherever the fixpoint is unique.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_INVALID 1e99   /* SCIP_INVALID, def.h:187 */

typedef struct { double hi, lo; } dd;

/* ------------------------------------------------------------------------------------------
 * numerics (def.h:191-204, set.c, misc.c:11162)
 * ---------------------------------------------------------------------------------------- */
#define MAXD(x, y) ((x) >= (y) ? (x) : (y))   /* def.h:229 */
#define MIND(x, y) ((x) <= (y) ? (x) : (y))   /* def.h:233 */

static inline int isInf(const orc_params* p, double x) { return x >= p->infinity; }
static inline int isHuge(const orc_params* p, double x) { return x >= p->hugeval; }
static inline int isEQ(const orc_params* p, double a, double b) { return fabs(a - b) <= p->epsilon; }
static inline int isLT(const orc_params* p, double a, double b) { return a - b < -p->epsilon; }
static inline int isLE(const orc_params* p, double a, double b) { return a - b <= p->epsilon; }
static inline int isGT(const orc_params* p, double a, double b) { return a - b > p->epsilon; }
static inline int isGE(const orc_params* p, double a, double b) { return a - b >= -p->epsilon; }
static inline int isPositive(const orc_params* p, double x) { return x > p->epsilon; }
static inline int isSumGT(const orc_params* p, double a, double b) { return a - b > p->sumepsilon; }
static inline double relDiff(double a, double b)
{
   double q = MAXD(MAXD(1.0, fabs(a)), fabs(b));
   return (a - b) / q;
}
static inline int isFeasEQ(const orc_params* p, double a, double b) { return fabs(relDiff(a, b)) <= p->feastol; }
static inline int isFeasLT(const orc_params* p, double a, double b) { return relDiff(a, b) < -p->feastol; }
static inline int isFeasGT(const orc_params* p, double a, double b) { return relDiff(a, b) > p->feastol; }
static inline int isFeasZero(const orc_params* p, double x) { return fabs(x) <= p->feastol; }
static inline double feasFloor(const orc_params* p, double x) { return floor(x + p->feastol); }
static inline double feasCeil(const orc_params* p, double x) { return ceil(x - p->feastol); }

/* set.c:7711 */
static inline int isLbBetter(const orc_params* p, double newlb, double oldlb, double oldub)
{
   double t;
   if( oldlb < 0.0 && newlb >= 0.0 )
      return 1;
   t = MIND(oldub - oldlb, fabs(oldlb));
   t = MAXD(t, 1e-3);
   return newlb - oldlb > p->boundstreps * t;
}

/* set.c:7735 */
static inline int isUbBetter(const orc_params* p, double newub, double oldlb, double oldub)
{
   double t;
   if( oldub > 0.0 && newub <= 0.0 )
      return 1;
   t = MIND(oldub - oldlb, fabs(oldub));
   t = MAXD(t, 1e-3);
   return newub - oldub < -(p->boundstreps * t);
}

/* set.c:8032 */
static inline int isUpdateUnreliable(const orc_params* p, double newvalue, double oldvalue)
{
   double quotient = fabs(oldvalue) / MAXD(fabs(newvalue), p->epsilon);
   return quotient >= p->recompfac;
}

/* var.c:1957 */
static inline double adjustedLb(const orc_params* p, int integral, double lb)
{
   if( lb < 0.0 && isInf(p, -lb) )
      return -p->infinity;
   else if( lb > 0.0 && isInf(p, lb) )
      return p->infinity;
   else if( integral )
      return feasCeil(p, lb);
   else if( lb > 0.0 && lb < p->epsilon )
      return 0.0;
   else
      return lb;
}

/* var.c:2006 */
static inline double adjustedUb(const orc_params* p, int integral, double ub)
{
   if( ub > 0.0 && isInf(p, ub) )
      return p->infinity;
   else if( ub < 0.0 && isInf(p, -ub) )
      return -p->infinity;
   else if( integral )
      return feasFloor(p, ub);
   else if( ub < 0.0 && ub > -p->epsilon )
      return 0.0;
   else
      return ub;
}

/* ------------------------------------------------------------------------------------------
 * double-double (dbldblarith.h:118-187)
 * ---------------------------------------------------------------------------------------- */
static inline void twoSum(double a, double b, double* rhi, double* rlo)
{
   double s = a + b;
   double bb = s - a;
   *rlo = (a - (s - bb)) + (b - bb);
   *rhi = s;
}

/* SCIPdbldblSum21: hi = TwoSum(ahi, b).hi, lo = TwoSum(ahi, b).lo + alo */
static inline dd sum21(dd a, double b)
{
   dd r;
   double h, l;
   twoSum(a.hi, b, &h, &l);
   r.lo = l + a.lo;
   r.hi = h;
   return r;
}

/* pairwise combination used by the canonical tree: hi = TwoSum(a.hi, b.hi).hi,
 * lo = TwoSum(a.hi, b.hi).lo + (a.lo + b.lo).  Commutative; (x)+(0,0) = x. */
static inline dd pairAdd(dd a, dd b)
{
   dd r;
   double h, l;
   twoSum(a.hi, b.hi, &h, &l);
   r.lo = l + (a.lo + b.lo);
   r.hi = h;
   return r;
}

static inline double ddVal(dd a) { return a.hi + a.lo; }  /* QUAD_TO_DBL */

void orc_dd_sum(double a, double b, double* rhi, double* rlo) { twoSum(a, b, rhi, rlo); }

void orc_dd_sum21(double ahi, double alo, double b, double* rhi, double* rlo)
{
   dd a = { ahi, alo };
   dd r = sum21(a, b);
   *rhi = r.hi;
   *rlo = r.lo;
}

void orc_dd_pairadd(double ahi, double alo, double bhi, double blo, double* rhi, double* rlo)
{
   dd a = { ahi, alo };
   dd b = { bhi, blo };
   dd r = pairAdd(a, b);
   *rhi = r.hi;
   *rlo = r.lo;
}

/* Dekker split with 2^27+1 and TwoProd without FMA (dbldblarith.h:118-138) */
static inline void split(double x, double* hi, double* lo)
{
   double t = 134217729.0 * x;
   *hi = t - (t - x);
   *lo = x - *hi;
}

void orc_dd_prod(double a, double b, double* rhi, double* rlo)
{
   double ahi, alo, bhi, blo, h;
   split(a, &ahi, &alo);
   split(b, &bhi, &blo);
   h = a * b;
   *rlo = alo * blo - (((h - ahi * bhi) - alo * bhi) - ahi * blo);
   *rhi = h;
}

void orc_dd_prod21(double ahi, double alo, double b, double* rhi, double* rlo)
{
   double h, l;
   orc_dd_prod(ahi, b, &h, &l);
   *rlo = alo * b + l;
   *rhi = h;
}

/* canonical activity sum: perfect binary tree over row positions 0..P-1 (P = power of two
 * >= n), leaf p = (x_p, 0), padding leaves (0, 0), internal node = pairAdd(left, right).
 * Any tiling into aligned power-of-two blocks reproduces it bit for bit. */
static dd canonicalSum(const double* x, int64_t n, dd* scratch)
{
   int64_t P = 1, w, q;
   dd zero = { 0.0, 0.0 };
   if( n <= 0 )
      return zero;
   while( P < n )
      P <<= 1;
   for( q = 0; q < P; ++q )
   {
      scratch[q].hi = (q < n) ? x[q] : 0.0;
      scratch[q].lo = 0.0;
   }
   for( w = 1; w < P; w <<= 1 )
      for( q = 0; q + w < P; q += 2 * w )
         scratch[q] = pairAdd(scratch[q], scratch[q + w]);
   return scratch[0];
}

void orc_canonical_sum(const double* x, int64_t n, double* rhi, double* rlo)
{
   int64_t P = 1;
   dd* s;
   dd r;
   while( P < n )
      P <<= 1;
   s = (dd*)malloc(sizeof(dd) * (size_t)(P > 0 ? P : 1));
   r = canonicalSum(x, n, s);
   free(s);
   *rhi = r.hi;
   *rlo = r.lo;
}

void orc_params_default(orc_params* p)
{
   p->infinity = 1e20;
   p->epsilon = 1e-9;
   p->sumepsilon = 1e-6;
   p->feastol = 1e-6;
   p->boundstreps = 0.05;
   p->hugeval = 1e15;
   p->recompfac = 1e6;
   p->maxeasyactivitydelta = 1e6;
   p->maxrounds = -1;
   p->rangedrows = 0;
   p->notighten = 0;
   p->pad = 0;
}

void orc_compute_binary(int64_t ncols, const uint8_t* integral, const double* glb, const double* gub,
   uint8_t* binary)
{
   int64_t j;
   for( j = 0; j < ncols; ++j )
      binary[j] = (uint8_t)(integral[j] && glb[j] >= 0.0 && gub[j] <= 1.0);
}

/* ------------------------------------------------------------------------------------------
 * activity counters: classification of one contribution a*b to an activity
 * (consdataUpdateActivities from oldbound 0, cons_linear.c:1773-1945)
 *   kind 0: finite, 1: +inf, 2: -inf, 3: +huge, 4: -huge
 * ---------------------------------------------------------------------------------------- */
static inline int classify(const orc_params* p, double a, double b, double* contrib)
{
   double c;
   if( isInf(p, fabs(b)) )
   {
      *contrib = 0.0;
      return ((b > 0.0) == (a > 0.0)) ? 1 : 2;
   }
   c = a * b;
   if( isHuge(p, fabs(c)) )
   {
      *contrib = 0.0;
      return c > 0.0 ? 3 : 4;
   }
   *contrib = c;
   return 0;
}

typedef struct
{
   dd  minact, maxact;
   int minposinf, minneginf, minposhuge, minneghuge;
   int maxposinf, maxneginf, maxposhuge, maxneghuge;
   double maxactdelta;
} rowstate;

static void countKind(int kind, int* pinf, int* ninf, int* phuge, int* nhuge)
{
   switch( kind )
   {
   case 1: ++(*pinf); break;
   case 2: ++(*ninf); break;
   case 3: ++(*phuge); break;
   case 4: ++(*nhuge); break;
   default: break;
   }
}

/* consdataRecomputeMaxActivityDelta (cons_linear.c:1542-1599) on current bounds */
static double rowMaxActDelta(const orc_matrix* m, const orc_params* p, int64_t i, const double* lb,
   const double* ub)
{
   int64_t s = m->rowptr[i], e = m->rowptr[i + 1], k;
   int allbin = 1;
   double mad = 0.0;
   for( k = s; k < e; ++k )
      if( !m->binary[m->colidx[k]] )
      {
         allbin = 0;
         break;
      }
   if( allbin )
   {
      for( k = e - 1; k >= s; --k )
      {
         int32_t j = m->colidx[k];
         if( lb[j] < 0.5 && ub[j] > 0.5 )
         {
            double d = fabs(m->vals[k]);
            if( d > mad )
               mad = d;
         }
      }
      return mad;
   }
   for( k = e - 1; k >= s; --k )
   {
      int32_t j = m->colidx[k];
      double d;
      if( isInf(p, -lb[j]) || isInf(p, ub[j]) )
         return p->infinity;
      d = fabs(m->vals[k]) * (ub[j] - lb[j]);
      if( d > mad )
         mad = d;
   }
   return mad;
}

/* from-scratch row state; canonical=1: pairwise tree over positions (device twin),
 * canonical=0: sequential Sum21 in position order (consdataCalcActivities) */
static void computeRowState(const orc_matrix* m, const orc_params* p, int64_t i, const double* lb,
   const double* ub, int canonical, double* xmin, double* xmax, dd* scratch, rowstate* rs)
{
   int64_t s = m->rowptr[i], e = m->rowptr[i + 1], k;
   int64_t n = e - s;
   memset(rs, 0, sizeof(*rs));
   for( k = s; k < e; ++k )
   {
      int32_t j = m->colidx[k];
      double a = m->vals[k];
      double c;
      int kind;
      kind = classify(p, a, a > 0.0 ? lb[j] : ub[j], &c);
      countKind(kind, &rs->minposinf, &rs->minneginf, &rs->minposhuge, &rs->minneghuge);
      xmin[k - s] = c;
      kind = classify(p, a, a > 0.0 ? ub[j] : lb[j], &c);
      countKind(kind, &rs->maxposinf, &rs->maxneginf, &rs->maxposhuge, &rs->maxneghuge);
      xmax[k - s] = c;
   }
   if( canonical )
   {
      rs->minact = canonicalSum(xmin, n, scratch);
      rs->maxact = canonicalSum(xmax, n, scratch);
   }
   else
   {
      dd z = { 0.0, 0.0 };
      rs->minact = z;
      rs->maxact = z;
      for( k = 0; k < n; ++k )
      {
         if( xmin[k] != 0.0 )
            rs->minact = sum21(rs->minact, xmin[k]);
         if( xmax[k] != 0.0 )
            rs->maxact = sum21(rs->maxact, xmax[k]);
      }
   }
   rs->maxactdelta = rowMaxActDelta(m, p, i, lb, ub);
}

static int64_t maxRowLen(const orc_matrix* m)
{
   int64_t i, mx = 1;
   for( i = 0; i < m->nrows; ++i )
      if( m->rowptr[i + 1] - m->rowptr[i] > mx )
         mx = m->rowptr[i + 1] - m->rowptr[i];
   return mx;
}

void orc_row_activities(const orc_matrix* m, const orc_params* p, const double* lb, const double* ub,
   int canonical, orc_rowact* out)
{
   int64_t L = maxRowLen(m), P = 1, i;
   double* xmin;
   double* xmax;
   dd* scratch;
   while( P < L )
      P <<= 1;
   xmin = (double*)malloc(sizeof(double) * (size_t)L);
   xmax = (double*)malloc(sizeof(double) * (size_t)L);
   scratch = (dd*)malloc(sizeof(dd) * (size_t)P);
   for( i = 0; i < m->nrows; ++i )
   {
      rowstate rs;
      computeRowState(m, p, i, lb, ub, canonical, xmin, xmax, scratch, &rs);
      out[i].minact_hi = rs.minact.hi;
      out[i].minact_lo = rs.minact.lo;
      out[i].maxact_hi = rs.maxact.hi;
      out[i].maxact_lo = rs.maxact.lo;
      out[i].minact = ddVal(rs.minact);
      out[i].maxact = ddVal(rs.maxact);
      out[i].minposinf = rs.minposinf;
      out[i].minneginf = rs.minneginf;
      out[i].minposhuge = rs.minposhuge;
      out[i].minneghuge = rs.minneghuge;
      out[i].maxposinf = rs.maxposinf;
      out[i].maxneginf = rs.maxneginf;
      out[i].maxposhuge = rs.maxposhuge;
      out[i].maxneghuge = rs.maxneghuge;
      out[i].maxactdelta = rs.maxactdelta;
   }
   free(xmin);
   free(xmax);
   free(scratch);
}

/* getMinActivity / getMaxActivity (cons_linear.c:2345-2529) */
static void getMinAct(const orc_params* p, dd act, int posinf, int neginf, int poshuge, int neghuge,
   double delta, int goodrelax, double* res, int* istight, int* issetinf)
{
   if( neginf > 0 )
   {
      *res = -p->infinity;
      *issetinf = 1;
      *istight = (posinf == 0);
   }
   else if( posinf > 0 )
   {
      *res = p->infinity;
      *issetinf = 1;
      *istight = 1;
   }
   else if( neghuge > 0 || (poshuge > 0 && !goodrelax) )
   {
      *res = -p->infinity;
      *issetinf = 1;
      *istight = 0;
   }
   else
   {
      dd t = sum21(act, -delta);
      if( poshuge > 0 )
      {
         t = sum21(t, poshuge * p->hugeval);
         *istight = 0;
      }
      else
         *istight = 1;
      *res = ddVal(t);
      *issetinf = 0;
   }
}

static void getMaxAct(const orc_params* p, dd act, int posinf, int neginf, int poshuge, int neghuge,
   double delta, int goodrelax, double* res, int* istight, int* issetinf)
{
   if( posinf > 0 )
   {
      *res = p->infinity;
      *issetinf = 1;
      *istight = (neginf == 0);
   }
   else if( neginf > 0 )
   {
      *res = -p->infinity;
      *issetinf = 1;
      *istight = 1;
   }
   else if( poshuge > 0 || (neghuge > 0 && !goodrelax) )
   {
      *res = p->infinity;
      *issetinf = 1;
      *istight = 0;
   }
   else
   {
      dd t = sum21(act, -delta);
      if( neghuge > 0 )
      {
         t = sum21(t, -neghuge * p->hugeval);
         *istight = 0;
      }
      else
         *istight = 1;
      *res = ddVal(t);
      *issetinf = 0;
   }
}

/* consdataGetActivityResiduals (cons_linear.c:2661-2806) for a variable with coefficient a
 * and local bounds [lb,ub], goodrelax as given */
static void residuals(const orc_params* p, const rowstate* rs, double a, double lb, double ub, int goodrelax,
   double* minres, double* maxres, int* mintight, int* maxtight, int* minsetinf, int* maxsetinf)
{
   double minb, maxb, absval;
   if( a > 0.0 )
   {
      minb = lb;
      maxb = ub;
      absval = a;
   }
   else
   {
      minb = -ub;
      maxb = -lb;
      absval = -a;
   }
   if( isInf(p, minb) )
      getMinAct(p, rs->minact, rs->minposinf - 1, rs->minneginf, rs->minposhuge, rs->minneghuge, 0.0, goodrelax,
         minres, mintight, minsetinf);
   else if( isInf(p, -minb) )
      getMinAct(p, rs->minact, rs->minposinf, rs->minneginf - 1, rs->minposhuge, rs->minneghuge, 0.0, goodrelax,
         minres, mintight, minsetinf);
   else if( isHuge(p, minb * absval) )
      getMinAct(p, rs->minact, rs->minposinf, rs->minneginf, rs->minposhuge - 1, rs->minneghuge, 0.0, goodrelax,
         minres, mintight, minsetinf);
   else if( isHuge(p, -minb * absval) )
      getMinAct(p, rs->minact, rs->minposinf, rs->minneginf, rs->minposhuge, rs->minneghuge - 1, 0.0, goodrelax,
         minres, mintight, minsetinf);
   else
      getMinAct(p, rs->minact, rs->minposinf, rs->minneginf, rs->minposhuge, rs->minneghuge, absval * minb,
         goodrelax, minres, mintight, minsetinf);

   if( isInf(p, -maxb) )
      getMaxAct(p, rs->maxact, rs->maxposinf, rs->maxneginf - 1, rs->maxposhuge, rs->maxneghuge, 0.0, goodrelax,
         maxres, maxtight, maxsetinf);
   else if( isInf(p, maxb) )
      getMaxAct(p, rs->maxact, rs->maxposinf - 1, rs->maxneginf, rs->maxposhuge, rs->maxneghuge, 0.0, goodrelax,
         maxres, maxtight, maxsetinf);
   else if( isHuge(p, absval * maxb) )
      getMaxAct(p, rs->maxact, rs->maxposinf, rs->maxneginf, rs->maxposhuge - 1, rs->maxneghuge, 0.0, goodrelax,
         maxres, maxtight, maxsetinf);
   else if( isHuge(p, -absval * maxb) )
      getMaxAct(p, rs->maxact, rs->maxposinf, rs->maxneginf, rs->maxposhuge, rs->maxneghuge - 1, 0.0, goodrelax,
         maxres, maxtight, maxsetinf);
   else
      getMaxAct(p, rs->maxact, rs->maxposinf, rs->maxneginf, rs->maxposhuge, rs->maxneghuge, absval * maxb,
         goodrelax, maxres, maxtight, maxsetinf);
}

/* ==========================================================================================
 * Jacobi twin
 *
 * Round r: every marked (non-redundant) row computes its state from the bounds at the start
 * of the round, then submits bound candidates exactly as tightenBounds would for each of its
 * variables if it were the only row; the candidates of a column are reduced to the tightest
 * regular and the tightest forced (single-variable row) candidate per side, and the column
 * applies SCIPinferVarLbCons then SCIPinferVarUbCons to them.  Rows containing a changed
 * column are marked for round r+1.  Candidate min/max use the order-preserving int64 image
 * of the double so -0.0 < +0.0, identical to the device atomics.
 * ======================================================================================== */
static inline int64_t dkey(double x)
{
   int64_t b;
   memcpy(&b, &x, 8);
   return b ^ ((b >> 63) & INT64_C(0x7FFFFFFFFFFFFFFF));
}

typedef struct
{
   const orc_matrix* m;
   const orc_params* p;
   const double* lb;
   const double* ub;
   double* clb;   /* regular lb candidates (max) */
   double* cub;   /* regular ub candidates (min) */
   double* flb;   /* forced lb candidates (max) */
   double* fub;   /* forced ub candidates (min) */
   uint8_t* touched;
   int cutoff;
   /* reasons: the smallest row that submitted each slot's best candidate (NULL: not kept),
    * the row being swept, and the side of a row's infeasibility (0: minimal activity
    * against rhs, 1: maximal activity against lhs) */
   int32_t* clbrow;
   int32_t* cubrow;
   int32_t* flbrow;
   int32_t* fubrow;
   int32_t currow;
   int32_t currule;   /* 0: residual activity, 1: ranged-row test; rows are kept as row * 2 + rule */
   int rowside;
   int64_t nrecomp;   /* reliable residual recomputes */
} jstate;

static inline void submitLb(jstate* js, int32_t j, double v, int force)
{
   double* t = force ? &js->flb[j] : &js->clb[j];
   int32_t* r = js->clbrow == NULL ? NULL : (force ? &js->flbrow[j] : &js->clbrow[j]);
   int32_t key = js->currow * 2 + js->currule;
   if( dkey(v) > dkey(*t) )
   {
      *t = v;
      if( r )
         *r = key;
   }
   else if( r && dkey(v) == dkey(*t) && key < *r )
      *r = key;
   js->touched[j] = 1;
}

static inline void submitUb(jstate* js, int32_t j, double v, int force)
{
   double* t = force ? &js->fub[j] : &js->cub[j];
   int32_t* r = js->cubrow == NULL ? NULL : (force ? &js->fubrow[j] : &js->cubrow[j]);
   int32_t key = js->currow * 2 + js->currule;
   if( dkey(v) < dkey(*t) )
   {
      *t = v;
      if( r )
         *r = key;
   }
   else if( r && dkey(v) == dkey(*t) && key < *r )
      *r = key;
   js->touched[j] = 1;
}

/* plain-double residual in position order (consdataGetReliableResidualActivity,
 * cons_linear.c:2580-2657) */
static double reliableResidual(const orc_matrix* m, int64_t i, int64_t skip, const double* lb, const double* ub,
   int isminres)
{
   int64_t k;
   double r = 0.0;
   for( k = m->rowptr[i]; k < m->rowptr[i + 1]; ++k )
   {
      int32_t j;
      double a;
      if( k == skip )
         continue;
      j = m->colidx[k];
      a = m->vals[k];
      if( a > 0.0 )
         r += a * (isminres ? lb[j] : ub[j]);
      else
         r += a * (isminres ? ub[j] : lb[j]);
   }
   return r;
}

/* tightenVarBoundsEasy (cons_linear.c:5380-5653) with start-of-round values */
static void jEasy(jstate* js, const rowstate* rs, int64_t i, int64_t k, int force)
{
   const orc_params* p = js->p;
   const orc_matrix* m = js->m;
   int32_t j = m->colidx[k];
   double a = m->vals[k];
   double lhs = m->lhs[i], rhs = m->rhs[i];
   double lb = js->lb[j], ub = js->ub[j];
   int integral = m->integral[j];
   double minact = ddVal(rs->minact), maxact = ddVal(rs->maxact);

   if( a > 0.0 )
   {
      if( !isInf(p, rhs) )
      {
         double slack, alpha;
         if( isFeasLT(p, rhs, minact) )
         {
            js->cutoff = 1;
            js->rowside = 0;
            return;
         }
         slack = rhs - minact;
         if( !isPositive(p, slack) )
            slack = 0.0;
         alpha = a * (ub - lb);
         if( isSumGT(p, alpha, slack) || (force && isGT(p, alpha, slack)) )
         {
            double newub = adjustedUb(p, integral, lb + (slack / a));
            if( force || isUbBetter(p, newub, lb, ub) )
               submitUb(js, j, newub, force);
         }
      }
      if( !isInf(p, -lhs) )
      {
         double slack, alpha;
         if( isFeasLT(p, maxact, lhs) )
         {
            js->cutoff = 1;
            js->rowside = 1;
            return;
         }
         slack = maxact - lhs;
         if( !isPositive(p, slack) )
            slack = 0.0;
         alpha = a * (ub - lb);
         if( isSumGT(p, alpha, slack) || (force && isGT(p, alpha, slack)) )
         {
            double newlb = adjustedLb(p, integral, ub - (slack / a));
            if( force || isLbBetter(p, newlb, lb, ub) )
               submitLb(js, j, newlb, force);
         }
      }
   }
   else
   {
      if( !isInf(p, rhs) )
      {
         double slack, alpha;
         if( isFeasLT(p, rhs, minact) )
         {
            js->cutoff = 1;
            js->rowside = 0;
            return;
         }
         slack = rhs - minact;
         if( !isPositive(p, slack) )
            slack = 0.0;
         alpha = a * (lb - ub);
         if( isSumGT(p, alpha, slack) || (force && isGT(p, alpha, slack)) )
         {
            double newlb = adjustedLb(p, integral, ub + slack / a);
            if( force || isLbBetter(p, newlb, lb, ub) )
               submitLb(js, j, newlb, force);
         }
      }
      if( !isInf(p, -lhs) )
      {
         double slack, alpha;
         if( isFeasLT(p, maxact, lhs) )
         {
            js->cutoff = 1;
            js->rowside = 1;
            return;
         }
         slack = maxact - lhs;
         if( !isPositive(p, slack) )
            slack = 0.0;
         alpha = a * (lb - ub);
         if( isSumGT(p, alpha, slack) || (force && isGT(p, alpha, slack)) )
         {
            double newub = adjustedUb(p, integral, lb - (slack / a));
            if( force || isUbBetter(p, newub, lb, ub) )
               submitUb(js, j, newub, force);
         }
      }
   }
}

/* tightenVarBounds (cons_linear.c:6700-6974) with start-of-round values */
static void jHard(jstate* js, const rowstate* rs, int64_t i, int64_t k, int force)
{
   const orc_params* p = js->p;
   const orc_matrix* m = js->m;
   int32_t j = m->colidx[k];
   double a = m->vals[k];
   double lhs = m->lhs[i], rhs = m->rhs[i];
   double lb = js->lb[j], ub = js->ub[j];
   int integral = m->integral[j];
   double minres, maxres;
   int mintight, maxtight, minsetinf, maxsetinf;
   double lastmin = ddVal(rs->minact), lastmax = ddVal(rs->maxact);

   residuals(p, rs, a, lb, ub, 0, &minres, &maxres, &mintight, &maxtight, &minsetinf, &maxsetinf);

   if( a > 0.0 )
   {
      if( !minsetinf && !isInf(p, rhs) && mintight && (isLT(p, fabs(rhs), 1.0) || !isEQ(p, minres / rhs, 1.0)) )
      {
         double newub = (rhs - minres) / a;
         if( !isInf(p, newub)
            && ((force && isLT(p, newub, ub)) || (integral && isFeasLT(p, newub, ub)) || isUbBetter(p, newub, lb, ub)) )
         {
            int unrel = isUpdateUnreliable(p, minres, lastmin);
            if( unrel )
            {
               minres = (js->nrecomp++, reliableResidual(m, i, k, js->lb, js->ub, 1));
               newub = (rhs - minres) / a;
               unrel = isInf(p, -minres)
                  || (!isUbBetter(p, newub, lb, ub) && (!isFeasLT(p, newub, ub) || !integral)
                     && (!force || !isLT(p, newub, ub)));
            }
            if( !unrel )
               submitUb(js, j, newub, force);
         }
      }
      if( !maxsetinf && !isInf(p, -lhs) && maxtight && (isLT(p, fabs(lhs), 1.0) || !isEQ(p, maxres / lhs, 1.0)) )
      {
         double newlb = (lhs - maxres) / a;
         if( !isInf(p, -newlb)
            && ((force && isGT(p, newlb, lb)) || (integral && isFeasGT(p, newlb, lb)) || isLbBetter(p, newlb, lb, ub)) )
         {
            if( isUpdateUnreliable(p, maxres, lastmax) )
            {
               maxres = (js->nrecomp++, reliableResidual(m, i, k, js->lb, js->ub, 0));
               newlb = (lhs - maxres) / a;
               if( isInf(p, maxres) || (!isLbBetter(p, newlb, lb, ub) && (!isFeasGT(p, newlb, lb) || !integral)
                     && (!force || !isGT(p, newlb, lb))) )
                  return;
            }
            submitLb(js, j, newlb, force);
         }
      }
   }
   else
   {
      if( !minsetinf && !isInf(p, rhs) && mintight && (isLT(p, fabs(rhs), 1.0) || !isEQ(p, minres / rhs, 1.0)) )
      {
         double newlb = (rhs - minres) / a;
         if( !isInf(p, -newlb)
            && ((force && isGT(p, newlb, lb)) || (integral && isFeasGT(p, newlb, lb)) || isLbBetter(p, newlb, lb, ub)) )
         {
            int unrel = isUpdateUnreliable(p, minres, lastmin);
            if( unrel )
            {
               minres = (js->nrecomp++, reliableResidual(m, i, k, js->lb, js->ub, 1));
               newlb = (rhs - minres) / a;
               unrel = isInf(p, -minres)
                  || (!isLbBetter(p, newlb, lb, ub) && (!isFeasGT(p, newlb, lb) || !integral)
                     && (!force || !isGT(p, newlb, lb)));
            }
            if( !unrel )
               submitLb(js, j, newlb, force);
         }
      }
      if( !maxsetinf && !isInf(p, -lhs) && maxtight && (isLT(p, fabs(lhs), 1.0) || !isEQ(p, maxres / lhs, 1.0)) )
      {
         double newub = (lhs - maxres) / a;
         if( !isInf(p, newub)
            && ((force && isLT(p, newub, ub)) || (integral && isFeasLT(p, newub, ub)) || isUbBetter(p, newub, lb, ub)) )
         {
            if( isUpdateUnreliable(p, maxres, lastmax) )
            {
               maxres = (js->nrecomp++, reliableResidual(m, i, k, js->lb, js->ub, 0));
               newub = (lhs - maxres) / a;
               /* the reference combines the last two tests with && here (cons_linear.c:6942-6944) */
               if( isInf(p, maxres) || (!isUbBetter(p, newub, lb, ub) && (!isFeasLT(p, newub, ub) && !integral)
                     && (!force || !isLT(p, newub, ub))) )
                  return;
            }
            submitUb(js, j, newub, force);
         }
      }
   }
}

/* ranged-row propagation of one row (rangedRowPropagation, cons_linear.c:5715-6697) on the
 * start-of-round bounds, without artificial constraints (cons_linear.c:5784, propagation after
 * the root): deductions go in as forced candidates (SCIPinferVarFixCons / SCIPinferVarLbCons
 * and UbCons with force, scip_var.c:6896-6945); the enumeration of check-group values gives up
 * after ORC_RANGED_STEPS steps (the engine's limit; SCIP has none) */
#define ORC_RANGED_STEPS (1 << 20)

static inline double epsFloor(const orc_params* p, double x) { return floor(x + p->epsilon); }  /* EPSFLOOR, def.h:200 */
static inline double epsCeil(const orc_params* p, double x) { return ceil(x - p->epsilon); }    /* EPSCEIL, def.h:201 */
static inline int epsIntegral(const orc_params* p, double x) { return x - floor(x + p->epsilon) <= p->epsilon; }

static long long gcdLL(long long a, long long b)
{
   while( b != 0 )
   {
      long long t = a % b;
      a = b;
      b = t;
   }
   return a;
}

static void jRanged(jstate* js, int64_t i)
{
   const orc_params* p = js->p;
   const orc_matrix* m = js->m;
   int64_t s = m->rowptr[i], e = m->rowptr[i + 1], n = e - s, k, nfixed = 0, nunfixed, ninf = 0, ncont = 0;
   int64_t first = -1, last = -1, w = -1, steps = 0;
   double fixedact = 0.0, lhs, rhs, minact = 0.0, maxact = 0.0, g, minvalue = 0.0, maxvalue = 0.0, value;
   int possiblegcd = 1, gcdisone = 1, nsols = 0;
   long long gcd, gcdinf = -1;
   uint8_t* inchk;

   if( n < 3 || isInf(p, -m->lhs[i]) || isInf(p, m->rhs[i]) )
      return;
   /* fixed activity, last nonzero first (cons_linear.c:5809-5818) */
   for( k = e - 1; k >= s; --k )
   {
      int32_t j = m->colidx[k];
      if( isEQ(p, js->lb[j], js->ub[j]) )
      {
         fixedact += js->lb[j] * m->vals[k];
         ++nfixed;
      }
   }
   if( isHuge(p, fabs(fixedact)) )
      return;
   lhs = m->lhs[i] - fixedact;
   rhs = m->rhs[i] - fixedact;
   nunfixed = n - nfixed;
   inchk = (uint8_t*)calloc((size_t)n, 1);

   /* partition (cons_linear.c:5835-5960) */
   for( k = s; k < e; ++k )
   {
      int32_t j = m->colidx[k];
      double a = m->vals[k];
      int integ = m->integral[j] != 0;
      int group = integ && epsIntegral(p, a) && !isEQ(p, fabs(a), 1.0);
      if( isEQ(p, js->lb[j], js->ub[j]) )
         continue;
      if( group )
         break;
      ncont += integ ? 0 : 1;
      gcdisone = gcdisone && isEQ(p, fabs(a), 1.0);
      possiblegcd = 0;
      inchk[k - s] = 1;
      ++ninf;
   }
   if( k == e || ncont + 2 > nunfixed )
      goto TERMINATE;
   gcd = (long long)(fabs(m->vals[k]) + p->feastol);
   for( ; k < e; ++k )
   {
      int32_t j = m->colidx[k];
      double a = m->vals[k];
      int integ = m->integral[j] != 0;
      if( isEQ(p, js->lb[j], js->ub[j]) )
         continue;
      if( !integ || !epsIntegral(p, a) || isEQ(p, fabs(a), 1.0) )
      {
         ncont += integ ? 0 : 1;
         gcdisone = gcdisone && isEQ(p, fabs(a), 1.0);
         possiblegcd = 0;
         inchk[k - s] = 1;
         ++ninf;
      }
      else
      {
         long long t = gcdLL(gcd, (long long)(fabs(a) + p->feastol));
         if( t == 1 )
         {
            inchk[k - s] = 1;
            ++ninf;
         }
         else
            gcd = t;
      }
   }
   if( ninf == 0 )
      goto TERMINATE;

   /* activity range of the check group, last first (cons_linear.c:5968-6012) */
   for( k = e - 1; k >= s; --k )
   {
      int32_t j;
      double a;
      int mininv = 0, maxinv = 0;
      if( !inchk[k - s] )
         continue;
      j = m->colidx[k];
      a = m->vals[k];
      if( last < 0 )
         last = k;
      first = k;
      if( isInf(p, -js->lb[j]) )
      {
         if( a < 0.0 )
            maxinv = 1;
         else
            mininv = 1;
      }
      else if( a < 0.0 )
         maxact += a * js->lb[j];
      else
         minact += a * js->lb[j];
      if( isInf(p, js->ub[j]) )
      {
         if( a > 0.0 )
            maxinv = 1;
         else
            mininv = 1;
      }
      else if( a > 0.0 )
         maxact += a * js->ub[j];
      else
         minact += a * js->ub[j];
      if( mininv || maxinv || isHuge(p, -minact) || isHuge(p, maxact) )
         goto TERMINATE;
   }
   g = (double)gcd;
   /* no value of the gcd part fits between the sides (cons_linear.c:6029-6040) */
   if( !epsIntegral(p, (lhs - maxact) / g) && isGT(p, epsCeil(p, (lhs - maxact) / g) * g, rhs - minact) )
   {
      js->cutoff = 1;
      js->rowside = 2;
      goto TERMINATE;
   }
   if( ncont != 0 )
      goto TERMINATE;
   if( possiblegcd )
   {
      gcdinf = (long long)(fabs(m->vals[last]) + p->feastol);
      for( k = last; k >= s && gcdinf >= 2; --k )
         if( inchk[k - s] )
            gcdinf = gcdLL(gcdinf, (long long)(fabs(m->vals[k]) + p->feastol));
   }
   else if( gcdisone )
      gcdinf = 1;
   if( gcdinf < 1 )
      goto TERMINATE;

   /* solutions of the ranged row (cons_linear.c:6058-6135) */
   for( value = epsCeil(p, minact - p->feastol); isLE(p, value, maxact); value += (double)gcdinf )
   {
      double value2;
      if( ++steps > ORC_RANGED_STEPS )
         goto TERMINATE;
      value2 = value + g * epsCeil(p, (lhs - value) / g);
      if( !isGE(p, value2, lhs) )
         value2 += g;
      if( isLE(p, value2, rhs) )
      {
         if( ++nsols == 3 )
            break;
         if( nsols == 1 )
            minvalue = value;
         maxvalue = value;
      }
   }
   if( nsols == 3 )
   {
      for( value = epsFloor(p, maxact + p->feastol); isGE(p, value, minact); value -= (double)gcdinf )
      {
         double value2;
         if( ++steps > 2 * (int64_t)ORC_RANGED_STEPS )
            goto TERMINATE;
         value2 = value + g * epsFloor(p, (rhs - value) / g);
         if( !isLE(p, value2, rhs) )
            value2 -= g;
         if( isGE(p, value2, lhs) )
         {
            maxvalue = value;
            break;
         }
      }
   }
   if( nsols == 0 )
   {
      js->cutoff = 1;
      js->rowside = 2;
      goto TERMINATE;
   }
   if( ninf == nunfixed - 1 )
      for( k = s; k < e && w < 0; ++k )
         if( !inchk[k - s] && !isEQ(p, js->lb[m->colidx[k]], js->ub[m->colidx[k]]) )
            w = k;
   if( ninf == 1 )
   {
      /* cons_linear.c:6154-6180, 6340-6388 */
      double a = m->vals[first];
      int32_t j = m->colidx[first];
      if( nsols == 1 )
      {
         submitLb(js, j, maxvalue / a, 1);
         submitUb(js, j, maxvalue / a, 1);
      }
      else
      {
         double nlb = a < 0.0 ? maxvalue / a : minvalue / a;
         double nub = a < 0.0 ? minvalue / a : maxvalue / a;
         if( nlb > js->lb[j] )
            submitLb(js, j, nlb, 1);
         if( nub < js->ub[j] )
            submitUb(js, j, nub, 1);
      }
   }
   else if( w >= 0 )
   {
      /* cons_linear.c:6184-6270, 6390-6505 */
      double a = m->vals[w];
      int32_t j = m->colidx[w];
      if( nsols == 1 )
      {
         double fix = a < 0.0 ? epsFloor(p, (lhs - maxvalue) / a) : epsCeil(p, (lhs - maxvalue) / a);
         submitLb(js, j, fix, 1);
         submitUb(js, j, fix, 1);
      }
      else
      {
         double nlb = a < 0.0 ? epsFloor(p, (rhs - minvalue) / a) : epsCeil(p, (lhs - maxvalue) / a);
         double nub = a < 0.0 ? epsFloor(p, (lhs - maxvalue) / a) : epsCeil(p, (rhs - minvalue) / a);
         if( nlb > js->lb[j] )
            submitLb(js, j, nlb, 1);
         if( nub < js->ub[j] )
            submitUb(js, j, nub, 1);
      }
   }
TERMINATE:
   free(inchk);
}

/* one row of one Jacobi round: tightenBounds + the infeasibility/redundancy test of
 * propagateCons; returns 1 if the row became redundant */
static int jRow(jstate* js, int64_t i, double* xmin, double* xmax, dd* scratch)
{
   const orc_params* p = js->p;
   const orc_matrix* m = js->m;
   rowstate rs;
   int64_t s = m->rowptr[i], e = m->rowptr[i + 1], k;
   int64_t nv = e - s;
   int infmin, infmax, tighten, force;
   double minact, maxact;
   int t1, t2, f1, f2;

   computeRowState(m, p, i, js->lb, js->ub, 1, xmin, xmax, scratch, &rs);

   infmin = rs.minneginf + rs.minposinf + rs.minneghuge + rs.minposhuge;
   infmax = rs.maxneginf + rs.maxposinf + rs.maxneghuge + rs.maxposhuge;
   tighten = !p->notighten && nv > 0 && !(infmin > 1 && infmax > 1);
   force = (nv == 1);
   if( tighten && isFeasZero(p, rs.maxactdelta) )
      tighten = 0;
   if( tighten && !isInf(p, rs.maxactdelta) )
   {
      double slack, surplus;
      getMinAct(p, rs.minact, rs.minposinf, rs.minneginf, rs.minposhuge, rs.minneghuge, 0.0, 0, &minact, &t1, &f1);
      getMaxAct(p, rs.maxact, rs.maxposinf, rs.maxneginf, rs.maxposhuge, rs.maxneghuge, 0.0, 0, &maxact, &t2, &f2);
      slack = (isInf(p, m->rhs[i]) || f1) ? p->infinity : (m->rhs[i] - minact);
      surplus = (isInf(p, -m->lhs[i]) || f2) ? p->infinity : (maxact - m->lhs[i]);
      if( isLE(p, rs.maxactdelta, MIND(slack, surplus)) )
         tighten = 0;
   }
   if( tighten )
   {
      int easy = isLT(p, rs.maxactdelta, p->maxeasyactivitydelta);
      for( k = s; k < e && !js->cutoff; ++k )
      {
         if( easy )
            jEasy(js, &rs, i, k, force);
         else
            jHard(js, &rs, i, k, force);
      }
   }
   /* propagateCons: ranged rows after the tightening, and only with it (cons_linear.c:7694-7700) */
   if( !js->cutoff && p->rangedrows && !p->notighten )
   {
      js->currule = 1;
      jRanged(js, i);
      js->currule = 0;
   }
   if( js->cutoff )
      return 0;

   getMinAct(p, rs.minact, rs.minposinf, rs.minneginf, rs.minposhuge, rs.minneghuge, 0.0, 1, &minact, &t1, &f1);
   getMaxAct(p, rs.maxact, rs.maxposinf, rs.maxneginf, rs.maxposhuge, rs.maxneghuge, 0.0, 1, &maxact, &t2, &f2);
   if( isFeasGT(p, minact, m->rhs[i]) || isFeasLT(p, maxact, m->lhs[i]) )
   {
      js->cutoff = 1;
      js->rowside = isFeasGT(p, minact, m->rhs[i]) ? 0 : 1;
      return 0;
   }
   if( isGE(p, minact, m->lhs[i]) && isLE(p, maxact, m->rhs[i]) )
      return 1;
   return 0;
}

/* column step: SCIPinferVarLbCons then SCIPinferVarUbCons (scip_var.c:6965-7174) on the
 * tightest forced and regular candidates; returns 1 if a bound changed, sets *cutoff */
static int jColumn(const orc_params* p, int integral, double* plb, double* pub, double clb, double flb,
   double cub, double fub, int* cutoff, int64_t* nchg, int* src)
{
   double lb0 = *plb, ub0 = *pub;
   double nlb = lb0, nub = ub0;
   int changed = 0;
   /* src[0] / src[1]: which candidate set the new lower / upper bound, or crossed the other
    * bound on a cutoff (0: none, 1: forced, 2: regular) */
   int dummy[2];
   if( src == NULL )
      src = dummy;
   src[0] = src[1] = 0;

   if( flb != -HUGE_VAL )
   {
      double t = adjustedLb(p, integral, flb);
      if( isInf(p, t) || isFeasGT(p, t, ub0) )
      {
         *cutoff = 1;
         src[0] = 1;
         return 0;
      }
      t = MIND(t, ub0);
      if( !isLE(p, t, lb0) && dkey(t) > dkey(nlb) )
      {
         nlb = t;
         src[0] = 1;
      }
   }
   if( clb != -HUGE_VAL )
   {
      double t = adjustedLb(p, integral, clb);
      if( isInf(p, t) || isFeasGT(p, t, ub0) )
      {
         *cutoff = 1;
         src[0] = 2;
         return 0;
      }
      t = MIND(t, ub0);
      if( isLbBetter(p, t, lb0, ub0) && dkey(t) > dkey(nlb) )
      {
         nlb = t;
         src[0] = 2;
      }
   }
   if( fub != HUGE_VAL )
   {
      double t = adjustedUb(p, integral, fub);
      if( isInf(p, -t) || isFeasLT(p, t, nlb) )
      {
         *cutoff = 1;
         src[1] = 1;
         return 0;
      }
      t = MAXD(t, nlb);
      if( !isGE(p, t, ub0) && dkey(t) < dkey(nub) )
      {
         nub = t;
         src[1] = 1;
      }
   }
   if( cub != HUGE_VAL )
   {
      double t = adjustedUb(p, integral, cub);
      if( isInf(p, -t) || isFeasLT(p, t, nlb) )
      {
         *cutoff = 1;
         src[1] = 2;
         return 0;
      }
      t = MAXD(t, nlb);
      if( isUbBetter(p, t, nlb, ub0) && dkey(t) < dkey(nub) )
      {
         nub = t;
         src[1] = 2;
      }
   }
   if( nlb != lb0 || dkey(nlb) != dkey(lb0) )
   {
      *plb = nlb;
      ++(*nchg);
      changed = 1;
   }
   else
      src[0] = 0;
   if( nub != ub0 || dkey(nub) != dkey(ub0) )
   {
      *pub = nub;
      ++(*nchg);
      changed = 1;
   }
   else
      src[1] = 0;
   return changed;
}

/* reason recording of the Jacobi twin: the history of accepted changes with the smallest row
 * that submitted each winning candidate, and the infeasibility certificate of a cutoff */
typedef struct
{
   int64_t cap;
   int64_t n;
   int32_t* round;
   int32_t* col;
   int32_t* bound;
   double* val;
   int32_t* row;
   int32_t* rule;
   orc_conflict* conflict;
} jreasons;

/* rowkey: row * 2 + rule */
static void histAdd(jreasons* R, int32_t round, int32_t j, int32_t bound, double val, int32_t rowkey)
{
   if( R->n < R->cap )
   {
      R->round[R->n] = round;
      R->col[R->n] = j;
      R->bound[R->n] = bound;
      R->val[R->n] = val;
      R->row[R->n] = rowkey >> 1;
      if( R->rule )
         R->rule[R->n] = rowkey & 1;
   }
   R->n++;
}

static int jacobi(const orc_matrix* m, const orc_params* p, double* lb, double* ub, int full_sweep,
   const uint8_t* mark0, orc_result* res, jreasons* R)
{
   int64_t nr = m->nrows, nc = m->ncols, i, j, k;
   int64_t L = maxRowLen(m), P = 1;
   uint8_t* marked = (uint8_t*)malloc((size_t)(nr > 0 ? nr : 1));
   uint8_t* redundant = (uint8_t*)calloc((size_t)(nr > 0 ? nr : 1), 1);
   uint8_t* colchg = (uint8_t*)calloc((size_t)(nc > 0 ? nc : 1), 1);
   double* xmin = (double*)malloc(sizeof(double) * (size_t)L);
   double* xmax = (double*)malloc(sizeof(double) * (size_t)L);
   dd* scratch;
   jstate js;
   int round = 0;
   int64_t nsave;

   while( P < L )
      P <<= 1;
   scratch = (dd*)malloc(sizeof(dd) * (size_t)P);
   memset(res, 0, sizeof(*res));
   memset(&js, 0, sizeof(js));
   js.m = m;
   js.p = p;
   js.lb = lb;
   js.ub = ub;
   js.clb = (double*)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1));
   js.cub = (double*)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1));
   js.flb = (double*)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1));
   js.fub = (double*)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1));
   js.touched = (uint8_t*)calloc((size_t)(nc > 0 ? nc : 1), 1);
   if( R != NULL )
   {
      js.clbrow = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nc > 0 ? nc : 1));
      js.cubrow = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nc > 0 ? nc : 1));
      js.flbrow = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nc > 0 ? nc : 1));
      js.fubrow = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nc > 0 ? nc : 1));
      memset(R->conflict, 0, sizeof(*R->conflict));
      R->conflict->row = R->conflict->col = R->conflict->rowlb = R->conflict->rowub = -1;
      R->n = 0;
   }
   for( j = 0; j < nc; ++j )
   {
      js.clb[j] = -HUGE_VAL;
      js.flb[j] = -HUGE_VAL;
      js.cub[j] = HUGE_VAL;
      js.fub[j] = HUGE_VAL;
   }
   if( mark0 != NULL )
      memcpy(marked, mark0, (size_t)nr);
   else
      memset(marked, 1, (size_t)(nr > 0 ? nr : 1));

   for( ;; )
   {
      int anymarked = 0, anychg = 0;
      for( i = 0; i < nr; ++i )
         if( marked[i] && !redundant[i] )
         {
            anymarked = 1;
            break;
         }
      if( !anymarked || (p->maxrounds >= 0 && round >= p->maxrounds) )
         break;
      ++round;

      /* with reasons every marked row of the round is swept, and the smallest infeasible row
       * is the certificate */
      for( i = 0; i < nr && (R != NULL || !js.cutoff); ++i )
      {
         int before = js.cutoff;
         if( redundant[i] || !(marked[i] || full_sweep) )
            continue;
         res->rows_swept++;
         res->nnz_swept += m->rowptr[i + 1] - m->rowptr[i];
         js.cutoff = 0;
         js.currow = (int32_t)i;
         if( jRow(&js, i, xmin, xmax, scratch) )
            redundant[i] = 1;
         if( js.cutoff && R != NULL && R->conflict->kind == 0 )
         {
            R->conflict->kind = ORC_CONFLICT_ROW;
            R->conflict->row = (int32_t)i;
            R->conflict->side = js.rowside;
            R->conflict->round = round;
         }
         js.cutoff |= before;
      }
      if( js.cutoff )
         break;

      /* a round ending infeasible applies nothing: its entries are dropped below */
      nsave = R != NULL ? R->n : 0;
      for( j = 0; j < nc; ++j )
      {
         int src[2], wascut = js.cutoff;
         colchg[j] = 0;
         if( !js.touched[j] )
            continue;
         js.touched[j] = 0;
         colchg[j] = (uint8_t)jColumn(p, m->integral[j], &lb[j], &ub[j], js.clb[j], js.flb[j], js.cub[j], js.fub[j],
            &js.cutoff, &res->nchgbds, src);
         if( R != NULL )
         {
            int32_t rl = src[0] == 1 ? js.flbrow[j] : src[0] == 2 ? js.clbrow[j] : -1;
            int32_t ru = src[1] == 1 ? js.fubrow[j] : src[1] == 2 ? js.cubrow[j] : -1;
            if( js.cutoff && !wascut && R->conflict->kind == 0 )
            {
               /* the first (smallest) column whose bounds crossed */
               R->conflict->kind = ORC_CONFLICT_COLUMN;
               R->conflict->col = (int32_t)j;
               R->conflict->rowlb = rl < 0 ? -1 : rl >> 1;
               R->conflict->rowub = ru < 0 ? -1 : ru >> 1;
               R->conflict->rules = (rl < 0 ? 0 : (rl & 1)) | (ru < 0 ? 0 : (ru & 1) << 1);
               R->conflict->round = round;
            }
            else if( colchg[j] )
            {
               if( src[0] )
                  histAdd(R, round, (int32_t)j, 0, lb[j], rl);
               if( src[1] )
                  histAdd(R, round, (int32_t)j, 1, ub[j], ru);
            }
         }
         js.clb[j] = -HUGE_VAL;
         js.flb[j] = -HUGE_VAL;
         js.cub[j] = HUGE_VAL;
         js.fub[j] = HUGE_VAL;
         anychg |= colchg[j];
      }
      if( js.cutoff && R != NULL )
         R->n = nsave;
      if( js.cutoff )
         break;
      for( i = 0; i < nr; ++i )
      {
         marked[i] = 0;
         for( k = m->rowptr[i]; k < m->rowptr[i + 1]; ++k )
            if( colchg[m->colidx[k]] )
            {
               marked[i] = 1;
               break;
            }
      }
      if( !anychg )
         break;
   }
   res->cutoff = js.cutoff;
   res->rounds = round;
   res->nrecompute = js.nrecomp;

   free(marked);
   free(redundant);
   free(colchg);
   free(xmin);
   free(xmax);
   free(scratch);
   free(js.clb);
   free(js.cub);
   free(js.flb);
   free(js.fub);
   free(js.touched);
   free(js.clbrow);
   free(js.cubrow);
   free(js.flbrow);
   free(js.fubrow);
   return 0;
}

int orc_propagate_jacobi(const orc_matrix* m, const orc_params* p, double* lb, double* ub, int full_sweep,
   orc_result* res)
{
   return jacobi(m, p, lb, ub, full_sweep, NULL, res, NULL);
}

int orc_propagate_jacobi_marked(const orc_matrix* m, const orc_params* p, double* lb, double* ub,
   const uint8_t* mark0, orc_result* res)
{
   return jacobi(m, p, lb, ub, 0, mark0, res, NULL);
}

int orc_propagate_jacobi_reasons(const orc_matrix* m, const orc_params* p, double* lb, double* ub, int64_t histcap,
   int64_t* nhist, int32_t* hround, int32_t* hcol, int32_t* hbound, double* hval, int32_t* hrow, int32_t* hrule,
   orc_conflict* conflict, orc_result* res)
{
   jreasons R;
   int rc;
   R.cap = histcap;
   R.n = 0;
   R.round = hround;
   R.col = hcol;
   R.bound = hbound;
   R.val = hval;
   R.row = hrow;
   R.rule = hrule;
   R.conflict = conflict;
   rc = jacobi(m, p, lb, ub, 0, NULL, res, &R);
   *nhist = R.n;
   if( !res->cutoff )
      conflict->kind = ORC_CONFLICT_NONE;
   return rc;
}

/* ==========================================================================================
 * SCIP-order restatement (Gauss-Seidel)
 *
 * Each linear row keeps the SCIP_ConsData fields used on the propagation path
 * (cons_linear.c:187-266): double-double activities updated incrementally by bound-change
 * events, last reliable activities, infinity/huge counters, maxactdelta and its variable,
 * boundstightened, the propagation sort order and nbinvars.  Bound changes go through
 * SCIPinferVar{Lb,Ub}Cons and immediately update every row containing the variable.
 *
 * Round order: a propagation round processes the rows marked at its start in ascending
 * row index; a row is marked for the next round when one of its bounds is tightened after
 * its own processing in the current round has begun (cons.c:7309-7366 delayed marking).
 * SCIP's propconss array is permuted by swaps and ages; that permutation is not
 * reproduced, which changes nothing for problems whose fixpoint is unique (pure binary)
 * and only the order-dependent 5% filter outcome otherwise.
 * ======================================================================================== */
typedef struct
{
   int64_t start;                   /* offset into sorted var arrays */
   int     nvars;
   int     nbinvars;
   int     hasnonbinvar;
   double  lhs, rhs;
   dd      minact, maxact;
   double  lastminact, lastmaxact;
   int     minposinf, minneginf, minposhuge, minneghuge;
   int     maxposinf, maxneginf, maxposhuge, maxneghuge;
   int     validminact, validmaxact;
   double  maxactdelta;
   int32_t maxactdeltavar;
   int     boundstightened;
   int     deleted;
} gsrow;

typedef struct
{
   const orc_matrix* m;
   const orc_params* p;
   double* lb;
   double* ub;
   gsrow* rows;
   int32_t* svar;        /* sorted var indices per row */
   double*  sval;        /* sorted coefficients per row */
   int64_t* cptr;        /* column -> (row, position) lists */
   int64_t* crow;
   int32_t* cpos;
   uint8_t* started;     /* processing of the row began in this round */
   uint8_t* nextmark;
   int64_t  nchg;
   int      cutoff;
   int64_t  nrecomp;     /* reliable residual recomputes */
} gsstate;

/* the qsort comparator's context; thread-local, so that tests may run the oracle on
 * several instances at once from threads */
static __thread const orc_params* g_sortp;
static __thread const gsstate* g_sortst;
static __thread int64_t g_sortrow;

/* consdataCompVarProp (cons_linear.c:3191-3257) on the row's original positions */
static int cmpVarProp(const void* pa, const void* pb)
{
   const orc_matrix* m = g_sortst->m;
   int64_t ka = *(const int64_t*)pa, kb = *(const int64_t*)pb;
   int32_t va = m->colidx[ka], vb = m->colidx[kb];
   int ba = m->binary[va], bb = m->binary[vb];
   (void)g_sortrow;
   if( ba != bb )
      return ba ? -1 : 1;
   if( ba )
   {
      double c1 = fabs(m->vals[ka]), c2 = fabs(m->vals[kb]);
      if( c1 - c2 > 1e-9 )
         return -1;
      if( c2 - c1 > 1e-9 )
         return 1;
      return (va > vb) - (va < vb);
   }
   else
   {
      int ia = m->integral[va], ib = m->integral[vb];
      /* SCIP_VARTYPE: integer before continuous */
      if( ia != ib )
         return ia ? -1 : 1;
      if( !ia )
         return (va > vb) - (va < vb);
      else
      {
         /* global domain widths enter via the bounds at the start of propagation */
         double c1 = fabs(m->vals[ka] * (g_sortst->ub[va] - g_sortst->lb[va]));
         double c2 = fabs(m->vals[kb] * (g_sortst->ub[vb] - g_sortst->lb[vb]));
         if( c1 - c2 > 1e-9 )
            return -1;
         if( c2 - c1 > 1e-9 )
            return 1;
         return (va > vb) - (va < vb);
      }
   }
}

/* consdataRecomputeMinactivity / Maxactivity (cons_linear.c:1307-1358): reverse order */
static void gsRecomputeMin(gsstate* st, gsrow* r)
{
   const orc_params* p = st->p;
   int v;
   r->minact.hi = 0.0;
   r->minact.lo = 0.0;
   for( v = r->nvars - 1; v >= 0; --v )
   {
      double a = st->sval[r->start + v];
      int32_t j = st->svar[r->start + v];
      double b = a > 0.0 ? st->lb[j] : st->ub[j];
      if( !isInf(p, b) && !isInf(p, -b) && !isHuge(p, a * b) && !isHuge(p, -a * b) )
         r->minact = sum21(r->minact, a * b);
   }
   r->validminact = 1;
   r->lastminact = ddVal(r->minact);
}

static void gsRecomputeMax(gsstate* st, gsrow* r)
{
   const orc_params* p = st->p;
   int v;
   r->maxact.hi = 0.0;
   r->maxact.lo = 0.0;
   for( v = r->nvars - 1; v >= 0; --v )
   {
      double a = st->sval[r->start + v];
      int32_t j = st->svar[r->start + v];
      double b = a > 0.0 ? st->ub[j] : st->lb[j];
      if( !isInf(p, b) && !isInf(p, -b) && !isHuge(p, a * b) && !isHuge(p, -a * b) )
         r->maxact = sum21(r->maxact, a * b);
   }
   r->validmaxact = 1;
   r->lastmaxact = ddVal(r->maxact);
}

/* consdataUpdateActivities (cons_linear.c:1604-1992), local activities only */
static void gsUpdateAct(gsstate* st, gsrow* r, double oldbound, double newbound, double val, int islower,
   int checkrel)
{
   const orc_params* p = st->p;
   dd* activity;
   double* lastactivity;
   int *aposinf, *aneginf, *aposhuge, *aneghuge;
   int validact;
   int* validflag;
   double oldc, newc;
   int finitenew, hugenew;
   dd delta = { 0.0, 0.0 };

   if( islower == (val > 0.0) )
   {
      activity = &r->minact;
      lastactivity = &r->lastminact;
      aposhuge = &r->minposhuge;
      aneghuge = &r->minneghuge;
      validflag = &r->validminact;
      if( val > 0.0 )
      {
         aposinf = &r->minposinf;
         aneginf = &r->minneginf;
      }
      else
      {
         aposinf = &r->minneginf;
         aneginf = &r->minposinf;
      }
   }
   else
   {
      activity = &r->maxact;
      lastactivity = &r->lastmaxact;
      aposhuge = &r->maxposhuge;
      aneghuge = &r->maxneghuge;
      validflag = &r->validmaxact;
      if( val > 0.0 )
      {
         aposinf = &r->maxposinf;
         aneginf = &r->maxneginf;
      }
      else
      {
         aposinf = &r->maxneginf;
         aneginf = &r->maxposinf;
      }
   }
   validact = *validflag;

   oldc = val * oldbound;
   newc = val * newbound;
   hugenew = isHuge(p, fabs(newc));
   finitenew = !isInf(p, fabs(newbound));

   if( isInf(p, fabs(oldbound)) )
   {
      if( oldbound > 0.0 )
      {
         if( finitenew || newbound < 0.0 )
         {
            (*aposinf)--;
            if( !finitenew && newbound < 0.0 )
               (*aneginf)++;
            else if( hugenew )
            {
               if( newc > 0.0 )
                  (*aposhuge)++;
               else
                  (*aneghuge)++;
            }
            else
            {
               delta.hi = newc;
               delta.lo = 0.0;
            }
         }
      }
      else
      {
         if( finitenew || newbound > 0.0 )
         {
            (*aneginf)--;
            if( !finitenew && newbound > 0.0 )
               (*aposinf)++;
            else if( hugenew )
            {
               if( newc > 0.0 )
                  (*aposhuge)++;
               else
                  (*aneghuge)++;
            }
            else
            {
               delta.hi = newc;
               delta.lo = 0.0;
            }
         }
      }
   }
   else if( isHuge(p, fabs(oldc)) )
   {
      if( oldc > 0.0 )
         (*aposhuge)--;
      else
         (*aneghuge)--;
      if( !finitenew )
      {
         if( newbound > 0.0 )
            (*aposinf)++;
         else
            (*aneginf)++;
      }
      else if( hugenew )
      {
         if( newc > 0.0 )
            (*aposhuge)++;
         else
            (*aneghuge)++;
      }
      else
      {
         delta.hi = newc;
         delta.lo = 0.0;
      }
   }
   else
   {
      if( !finitenew )
      {
         if( newbound > 0.0 )
            (*aposinf)++;
         else
            (*aneginf)++;
         delta.hi = -oldc;
         delta.lo = 0.0;
      }
      else if( hugenew )
      {
         if( newc > 0.0 )
            (*aposhuge)++;
         else
            (*aneghuge)++;
         delta.hi = -oldc;
         delta.lo = 0.0;
      }
      else
      {
         delta.hi = newc;
         delta.lo = 0.0;
         delta = sum21(delta, -oldc);
      }
   }

   if( validact && ddVal(delta) != 0.0 )
   {
      double cur;
      *activity = sum21(*activity, ddVal(delta));
      cur = ddVal(*activity);
      if( fabs(*lastactivity) < fabs(cur) )
         *lastactivity = cur;
      else if( checkrel && isUpdateUnreliable(p, cur, *lastactivity) )
         *validflag = 0;
   }
}

/* consdataCalcActivities (cons_linear.c:2283-2339) in the row's current var order */
static void gsCalcActivities(gsstate* st, gsrow* r)
{
   int v;
   r->minact.hi = r->minact.lo = 0.0;
   r->maxact.hi = r->maxact.lo = 0.0;
   r->lastminact = r->lastmaxact = 0.0;
   r->minposinf = r->minneginf = r->minposhuge = r->minneghuge = 0;
   r->maxposinf = r->maxneginf = r->maxposhuge = r->maxneghuge = 0;
   r->validminact = r->validmaxact = 1;
   for( v = 0; v < r->nvars; ++v )
   {
      int32_t j = st->svar[r->start + v];
      double a = st->sval[r->start + v];
      gsUpdateAct(st, r, 0.0, st->lb[j], a, 1, 0);
      gsUpdateAct(st, r, 0.0, st->ub[j], a, 0, 0);
   }
   r->lastminact = ddVal(r->minact);
   r->lastmaxact = ddVal(r->maxact);
}

static void gsRecomputeMaxActDelta(gsstate* st, gsrow* r)
{
   const orc_params* p = st->p;
   int v;
   r->maxactdelta = 0.0;
   if( !r->hasnonbinvar )
   {
      for( v = r->nvars - 1; v >= 0; --v )
      {
         int32_t j = st->svar[r->start + v];
         if( st->lb[j] < 0.5 && st->ub[j] > 0.5 )
         {
            double d = fabs(st->sval[r->start + v]);
            if( d > r->maxactdelta )
            {
               r->maxactdelta = d;
               r->maxactdeltavar = j;
            }
         }
      }
      return;
   }
   for( v = r->nvars - 1; v >= 0; --v )
   {
      int32_t j = st->svar[r->start + v];
      double d;
      if( isInf(p, -st->lb[j]) || isInf(p, st->ub[j]) )
      {
         r->maxactdelta = p->infinity;
         r->maxactdeltavar = j;
         break;
      }
      d = fabs(st->sval[r->start + v]) * (st->ub[j] - st->lb[j]);
      if( d > r->maxactdelta )
      {
         r->maxactdelta = d;
         r->maxactdeltavar = j;
      }
   }
}

/* eventExecLinear (cons_linear.c:17162-17279) for a tightening of var j */
static void gsBoundEvent(gsstate* st, int32_t j, double oldb, double newb, int islower, int64_t currow)
{
   const orc_params* p = st->p;
   int64_t q;
   for( q = st->cptr[j]; q < st->cptr[j + 1]; ++q )
   {
      int64_t ri = st->crow[q];
      gsrow* r = &st->rows[ri];
      double val = st->sval[r->start + st->cpos[q]];
      if( r->deleted )
         continue;
      gsUpdateAct(st, r, oldb, newb, val, islower, 1);
      /* SCIPmarkConsPropagate with delayed updates: only a row whose processing began in
       * this round (or the row being processed) ends up marked for the next round */
      if( st->started[ri] || ri == currow )
         st->nextmark[ri] = 1;
      if( r->maxactdeltavar == j )
      {
         r->maxactdelta = ORC_INVALID;
         r->maxactdeltavar = -1;
      }
      if( r->boundstightened > 0 )
      {
         if( islower )
         {
            if( val > 0.0 ? !isInf(p, r->rhs) : !isInf(p, -r->lhs) )
               r->boundstightened = 0;
         }
         else
         {
            if( val > 0.0 ? !isInf(p, -r->lhs) : !isInf(p, r->rhs) )
               r->boundstightened = 0;
         }
      }
   }
}

/* SCIPinferVarUbCons / SCIPinferVarLbCons (scip_var.c:6965-7174) */
static void gsInferUb(gsstate* st, int32_t j, double newbound, int force, int* infeasible, int* tightened,
   int64_t currow)
{
   const orc_params* p = st->p;
   double lb, ub;
   *infeasible = 0;
   *tightened = 0;
   newbound = adjustedUb(p, st->m->integral[j], newbound);
   lb = st->lb[j];
   ub = st->ub[j];
   if( isInf(p, -newbound) || isFeasLT(p, newbound, lb) )
   {
      *infeasible = 1;
      return;
   }
   newbound = MAXD(newbound, lb);
   if( (force && isGE(p, newbound, ub)) || (!force && !isUbBetter(p, newbound, lb, ub)) )
      return;
   if( newbound < ub )
   {
      st->ub[j] = newbound;
      *tightened = 1;
      st->nchg++;
      gsBoundEvent(st, j, ub, newbound, 0, currow);
   }
}

static void gsInferLb(gsstate* st, int32_t j, double newbound, int force, int* infeasible, int* tightened,
   int64_t currow)
{
   const orc_params* p = st->p;
   double lb, ub;
   *infeasible = 0;
   *tightened = 0;
   newbound = adjustedLb(p, st->m->integral[j], newbound);
   lb = st->lb[j];
   ub = st->ub[j];
   if( isInf(p, newbound) || isFeasGT(p, newbound, ub) )
   {
      *infeasible = 1;
      return;
   }
   newbound = MIND(newbound, ub);
   if( (force && isLE(p, newbound, lb)) || (!force && !isLbBetter(p, newbound, lb, ub)) )
      return;
   if( newbound > lb )
   {
      st->lb[j] = newbound;
      *tightened = 1;
      st->nchg++;
      gsBoundEvent(st, j, lb, newbound, 1, currow);
   }
}

/* tightenVarUb / tightenVarLb (cons_linear.c:5242-5376) */
static void gsTightenVarUb(gsstate* st, int32_t j, double newub, double oldub, int force, int64_t ri)
{
   int inf, t;
   double lb = st->lb[j];
   newub = adjustedUb(st->p, st->m->integral[j], newub);
   if( force || isUbBetter(st->p, newub, lb, oldub) )
   {
      gsInferUb(st, j, newub, force, &inf, &t, ri);
      if( inf )
         st->cutoff = 1;
   }
}

static void gsTightenVarLb(gsstate* st, int32_t j, double newlb, double oldlb, int force, int64_t ri)
{
   int inf, t;
   double ub = st->ub[j];
   newlb = adjustedLb(st->p, st->m->integral[j], newlb);
   if( force || isLbBetter(st->p, newlb, oldlb, ub) )
   {
      gsInferLb(st, j, newlb, force, &inf, &t, ri);
      if( inf )
         st->cutoff = 1;
   }
}

static rowstate gsRowState(const gsrow* r)
{
   rowstate rs;
   rs.minact = r->minact;
   rs.maxact = r->maxact;
   rs.minposinf = r->minposinf;
   rs.minneginf = r->minneginf;
   rs.minposhuge = r->minposhuge;
   rs.minneghuge = r->minneghuge;
   rs.maxposinf = r->maxposinf;
   rs.maxneginf = r->maxneginf;
   rs.maxposhuge = r->maxposhuge;
   rs.maxneghuge = r->maxneghuge;
   rs.maxactdelta = r->maxactdelta;
   return rs;
}

/* reliable residual over the row's current var order (cons_linear.c:2580-2657) */
static double gsReliableResidual(gsstate* st, gsrow* r, int skip, int isminres)
{
   int v;
   double res = 0.0;
   st->nrecomp++;
   for( v = 0; v < r->nvars; ++v )
   {
      int32_t j;
      double a;
      if( v == skip )
         continue;
      j = st->svar[r->start + v];
      a = st->sval[r->start + v];
      if( a > 0.0 )
         res += a * (isminres ? st->lb[j] : st->ub[j]);
      else
         res += a * (isminres ? st->ub[j] : st->lb[j]);
   }
   return res;
}

/* tightenVarBoundsEasy with live values (cons_linear.c:5380-5653) */
static void gsEasy(gsstate* st, int64_t ri, int v, int force)
{
   const orc_params* p = st->p;
   gsrow* r = &st->rows[ri];
   int32_t j = st->svar[r->start + v];
   double a = st->sval[r->start + v];
   double lhs = r->lhs, rhs = r->rhs;
   double lb = st->lb[j], ub = st->ub[j];

   if( !r->validminact )
      gsRecomputeMin(st, r);
   if( a > 0.0 )
   {
      if( !isInf(p, rhs) )
      {
         double slack, alpha;
         if( isFeasLT(p, rhs, ddVal(r->minact)) )
         {
            st->cutoff = 1;
            return;
         }
         slack = rhs - ddVal(r->minact);
         if( !isPositive(p, slack) )
            slack = 0.0;
         alpha = a * (ub - lb);
         if( isSumGT(p, alpha, slack) || (force && isGT(p, alpha, slack)) )
         {
            gsTightenVarUb(st, j, lb + (slack / a), ub, force, ri);
            if( st->cutoff )
               return;
            ub = st->ub[j];
         }
      }
      if( !isInf(p, -lhs) )
      {
         double slack, alpha;
         if( !r->validmaxact )
            gsRecomputeMax(st, r);
         if( isFeasLT(p, ddVal(r->maxact), lhs) )
         {
            st->cutoff = 1;
            return;
         }
         slack = ddVal(r->maxact) - lhs;
         if( !isPositive(p, slack) )
            slack = 0.0;
         alpha = a * (ub - lb);
         if( isSumGT(p, alpha, slack) || (force && isGT(p, alpha, slack)) )
         {
            gsTightenVarLb(st, j, ub - (slack / a), lb, force, ri);
            if( st->cutoff )
               return;
         }
      }
   }
   else
   {
      if( !isInf(p, rhs) )
      {
         double slack, alpha;
         if( isFeasLT(p, rhs, ddVal(r->minact)) )
         {
            st->cutoff = 1;
            return;
         }
         slack = rhs - ddVal(r->minact);
         if( !isPositive(p, slack) )
            slack = 0.0;
         alpha = a * (lb - ub);
         if( isSumGT(p, alpha, slack) || (force && isGT(p, alpha, slack)) )
         {
            gsTightenVarLb(st, j, ub + slack / a, lb, force, ri);
            if( st->cutoff )
               return;
            lb = st->lb[j];
         }
      }
      if( !isInf(p, -lhs) )
      {
         double slack, alpha;
         if( !r->validmaxact )
            gsRecomputeMax(st, r);
         if( isFeasLT(p, ddVal(r->maxact), lhs) )
         {
            st->cutoff = 1;
            return;
         }
         slack = ddVal(r->maxact) - lhs;
         if( !isPositive(p, slack) )
            slack = 0.0;
         alpha = a * (lb - ub);
         if( isSumGT(p, alpha, slack) || (force && isGT(p, alpha, slack)) )
         {
            gsTightenVarUb(st, j, lb - (slack / a), ub, force, ri);
            if( st->cutoff )
               return;
         }
      }
   }
}

/* getMinActivity/getMaxActivity recompute an invalid activity only when they reach the
 * finite branch (cons_linear.c:2391-2405, 2486-2500) */
static void gsEnsureValidFor(gsstate* st, gsrow* r, int minpinf, int minninf, int minphuge, int minnhuge,
   int maxpinf, int maxninf, int maxphuge, int maxnhuge, int goodrelax)
{
   if( !r->validminact && !(minninf > 0 || minpinf > 0 || minnhuge > 0 || (minphuge > 0 && !goodrelax)) )
      gsRecomputeMin(st, r);
   if( !r->validmaxact && !(maxpinf > 0 || maxninf > 0 || maxphuge > 0 || (maxnhuge > 0 && !goodrelax)) )
      gsRecomputeMax(st, r);
}

static void gsEnsureValid(gsstate* st, gsrow* r, int goodrelax)
{
   gsEnsureValidFor(st, r, r->minposinf, r->minneginf, r->minposhuge, r->minneghuge, r->maxposinf, r->maxneginf,
      r->maxposhuge, r->maxneghuge, goodrelax);
}

/* counters consdataGetActivityResiduals passes on for the variable (cons_linear.c:2710-2806) */
static void gsEnsureValidResidual(gsstate* st, gsrow* r, double a, double lb, double ub)
{
   const orc_params* p = st->p;
   double minb = a > 0.0 ? lb : -ub, maxb = a > 0.0 ? ub : -lb, absval = fabs(a);
   int c[8] = { r->minposinf, r->minneginf, r->minposhuge, r->minneghuge,
                r->maxposinf, r->maxneginf, r->maxposhuge, r->maxneghuge };
   if( isInf(p, minb) )
      c[0]--;
   else if( isInf(p, -minb) )
      c[1]--;
   else if( isHuge(p, minb * absval) )
      c[2]--;
   else if( isHuge(p, -minb * absval) )
      c[3]--;
   if( isInf(p, -maxb) )
      c[5]--;
   else if( isInf(p, maxb) )
      c[4]--;
   else if( isHuge(p, absval * maxb) )
      c[6]--;
   else if( isHuge(p, -absval * maxb) )
      c[7]--;
   gsEnsureValidFor(st, r, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], 0);
}

/* tightenVarBounds with live values (cons_linear.c:6700-6974) */
static void gsHard(gsstate* st, int64_t ri, int v, int force)
{
   const orc_params* p = st->p;
   gsrow* r = &st->rows[ri];
   int32_t j = st->svar[r->start + v];
   double a = st->sval[r->start + v];
   double lhs = r->lhs, rhs = r->rhs;
   double lb, ub, minres, maxres;
   int mintight, maxtight, minsetinf, maxsetinf, inf, t;
   rowstate rs;

   lb = st->lb[j];
   ub = st->ub[j];
   gsEnsureValidResidual(st, r, a, lb, ub);
   rs = gsRowState(r);
   residuals(p, &rs, a, lb, ub, 0, &minres, &maxres, &mintight, &maxtight, &minsetinf, &maxsetinf);
   {
      int integral = st->m->integral[j];
      if( a > 0.0 )
      {
         if( !minsetinf && !isInf(p, rhs) && mintight && (isLT(p, fabs(rhs), 1.0) || !isEQ(p, minres / rhs, 1.0)) )
         {
            double newub = (rhs - minres) / a;
            if( !isInf(p, newub)
               && ((force && isLT(p, newub, ub)) || (integral && isFeasLT(p, newub, ub)) || isUbBetter(p, newub, lb, ub)) )
            {
               int unrel = isUpdateUnreliable(p, minres, r->lastminact);
               if( unrel )
               {
                  minres = gsReliableResidual(st, r, v, 1);
                  newub = (rhs - minres) / a;
                  unrel = isInf(p, -minres)
                     || (!isUbBetter(p, newub, lb, ub) && (!isFeasLT(p, newub, ub) || !integral)
                        && (!force || !isLT(p, newub, ub)));
               }
               if( !unrel )
               {
                  gsInferUb(st, j, newub, force, &inf, &t, ri);
                  if( inf )
                  {
                     st->cutoff = 1;
                     return;
                  }
                  if( t )
                     ub = st->ub[j];
               }
            }
         }
         if( !maxsetinf && !isInf(p, -lhs) && maxtight && (isLT(p, fabs(lhs), 1.0) || !isEQ(p, maxres / lhs, 1.0)) )
         {
            double newlb = (lhs - maxres) / a;
            if( !isInf(p, -newlb)
               && ((force && isGT(p, newlb, lb)) || (integral && isFeasGT(p, newlb, lb)) || isLbBetter(p, newlb, lb, ub)) )
            {
               if( isUpdateUnreliable(p, maxres, r->lastmaxact) )
               {
                  maxres = gsReliableResidual(st, r, v, 0);
                  newlb = (lhs - maxres) / a;
                  if( isInf(p, maxres) || (!isLbBetter(p, newlb, lb, ub) && (!isFeasGT(p, newlb, lb) || !integral)
                        && (!force || !isGT(p, newlb, lb))) )
                     return;
               }
               gsInferLb(st, j, newlb, force, &inf, &t, ri);
               if( inf )
               {
                  st->cutoff = 1;
                  return;
               }
            }
         }
      }
      else
      {
         if( !minsetinf && !isInf(p, rhs) && mintight && (isLT(p, fabs(rhs), 1.0) || !isEQ(p, minres / rhs, 1.0)) )
         {
            double newlb = (rhs - minres) / a;
            if( !isInf(p, -newlb)
               && ((force && isGT(p, newlb, lb)) || (integral && isFeasGT(p, newlb, lb)) || isLbBetter(p, newlb, lb, ub)) )
            {
               int unrel = isUpdateUnreliable(p, minres, r->lastminact);
               if( unrel )
               {
                  minres = gsReliableResidual(st, r, v, 1);
                  newlb = (rhs - minres) / a;
                  unrel = isInf(p, -minres)
                     || (!isLbBetter(p, newlb, lb, ub) && (!isFeasGT(p, newlb, lb) || !integral)
                        && (!force || !isGT(p, newlb, lb)));
               }
               if( !unrel )
               {
                  gsInferLb(st, j, newlb, force, &inf, &t, ri);
                  if( inf )
                  {
                     st->cutoff = 1;
                     return;
                  }
                  if( t )
                     lb = st->lb[j];
               }
            }
         }
         if( !maxsetinf && !isInf(p, -lhs) && maxtight && (isLT(p, fabs(lhs), 1.0) || !isEQ(p, maxres / lhs, 1.0)) )
         {
            double newub = (lhs - maxres) / a;
            if( !isInf(p, newub)
               && ((force && isLT(p, newub, ub)) || (integral && isFeasLT(p, newub, ub)) || isUbBetter(p, newub, lb, ub)) )
            {
               if( isUpdateUnreliable(p, maxres, r->lastmaxact) )
               {
                  maxres = gsReliableResidual(st, r, v, 0);
                  newub = (lhs - maxres) / a;
                  if( isInf(p, maxres) || (!isUbBetter(p, newub, lb, ub) && (!isFeasLT(p, newub, ub) && !integral)
                        && (!force || !isLT(p, newub, ub))) )
                     return;
               }
               gsInferUb(st, j, newub, force, &inf, &t, ri);
               if( inf )
               {
                  st->cutoff = 1;
                  return;
               }
            }
         }
      }
   }
}

#define ORC_MAXTIGHTENROUNDS 10   /* cons_linear.c:6976 */

/* tightenBounds (cons_linear.c:6980-7154), tightenmode 1 */
static void gsTightenBounds(gsstate* st, int64_t ri)
{
   const orc_params* p = st->p;
   gsrow* r = &st->rows[ri];
   int infmin, infmax, force, easy, nrounds, lastchange, v;
   const int tightenmode = 1;

   infmin = r->minneginf + r->minposinf + r->minneghuge + r->minposhuge;
   infmax = r->maxneginf + r->maxposinf + r->maxneghuge + r->maxposhuge;
   if( infmin > 1 && infmax > 1 )
      return;
   force = (r->nvars == 1);
   if( !force && r->boundstightened >= tightenmode )
      return;
   if( r->maxactdelta == ORC_INVALID )
      gsRecomputeMaxActDelta(st, r);
   if( isFeasZero(p, r->maxactdelta) )
      return;
   if( !isInf(p, r->maxactdelta) )
   {
      double minact, maxact, slack, surplus;
      int t1, t2, f1, f2;
      gsEnsureValid(st, r, 0);
      getMinAct(p, r->minact, r->minposinf, r->minneginf, r->minposhuge, r->minneghuge, 0.0, 0, &minact, &t1, &f1);
      getMaxAct(p, r->maxact, r->maxposinf, r->maxneginf, r->maxposhuge, r->maxneghuge, 0.0, 0, &maxact, &t2, &f2);
      slack = (isInf(p, r->rhs) || f1) ? p->infinity : (r->rhs - minact);
      surplus = (isInf(p, -r->lhs) || f2) ? p->infinity : (maxact - r->lhs);
      if( isLE(p, r->maxactdelta, MIND(slack, surplus)) )
         return;
   }
   easy = isLT(p, r->maxactdelta, p->maxeasyactivitydelta);
   lastchange = -1;
   for( nrounds = 0; (force || r->boundstightened < tightenmode) && nrounds < ORC_MAXTIGHTENROUNDS; ++nrounds )
   {
      r->boundstightened = tightenmode;
      v = 0;
      while( v < r->nvars && v != lastchange && !st->cutoff )
      {
         int64_t old = st->nchg;
         if( easy )
            gsEasy(st, ri, v, force);
         else
            gsHard(st, ri, v, force);
         if( st->nchg > old )
         {
            lastchange = v;
            ++v;
         }
         else
         {
            int32_t j = st->svar[r->start + v];
            if( v < r->nbinvars - 1 && !isFeasEQ(p, st->ub[j], st->lb[j]) )
               v = r->nbinvars;
            else
               ++v;
         }
      }
   }
}

/* propagateCons (cons_linear.c:7621-7758) without ranged-row propagation */
static void gsPropagateCons(gsstate* st, int64_t ri)
{
   const orc_params* p = st->p;
   gsrow* r = &st->rows[ri];
   double minact, maxact;
   int t1, t2, f1, f2;

   /* tightenbounds FALSE (consPropLinear, cons_linear.c:16147-16161): only the final test */
   if( !p->notighten )
      gsTightenBounds(st, ri);
   if( st->cutoff )
      return;
   gsEnsureValid(st, r, 1);
   getMinAct(p, r->minact, r->minposinf, r->minneginf, r->minposhuge, r->minneghuge, 0.0, 1, &minact, &t1, &f1);
   getMaxAct(p, r->maxact, r->maxposinf, r->maxneginf, r->maxposhuge, r->maxneghuge, 0.0, 1, &maxact, &t2, &f2);
   if( isFeasGT(p, minact, r->rhs) || isFeasLT(p, maxact, r->lhs) )
      st->cutoff = 1;
   else if( isGE(p, minact, r->lhs) && isLE(p, maxact, r->rhs) )
      r->deleted = 1;
}

int orc_propagate_scip(const orc_matrix* m, const orc_params* p, double* lb, double* ub, orc_result* res)
{
   int64_t nr = m->nrows, nc = m->ncols, nnz = m->rowptr[nr], i, j, k;
   gsstate st;
   int64_t* order;
   int64_t* cfill;
   uint8_t* marked;
   int round = 0;

   memset(res, 0, sizeof(*res));
   memset(&st, 0, sizeof(st));
   st.m = m;
   st.p = p;
   st.lb = lb;
   st.ub = ub;
   st.rows = (gsrow*)calloc((size_t)(nr > 0 ? nr : 1), sizeof(gsrow));
   st.svar = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
   st.sval = (double*)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
   st.cptr = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));
   st.crow = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz > 0 ? nnz : 1));
   st.cpos = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
   st.started = (uint8_t*)calloc((size_t)(nr > 0 ? nr : 1), 1);
   st.nextmark = (uint8_t*)calloc((size_t)(nr > 0 ? nr : 1), 1);
   marked = (uint8_t*)malloc((size_t)(nr > 0 ? nr : 1));
   order = (int64_t*)malloc(sizeof(int64_t) * (size_t)(maxRowLen(m)));
   cfill = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));

   /* sorted variable order per row (consdataSort at INITSOLVE) */
   g_sortp = p;
   g_sortst = &st;
   for( i = 0; i < nr; ++i )
   {
      gsrow* r = &st.rows[i];
      int64_t s = m->rowptr[i], n = m->rowptr[i + 1] - s;
      r->start = s;
      r->nvars = (int)n;
      r->lhs = m->lhs[i];
      r->rhs = m->rhs[i];
      for( k = 0; k < n; ++k )
         order[k] = s + k;
      g_sortrow = i;
      if( n > 1 )
         qsort(order, (size_t)n, sizeof(int64_t), cmpVarProp);
      r->nbinvars = 0;
      r->hasnonbinvar = 0;
      for( k = 0; k < n; ++k )
      {
         st.svar[s + k] = m->colidx[order[k]];
         st.sval[s + k] = m->vals[order[k]];
         if( !m->binary[st.svar[s + k]] )
            r->hasnonbinvar = 1;
      }
      for( k = 0; k < n && m->binary[st.svar[s + k]]; ++k )
         r->nbinvars++;
      r->maxactdelta = ORC_INVALID;
      r->maxactdeltavar = -1;
      gsCalcActivities(&st, r);
   }
   for( k = 0; k < nnz; ++k )
      st.cptr[st.svar[k] + 1]++;
   for( j = 0; j < nc; ++j )
      st.cptr[j + 1] += st.cptr[j];
   for( i = 0; i < nr; ++i )
      for( k = m->rowptr[i]; k < m->rowptr[i + 1]; ++k )
      {
         int32_t c = st.svar[k];
         int64_t q = st.cptr[c] + cfill[c]++;
         st.crow[q] = i;
         st.cpos[q] = (int32_t)(k - m->rowptr[i]);
      }

   memset(marked, 1, (size_t)(nr > 0 ? nr : 1));
   for( ;; )
   {
      int any = 0;
      for( i = 0; i < nr; ++i )
         if( marked[i] && !st.rows[i].deleted )
         {
            any = 1;
            break;
         }
      if( !any || (p->maxrounds >= 0 && round >= p->maxrounds) )
         break;
      ++round;
      memset(st.started, 0, (size_t)(nr > 0 ? nr : 1));
      memset(st.nextmark, 0, (size_t)(nr > 0 ? nr : 1));
      /* rows not in this round's marked set are marked by any event */
      for( i = 0; i < nr; ++i )
         if( !marked[i] )
            st.started[i] = 1;
      for( i = 0; i < nr && !st.cutoff; ++i )
      {
         if( !marked[i] || st.rows[i].deleted )
            continue;
         st.started[i] = 1;
         res->rows_swept++;
         res->nnz_swept += st.rows[i].nvars;
         gsPropagateCons(&st, i);
      }
      if( st.cutoff )
         break;
      memcpy(marked, st.nextmark, (size_t)(nr > 0 ? nr : 1));
   }
   res->cutoff = st.cutoff;
   res->rounds = round;
   res->nchgbds = st.nchg;
   res->nrecompute = st.nrecomp;

   free(st.rows);
   free(st.svar);
   free(st.sval);
   free(st.cptr);
   free(st.crow);
   free(st.cpos);
   free(st.started);
   free(st.nextmark);
   free(marked);
   free(order);
   free(cfill);
   return 0;
}
