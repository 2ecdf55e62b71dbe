// devnum.h -- device-side SCIP numerics and double-double arithmetic for the propagation
// kernels.  Every function is the IEEE FP64 operation sequence of the reference predicate
// (compiled with -ffp-contract=off, no fast-math), so results are bit-identical to a
// host build of the same expressions.
//
//   def.h:191-204 EPS* macros; set.c:6824-7753 SCIPsetIs*; set.c:7711/7735 IsLb/UbBetter;
//   set.c:8032 IsUpdateUnreliable; misc.c:11162 SCIPrelDiff; var.c:1957/2006 adjustedLb/Ub;
//   dbldblarith.h:154-187 TwoSum and Sum21.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sap {

struct Params
{
   double  infinity;
   double  epsilon;
   double  sumepsilon;
   double  feastol;
   double  boundstreps;
   double  hugeval;
   double  recompfac;
   double  maxeasy;
   int32_t maxrounds;
   int32_t flags;
};

struct DD
{
   double hi;
   double lo;
};

__device__ __forceinline__ double dmax(double x, double y) { return x >= y ? x : y; }
__device__ __forceinline__ double dmin(double x, double y) { return x <= y ? x : y; }

__device__ __forceinline__ bool isInf(const Params& p, double x) { return x >= p.infinity; }
__device__ __forceinline__ bool isHuge(const Params& p, double x) { return x >= p.hugeval; }
__device__ __forceinline__ bool isEQ(const Params& p, double a, double b) { return fabs(a - b) <= p.epsilon; }
__device__ __forceinline__ bool isLT(const Params& p, double a, double b) { return a - b < -p.epsilon; }
__device__ __forceinline__ bool isLE(const Params& p, double a, double b) { return a - b <= p.epsilon; }
__device__ __forceinline__ bool isGT(const Params& p, double a, double b) { return a - b > p.epsilon; }
__device__ __forceinline__ bool isGE(const Params& p, double a, double b) { return a - b >= -p.epsilon; }
__device__ __forceinline__ bool isPositive(const Params& p, double x) { return x > p.epsilon; }
__device__ __forceinline__ bool isSumGT(const Params& p, double a, double b) { return a - b > p.sumepsilon; }

__device__ __forceinline__ double relDiff(double a, double b)
{
   double q = dmax(dmax(1.0, fabs(a)), fabs(b));
   return (a - b) / q;
}
__device__ __forceinline__ bool isFeasEQ(const Params& p, double a, double b) { return fabs(relDiff(a, b)) <= p.feastol; }
__device__ __forceinline__ bool isFeasLT(const Params& p, double a, double b) { return relDiff(a, b) < -p.feastol; }
__device__ __forceinline__ bool isFeasGT(const Params& p, double a, double b) { return relDiff(a, b) > p.feastol; }
__device__ __forceinline__ bool isFeasZero(const Params& p, double x) { return fabs(x) <= p.feastol; }

__device__ __forceinline__ bool isLbBetter(const Params& p, double newlb, double oldlb, double oldub)
{
   if( oldlb < 0.0 && newlb >= 0.0 )
      return true;
   double t = dmin(oldub - oldlb, fabs(oldlb));
   t = dmax(t, 1e-3);
   return newlb - oldlb > p.boundstreps * t;
}

__device__ __forceinline__ bool isUbBetter(const Params& p, double newub, double oldlb, double oldub)
{
   if( oldub > 0.0 && newub <= 0.0 )
      return true;
   double t = dmin(oldub - oldlb, fabs(oldub));
   t = dmax(t, 1e-3);
   return newub - oldub < -(p.boundstreps * t);
}

__device__ __forceinline__ bool isUpdateUnreliable(const Params& p, double newv, double oldv)
{
   double q = fabs(oldv) / dmax(fabs(newv), p.epsilon);
   return q >= p.recompfac;
}

__device__ __forceinline__ double adjustedLb(const Params& p, bool integral, double lb)
{
   if( lb < 0.0 && isInf(p, -lb) )
      return -p.infinity;
   if( lb > 0.0 && isInf(p, lb) )
      return p.infinity;
   if( integral )
      return ceil(lb - p.feastol);
   if( lb > 0.0 && lb < p.epsilon )
      return 0.0;
   return lb;
}

__device__ __forceinline__ double adjustedUb(const Params& p, bool integral, double ub)
{
   if( ub > 0.0 && isInf(p, ub) )
      return p.infinity;
   if( ub < 0.0 && isInf(p, -ub) )
      return -p.infinity;
   if( integral )
      return floor(ub + p.feastol);
   if( ub < 0.0 && ub > -p.epsilon )
      return 0.0;
   return ub;
}

// Knuth TwoSum (dbldblarith.h:154-163)
__device__ __forceinline__ void twoSum(double a, double b, double& s, double& e)
{
   s = a + b;
   double bb = s - a;
   e = (a - (s - bb)) + (b - bb);
}

// SCIPdbldblSum21 (dbldblarith.h:180-187): unnormalized, hi = TwoSum.hi, lo = TwoSum.lo + alo
__device__ __forceinline__ DD sum21(DD a, double b)
{
   double h, l;
   twoSum(a.hi, b, h, l);
   DD r;
   r.lo = l + a.lo;
   r.hi = h;
   return r;
}

// node of the canonical pairwise activity tree: commutative, (x) + (0,0) = x
__device__ __forceinline__ DD pairAdd(DD a, DD b)
{
   double h, l;
   twoSum(a.hi, b.hi, h, l);
   DD r;
   r.lo = l + (a.lo + b.lo);
   r.hi = h;
   return r;
}

__device__ __forceinline__ double ddVal(DD a) { return a.hi + a.lo; }

// order-preserving int64 image of a double (total order, -0.0 < +0.0)
__device__ __forceinline__ long long dkey(double x)
{
   long long b = __double_as_longlong(x);
   return b ^ ((b >> 63) & 0x7FFFFFFFFFFFFFFFLL);
}
__device__ __forceinline__ double keyToDouble(long long k)
{
   long long b = k ^ ((k >> 63) & 0x7FFFFFFFFFFFFFFFLL);
   return __longlong_as_double(b);
}

}  // namespace sap
