// engine.hip -- MI355X (gfx950) linear-constraint domain propagation engine.
//
// Device algorithm (one propagation round, all bound vectors of a sap_bounds at once):
//   k_round_begin : advance the round id, stop when the worklists are empty / cutoff /
//                   maxrounds, reset the next round's counters (in a plain call the call's
//                   reset kernel does this for the first round).
//   k_sweep       : one launch over all row-length bins.  Every worklist item (row, vector)
//                   computes the row's min/max activity (canonical pairwise double-double
//                   sum over row positions), infinity/huge counters and maxactdelta from the
//                   round's bounds, then each nonzero computes its bound candidates exactly
//                   as tightenVarBoundsEasy / tightenVarBounds (cons_linear.c:5380, :6700)
//                   and reduces them into per-column candidate slots with 64-bit atomic
//                   min/max on the order-preserving integer image of the FP64 value.  The
//                   row's infeasibility / redundancy test (propagateCons, :7721-7753) runs
//                   on the same state.
//   k_columns     : every touched column applies SCIPinferVarLbCons / SCIPinferVarUbCons
//                   (scip_var.c:6965-7174) to its tightest regular and forced candidates and
//                   writes the new bounds.
//   k_mark        : the rows of changed columns (CSC) are marked for the next round (dense
//                   mode) or appended to its per-bin worklists (sparse mode).
// Rounds run in groups of 1 (enqueued directly), then 2, 4 and 8 (captured graphs); the host
// applies the same stop test as k_round_begin to the status it reads back after each group.
//
// Bins: in the dense mode rows of up to 128 nonzeros are swept in tiles of 64 lanes holding
// 1, 2 or 4 nonzeros each (rows packed first-fit decreasing); in the sparse mode rows of
// length <= 4/8/16/32/64 are swept by groups of that many lanes; longer rows by one wavefront
// (64-element chunks), the longest by a whole workgroup.  All tilings compute the same
// canonical tree, so the result does not depend on the binning.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <memory>
#include <thread>
#include <vector>

#include "devnum.h"
#include "scip_amd_prop.h"

using namespace sap;

#define SAP_NBINS 9
#define SAP_NKINDS (SAP_NBINS + 1)
#define SAP_BLOCK 256
#define SAP_WAVE 64
#define SAP_MAXLEVELS 40
#define SAP_ROUND_CAP 10000000
#define SAP_STAT_SLOTS 256

// row-length bins: 0 empty rows; 1..5 at most 4,8,16,32,64 nonzeros (tiles in dense mode,
// groups of 4,8,16,32,64 lanes in sparse mode); 6 at most 128 nonzeros (tiles of width 2 in
// dense mode, else a wavefront per row); 7 wavefront per row; 8 workgroup per row (rows longer
// than the tuning's longrow, from any of bins 6..8)
static const int32_t kBinMaxLen[SAP_NBINS - 2] = { 0, 4, 8, 16, 32, 64, 128 };

struct Status
{
   int32_t round;       // global round id, monotone over the life of a sap_bounds
   int32_t rlocal;      // rounds executed in the current call
   int32_t done;
   int32_t cutoff;      // single-vector mode: infeasibility detected
   int32_t firstall;    // dense mode: the first round of the call sweeps every row
   int32_t anymark[2];  // dense mode: some row is marked for the round of this parity
   int32_t rgenm1;      // generation of the locally-redundant row flags, minus one (redGen)
   uint32_t wlcount[2][SAP_NBINS];   // sparse mode worklist lengths
   uint32_t touched;                 // sparse mode touched-column count
   uint32_t pad2;
   unsigned long long rows_swept;
   unsigned long long nnz_swept;
   unsigned long long nchg;
   // per-workgroup statistics land in one of SAP_STAT_SLOTS slots (spread over cache lines
   // so workgroups do not serialise on one address); the host sums them after the call
   unsigned long long slot_nnz[SAP_STAT_SLOTS];
   unsigned long long slot_rows[SAP_STAT_SLOTS];
   unsigned long long slot_chg[SAP_STAT_SLOTS];
   // sweep launch duration from the constant-rate wall clock: earliest workgroup start and
   // latest workgroup end per slot, folded at the end of the round
   unsigned long long tslot0[SAP_STAT_SLOTS];
   unsigned long long tslot1[SAP_STAT_SLOTS];
   unsigned long long sweep_ticks;
   unsigned long long sweep_count;
   int32_t cont;        // continuous batching: no global round cap (per-vector caps apply)
   int32_t pad3;
   // reason recording (one vector): infeasibility certificates (smallest caller row * 8 + the
   // row's test code, see RowDecision; smallest column), candidate-log and history lengths
   int32_t cutrow;
   int32_t cutcol;
   uint32_t clogn;
   uint32_t histn;
   int32_t clogovf;
   int32_t pad4;
   // residual activities recomputed in plain double because the incremental one was
   // unreliable (consdataGetReliableResidualActivity's path, counted per call)
   unsigned long long nrecomp;
};

struct Mat
{
   int64_t nrows;
   int64_t ncols;
   const int64_t* rowptr;   // rows in bin order (stable): bin b is [binstart[b], binstart[b+1]),
                            // except that the rows of the tiled bins are in tile order
   const int32_t* colidx;
   const double* vals;
   const double2* sides;    // (lhs, rhs)
   const uint8_t* colflags; // bit0 integral, bit1 binary
   const uint8_t* nzflags;  // colflags of each nonzero's column, in nonzero order (read coalesced)
   const uint8_t* rowflags; // bit0 all-binary row
   const uint8_t* rowbin;
   const int32_t* binrows;  // rows of bin b at [binstart[b], binstart[b+1]) (the tiled bins are
                            // not contiguous ranges: their rows are in tile order)
   const int64_t* colptr;   // CSC (device row ids) for row marking
   const int32_t* rowidx;
   int32_t binstart[SAP_NBINS + 1];
   // tiles: the rows of the tiled bins packed first-fit decreasing into runs of consecutive
   // rows of at most 64 lanes (tile width 1: a nonzero per lane; width 2: two, rows starting at
   // even positions); tile t holds nonzeros [tstart[t], tstart[t+1]) and rows trow0[t]..,
   // with bit i of tmask[t] set where lane i starts a row
   int32_t ntiles;
   int32_t pad;
   const int64_t* tstart;
   const unsigned long long* tmask;
   const int32_t* trow0;
   const uint8_t* tlev;     // levels of the shift tree: ceil(log2(longest row of the tile))
   const unsigned long long* tpad;   // wide tiles: two masks per tile; bit i of tpad[2t] + 2 * bit i
                                     // of tpad[2t+1] = empty positions of the row starting at lane i
   int32_t twidth;          // nonzeros per lane in a tile (1, 2 or 4)
   int32_t pad3;
   // independent instances packed block-diagonally: rows and columns carry their instance
   // (segment); verdicts, cutoffs and round counts are kept per (vector, segment) unit
   int32_t nseg;
   int32_t rowbins;         // bins 1..rowbins are swept one row per lane
   const int32_t* rowseg;   // NULL when nseg == 1
   const int32_t* colseg;
   const int32_t* perm;     // device row position -> caller's row index
   const int32_t* rlist;    // rows with two finite sides and at least 3 nonzeros (rowflags bit 1)
   int32_t nranged;
   int32_t pad2;
};

struct Vec
{
   int32_t nvec;
   int32_t single;          // 1: one vector, a cutoff stops the whole call
   int32_t hasforced;
   int32_t dense;           // 1: round-tagged marks scanned densely; 0: worklists
   double2* bnd;            // nvec*ncols (lb, ub)
   long long* clb;          // regular lb candidates (max)
   long long* cub;          // regular ub candidates (min)
   long long* flb;          // forced lb candidates (single-variable rows)
   long long* fub;
   int32_t* coltag;         // last round a column got a candidate
   int32_t* chgtag;         // last round a column's bounds changed (its rows are marked after)
   int32_t* rowmark;        // round a row is marked for
   uint8_t* redundant;      // nvec*nrows
   int32_t* vcutoff;        // per unit (vector, segment): infeasibility detected
   int32_t* vrounds;        // per unit: last round id in which one of its rows was swept
   int32_t* vstart;         // per vector: round id of its first round
   int32_t* vpend;          // per vector: latest round id for which one of its rows is marked
   int32_t* vlogn;          // per vector: columns with a changed bound (change log length)
   int32_t* log;            // nvec*ncols: changed columns of vector v at [v*ncols, v*ncols+vlogn[v])
   uint32_t* logbits;       // nvec*ncols bits: column already in the change log
   int32_t* ovl;            // nvec*ncols: overlay columns of vector v at [v*ncols, v*ncols+ovln[v])
   int32_t* ovln;           // per vector
   int32_t logchg;          // record changed columns
   int32_t nseg;            // segments of the matrix (units per vector)
   int2* wl[2][SAP_NBINS];  // sparse mode: (row, vector) worklists, double-buffered by round parity
   int2* touched;           // sparse mode: (column, vector)
   Status* st;
   // reason recording (one vector, one instance): every candidate that reaches or ties the
   // running best of its column goes to the round's candidate log; the column step notes
   // which key won each accepted change (rflag/rkey), k_reason finds the smallest row that
   // submitted it (rrow) and k_history appends the change with its row to the history
   int32_t reasons;
   uint32_t clogcap;
   struct CandRec* clog;
   uint8_t* rflag;          // per column: RF_* bits of this round's outcome
   long long* rkey;         // 2 per column: winning key of the lb / ub change
   int32_t* rrow;           // 2 per column: smallest caller row * 2 + rule that submitted it
   struct HistRec* hist;
   uint32_t histcap;
   uint32_t pad5;
   uint8_t* rgflag;         // ranged-row propagation: per (vector, nonzero) check-group flags
};

struct CandRec
{
   long long key;
   int32_t col;
   int32_t rowside;         // caller row * 8 + rule * 4 + side (0 lb, 1 ub) * 2 + forced
};

struct HistRec
{
   double val;
   int32_t round;
   int32_t colside;         // column * 2 + side (0 lb, 1 ub)
   int32_t row;
   int32_t pad;
};

#define RF_LB      1   // the lower bound changed (or, with RF_CUT, the lower side is a row's)
#define RF_LBFORCE 2   // ... by a forced candidate (single-variable row)
#define RF_UB      4
#define RF_UBFORCE 8
#define RF_CUT     16  // the column's bounds crossed

// block ranges of one sweep launch: kind 0 = tiles (dense mode), kind 1+b = bin b
struct Layout
{
   int32_t blockstart[SAP_NKINDS + 1];
};

#define KEY_NEG_INF ((long long)0x800FFFFFFFFFFFFFLL)   // dkey(-inf)
#define KEY_POS_INF ((long long)0x7FF0000000000000LL)   // dkey(+inf)

// ------------------------------------------------------------------------------------------
// row state
// ------------------------------------------------------------------------------------------
struct RowState
{
   DD minact;
   DD maxact;
   int32_t c[8];   // minposinf, minneginf, minposhuge, minneghuge, maxposinf, maxneginf, maxposhuge, maxneghuge
   double mad;     // maxactdelta
};

// classification of a contribution a*b (consdataUpdateActivities from bound 0,
// cons_linear.c:1773-1945): 0 finite, 1 +inf, 2 -inf, 3 +huge, 4 -huge
__device__ __forceinline__ int classify(const Params& p, double a, double b, double& c)
{
   if( isInf(p, fabs(b)) )
   {
      c = 0.0;
      return ((b > 0.0) == (a > 0.0)) ? 1 : 2;
   }
   double x = a * b;
   if( isHuge(p, fabs(x)) )
   {
      c = 0.0;
      return x > 0.0 ? 3 : 4;
   }
   c = x;
   return 0;
}

// per-element contribution to the row state
struct Elem
{
   double a;
   double lb;
   double ub;
   int32_t j;
   uint8_t cf;
};

__device__ __forceinline__ void elemContrib(const Params& p, const Elem& e, bool allbin, double& cmin, double& cmax,
   int& kmin, int& kmax, double& mad, bool& madinf)
{
   kmin = classify(p, e.a, e.a > 0.0 ? e.lb : e.ub, cmin);
   kmax = classify(p, e.a, e.a > 0.0 ? e.ub : e.lb, cmax);
   if( allbin )
   {
      mad = (e.lb < 0.5 && e.ub > 0.5) ? fabs(e.a) : 0.0;
      madinf = false;
   }
   else
   {
      madinf = isInf(p, -e.lb) || isInf(p, e.ub);
      mad = madinf ? 0.0 : fabs(e.a) * (e.ub - e.lb);
   }
}

// packed counters: 8 fields of 8 bits (rows of at most 64 nonzeros per group)
__device__ __forceinline__ unsigned long long packKinds(int kmin, int kmax)
{
   unsigned long long v = 0;
   if( kmin > 0 )
      v |= 1ULL << (8 * (kmin - 1));
   if( kmax > 0 )
      v |= 1ULL << (8 * (kmax + 3));
   return v;
}

__device__ __forceinline__ double shflXorD(double x, int m) { return __shfl_xor(x, m, SAP_WAVE); }
__device__ __forceinline__ DD shflXorDD(DD x, int m)
{
   DD r;
   r.hi = __shfl_xor(x.hi, m, SAP_WAVE);
   r.lo = __shfl_xor(x.lo, m, SAP_WAVE);
   return r;
}
__device__ __forceinline__ DD shflDD(DD x, int src)
{
   DD r;
   r.hi = __shfl(x.hi, src, SAP_WAVE);
   r.lo = __shfl(x.lo, src, SAP_WAVE);
   return r;
}

// ------------------------------------------------------------------------------------------
// activities (getMinActivity / getMaxActivity, cons_linear.c:2345-2529)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void getMinAct(const Params& p, DD act, int pinf, int ninf, int phuge, int nhuge,
   double delta, bool goodrelax, double& res, bool& tight, bool& setinf)
{
   if( ninf > 0 )
   {
      res = -p.infinity;
      setinf = true;
      tight = (pinf == 0);
   }
   else if( pinf > 0 )
   {
      res = p.infinity;
      setinf = true;
      tight = true;
   }
   else if( nhuge > 0 || (phuge > 0 && !goodrelax) )
   {
      res = -p.infinity;
      setinf = true;
      tight = false;
   }
   else
   {
      DD t = sum21(act, -delta);
      if( phuge > 0 )
      {
         t = sum21(t, phuge * p.hugeval);
         tight = false;
      }
      else
         tight = true;
      res = ddVal(t);
      setinf = false;
   }
}

__device__ __forceinline__ void getMaxAct(const Params& p, DD act, int pinf, int ninf, int phuge, int nhuge,
   double delta, bool goodrelax, double& res, bool& tight, bool& setinf)
{
   if( pinf > 0 )
   {
      res = p.infinity;
      setinf = true;
      tight = (ninf == 0);
   }
   else if( ninf > 0 )
   {
      res = -p.infinity;
      setinf = true;
      tight = true;
   }
   else if( phuge > 0 || (nhuge > 0 && !goodrelax) )
   {
      res = p.infinity;
      setinf = true;
      tight = false;
   }
   else
   {
      DD t = sum21(act, -delta);
      if( nhuge > 0 )
      {
         t = sum21(t, -nhuge * p.hugeval);
         tight = false;
      }
      else
         tight = true;
      res = ddVal(t);
      setinf = false;
   }
}

// consdataGetActivityResiduals (cons_linear.c:2661-2806), goodrelax = FALSE
__device__ __forceinline__ void residuals(const Params& p, const RowState& rs, double a, double lb, double ub,
   double& minres, double& maxres, bool& mintight, bool& maxtight, bool& minsetinf, bool& maxsetinf)
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
   int c0 = rs.c[0], c1 = rs.c[1], c2 = rs.c[2], c3 = rs.c[3];
   double delta = 0.0;
   if( isInf(p, minb) )
      c0 -= 1;
   else if( isInf(p, -minb) )
      c1 -= 1;
   else if( isHuge(p, minb * absval) )
      c2 -= 1;
   else if( isHuge(p, -minb * absval) )
      c3 -= 1;
   else
      delta = absval * minb;
   getMinAct(p, rs.minact, c0, c1, c2, c3, delta, false, minres, mintight, minsetinf);

   int d0 = rs.c[4], d1 = rs.c[5], d2 = rs.c[6], d3 = rs.c[7];
   delta = 0.0;
   if( isInf(p, -maxb) )
      d1 -= 1;
   else if( isInf(p, maxb) )
      d0 -= 1;
   else if( isHuge(p, absval * maxb) )
      d2 -= 1;
   else if( isHuge(p, -absval * maxb) )
      d3 -= 1;
   else
      delta = absval * maxb;
   getMaxAct(p, rs.maxact, d0, d1, d2, d3, delta, false, maxres, maxtight, maxsetinf);
}

// ------------------------------------------------------------------------------------------
// wavefront-aggregated appends: one atomic per wavefront (and bin) instead of one per lane;
// valid in divergent code, the currently active lanes cooperate
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t laneMaskLt()
{
   const int lane = (int)threadIdx.x & (SAP_WAVE - 1);
   return lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
}

__device__ __forceinline__ void waveAppendTouched(Status* st, int2* list, bool pred, int2 item)
{
   const uint64_t m = __ballot(pred);
   if( m == 0 )
      return;
   const int leader = __ffsll((unsigned long long)m) - 1;
   const int lane = (int)threadIdx.x & (SAP_WAVE - 1);
   uint32_t base = 0;
   if( lane == leader )
      base = atomicAdd(&st->touched, (uint32_t)__popcll(m));
   base = __shfl(base, leader, SAP_WAVE);
   if( pred )
      list[base + __popcll(m & laneMaskLt())] = item;
}

__device__ __forceinline__ void waveAppendRow(Status* st, int2* const* wl, int npar, bool pred, int bin, int2 item)
{
   uint64_t pend = __ballot(pred);
   const int lane = (int)threadIdx.x & (SAP_WAVE - 1);
   while( pend )
   {
      const int leader = __ffsll((unsigned long long)pend) - 1;
      const int lbin = __shfl(bin, leader, SAP_WAVE);
      const bool mine = pred && bin == lbin;
      const uint64_t grp = __ballot(mine);
      uint32_t base = 0;
      if( lane == leader )
         base = atomicAdd(&st->wlcount[npar][lbin], (uint32_t)__popcll(grp));
      base = __shfl(base, leader, SAP_WAVE);
      if( mine )
      {
         wl[lbin][base + __popcll(grp & laneMaskLt())] = item;
         pred = false;
      }
      pend &= ~grp;
   }
}

// one pass of (row, vector) worklist appends by a whole workgroup of SAP_BLOCK threads: the
// fresh rows are counted per length bin in LDS (scnt, sbase: SAP_NBINS entries each) and
// appended with one global atomic per bin.  Every thread of the workgroup calls it (barriers
// inside).  scnt[t] is cleared by thread t, which also was its only reader in the previous
// pass; sbase is rewritten only after the second barrier, which every reader of the previous
// pass has passed.
__device__ __forceinline__ void blockAppendRows(Status* st, int2* const* wl, int npar, bool fresh, int bin,
   int2 item, uint32_t* scnt, uint32_t* sbase)
{
   const int t = (int)threadIdx.x;
   if( t < SAP_NBINS )
      scnt[t] = 0;
   __syncthreads();
   uint32_t pos = 0;
   if( fresh )
      pos = atomicAdd(&scnt[bin], 1u);
   __syncthreads();
   if( t < SAP_NBINS && scnt[t] > 0 )
      sbase[t] = atomicAdd(&st->wlcount[npar][t], scnt[t]);
   __syncthreads();
   if( fresh )
      wl[bin][sbase[bin] + pos] = item;
}

// ------------------------------------------------------------------------------------------
// candidates
// ------------------------------------------------------------------------------------------
// R: reason recording compiled in (a separate instantiation of the sweep, so the plain one
// carries none of its code)
template <bool R>
struct CtxT
{
   int32_t rule = 0;   // reason rule of its candidates: 0 residual activity, 1 ranged row
   const Mat* M;
   const Vec* V;
   const Params* P;
   int32_t v;
   int32_t round;
   int32_t row;
};

// A row is locally redundant in the current call when its flag holds the call's generation.
// Generations run 1..255. A call that starts generation 1 first clears the flags of every
// vector to 0, and a slot admission clears its slot's flags, so no flag read in a call holds
// the generation of an earlier one. A zero-filled status block reads as generation 1.
__device__ __forceinline__ uint8_t redGen(const Vec& V)
{
   return (uint8_t)(V.st->rgenm1 + 1);
}

__device__ __forceinline__ bool rowRedundant(const Vec& V, size_t rflat)
{
   return V.redundant[rflat] == redGen(V);
}

__device__ __forceinline__ int32_t unitOfRow(const Mat& M, int32_t v, int32_t row)
{
   return M.rowseg == NULL ? v : v * M.nseg + M.rowseg[row];
}

__device__ __forceinline__ int32_t unitOfCol(const Mat& M, int32_t v, int32_t j)
{
   return M.colseg == NULL ? v : v * M.nseg + M.colseg[j];
}

template <bool R>
__device__ __forceinline__ void touch(const CtxT<R>& C, size_t flat, int32_t j)
{
   if( C.V->dense )
   {
      // round tag scanned by k_columns; plain store (not read first), every writer stores the
      // same value
      C.V->coltag[flat] = C.round;
   }
   else
   {
      const bool first = atomicMax(&C.V->coltag[flat], C.round) < C.round;
      waveAppendTouched(C.V->st, C.V->touched, first, make_int2(j, C.v));
   }
}

// reason recording: a candidate that reaches or ties the running best of its slot goes to the
// round's log.  The best only improves, so every candidate equal to the final best is logged.
template <bool R>
__device__ __forceinline__ void logCand(const CtxT<R>& C, int32_t j, long long k, int side, bool force)
{
   Status* st = C.V->st;
   const uint32_t at = atomicAdd(&st->clogn, 1u);
   if( at < C.V->clogcap )
   {
      CandRec r;
      r.key = k;
      r.col = j;
      r.rowside = C.M->perm[C.row] * 8 + C.rule * 4 + side * 2 + (force ? 1 : 0);
      C.V->clog[at] = r;
   }
   else if( st->clogovf == 0 )
      st->clogovf = 1;
}

template <bool R>
__device__ __forceinline__ void submitUb(const CtxT<R>& C, int32_t j, double val, bool force)
{
   size_t flat = (size_t)C.v * (size_t)C.M->ncols + (size_t)j;
   long long k = dkey(val);
   long long* slot = force ? &C.V->fub[flat] : &C.V->cub[flat];
   if constexpr( R )
   {
      if( k <= *slot && k <= atomicMin(slot, k) )
         logCand(C, j, k, 1, force);
   }
   else
      atomicMin(slot, k);   // result unused: the wavefront does not wait for it
   touch(C, flat, j);
}

template <bool R>
__device__ __forceinline__ void submitLb(const CtxT<R>& C, int32_t j, double val, bool force)
{
   size_t flat = (size_t)C.v * (size_t)C.M->ncols + (size_t)j;
   long long k = dkey(val);
   long long* slot = force ? &C.V->flb[flat] : &C.V->clb[flat];
   if constexpr( R )
   {
      if( k >= *slot && k >= atomicMax(slot, k) )
         logCand(C, j, k, 0, force);
   }
   else
      atomicMax(slot, k);
   touch(C, flat, j);
}

template <bool R>
__device__ __forceinline__ void setCutoff(const CtxT<R>& C)
{
   C.V->vcutoff[unitOfRow(*C.M, C.v, C.row)] = 1;
   if( C.V->single )
      C.V->st->cutoff = 1;
}

// tightenVarBoundsEasy (cons_linear.c:5380-5653) on start-of-round values
template <bool R>
__device__ __forceinline__ void candEasy(const CtxT<R>& C, const RowState& rs, double lhs, double rhs, const Elem& e, bool force)
{
   const Params& p = *C.P;
   const bool integral = (e.cf & 1) != 0;
   const double a = e.a, lb = e.lb, ub = e.ub;
   const double minact = ddVal(rs.minact), maxact = ddVal(rs.maxact);

   if( a > 0.0 )
   {
      if( !isInf(p, rhs) )
      {
         if( isFeasLT(p, rhs, minact) )
         {
            setCutoff(C);
            return;
         }
         double slack = rhs - minact;
         if( !isPositive(p, slack) )
            slack = 0.0;
         double alpha = a * (ub - lb);
         if( isSumGT(p, alpha, slack) || (force && isGT(p, alpha, slack)) )
         {
            double newub = adjustedUb(p, integral, lb + (slack / a));
            if( force || isUbBetter(p, newub, lb, ub) )
               submitUb(C, e.j, newub, force);
         }
      }
      if( !isInf(p, -lhs) )
      {
         if( isFeasLT(p, maxact, lhs) )
         {
            setCutoff(C);
            return;
         }
         double slack = maxact - lhs;
         if( !isPositive(p, slack) )
            slack = 0.0;
         double alpha = a * (ub - lb);
         if( isSumGT(p, alpha, slack) || (force && isGT(p, alpha, slack)) )
         {
            double newlb = adjustedLb(p, integral, ub - (slack / a));
            if( force || isLbBetter(p, newlb, lb, ub) )
               submitLb(C, e.j, newlb, force);
         }
      }
   }
   else
   {
      if( !isInf(p, rhs) )
      {
         if( isFeasLT(p, rhs, minact) )
         {
            setCutoff(C);
            return;
         }
         double slack = rhs - minact;
         if( !isPositive(p, slack) )
            slack = 0.0;
         double alpha = a * (lb - ub);
         if( isSumGT(p, alpha, slack) || (force && isGT(p, alpha, slack)) )
         {
            double newlb = adjustedLb(p, integral, ub + slack / a);
            if( force || isLbBetter(p, newlb, lb, ub) )
               submitLb(C, e.j, newlb, force);
         }
      }
      if( !isInf(p, -lhs) )
      {
         if( isFeasLT(p, maxact, lhs) )
         {
            setCutoff(C);
            return;
         }
         double slack = maxact - lhs;
         if( !isPositive(p, slack) )
            slack = 0.0;
         double alpha = a * (lb - ub);
         if( isSumGT(p, alpha, slack) || (force && isGT(p, alpha, slack)) )
         {
            double newub = adjustedUb(p, integral, lb - (slack / a));
            if( force || isUbBetter(p, newub, lb, ub) )
               submitUb(C, e.j, newub, force);
         }
      }
   }
}

// recomputes of this workgroup (consdataGetReliableResidualActivity's path), added to the
// status once at the end of the sweep: a counter in LDS keeps global atomics out of the sweep
__shared__ unsigned int s_nrecomp;

// consdataGetReliableResidualActivity (cons_linear.c:2580-2657): plain double sum in
// position order over the row, skipping position skip
template <bool R>
__device__ __forceinline__ double reliableResidual(const CtxT<R>& C, int64_t s, int64_t n, int64_t skip, bool isminres)
{
   const double2* bnd = C.V->bnd + (size_t)C.v * (size_t)C.M->ncols;
   atomicAdd(&s_nrecomp, 1u);
   double r = 0.0;
   for( int64_t q = 0; q < n; ++q )
   {
      if( q == skip )
         continue;
      double a = C.M->vals[s + q];
      double2 b = bnd[C.M->colidx[s + q]];
      if( a > 0.0 )
         r += a * (isminres ? b.x : b.y);
      else
         r += a * (isminres ? b.y : b.x);
   }
   return r;
}

// tightenVarBounds (cons_linear.c:6700-6974) on start-of-round values
template <bool R>
__device__ __forceinline__ void candHard(const CtxT<R>& C, const RowState& rs, double lhs, double rhs, const Elem& e, bool force,
   int64_t s, int64_t n, int64_t pos)
{
   const Params& p = *C.P;
   const bool integral = (e.cf & 1) != 0;
   const double a = e.a, lb = e.lb, ub = e.ub;
   const double lastmin = ddVal(rs.minact), lastmax = ddVal(rs.maxact);
   double minres, maxres;
   bool mintight, maxtight, minsetinf, maxsetinf;
   residuals(p, rs, a, lb, ub, minres, maxres, mintight, maxtight, minsetinf, maxsetinf);

   if( a > 0.0 )
   {
      if( !minsetinf && !isInf(p, rhs) && mintight && (isLT(p, fabs(rhs), 1.0) || !isEQ(p, minres / rhs, 1.0)) )
      {
         double newub = (rhs - minres) / a;
         if( !isInf(p, newub)
            && ((force && isLT(p, newub, ub)) || (integral && isFeasLT(p, newub, ub)) || isUbBetter(p, newub, lb, ub)) )
         {
            bool unrel = isUpdateUnreliable(p, minres, lastmin);
            if( unrel )
            {
               minres = reliableResidual(C, s, n, pos, true);
               newub = (rhs - minres) / a;
               unrel = isInf(p, -minres)
                  || (!isUbBetter(p, newub, lb, ub) && (!isFeasLT(p, newub, ub) || !integral)
                     && (!force || !isLT(p, newub, ub)));
            }
            if( !unrel )
               submitUb(C, e.j, newub, force);
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
               maxres = reliableResidual(C, s, n, pos, false);
               newlb = (lhs - maxres) / a;
               if( isInf(p, maxres) || (!isLbBetter(p, newlb, lb, ub) && (!isFeasGT(p, newlb, lb) || !integral)
                     && (!force || !isGT(p, newlb, lb))) )
                  return;
            }
            submitLb(C, e.j, newlb, force);
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
            bool unrel = isUpdateUnreliable(p, minres, lastmin);
            if( unrel )
            {
               minres = reliableResidual(C, s, n, pos, true);
               newlb = (rhs - minres) / a;
               unrel = isInf(p, -minres)
                  || (!isLbBetter(p, newlb, lb, ub) && (!isFeasGT(p, newlb, lb) || !integral)
                     && (!force || !isGT(p, newlb, lb)));
            }
            if( !unrel )
               submitLb(C, e.j, newlb, force);
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
               maxres = reliableResidual(C, s, n, pos, false);
               newub = (lhs - maxres) / a;
               // the reference joins the last two tests with && here (cons_linear.c:6942-6944)
               if( isInf(p, maxres) || (!isUbBetter(p, newub, lb, ub) && (!isFeasLT(p, newub, ub) && !integral)
                     && (!force || !isLT(p, newub, ub))) )
                  return;
            }
            submitUb(C, e.j, newub, force);
         }
      }
   }
}

// row-level decisions of tightenBounds (cons_linear.c:7014-7086)
struct RowDecision
{
   bool tighten;
   bool easy;
   bool force;
   bool cut;         // infeasible (propagateCons / tightenVarBoundsEasy tests)
   int cutcode;      // ... found by, in the row's own order of tests: 0 / 1 the tightening loop's
                     // minimal-activity / maximal-activity test, 3 / 4 the final test's (2 is
                     // the ranged-row test, which runs between them)
   bool redundant;   // locally redundant
};

// minimal / maximal activity with delta 0 for goodrelax FALSE and TRUE (getMinActivity /
// getMaxActivity above); the double-double part is shared, the cases as in those functions
__device__ __forceinline__ void minActBoth(const Params& p, DD act, int pinf, int ninf, int phuge, int nhuge,
   double& resF, bool& infF, double& resT)
{
   const DD t = sum21(act, -0.0);
   const double plain = ddVal(t);
   const double relaxed = phuge > 0 ? ddVal(sum21(t, phuge * p.hugeval)) : plain;
   if( ninf > 0 || (pinf == 0 && nhuge > 0) )
   {
      resF = resT = -p.infinity;
      infF = true;
   }
   else if( pinf > 0 )
   {
      resF = resT = p.infinity;
      infF = true;
   }
   else
   {
      resT = relaxed;
      infF = phuge > 0;
      resF = infF ? -p.infinity : plain;
   }
}

__device__ __forceinline__ void maxActBoth(const Params& p, DD act, int pinf, int ninf, int phuge, int nhuge,
   double& resF, bool& infF, double& resT)
{
   const DD t = sum21(act, -0.0);
   const double plain = ddVal(t);
   const double relaxed = nhuge > 0 ? ddVal(sum21(t, -nhuge * p.hugeval)) : plain;
   if( pinf > 0 || (ninf == 0 && phuge > 0) )
   {
      resF = resT = p.infinity;
      infF = true;
   }
   else if( ninf > 0 )
   {
      resF = resT = -p.infinity;
      infF = true;
   }
   else
   {
      resT = relaxed;
      infF = nhuge > 0;
      resF = infF ? p.infinity : plain;
   }
}

// row-level decisions of tightenBounds (cons_linear.c:7014-7086) together with the row's
// verdict: the cutoff tests of tightenVarBoundsEasy (cons_linear.c:5450, :5504) and the
// infeasibility / redundancy test at the end of propagateCons (cons_linear.c:7717-7753)
__device__ __forceinline__ RowDecision decideRow(const Params& p, const RowState& rs, double lhs, double rhs, int64_t n)
{
   RowDecision d;
   int infmin = rs.c[0] + rs.c[1] + rs.c[2] + rs.c[3];
   int infmax = rs.c[4] + rs.c[5] + rs.c[6] + rs.c[7];
   double minF, maxF, minT, maxT;
   bool f1, f2;
   minActBoth(p, rs.minact, rs.c[0], rs.c[1], rs.c[2], rs.c[3], minF, f1, minT);
   maxActBoth(p, rs.maxact, rs.c[4], rs.c[5], rs.c[6], rs.c[7], maxF, f2, maxT);
   d.force = (n == 1);
   d.tighten = n > 0 && !(p.flags & SAP_FLAG_NOTIGHTEN) && !(infmin > 1 && infmax > 1);
   if( d.tighten && isFeasZero(p, rs.mad) )
      d.tighten = false;
   if( d.tighten && !isInf(p, rs.mad) )
   {
      double slack = (isInf(p, rhs) || f1) ? p.infinity : (rhs - minF);
      double surplus = (isInf(p, -lhs) || f2) ? p.infinity : (maxF - lhs);
      if( isLE(p, rs.mad, dmin(slack, surplus)) )
         d.tighten = false;
   }
   d.easy = isLT(p, rs.mad, p.maxeasy);
   const bool emin = d.tighten && d.easy && !isInf(p, rhs) && isFeasLT(p, rhs, ddVal(rs.minact));
   const bool emax = d.tighten && d.easy && !isInf(p, -lhs) && isFeasLT(p, ddVal(rs.maxact), lhs);
   const bool fmin = isFeasGT(p, minT, rhs);
   d.cut = emin || emax || fmin || isFeasLT(p, maxT, lhs);
   // the tightening loop tests rhs before lhs at every nonzero; the final test checks the
   // minimal activity first
   d.cutcode = (emin || emax) ? (emin ? 0 : 1) : (fmin ? 3 : 4);
   d.redundant = !d.cut && isGE(p, minT, lhs) && isLE(p, maxT, rhs);
   return d;
}

// the row's verdict (computed by decideRow) applied by one lane of the row
template <bool R>
__device__ __forceinline__ void rowOutcome(const CtxT<R>& C, const RowDecision& d)
{
   if constexpr( R )
      if( d.cut )
         atomicMin(&C.V->st->cutrow, C.M->perm[C.row] * 8 + d.cutcode);
   if( d.cut )
      setCutoff(C);
   else if( d.redundant )
      C.V->redundant[(size_t)C.v * (size_t)C.M->nrows + (size_t)C.row] = redGen(*C.V);
}


__device__ __forceinline__ Elem loadElem(const Mat& M, const double2* bnd, int64_t k)
{
   Elem e;
   e.a = M.vals[k];
   e.j = M.colidx[k];
   double2 b = bnd[e.j];
   e.lb = b.x;
   e.ub = b.y;
   e.cf = M.nzflags[k];
   return e;
}

// ------------------------------------------------------------------------------------------
// screening and the per-wavefront candidate queue
//
// Most nonzeros of a tightening row yield no candidate.  A cheap exact screen decides per
// nonzero whether tightenVarBoundsEasy / tightenVarBounds could submit one; only those go to
// a per-wavefront queue in LDS, which is drained 64 at a time through candEasy / candHard, so
// the division-heavy exact code runs on full wavefronts of real work.
//   easy rows: the alpha > slack tests of tightenVarBoundsEasy (cons_linear.c:5467, :5521, ...)
//   hard rows: every acceptance test of a candidate ub = x/a requires ub_new < ub; with
//     correctly rounded division, fl(x/a) >= ub whenever x/a >= ub exactly, which a single
//     rounding fma(a, ub, -x) decides by its sign when it is strictly nonzero; symmetric for
//     lower bounds.  Equality is left to the exact path.
// ------------------------------------------------------------------------------------------
#define SAP_QCAP 128
// the queue is a ring indexed modulo SAP_QCAP: a drain leaves fewer than a wavefront's worth
// of entries and an append adds at most one per lane, so it never holds more than
// 2 * SAP_WAVE - 1 entries
static_assert((SAP_QCAP & (SAP_QCAP - 1)) == 0 && SAP_QCAP >= 2 * SAP_WAVE, "queue capacity");

struct WaveQueue
{
   int64_t* k;              // nonzero index
   int32_t* row;
   int32_t* v;              // vector index (< 2^30); bit 30 easy, bit 31 force
   double* mnhi;
   double* mnlo;
   double* mxhi;
   double* mxlo;
   unsigned long long* cnt; // eight counters, 8 bits each, saturated at 255
   double* lb;              // the nonzero's column bounds (start of round) and column
   double* ub;
   int32_t* j;
   int head;
   int count;
};

__device__ __forceinline__ unsigned long long packCounts(const RowState& rs)
{
   unsigned long long v = 0;
#pragma unroll
   for( int q = 0; q < 8; ++q )
   {
      const unsigned long long c = (unsigned long long)(rs.c[q] > 255 ? 255 : rs.c[q]);
      v |= c << (8 * q);
   }
   return v;
}

// true if the nonzero can produce a candidate (necessary condition, exact)
__device__ __forceinline__ bool screenElem(const Params& p, const RowState& rs, const RowDecision& d, double lhs,
   double rhs, const Elem& e)
{
   const double a = e.a, lb = e.lb, ub = e.ub;
   if( d.easy )
   {
      const double minact = ddVal(rs.minact), maxact = ddVal(rs.maxact);
      const double alpha = a > 0.0 ? a * (ub - lb) : a * (lb - ub);
      bool pass = false;
      if( !isInf(p, rhs) )
      {
         double slack = rhs - minact;
         if( !isPositive(p, slack) )
            slack = 0.0;
         pass |= isSumGT(p, alpha, slack) || (d.force && isGT(p, alpha, slack));
      }
      if( !isInf(p, -lhs) )
      {
         double slack = maxact - lhs;
         if( !isPositive(p, slack) )
            slack = 0.0;
         pass |= isSumGT(p, alpha, slack) || (d.force && isGT(p, alpha, slack));
      }
      return pass;
   }
   double minres, maxres;
   bool mintight, maxtight, minsetinf, maxsetinf;
   residuals(p, rs, a, lb, ub, minres, maxres, mintight, maxtight, minsetinf, maxsetinf);
   bool pass = false;
   if( !minsetinf && !isInf(p, rhs) && mintight )
   {
      const double x = rhs - minres;
      // a > 0: ub candidate x/a must be < ub;  a < 0: lb candidate x/a must be > lb
      pass |= a > 0.0 ? !(fma(a, ub, -x) < 0.0) : !(fma(a, lb, -x) < 0.0);
   }
   if( !maxsetinf && !isInf(p, -lhs) && maxtight )
   {
      const double y = lhs - maxres;
      // a > 0: lb candidate y/a must be > lb;  a < 0: ub candidate y/a must be < ub
      pass |= a > 0.0 ? !(fma(a, lb, -y) > 0.0) : !(fma(a, ub, -y) > 0.0);
   }
   return pass;
}

// exact candidates of queue entries [head, head+n) for lanes < n
template <bool R>
__device__ __forceinline__ void drainQueue(const Mat& M, const Vec& V, const Params& P, WaveQueue& Q, int n, int32_t round)
{
   const int lane = (int)threadIdx.x & (SAP_WAVE - 1);
   __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
   if( lane < n )
   {
      const int q = (Q.head + lane) & (SAP_QCAP - 1);
      const int64_t k = Q.k[q];
      const int32_t row = Q.row[q];
      const uint32_t vf = (uint32_t)Q.v[q];
      const int32_t v = (int32_t)(vf & 0x3FFFFFFFu);
      const uint32_t fl = vf >> 30;
      RowState rs;
      rs.minact.hi = Q.mnhi[q];
      rs.minact.lo = Q.mnlo[q];
      rs.maxact.hi = Q.mxhi[q];
      rs.maxact.lo = Q.mxlo[q];
      const unsigned long long c = Q.cnt[q];
#pragma unroll
      for( int i = 0; i < 8; ++i )
         rs.c[i] = (int)((c >> (8 * i)) & 0xFF);
      rs.mad = 0.0;
      CtxT<R> C;
      C.M = &M;
      C.V = &V;
      C.P = &P;
      C.v = v;
      C.round = round;
      C.row = row;
      Elem e;
      e.a = M.vals[k];
      e.cf = M.nzflags[k];
      e.j = Q.j[q];
      e.lb = Q.lb[q];
      e.ub = Q.ub[q];
      const double2 sides = M.sides[row];
      if( fl & 1 )
         candEasy(C, rs, sides.x, sides.y, e, (fl & 2) != 0);
      else
      {
         const int64_t s0 = M.rowptr[row];
         const int64_t n0 = M.rowptr[row + 1] - s0;
         candHard(C, rs, sides.x, sides.y, e, (fl & 2) != 0, s0, n0, k - s0);
      }
   }
   Q.head = (Q.head + n) & (SAP_QCAP - 1);
   Q.count -= n;
}

// append the lanes with pred; drain a full wavefront of work when one is available
__device__ __forceinline__ void pushQueue(const Mat& M, const Vec& V, const Params& P, WaveQueue& Q, bool pred,
   int64_t k, const Elem& e, int32_t row, int32_t v, const RowState& rs, const RowDecision& d, int32_t round)
{
   const uint64_t m = __ballot(pred);
   if( m == 0 )
      return;
   if( pred )
   {
      const int q = (Q.head + Q.count + __popcll(m & laneMaskLt())) & (SAP_QCAP - 1);
      Q.k[q] = k;
      Q.row[q] = row;
      Q.v[q] = (int32_t)((uint32_t)v | (d.easy ? 1u << 30 : 0u) | (d.force ? 1u << 31 : 0u));
      Q.lb[q] = e.lb;
      Q.ub[q] = e.ub;
      Q.j[q] = e.j;
      Q.mnhi[q] = rs.minact.hi;
      Q.mnlo[q] = rs.minact.lo;
      Q.mxhi[q] = rs.maxact.hi;
      Q.mxlo[q] = rs.maxact.lo;
      Q.cnt[q] = packCounts(rs);
   }
   Q.count += __popcll(m);
}

// drain one full wavefront of queued work if available (end of a loop iteration, where few
// values are live; the queue then holds at most 63 entries before the next push)
template <bool R>
__device__ __forceinline__ void maybeDrain(const Mat& M, const Vec& V, const Params& P, WaveQueue& Q, int32_t round)
{
   if( Q.count >= SAP_WAVE )
      drainQueue<R>(M, V, P, Q, SAP_WAVE, round);
}


// ------------------------------------------------------------------------------------------
// work items of one bin: dense mode enumerates (vector, row) over the bin's row range and
// keeps the rows marked for this round; sparse mode reads the round's worklist
// ------------------------------------------------------------------------------------------
struct Items
{
   const int2* list;
   uint32_t n;
   int32_t binbase;
   int32_t binlen;
   int32_t round;
   int32_t rlocal;
   int32_t firstall;
};

// the unit of (v, row) swept a row in this round
__device__ __forceinline__ void noteRound(const Mat& M, const Vec& V, int32_t v, int32_t row, int32_t round)
{
   const int32_t u = unitOfRow(M, v, row);
   if( V.vrounds[u] < round )
      atomicMax(&V.vrounds[u], round);
}

template <bool R, bool CUTCHECK = true>
__device__ __forceinline__ bool fetchItem(const Mat& M, const Vec& V, const Items& I, uint32_t it, int32_t& row,
   int32_t& v)
{
   if( V.dense )
   {
      v = (int32_t)(it / (uint32_t)I.binlen);
      row = I.binbase + (int32_t)(it - (uint32_t)v * (uint32_t)I.binlen);
      const size_t rflat = (size_t)v * (size_t)M.nrows + (size_t)row;
      if( !I.firstall && (V.rowmark[rflat] != I.round || rowRedundant(V, rflat)) )
         return false;
   }
   else
   {
      const int2 item = I.list[it];
      row = item.x;
      v = item.y;
   }
   // with reason recording every marked row of the round is swept, so that the smallest
   // infeasible row does not depend on timing
   return R || !CUTCHECK || V.vcutoff[unitOfRow(M, v, row)] == 0;
}

// ------------------------------------------------------------------------------------------
// sweep: groups of G lanes, one row per group, one nonzero per lane (rows of <= G nonzeros)
// ------------------------------------------------------------------------------------------
template <int G, bool R>
__device__ __forceinline__ void sweepGroups(const Mat& M, const Vec& V, const Params& P, const Items& I, int lblock,
   int nblocks, unsigned long long& nnzacc, unsigned long long& rowsacc, WaveQueue& Q)
{
   const int tid = lblock * SAP_BLOCK + (int)threadIdx.x;
   const int glane = (int)threadIdx.x & (G - 1);
   const uint32_t ngroups = (uint32_t)nblocks * (SAP_BLOCK / G);
   const int32_t rlocal = I.rlocal;
   // the loop runs a wavefront-uniform number of times so that queue pushes stay convergent
   const uint32_t base = (uint32_t)((tid & ~(SAP_WAVE - 1)) / G);
   for( uint32_t it0 = base; it0 < I.n; it0 += ngroups )
   {
      const uint32_t it = it0 + (uint32_t)((tid & (SAP_WAVE - 1)) / G);
      int32_t row = 0, v = 0;
      const bool live = it < I.n && fetchItem<R>(M, V, I, it, row, v);
      CtxT<R> C;
      C.M = &M;
      C.V = &V;
      C.P = &P;
      C.v = v;
      C.round = I.round;
      C.row = row;
      const double2* bnd = V.bnd + (size_t)v * (size_t)M.ncols;
      const int64_t s = live ? M.rowptr[row] : 0;
      const int64_t n = live ? M.rowptr[row + 1] - s : 0;
      const double2 sides = live ? M.sides[row] : make_double2(0.0, 0.0);
      const bool allbin = live && (M.rowflags[row] & 1) != 0;
      const bool has = live && glane < n;
      Elem e;
      e.a = 0.0;
      e.lb = 0.0;
      e.ub = 0.0;
      e.j = 0;
      e.cf = 0;
      if( has )
         e = loadElem(M, bnd, s + glane);

      double cmin = 0.0, cmax = 0.0, mad = 0.0;
      int kmin = 0, kmax = 0;
      bool madinf = false;
      if( has )
         elemContrib(P, e, allbin, cmin, cmax, kmin, kmax, mad, madinf);
      unsigned long long cnt = packKinds(kmin, kmax);
      DD dmn = { cmin, 0.0 };
      DD dmx = { cmax, 0.0 };
      int inf = madinf ? 1 : 0;
#pragma unroll
      for( int off = 1; off < G; off <<= 1 )
      {
         dmn = pairAdd(dmn, shflXorDD(dmn, off));
         dmx = pairAdd(dmx, shflXorDD(dmx, off));
         cnt += (unsigned long long)__shfl_xor((long long)cnt, off, SAP_WAVE);
         mad = dmax(mad, shflXorD(mad, off));
         inf |= __shfl_xor(inf, off, SAP_WAVE);
      }
      RowState rs;
      rs.minact = dmn;
      rs.maxact = dmx;
#pragma unroll
      for( int q = 0; q < 8; ++q )
         rs.c[q] = (int)((cnt >> (8 * q)) & 0xFF);
      rs.mad = inf ? P.infinity : mad;

      RowDecision d = decideRow(P, rs, sides.x, sides.y, n);
      const bool pass = has && d.tighten && screenElem(P, rs, d, sides.x, sides.y, e);
      pushQueue(M, V, P, Q, pass, s + glane, e, row, v, rs, d, I.round);
      if( live && glane == 0 )
      {
         rowOutcome(C, d);
         nnzacc += (unsigned long long)n;
         rowsacc += 1;
         if( !V.single )
            noteRound(M, V, v, row, I.round);
      }
      maybeDrain<R>(M, V, P, Q, I.round);
   }
}

// ------------------------------------------------------------------------------------------
// sweep: one row per lane (rows of <= MAXL nonzeros).  The canonical tree of the row is formed
// in registers by a binary counter over positions 0..MAXL-1: position k closes the subtrees
// of the trailing one bits of k; positions past the row contribute the identity (0, 0).  The
// row decision runs once per row; the screening pass reloads the row's nonzeros.
// ------------------------------------------------------------------------------------------
template <int MAXL, bool R>
__device__ __forceinline__ void sweepRows(const Mat& M, const Vec& V, const Params& P, const Items& I, int lblock,
   int nblocks, unsigned long long& nnzacc, unsigned long long& rowsacc, WaveQueue& Q)
{
   constexpr int LV = (MAXL <= 4) ? 2 : ((MAXL <= 8) ? 3 : 4);
   const int tid = lblock * SAP_BLOCK + (int)threadIdx.x;
   const uint32_t nthreads = (uint32_t)nblocks * SAP_BLOCK;
   const int lane = (int)threadIdx.x & (SAP_WAVE - 1);
   // wavefront-uniform trip count so that queue pushes stay convergent
   for( uint32_t it0 = (uint32_t)(tid - lane); it0 < I.n; it0 += nthreads )
   {
      const uint32_t it = it0 + (uint32_t)lane;
      int32_t row = 0, v = 0;
      const bool live = it < I.n && fetchItem<R>(M, V, I, it, row, v);
      const double2* bnd = V.bnd + (size_t)v * (size_t)M.ncols;
      const int64_t s = live ? M.rowptr[row] : 0;
      const int n = live ? (int)(M.rowptr[row + 1] - s) : 0;
      const double2 sides = live ? M.sides[row] : make_double2(0.0, 0.0);
      const bool allbin = live && (M.rowflags[row] & 1) != 0;
      DD stmn[LV], stmx[LV];
      DD rmn = { 0.0, 0.0 }, rmx = { 0.0, 0.0 };
      unsigned long long cnt = 0;
      double mad = 0.0;
      bool madinf = false;
#pragma unroll
      for( int k = 0; k < MAXL; ++k )
      {
         double cmin = 0.0, cmax = 0.0;
         if( k < n )
         {
            const Elem e = loadElem(M, bnd, s + k);
            int kmin, kmax;
            double m1;
            bool mi;
            elemContrib(P, e, allbin, cmin, cmax, kmin, kmax, m1, mi);
            cnt += packKinds(kmin, kmax);
            mad = dmax(mad, m1);
            madinf = madinf || mi;
         }
         DD nmn = { cmin, 0.0 };
         DD nmx = { cmax, 0.0 };
         int lvl = 0;
#pragma unroll
         for( int l = 0; l < LV; ++l )
         {
            if( lvl == l && ((k >> l) & 1) != 0 )
            {
               nmn = pairAdd(stmn[l], nmn);
               nmx = pairAdd(stmx[l], nmx);
               lvl = l + 1;
            }
         }
         if( k == MAXL - 1 )
         {
            rmn = nmn;
            rmx = nmx;
         }
         else
         {
#pragma unroll
            for( int l = 0; l < LV; ++l )
               if( l == lvl )
               {
                  stmn[l] = nmn;
                  stmx[l] = nmx;
               }
         }
      }
      RowState rs;
      rs.minact = rmn;
      rs.maxact = rmx;
#pragma unroll
      for( int q = 0; q < 8; ++q )
         rs.c[q] = (int)((cnt >> (8 * q)) & 0xFF);
      rs.mad = madinf ? P.infinity : mad;
      const RowDecision d = decideRow(P, rs, sides.x, sides.y, n);
      if( live )
      {
         CtxT<R> C;
         C.M = &M;
         C.V = &V;
         C.P = &P;
         C.v = v;
         C.round = I.round;
         C.row = row;
         rowOutcome(C, d);
         nnzacc += (unsigned long long)n;
         rowsacc += 1;
         if( !V.single )
            noteRound(M, V, v, row, I.round);
      }
      const bool tight = live && d.tighten;
#pragma unroll 1
      for( int k = 0; k < MAXL; ++k )
      {
         const bool act = tight && k < n;
         if( __ballot(act) == 0 )
            break;
         Elem e;
         e.a = 0.0;
         e.lb = 0.0;
         e.ub = 0.0;
         e.j = 0;
         e.cf = 0;
         if( act )
            e = loadElem(M, bnd, s + k);
         const bool pass = act && screenElem(P, rs, d, sides.x, sides.y, e);
         pushQueue(M, V, P, Q, pass, s + k, e, row, v, rs, d, I.round);
         maybeDrain<R>(M, V, P, Q, I.round);
      }
   }
}

// ------------------------------------------------------------------------------------------
// sweep of 64-nonzero tiles (dense mode, rows of 1..64 nonzeros): a wavefront loads one tile,
// lane i holds nonzero tstart+i at position p of its row; the canonical tree of every row is
// formed by shifts: at width w a lane at a position aligned to 2w adds the partial held by the
// lane w above it when that position is still inside the row
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ DD shflDownDD(DD x, int w)
{
   DD r;
   r.hi = __shfl_down(x.hi, w, SAP_WAVE);
   r.lo = __shfl_down(x.lo, w, SAP_WAVE);
   return r;
}

template <bool R>
__device__ __forceinline__ void sweepTiles(const Mat& M, const Vec& V, const Params& P, const Items& I, int lblock,
   int nblocks, unsigned long long& nnzacc, unsigned long long& rowsacc, WaveQueue& Q)
{
   const int lane = (int)threadIdx.x & (SAP_WAVE - 1);
   const int wid = lblock * (SAP_BLOCK / SAP_WAVE) + (int)threadIdx.x / SAP_WAVE;
   const uint32_t nw = (uint32_t)nblocks * (SAP_BLOCK / SAP_WAVE);
   const uint32_t ntiles = (uint32_t)M.ntiles;
   const uint32_t ntot = (uint32_t)V.nvec * ntiles;
   const uint64_t lemask = (lane == 63) ? ~0ULL : ((2ULL << lane) - 1);
   for( uint32_t t = (uint32_t)wid; t < ntot; t += nw )
   {
      const uint32_t tu = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
      const int32_t v = (int32_t)(tu / ntiles);
      const int32_t tile = (int32_t)(tu - (uint32_t)v * ntiles);
      if( M.rowseg == NULL && V.vcutoff[v] && !R )
         continue;
      const int64_t e0 = M.tstart[tile];
      const int cnt = (int)(M.tstart[tile + 1] - e0);
      const uint64_t mask = M.tmask[tile];
      const int32_t row0 = M.trow0[tile];
      const int lev = M.tlev[tile];
      const bool has = lane < cnt;
      // my row: start lane s, end lane ns (exclusive), position p
      const uint64_t le = mask & lemask;
      const int k = __popcll(le) - 1;
      const int sl = 63 - __clzll((long long)le);
      const uint64_t above = (sl >= 63) ? 0ULL : (mask >> (sl + 1));
      const int ns = above ? (sl + 1 + __ffsll((unsigned long long)above) - 1) : cnt;
      const int L = ns - sl;
      const int p = lane - sl;
      const int32_t row = row0 + k;
      const bool marked = has
         && (I.firstall
            || (V.rowmark[(size_t)v * (size_t)M.nrows + (size_t)row] == I.round
               && !rowRedundant(V, (size_t)v * (size_t)M.nrows + (size_t)row)))
         && (M.rowseg == NULL || V.vcutoff[unitOfRow(M, v, row)] == 0);
      if( __ballot(marked) == 0 )
         continue;

      const double2* bnd = V.bnd + (size_t)v * (size_t)M.ncols;
      Elem e;
      e.a = 0.0;
      e.lb = 0.0;
      e.ub = 0.0;
      e.j = 0;
      e.cf = 0;
      bool allbin = false;
      double2 sides = make_double2(0.0, 0.0);
      if( marked )
      {
         e = loadElem(M, bnd, e0 + lane);
         allbin = (M.rowflags[row] & 1) != 0;
         sides = M.sides[row];
      }
      double cmin = 0.0, cmax = 0.0, mad = 0.0;
      int kmin = 0, kmax = 0;
      bool madinf = false;
      if( marked )
         elemContrib(P, e, allbin, cmin, cmax, kmin, kmax, mad, madinf);
      DD dmn = { cmin, 0.0 };
      DD dmx = { cmax, 0.0 };
      for( int w = 1; w < (1 << lev); w <<= 1 )
      {
         const DD omn = shflDownDD(dmn, w);
         const DD omx = shflDownDD(dmx, w);
         const double omad = __shfl_down(mad, w, SAP_WAVE);
         if( (p & (2 * w - 1)) == 0 && p + w < L )
         {
            dmn = pairAdd(dmn, omn);
            dmx = pairAdd(dmx, omx);
            mad = dmax(mad, omad);
         }
      }
      RowState rs;
      rs.minact = shflDD(dmn, sl);
      rs.maxact = shflDD(dmx, sl);
      rs.mad = __shfl(mad, sl, SAP_WAVE);
      const uint64_t rowbits = ((L >= 64) ? ~0ULL : ((1ULL << L) - 1)) << sl;
      rs.c[0] = __popcll(__ballot(kmin == 1) & rowbits);
      rs.c[1] = __popcll(__ballot(kmin == 2) & rowbits);
      rs.c[2] = __popcll(__ballot(kmin == 3) & rowbits);
      rs.c[3] = __popcll(__ballot(kmin == 4) & rowbits);
      rs.c[4] = __popcll(__ballot(kmax == 1) & rowbits);
      rs.c[5] = __popcll(__ballot(kmax == 2) & rowbits);
      rs.c[6] = __popcll(__ballot(kmax == 3) & rowbits);
      rs.c[7] = __popcll(__ballot(kmax == 4) & rowbits);
      if( (__ballot(madinf) & rowbits) != 0 )
         rs.mad = P.infinity;
      CtxT<R> C;
      C.M = &M;
      C.V = &V;
      C.P = &P;
      C.v = v;
      C.round = I.round;
      C.row = row;
      const RowDecision d = decideRow(P, rs, sides.x, sides.y, L);
      const bool pass = marked && d.tighten && screenElem(P, rs, d, sides.x, sides.y, e);
      pushQueue(M, V, P, Q, pass, e0 + lane, e, row, v, rs, d, I.round);
      if( marked && p == 0 )
      {
         rowOutcome(C, d);
         nnzacc += (unsigned long long)L;
         rowsacc += 1;
         if( !V.single )
            noteRound(M, V, v, row, I.round);
      }
      maybeDrain<R>(M, V, P, Q, I.round);
   }
}

// ------------------------------------------------------------------------------------------
// sweep of wide tiles (dense mode, tile width W = 2 or 4): lane i holds positions W*i ..
// W*i+W-1 of its tile, every row starts at a position divisible by W (a row whose length is
// not leaves the rest of its last lane empty), so all positions of a lane belong to one row.
// The lowest log2(W) levels of the canonical tree are added inside the lane (positions 2p and
// 2p+1, then 4p and 4p+2, each when the right one is still inside the row); the levels above
// are the shift tree of the width-1 tiles over lanes: lane q of a row adds lane q+w when q is
// aligned to 2w and q+w is still inside the row.  Per 64 lanes the row decision, the tree
// levels and the tile bookkeeping thus serve up to 64*W nonzeros.
// ------------------------------------------------------------------------------------------
template <bool R, int W>
__device__ __forceinline__ void sweepTilesW(const Mat& M, const Vec& V, const Params& P, const Items& I, int lblock,
   int nblocks, unsigned long long& nnzacc, unsigned long long& rowsacc, WaveQueue& Q)
{
   static_assert(W == 2 || W == 4, "tile width");
   constexpr int LW = W == 2 ? 1 : 2;
   const int lane = (int)threadIdx.x & (SAP_WAVE - 1);
   const int wid = lblock * (SAP_BLOCK / SAP_WAVE) + (int)threadIdx.x / SAP_WAVE;
   const uint32_t nw = (uint32_t)nblocks * (SAP_BLOCK / SAP_WAVE);
   const uint32_t ntiles = (uint32_t)M.ntiles;
   const uint32_t ntot = (uint32_t)V.nvec * ntiles;
   const uint64_t lemask = (lane == 63) ? ~0ULL : ((2ULL << lane) - 1);
   for( uint32_t t = (uint32_t)wid; t < ntot; t += nw )
   {
      const uint32_t tu = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
      const int32_t v = (int32_t)(tu / ntiles);
      const int32_t tile = (int32_t)(tu - (uint32_t)v * ntiles);
      if( M.rowseg == NULL && V.vcutoff[v] && !R )
         continue;
      const int64_t e0 = M.tstart[tile];
      const uint64_t mask = M.tmask[tile];
      // empty positions at the end of each row, by the row's first lane: pad0 + 2 * pad1
      const uint64_t pad0 = M.tpad[2 * (size_t)tile];
      const uint64_t pad1 = M.tpad[2 * (size_t)tile + 1];
      const int lanes = ((int)(M.tstart[tile + 1] - e0) + __popcll(pad0) + 2 * __popcll(pad1)) >> LW;
      const int32_t row0 = M.trow0[tile];
      const int lev = M.tlev[tile];
      const bool has = lane < lanes;
      // my row: start lane sl, end lane ns (exclusive), lane position q, length L
      const uint64_t le = mask & lemask;
      const int k = __popcll(le) - 1;
      const int sl = 63 - __clzll((long long)le);
      const uint64_t above = (sl >= 63) ? 0ULL : (mask >> (sl + 1));
      const int ns = above ? (sl + 1 + __ffsll((unsigned long long)above) - 1) : lanes;
      const int Ll = ns - sl;
      const int L = W * Ll - (int)((pad0 >> sl) & 1) - 2 * (int)((pad1 >> sl) & 1);
      const int q = lane - sl;
      const uint64_t below = (1ULL << sl) - 1;
      const int64_t k0 = e0 + W * lane - __popcll(pad0 & below) - 2 * __popcll(pad1 & below);
      const int nin = L - W * q;   // positions of this lane inside the row (>= W: all)
      const int32_t row = row0 + k;
      const bool marked = has
         && (I.firstall
            || (V.rowmark[(size_t)v * (size_t)M.nrows + (size_t)row] == I.round
               && !rowRedundant(V, (size_t)v * (size_t)M.nrows + (size_t)row)))
         && (M.rowseg == NULL || V.vcutoff[unitOfRow(M, v, row)] == 0);
      if( __ballot(marked) == 0 )
         continue;

      const double2* bnd = V.bnd + (size_t)v * (size_t)M.ncols;
      Elem e[W];
#pragma unroll
      for( int i = 0; i < W; ++i )
      {
         e[i].a = 0.0;
         e[i].lb = 0.0;
         e[i].ub = 0.0;
         e[i].j = 0;
         e[i].cf = 0;
      }
      bool allbin = false;
      double2 sides = make_double2(0.0, 0.0);
      if( marked )
      {
#pragma unroll
         for( int i = 0; i < W; ++i )
            if( i < nin )
               e[i] = loadElem(M, bnd, k0 + i);
         allbin = (M.rowflags[row] & 1) != 0;
         sides = M.sides[row];
      }
      DD dmn[W], dmx[W];
      int kmin[W], kmax[W];
      double mad = 0.0;
      bool madinf = false;
#pragma unroll
      for( int i = 0; i < W; ++i )
      {
         double cmin = 0.0, cmax = 0.0, md = 0.0;
         bool mi = false;
         kmin[i] = 0;
         kmax[i] = 0;
         if( marked && i < nin )
         {
            elemContrib(P, e[i], allbin, cmin, cmax, kmin[i], kmax[i], md, mi);
            mad = i == 0 ? md : dmax(mad, md);
            madinf = madinf || mi;
         }
         dmn[i] = { cmin, 0.0 };
         dmx[i] = { cmax, 0.0 };
      }
      // levels inside the lane
#pragma unroll
      for( int h = 1; h < W; h <<= 1 )
#pragma unroll
         for( int i = 0; i < W; i += 2 * h )
            if( i + h < nin )
            {
               dmn[i] = pairAdd(dmn[i], dmn[i + h]);
               dmx[i] = pairAdd(dmx[i], dmx[i + h]);
            }
      DD smn = dmn[0], smx = dmx[0];
      for( int w = 1; w < ((1 << lev) >> LW); w <<= 1 )
      {
         const DD omn = shflDownDD(smn, w);
         const DD omx = shflDownDD(smx, w);
         const double omad = __shfl_down(mad, w, SAP_WAVE);
         if( (q & (2 * w - 1)) == 0 && q + w < Ll )
         {
            smn = pairAdd(smn, omn);
            smx = pairAdd(smx, omx);
            mad = dmax(mad, omad);
         }
      }
      RowState rs;
      rs.minact = shflDD(smn, sl);
      rs.maxact = shflDD(smx, sl);
      rs.mad = __shfl(mad, sl, SAP_WAVE);
      const uint64_t rowbits = ((Ll >= 64) ? ~0ULL : ((1ULL << Ll) - 1)) << sl;
#pragma unroll
      for( int c = 0; c < 4; ++c )
      {
         int a = 0, b = 0;
#pragma unroll
         for( int i = 0; i < W; ++i )
         {
            a += __popcll(__ballot(kmin[i] == c + 1) & rowbits);
            b += __popcll(__ballot(kmax[i] == c + 1) & rowbits);
         }
         rs.c[c] = a;
         rs.c[4 + c] = b;
      }
      if( (__ballot(madinf) & rowbits) != 0 )
         rs.mad = P.infinity;
      const RowDecision d = decideRow(P, rs, sides.x, sides.y, L);
      const bool scr = marked && d.tighten;
      uint32_t pass = 0;
#pragma unroll
      for( int i = 0; i < W; ++i )
         if( scr && i < nin && screenElem(P, rs, d, sides.x, sides.y, e[i]) )
            pass |= 1u << i;
      if( marked && q == 0 )
      {
         CtxT<R> C;
         C.M = &M;
         C.V = &V;
         C.P = &P;
         C.v = v;
         C.round = I.round;
         C.row = row;
         rowOutcome(C, d);
         nnzacc += (unsigned long long)L;
         rowsacc += 1;
         if( !V.single )
            noteRound(M, V, v, row, I.round);
      }
      // one push (and one drain site) per position of the lane
#pragma unroll 1
      for( int i = 0; i < W; ++i )
      {
         Elem x = e[0];
#pragma unroll
         for( int u = 1; u < W; ++u )
            if( i == u )
               x = e[u];
         pushQueue(M, V, P, Q, ((pass >> i) & 1) != 0, k0 + i, x, row, v, rs, d, I.round);
         maybeDrain<R>(M, V, P, Q, I.round);
      }
   }
}

// ------------------------------------------------------------------------------------------
// sweep of one long row by one wavefront: the chunks [c0, c0+nch) of 64 positions form an
// aligned subtree of the canonical tree; the wave returns its root (all lanes) and the
// per-lane counters / maxactdelta contributions.
// ------------------------------------------------------------------------------------------
struct WavePart
{
   DD minact;
   DD maxact;
   int32_t c[8];
   double mad;
   int inf;
};

__device__ __forceinline__ WavePart waveSubtree(const Mat& M, const Params& P, const double2* bnd, int64_t s, int64_t n, int64_t c0,
   int64_t nch, bool allbin)
{
   const int lane = (int)threadIdx.x & (SAP_WAVE - 1);
   WavePart w;
   DD stmin = { 0.0, 0.0 }, stmax = { 0.0, 0.0 };   // level `lane` of the binary-counter stack
   int32_t k0 = 0, k1 = 0, k2 = 0, k3 = 0, k4 = 0, k5 = 0, k6 = 0, k7 = 0;
   double mad = 0.0;
   int inf = 0;
   int64_t cnt = 0;
   for( int64_t c = c0; c < c0 + nch; ++c )
   {
      const int64_t pos = c * SAP_WAVE + lane;
      double cmin = 0.0, cmax = 0.0;
      if( pos < n )
      {
         Elem e = loadElem(M, bnd, s + pos);
         int kmin, kmax;
         double md;
         bool mi;
         elemContrib(P, e, allbin, cmin, cmax, kmin, kmax, md, mi);
         k0 += (kmin == 1);
         k1 += (kmin == 2);
         k2 += (kmin == 3);
         k3 += (kmin == 4);
         k4 += (kmax == 1);
         k5 += (kmax == 2);
         k6 += (kmax == 3);
         k7 += (kmax == 4);
         mad = dmax(mad, md);
         inf |= mi ? 1 : 0;
      }
      DD rmin = { cmin, 0.0 }, rmax = { cmax, 0.0 };
#pragma unroll
      for( int off = 1; off < SAP_WAVE; off <<= 1 )
      {
         rmin = pairAdd(rmin, shflXorDD(rmin, off));
         rmax = pairAdd(rmax, shflXorDD(rmax, off));
      }
      // push the chunk root on the binary-counter stack (level l lives in lane l)
      int64_t cc = cnt;
      int lvl = 0;
      while( cc & 1 )
      {
         rmin = pairAdd(shflDD(stmin, lvl), rmin);
         rmax = pairAdd(shflDD(stmax, lvl), rmax);
         cc >>= 1;
         ++lvl;
      }
      if( lane == lvl )
      {
         stmin = rmin;
         stmax = rmax;
      }
      ++cnt;
   }
   // fold the stack from the lowest set level upwards
   bool first = true;
   DD amin = { 0.0, 0.0 }, amax = { 0.0, 0.0 };
   for( int lvl = 0; lvl < SAP_MAXLEVELS && (cnt >> lvl) != 0; ++lvl )
   {
      if( (cnt >> lvl) & 1 )
      {
         DD emin = shflDD(stmin, lvl), emax = shflDD(stmax, lvl);
         if( first )
         {
            amin = emin;
            amax = emax;
            first = false;
         }
         else
         {
            amin = pairAdd(emin, amin);
            amax = pairAdd(emax, amax);
         }
      }
   }
   w.minact = amin;
   w.maxact = amax;
   // counters and maxactdelta reduced over the wave (pairs of 32-bit counters per 64-bit lane word)
   long long p01 = ((long long)k1 << 32) | (unsigned)k0, p23 = ((long long)k3 << 32) | (unsigned)k2;
   long long p45 = ((long long)k5 << 32) | (unsigned)k4, p67 = ((long long)k7 << 32) | (unsigned)k6;
#pragma unroll
   for( int off = 1; off < SAP_WAVE; off <<= 1 )
   {
      p01 += __shfl_xor(p01, off, SAP_WAVE);
      p23 += __shfl_xor(p23, off, SAP_WAVE);
      p45 += __shfl_xor(p45, off, SAP_WAVE);
      p67 += __shfl_xor(p67, off, SAP_WAVE);
   }
   w.c[0] = (int32_t)(p01 & 0xFFFFFFFF);
   w.c[1] = (int32_t)(p01 >> 32);
   w.c[2] = (int32_t)(p23 & 0xFFFFFFFF);
   w.c[3] = (int32_t)(p23 >> 32);
   w.c[4] = (int32_t)(p45 & 0xFFFFFFFF);
   w.c[5] = (int32_t)(p45 >> 32);
   w.c[6] = (int32_t)(p67 & 0xFFFFFFFF);
   w.c[7] = (int32_t)(p67 >> 32);
#pragma unroll
   for( int off = 1; off < SAP_WAVE; off <<= 1 )
   {
      mad = dmax(mad, shflXorD(mad, off));
      inf |= __shfl_xor(inf, off, SAP_WAVE);
   }
   w.mad = mad;
   w.inf = inf;
   return w;
}

template <bool R>
__device__ __forceinline__ void candidatesRange(const Mat& M, const Vec& V, const Params& P, WaveQueue& Q,
   const RowState& rs, const RowDecision& d, double lhs, double rhs, const double2* bnd, int64_t s, int64_t n,
   int64_t c0, int64_t nch, int32_t row, int32_t v, int32_t round)
{
   const int lane = (int)threadIdx.x & (SAP_WAVE - 1);
   for( int64_t c = c0; c < c0 + nch; ++c )
   {
      const int64_t pos = c * SAP_WAVE + lane;
      const bool has = pos < n;
      Elem e;
      e.a = 1.0;
      e.lb = 0.0;
      e.ub = 0.0;
      e.j = 0;
      e.cf = 0;
      if( has )
         e = loadElem(M, bnd, s + pos);
      const bool pass = has && screenElem(P, rs, d, lhs, rhs, e);
      pushQueue(M, V, P, Q, pass, s + pos, e, row, v, rs, d, round);
      maybeDrain<R>(M, V, P, Q, round);
   }
}

template <bool R>
__device__ __forceinline__ void sweepWaves(const Mat& M, const Vec& V, const Params& P, const Items& I, int lblock,
   int nblocks, unsigned long long& nnzacc, unsigned long long& rowsacc, WaveQueue& Q)
{
   const int lane = (int)threadIdx.x & (SAP_WAVE - 1);
   const int wid = lblock * (SAP_BLOCK / SAP_WAVE) + (int)threadIdx.x / SAP_WAVE;
   const uint32_t nw = (uint32_t)nblocks * (SAP_BLOCK / SAP_WAVE);
   const int32_t rlocal = I.rlocal;
   for( uint32_t it = (uint32_t)wid; it < I.n; it += nw )
   {
      int32_t row, v;
      if( !fetchItem<R>(M, V, I, it, row, v) )
         continue;
      CtxT<R> C;
      C.M = &M;
      C.V = &V;
      C.P = &P;
      C.v = v;
      C.round = I.round;
      C.row = row;
      const double2* bnd = V.bnd + (size_t)v * (size_t)M.ncols;
      const int64_t s = M.rowptr[row];
      const int64_t n = M.rowptr[row + 1] - s;
      const int64_t nch = (n + SAP_WAVE - 1) / SAP_WAVE;
      const double2 sides = M.sides[row];
      const bool allbin = (M.rowflags[row] & 1) != 0;
      WavePart w = waveSubtree(M, P, bnd, s, n, 0, nch, allbin);
      RowState rs;
      rs.minact = w.minact;
      rs.maxact = w.maxact;
#pragma unroll
      for( int q = 0; q < 8; ++q )
         rs.c[q] = w.c[q];
      rs.mad = w.inf ? P.infinity : w.mad;
      RowDecision d = decideRow(P, rs, sides.x, sides.y, n);
      if( d.tighten )
         candidatesRange<R>(M, V, P, Q, rs, d, sides.x, sides.y, bnd, s, n, 0, nch, row, v, I.round);
      if( lane == 0 )
      {
         rowOutcome(C, d);
         nnzacc += (unsigned long long)n;
         rowsacc += 1;
         if( !V.single )
            noteRound(M, V, v, row, I.round);
      }
   }
}

// workgroup per row: wave w reduces the aligned chunk subtree [w*span, (w+1)*span)
template <bool R>
__device__ __forceinline__ void sweepBlocks(const Mat& M, const Vec& V, const Params& P, const Items& I, int lblock,
   int nblocks, unsigned long long& nnzacc, unsigned long long& rowsacc, WaveQueue& Q)
{
   constexpr int NW = SAP_BLOCK / SAP_WAVE;
   __shared__ WavePart parts[NW];
   __shared__ RowState srs;
   const int lane = (int)threadIdx.x & (SAP_WAVE - 1);
   const int wave = (int)threadIdx.x / SAP_WAVE;
   const int32_t rlocal = I.rlocal;
   for( uint32_t it = (uint32_t)lblock; it < I.n; it += (uint32_t)nblocks )
   {
      // every wavefront must take the same branch around the barriers below: the item is
      // fetched without the test of its unit's cutoff flag, which another workgroup may set
      // while the wavefronts read it (a row of a unit cut off in this round is then swept
      // anyway; its unit's results are void)
      int32_t row, v;
      if( !fetchItem<R, false>(M, V, I, it, row, v) )
         continue;   // workgroup-uniform
      CtxT<R> C;
      C.M = &M;
      C.V = &V;
      C.P = &P;
      C.v = v;
      C.round = I.round;
      C.row = row;
      const double2* bnd = V.bnd + (size_t)v * (size_t)M.ncols;
      const int64_t s = M.rowptr[row];
      const int64_t n = M.rowptr[row + 1] - s;
      const int64_t nch = (n + SAP_WAVE - 1) / SAP_WAVE;
      int64_t P2 = 1;
      while( P2 < nch )
         P2 <<= 1;
      const int64_t span = P2 >= NW ? P2 / NW : 1;
      const int64_t c0 = (int64_t)wave * span;
      const int64_t mine = c0 < nch ? (c0 + span <= nch ? span : nch - c0) : 0;
      const double2 sides = M.sides[row];
      const bool allbin = (M.rowflags[row] & 1) != 0;
      WavePart w = waveSubtree(M, P, bnd, s, n, c0, mine, allbin);
      if( lane == 0 )
         parts[wave] = w;
      __syncthreads();
      if( threadIdx.x == 0 )
      {
         // pairwise combination of the four subtree roots (wave order = position order)
         static_assert(NW == 4, "workgroup combine is written for four wavefronts");
         RowState rs;
         rs.minact = pairAdd(pairAdd(parts[0].minact, parts[1].minact), pairAdd(parts[2].minact, parts[3].minact));
         rs.maxact = pairAdd(pairAdd(parts[0].maxact, parts[1].maxact), pairAdd(parts[2].maxact, parts[3].maxact));
#pragma unroll
         for( int k = 0; k < 8; ++k )
            rs.c[k] = parts[0].c[k] + parts[1].c[k] + parts[2].c[k] + parts[3].c[k];
         const double mad = dmax(dmax(parts[0].mad, parts[1].mad), dmax(parts[2].mad, parts[3].mad));
         const int inf = parts[0].inf | parts[1].inf | parts[2].inf | parts[3].inf;
         rs.mad = inf ? P.infinity : mad;
         srs = rs;
      }
      __syncthreads();
      RowState rs = srs;
      RowDecision d = decideRow(P, rs, sides.x, sides.y, n);
      if( d.tighten )
         candidatesRange<R>(M, V, P, Q, rs, d, sides.x, sides.y, bnd, s, n, c0, mine, row, v, I.round);
      if( threadIdx.x == 0 )
      {
         rowOutcome(C, d);
         nnzacc += (unsigned long long)n;
         rowsacc += 1;
         if( !V.single )
            noteRound(M, V, v, row, I.round);
      }
      __syncthreads();
   }
}

// workgroup sum of two per-lane counters, added into the slot of this workgroup
__device__ __forceinline__ void blockStats(unsigned long long a, unsigned long long b, unsigned long long* slota,
   unsigned long long* slotb)
{
   __shared__ unsigned long long red[2][SAP_BLOCK / SAP_WAVE];
#pragma unroll
   for( int off = 1; off < SAP_WAVE; off <<= 1 )
   {
      a += (unsigned long long)__shfl_xor((long long)a, off, SAP_WAVE);
      b += (unsigned long long)__shfl_xor((long long)b, off, SAP_WAVE);
   }
   const int wave = (int)threadIdx.x / SAP_WAVE;
   if( ((int)threadIdx.x & (SAP_WAVE - 1)) == 0 )
   {
      red[0][wave] = a;
      red[1][wave] = b;
   }
   __syncthreads();
   if( threadIdx.x == 0 )
   {
      unsigned long long sa = 0, sb = 0;
      for( int w = 0; w < SAP_BLOCK / SAP_WAVE; ++w )
      {
         sa += red[0][w];
         sb += red[1][w];
      }
      const int slot = (int)(blockIdx.x & (SAP_STAT_SLOTS - 1));
      if( sa != 0 )
         atomicAdd(&slota[slot], sa);
      if( sb != 0 && slotb != NULL )
         atomicAdd(&slotb[slot], sb);
   }
}

template <bool R>
__global__ __launch_bounds__(SAP_BLOCK, 4) void k_sweep(Mat M, Vec V, Params P, Layout L)
{
   Status* st = V.st;
   if( st->done )
      return;
   const unsigned long long tstart = (unsigned long long)wall_clock64();
   if( threadIdx.x == 0 )
      s_nrecomp = 0;
   __syncthreads();
   const int blk = (int)blockIdx.x;
   int kind = 0;
   while( kind < SAP_NKINDS - 1 && blk >= L.blockstart[kind + 1] )
      ++kind;
   const int lblock = blk - L.blockstart[kind];
   const int nblocks = L.blockstart[kind + 1] - L.blockstart[kind];
   const int b = kind - 1;   // row-length bin (kind 0: tiles)
   Items I;
   I.round = st->round;
   I.rlocal = st->rlocal;
   I.firstall = (V.dense && st->firstall && I.rlocal == 1) ? 1 : 0;
   I.binbase = b >= 0 ? M.binstart[b] : 0;
   I.binlen = b >= 0 ? M.binstart[b + 1] - M.binstart[b] : 0;
   if( V.dense )
   {
      I.list = NULL;
      I.n = (uint32_t)V.nvec * (uint32_t)I.binlen;
   }
   else
   {
      I.list = b >= 0 ? V.wl[I.round & 1][b] : NULL;
      I.n = b >= 0 ? st->wlcount[I.round & 1][b] : 0;
   }
   unsigned long long nnzacc = 0, rowsacc = 0;
   constexpr int NW = SAP_BLOCK / SAP_WAVE;
   __shared__ int64_t qk[NW][SAP_QCAP];
   __shared__ int32_t qrow[NW][SAP_QCAP];
   __shared__ int32_t qv[NW][SAP_QCAP];
   __shared__ double qmnhi[NW][SAP_QCAP];
   __shared__ double qmnlo[NW][SAP_QCAP];
   __shared__ double qmxhi[NW][SAP_QCAP];
   __shared__ double qmxlo[NW][SAP_QCAP];
   __shared__ unsigned long long qcnt[NW][SAP_QCAP];
   __shared__ double qlb[NW][SAP_QCAP];
   __shared__ double qub[NW][SAP_QCAP];
   __shared__ int32_t qj[NW][SAP_QCAP];
   const int wave = (int)threadIdx.x / SAP_WAVE;
   WaveQueue Q;
   Q.k = qk[wave];
   Q.row = qrow[wave];
   Q.v = qv[wave];
   Q.mnhi = qmnhi[wave];
   Q.mnlo = qmnlo[wave];
   Q.mxhi = qmxhi[wave];
   Q.mxlo = qmxlo[wave];
   Q.cnt = qcnt[wave];
   Q.lb = qlb[wave];
   Q.ub = qub[wave];
   Q.j = qj[wave];
   Q.head = 0;
   Q.count = 0;
   switch( kind )
   {
   case 0:
      if( M.twidth == 2 )
         sweepTilesW<R, 2>(M, V, P, I, lblock, nblocks, nnzacc, rowsacc, Q);
      else if( M.twidth == 4 )
         sweepTilesW<R, 4>(M, V, P, I, lblock, nblocks, nnzacc, rowsacc, Q);
      else
         sweepTiles<R>(M, V, P, I, lblock, nblocks, nnzacc, rowsacc, Q);
      // empty rows (first round only; they are never marked again): activity 0 against the
      // sides, propagateCons's infeasibility test (cons_linear.c:7721-7742)
      if( I.firstall && lblock == 0 )
      {
         const int32_t nempty = M.binstart[1] - M.binstart[0];
         for( int32_t r = M.binstart[0] + (int32_t)threadIdx.x; r < M.binstart[1]; r += SAP_BLOCK )
         {
            const double2 sd = M.sides[r];
            if( (isFeasGT(P, 0.0, sd.y) || isFeasLT(P, 0.0, sd.x)) && R )
               atomicMin(&st->cutrow, M.perm[r] * 8 + (isFeasGT(P, 0.0, sd.y) ? 3 : 4));
            if( isFeasGT(P, 0.0, sd.y) || isFeasLT(P, 0.0, sd.x) )
               for( int32_t v = 0; v < V.nvec; ++v )
               {
                  V.vcutoff[unitOfRow(M, v, r)] = 1;
                  if( V.single )
                     st->cutoff = 1;
               }
         }
         if( threadIdx.x == 0 )
            rowsacc += (unsigned long long)V.nvec * (unsigned long long)nempty;
      }
      break;
   case 1: sweepGroups<4, R>(M, V, P, I, lblock, nblocks, nnzacc, rowsacc, Q); break;
   case 2:
      if( M.rowbins >= 1 )
         sweepRows<4, R>(M, V, P, I, lblock, nblocks, nnzacc, rowsacc, Q);
      else
         sweepGroups<4, R>(M, V, P, I, lblock, nblocks, nnzacc, rowsacc, Q);
      break;
   case 3:
      if( M.rowbins >= 2 )
         sweepRows<8, R>(M, V, P, I, lblock, nblocks, nnzacc, rowsacc, Q);
      else
         sweepGroups<8, R>(M, V, P, I, lblock, nblocks, nnzacc, rowsacc, Q);
      break;
   case 4:
      if( M.rowbins >= 3 )
         sweepRows<16, R>(M, V, P, I, lblock, nblocks, nnzacc, rowsacc, Q);
      else
         sweepGroups<16, R>(M, V, P, I, lblock, nblocks, nnzacc, rowsacc, Q);
      break;
   case 5: sweepGroups<32, R>(M, V, P, I, lblock, nblocks, nnzacc, rowsacc, Q); break;
   case 6: sweepGroups<64, R>(M, V, P, I, lblock, nblocks, nnzacc, rowsacc, Q); break;
   case 7:
   case 8: sweepWaves<R>(M, V, P, I, lblock, nblocks, nnzacc, rowsacc, Q); break;
   default: sweepBlocks<R>(M, V, P, I, lblock, nblocks, nnzacc, rowsacc, Q); break;
   }
   // remaining queued candidates of this wavefront
   if( Q.count > 0 )
      drainQueue<R>(M, V, P, Q, Q.count, I.round);
   blockStats(nnzacc, rowsacc, st->slot_nnz, st->slot_rows);   // ends with a barrier
   if( threadIdx.x == 0 )
   {
      if( s_nrecomp != 0 )
         atomicAdd(&st->nrecomp, (unsigned long long)s_nrecomp);
      const int slot = (int)(blockIdx.x & (SAP_STAT_SLOTS - 1));
      atomicMin(&st->tslot0[slot], tstart);
      atomicMax(&st->tslot1[slot], (unsigned long long)wall_clock64());
   }
}

// the column step for one (column, vector) whose candidate slots may hold values;
// returns the number of changed bounds
__device__ __forceinline__ int columnStep(const Mat& M, const Vec& V, const Params& P, int32_t j, int32_t v,
   int32_t next)
{
   Status* st = V.st;
   const size_t flat = (size_t)v * (size_t)M.ncols + (size_t)j;
   const long long kcl = V.clb[flat], kcu = V.cub[flat];
   V.clb[flat] = KEY_NEG_INF;
   V.cub[flat] = KEY_POS_INF;
   long long kfl = KEY_NEG_INF, kfu = KEY_POS_INF;
   if( V.hasforced )
   {
      kfl = V.flb[flat];
      kfu = V.fub[flat];
      V.flb[flat] = KEY_NEG_INF;
      V.fub[flat] = KEY_POS_INF;
   }
   const int32_t unit = unitOfCol(M, v, j);
   // with reason recording every column of a round without a row cutoff is stepped, so that
   // the smallest crossing column does not depend on timing
   if( V.reasons ? st->cutrow != INT32_MAX : V.vcutoff[unit] != 0 )
      return 0;
   const bool integral = (M.colflags[j] & 1) != 0;
   const double2 b0 = V.bnd[flat];
   const double lb0 = b0.x, ub0 = b0.y;
   double nlb = lb0, nub = ub0;
   bool cut = false;
   // which candidate set each new bound (RF_* bits) and its key; on a cut, the crossing one
   int lsrc = 0, usrc = 0;
   long long lkey = 0, ukey = 0;
   // SCIPinferVarLbCons (scip_var.c:6965-7070): forced, then regular candidate
   if( kfl != KEY_NEG_INF )
   {
      double tt = adjustedLb(P, integral, keyToDouble(kfl));
      if( isInf(P, tt) || isFeasGT(P, tt, ub0) )
      {
         cut = true;
         lsrc = RF_LB | RF_LBFORCE;
         lkey = kfl;
      }
      else
      {
         tt = dmin(tt, ub0);
         if( !isLE(P, tt, lb0) && dkey(tt) > dkey(nlb) )
         {
            nlb = tt;
            lsrc = RF_LB | RF_LBFORCE;
            lkey = kfl;
         }
      }
   }
   if( !cut && kcl != KEY_NEG_INF )
   {
      double tt = adjustedLb(P, integral, keyToDouble(kcl));
      if( isInf(P, tt) || isFeasGT(P, tt, ub0) )
      {
         cut = true;
         lsrc = RF_LB;
         lkey = kcl;
      }
      else
      {
         tt = dmin(tt, ub0);
         if( isLbBetter(P, tt, lb0, ub0) && dkey(tt) > dkey(nlb) )
         {
            nlb = tt;
            lsrc = RF_LB;
            lkey = kcl;
         }
      }
   }
   // SCIPinferVarUbCons (scip_var.c:7071-7174) against the new lower bound
   if( !cut && kfu != KEY_POS_INF )
   {
      double tt = adjustedUb(P, integral, keyToDouble(kfu));
      if( isInf(P, -tt) || isFeasLT(P, tt, nlb) )
      {
         cut = true;
         usrc = RF_UB | RF_UBFORCE;
         ukey = kfu;
      }
      else
      {
         tt = dmax(tt, nlb);
         if( !isGE(P, tt, ub0) && dkey(tt) < dkey(nub) )
         {
            nub = tt;
            usrc = RF_UB | RF_UBFORCE;
            ukey = kfu;
         }
      }
   }
   if( !cut && kcu != KEY_POS_INF )
   {
      double tt = adjustedUb(P, integral, keyToDouble(kcu));
      if( isInf(P, -tt) || isFeasLT(P, tt, nlb) )
      {
         cut = true;
         usrc = RF_UB;
         ukey = kcu;
      }
      else
      {
         tt = dmax(tt, nlb);
         if( isUbBetter(P, tt, nlb, ub0) && dkey(tt) < dkey(nub) )
         {
            nub = tt;
            usrc = RF_UB;
            ukey = kcu;
         }
      }
   }
   if( cut )
   {
      if( V.reasons )
      {
         // a lower-bound candidate crossed the upper bound of the round start, or an
         // upper-bound candidate crossed the (possibly new) lower bound: the sides without
         // a candidate are the column's own bounds
         V.rflag[j] = (uint8_t)(RF_CUT | lsrc | usrc);
         V.rkey[2 * (size_t)j] = lkey;
         V.rkey[2 * (size_t)j + 1] = ukey;
         V.rrow[2 * (size_t)j] = INT32_MAX;
         V.rrow[2 * (size_t)j + 1] = INT32_MAX;
         atomicMin(&st->cutcol, j);
      }
      V.vcutoff[unit] = 1;
      if( V.single )
         st->cutoff = 1;
      return 0;
   }
   const bool chl = dkey(nlb) != dkey(lb0);
   const bool chu = dkey(nub) != dkey(ub0);
   if( !chl && !chu )
      return 0;
   if( V.reasons )
   {
      V.rflag[j] = (uint8_t)((chl ? lsrc : 0) | (chu ? usrc : 0));
      V.rkey[2 * (size_t)j] = lkey;
      V.rkey[2 * (size_t)j + 1] = ukey;
      V.rrow[2 * (size_t)j] = INT32_MAX;
      V.rrow[2 * (size_t)j + 1] = INT32_MAX;
   }
   V.bnd[flat] = make_double2(nlb, nub);
   if( V.logchg )
   {
      // first change of this column in the vector's life: append it to the vector's log
      const uint32_t bit = 1u << (flat & 31);
      if( (atomicOr(&V.logbits[flat >> 5], bit) & bit) == 0 )
      {
         const int32_t at = atomicAdd(&V.vlogn[v], 1);
         V.log[(size_t)v * (size_t)M.ncols + (size_t)at] = j;
      }
   }
   // a vector stops after its own maxrounds rounds: rows are not marked beyond
   if( !(P.maxrounds >= 0 && next - V.vstart[v] + 1 > P.maxrounds) )
      V.chgtag[flat] = next - 1;   // k_mark marks its rows unless the unit is cut off
   return (chl ? 1 : 0) + (chu ? 1 : 0);
}

// fold the sweep's per-slot wall-clock stamps into the launch time (all SAP_STAT_SLOTS threads
// of one workgroup)
__device__ __forceinline__ void foldStamps(Status* st)
{
   __shared__ unsigned long long tmin[SAP_STAT_SLOTS], tmax[SAP_STAT_SLOTS];
   const int t = (int)threadIdx.x;
   tmin[t] = st->tslot0[t];
   tmax[t] = st->tslot1[t];
   st->tslot0[t] = ~0ULL;
   st->tslot1[t] = 0;
   __syncthreads();
   for( int w = SAP_STAT_SLOTS / 2; w > 0; w >>= 1 )
   {
      if( t < w )
      {
         tmin[t] = tmin[t] < tmin[t + w] ? tmin[t] : tmin[t + w];
         tmax[t] = tmax[t] > tmax[t + w] ? tmax[t] : tmax[t + w];
      }
      __syncthreads();
   }
   if( t == 0 && tmin[0] != ~0ULL && tmax[0] >= tmin[0] )
   {
      st->sweep_ticks += tmax[0] - tmin[0];
      st->sweep_count += 1;
   }
}

// the stop test of the round after the current one: nothing marked (queued), cutoff of the
// single vector, or the call's round cap.  k_round_begin applies it at the start of a round;
// the host applies it to the status it reads back after a graph group, so that a call whose
// last round ended the group launches no further group.
__host__ __device__ __forceinline__ bool stopsNext(const Status& st, bool dense, int32_t maxrounds)
{
   const int32_t r = st.round + 1;
   bool more;
   if( dense )
      more = (st.rlocal == 0 && st.firstall) || st.anymark[r & 1];
   else
   {
      uint32_t in = 0;
      for( int b = 0; b < SAP_NBINS; ++b )
         in += st.wlcount[r & 1][b];
      more = in != 0;
   }
   return !more || st.cutoff || (!st.cont && maxrounds >= 0 && st.rlocal >= maxrounds);
}

// mark the rows of a changed column for the next round (eventExecLinear marks them)
__device__ __forceinline__ void markColumn(const Mat& M, const Vec& V, int32_t j, int32_t v, int32_t next, int npar,
   bool& anymarked)
{
   Status* st = V.st;
   if( V.dense )
   {
      // Round tags are stored without being read first, also on locally redundant rows, which
      // the dense sweep skips.  Whether the vector has a non-redundant row marked for the
      // round (its pending round, the round's "any row marked") reads the redundant flags
      // only until such a row is known.
      bool known = V.vpend[v] == next;
      for( int64_t q = M.colptr[j]; q < M.colptr[j + 1]; ++q )
      {
         const size_t rflat = (size_t)v * (size_t)M.nrows + (size_t)M.rowidx[q];
         V.rowmark[rflat] = next;
         if( !known && !rowRedundant(V, rflat) )
            known = true;
      }
      if( known )
      {
         if( V.vpend[v] != next )
            V.vpend[v] = next;   // every writer of this round stores the same value
         anymarked = true;
      }
      return;
   }
   for( int64_t q = M.colptr[j]; q < M.colptr[j + 1]; ++q )
   {
      const int32_t row = M.rowidx[q];
      const size_t rflat = (size_t)v * (size_t)M.nrows + (size_t)row;
      if( rowRedundant(V, rflat) )
         continue;
      if( V.vpend[v] != next )
         V.vpend[v] = next;   // every writer of this round stores the same value
      const bool first = atomicMax(&V.rowmark[rflat], next) < next;
      waveAppendRow(st, V.wl[npar], npar, first, M.rowbin[row], make_int2(row, v));
   }
}

__global__ __launch_bounds__(SAP_BLOCK) void k_columns(Mat M, Vec V, Params P)
{
   Status* st = V.st;
   if( st->done )
      return;
   const int32_t round = st->round;
   const int32_t next = round + 1;
   unsigned long long nchg = 0;
   if( V.dense )
   {
      const size_t n = (size_t)V.nvec * (size_t)M.ncols;
      for( size_t flat = blockIdx.x * (size_t)SAP_BLOCK + threadIdx.x; flat < n; flat += (size_t)gridDim.x * SAP_BLOCK )
      {
         if( V.coltag[flat] != round )
            continue;
         const int32_t v = (int32_t)(flat / (size_t)M.ncols);
         const int32_t j = (int32_t)(flat - (size_t)v * (size_t)M.ncols);
         nchg += (unsigned long long)columnStep(M, V, P, j, v, next);
      }
   }
   else
   {
      const uint32_t n = st->touched;
      for( uint32_t idx = blockIdx.x * SAP_BLOCK + threadIdx.x; idx < n; idx += gridDim.x * SAP_BLOCK )
      {
         const int2 t = V.touched[idx];
         nchg += (unsigned long long)columnStep(M, V, P, t.x, t.y, next);
      }
   }
   blockStats(nchg, 0, st->slot_chg, NULL);
}

// second half of the column step: rows of the columns changed in this round are marked for the
// next one, unless their unit was cut off in the round (acceptance and marking are separate
// launches so that a round ending in infeasibility marks nothing)
__device__ __forceinline__ void markRound(const Mat& M, const Vec& V)
{
   Status* st = V.st;
   const int32_t round = st->round;
   const int32_t next = round + 1;
   const int npar = next & 1;
   bool anymarked = false;
   if( V.dense )
   {
      const size_t n = (size_t)V.nvec * (size_t)M.ncols;
      for( size_t flat = blockIdx.x * (size_t)SAP_BLOCK + threadIdx.x; flat < n; flat += (size_t)gridDim.x * SAP_BLOCK )
      {
         if( V.chgtag[flat] != round )
            continue;
         const int32_t v = (int32_t)(flat / (size_t)M.ncols);
         const int32_t j = (int32_t)(flat - (size_t)v * (size_t)M.ncols);
         if( V.vcutoff[unitOfCol(M, v, j)] == 0 )
            markColumn(M, V, j, v, next, npar, anymarked);
      }
   }
   else
   {
      // The touched entries are taken SAP_BLOCK at a time.  The rows of the changed columns
      // among them are numbered across the workgroup (a scan of the column lengths in LDS),
      // so a wavefront appends up to 64 rows to the worklists at once instead of one append
      // per row of its longest column.  st->touched is not written in this launch, so the
      // loop around the barriers is workgroup-uniform.
      __shared__ int64_t soff[SAP_BLOCK + 1];   // soff[k]: rows of the chunk's columns before entry k
      __shared__ int64_t sbeg[SAP_BLOCK];       // colptr of entry k's column
      __shared__ int32_t svec[SAP_BLOCK];       // entry k's vector
      __shared__ uint32_t scnt[SAP_NBINS];
      __shared__ uint32_t sbase[SAP_NBINS];
      const int t = (int)threadIdx.x;
      const uint32_t n = st->touched;
      for( uint32_t c0 = blockIdx.x * SAP_BLOCK; c0 < n; c0 += gridDim.x * SAP_BLOCK )
      {
         int64_t len = 0, beg = 0;
         int32_t v = 0;
         if( c0 + (uint32_t)t < n )
         {
            const int2 e = V.touched[c0 + (uint32_t)t];
            const size_t flat = (size_t)e.y * (size_t)M.ncols + (size_t)e.x;
            if( V.chgtag[flat] == round && V.vcutoff[unitOfCol(M, e.y, e.x)] == 0 )
            {
               beg = M.colptr[e.x];
               len = M.colptr[e.x + 1] - beg;
               v = e.y;
            }
         }
         if( t == 0 )
            soff[0] = 0;
         soff[t + 1] = len;
         sbeg[t] = beg;
         svec[t] = v;
         __syncthreads();
         for( int d = 1; d < SAP_BLOCK; d <<= 1 )
         {
            const int64_t add = t >= d ? soff[t + 1 - d] : 0;
            __syncthreads();
            soff[t + 1] += add;
            __syncthreads();
         }
         const int64_t total = soff[SAP_BLOCK];
         // passes of SAP_BLOCK rows (workgroup-uniform: total is read from LDS), appended per
         // workgroup
         for( int64_t f0 = 0; f0 < total; f0 += SAP_BLOCK )
         {
            const int64_t f = f0 + t;
            bool fresh = false;
            int bin = 0;
            int2 item = make_int2(0, 0);
            if( f < total )
            {
               // the entry of row f: the largest k with soff[k] <= f, which has soff[k + 1] > f
               int lo = 0, hi = SAP_BLOCK - 1;
               while( lo < hi )
               {
                  const int mid = (lo + hi + 1) >> 1;
                  if( soff[mid] <= f )
                     lo = mid;
                  else
                     hi = mid - 1;
               }
               const int32_t rv = svec[lo];
               const int32_t row = M.rowidx[sbeg[lo] + (f - soff[lo])];
               const size_t rflat = (size_t)rv * (size_t)M.nrows + (size_t)row;
               if( !rowRedundant(V, rflat) )
               {
                  if( V.vpend[rv] != next )
                     V.vpend[rv] = next;   // every writer of this round stores the same value
                  fresh = atomicMax(&V.rowmark[rflat], next) < next;
                  bin = M.rowbin[row];
                  item = make_int2(row, rv);
               }
            }
            blockAppendRows(st, V.wl[npar], npar, fresh, bin, item, scnt, sbase);
         }
         __syncthreads();   // soff, sbeg and svec are rewritten by the next chunk
      }
   }
   if( __ballot(anymarked) != 0 && ((int)threadIdx.x & (SAP_WAVE - 1)) == 0 && st->anymark[npar] == 0 )
      st->anymark[npar] = 1;
}

__global__ __launch_bounds__(SAP_BLOCK) void k_mark(Mat M, Vec V)
{
   Status* st = V.st;
   if( !(st->done || (V.single && st->cutoff)) )
      markRound(M, V);
}

// reason recording, after the column step: every logged candidate whose key won its column's
// change (or crossing) offers its row; the smallest caller row is kept
__global__ __launch_bounds__(SAP_BLOCK) void k_reason(Vec V)
{
   Status* st = V.st;
   if( st->done )
      return;
   const uint32_t n = st->clogn < V.clogcap ? st->clogn : V.clogcap;
   for( uint32_t i = blockIdx.x * SAP_BLOCK + threadIdx.x; i < n; i += gridDim.x * SAP_BLOCK )
   {
      const CandRec r = V.clog[i];
      const int side = (r.rowside >> 1) & 1;
      const int forced = r.rowside & 1;
      const int f = V.rflag[r.col];
      const int have = side ? RF_UB : RF_LB;
      const int force = side ? RF_UBFORCE : RF_LBFORCE;
      if( (f & have) && ((f & force) != 0) == (forced != 0) && V.rkey[2 * (size_t)r.col + side] == r.key )
         atomicMin(&V.rrow[2 * (size_t)r.col + side], (r.rowside >> 3) * 2 + ((r.rowside >> 2) & 1));
   }
}

// reason recording: the changes accepted in this round join the history with their rows
// (not in a round that ended infeasible: those changes are not applied)
__global__ __launch_bounds__(SAP_BLOCK) void k_history(Mat M, Vec V)
{
   Status* st = V.st;
   if( st->done || st->cutoff )
      return;
   const int32_t round = st->rlocal;
   for( int64_t j = blockIdx.x * (int64_t)SAP_BLOCK + threadIdx.x; j < M.ncols; j += (int64_t)gridDim.x * SAP_BLOCK )
   {
      const int f = V.rflag[j];
      if( f == 0 )
         continue;
      V.rflag[j] = 0;
      const double2 b = V.bnd[j];
      for( int side = 0; side < 2; ++side )
      {
         if( !(f & (side ? RF_UB : RF_LB)) )
            continue;
         const uint32_t at = atomicAdd(&st->histn, 1u);
         if( at < V.histcap )
         {
            HistRec h;
            h.val = side ? b.y : b.x;
            h.round = round;
            h.colside = (int32_t)j * 2 + side;
            h.row = V.rrow[2 * (size_t)j + side];
            h.pad = 0;
            V.hist[at] = h;
         }
      }
   }
}

// ------------------------------------------------------------------------------------------
// ranged-row propagation (rangedRowPropagation, cons_linear.c:5715-6697), as propagation runs
// it after the node's root (no artificial constraints: cons_linear.c:5784).  A ranged row or
// equation with at least 3 nonzeros is split into the unfixed integral columns whose integral
// coefficients share a gcd >= 2 and the rest (the check group); the check group's activity
// range must meet a value that the gcd part can complement into [lhs, rhs].  No such value:
// infeasible; one or two check-group values or a single column outside the check group give
// fixings and bounds, submitted as forced candidates (SCIPinferVarFixCons / SCIPinferVarLbCons
// with force).  One thread per (vector, row), on the bounds at the start of the round.  The
// enumeration of check-group values stops after SAP_RANGED_STEPS steps, and the row then
// yields nothing.
// ------------------------------------------------------------------------------------------
#define SAP_RANGED_STEPS (1 << 20)

template <bool R>
__device__ __forceinline__ void rangedCut(const CtxT<R>& C)
{
   if constexpr( R )
      atomicMin(&C.V->st->cutrow, C.M->perm[C.row] * 8 + 2);
   setCutoff(C);
}

__device__ __forceinline__ double epsFloor(const Params& p, double x) { return floor(x + p.epsilon); }
__device__ __forceinline__ double epsCeil(const Params& p, double x) { return ceil(x - p.epsilon); }
__device__ __forceinline__ bool epsIntegral(const Params& p, double x) { return x - floor(x + p.epsilon) <= p.epsilon; }

__device__ __forceinline__ long long gcdLL(long long a, long long b)
{
   while( b != 0 )
   {
      const long long t = a % b;
      a = b;
      b = t;
   }
   return a;
}

template <bool R>
__device__ void rangedRow(const Mat& M, const Vec& V, const Params& P, int32_t v, int32_t row)
{
   const double2 sd = M.sides[row];
   const int64_t s = M.rowptr[row], e = M.rowptr[row + 1];
   if( e - s < 3 || isInf(P, -sd.x) || isInf(P, sd.y) )
      return;
   const double2* bnd = V.bnd + (size_t)v * (size_t)M.ncols;
   uint8_t* inchk = V.rgflag + (size_t)v * (size_t)M.rowptr[M.nrows];
   CtxT<R> C;
   C.rule = 1;
   C.M = &M;
   C.V = &V;
   C.P = &P;
   C.v = v;
   C.round = V.st->round;
   C.row = row;

   // activity of the fixed columns, summed from the last nonzero to the first
   double fixedact = 0.0;
   int64_t nfixed = 0;
   for( int64_t k = e - 1; k >= s; --k )
   {
      const double2 b = bnd[M.colidx[k]];
      if( isEQ(P, b.x, b.y) )
      {
         fixedact += b.x * M.vals[k];
         ++nfixed;
      }
   }
   if( isHuge(P, fabs(fixedact)) )
      return;
   const double lhs = sd.x - fixedact, rhs = sd.y - fixedact;
   const int64_t nunfixed = (e - s) - nfixed;

   // partition: up to the first unfixed gcd-group column, every other unfixed column joins the
   // check group; from there, a gcd-group column joins it when its gcd with the running gcd is 1
   bool possiblegcd = true, gcdisone = true;
   int64_t ninf = 0, ncont = 0, k = s;
   for( ; k < e; ++k )
   {
      const double a = M.vals[k];
      const double2 b = bnd[M.colidx[k]];
      const bool fixed = isEQ(P, b.x, b.y);
      const bool integ = (M.nzflags[k] & 1) != 0;
      const bool group = integ && epsIntegral(P, a) && !isEQ(P, fabs(a), 1.0);
      inchk[k] = 0;
      if( fixed )
         continue;
      if( group )
         break;
      ncont += integ ? 0 : 1;
      gcdisone = gcdisone && isEQ(P, fabs(a), 1.0);
      possiblegcd = false;
      inchk[k] = 1;
      ++ninf;
   }
   if( k == e || ncont + 2 > nunfixed )
      return;
   long long gcd = (long long)(fabs(M.vals[k]) + P.feastol);
   for( ; k < e; ++k )
   {
      const double a = M.vals[k];
      const double2 b = bnd[M.colidx[k]];
      const bool integ = (M.nzflags[k] & 1) != 0;
      inchk[k] = 0;
      if( isEQ(P, b.x, b.y) )
         continue;
      if( !integ || !epsIntegral(P, a) || isEQ(P, fabs(a), 1.0) )
      {
         ncont += integ ? 0 : 1;
         gcdisone = gcdisone && isEQ(P, fabs(a), 1.0);
         possiblegcd = false;
         inchk[k] = 1;
         ++ninf;
      }
      else
      {
         const long long g = gcdLL(gcd, (long long)(fabs(a) + P.feastol));
         if( g == 1 )
         {
            inchk[k] = 1;
            ++ninf;
         }
         else
            gcd = g;
      }
   }
   if( ninf == 0 )
      return;

   // activity range of the check group, from its last column to its first
   double minact = 0.0, maxact = 0.0;
   int64_t last = -1, first = -1;
   for( k = e - 1; k >= s; --k )
   {
      if( !inchk[k] )
         continue;
      last = last < 0 ? k : last;
      first = k;
      const double a = M.vals[k];
      const double2 b = bnd[M.colidx[k]];
      bool mininv = false, maxinv = false;
      if( isInf(P, -b.x) )
         (a < 0.0 ? maxinv : mininv) = true;
      else if( a < 0.0 )
         maxact += a * b.x;
      else
         minact += a * b.x;
      if( isInf(P, b.y) )
         (a > 0.0 ? maxinv : mininv) = true;
      else if( a > 0.0 )
         maxact += a * b.y;
      else
         minact += a * b.y;
      if( mininv || maxinv || isHuge(P, -minact) || isHuge(P, maxact) )
         return;
   }
   const double g = (double)gcd;
   if( !epsIntegral(P, (lhs - maxact) / g) && isGT(P, epsCeil(P, (lhs - maxact) / g) * g, rhs - minact) )
   {
      rangedCut(C);
      return;
   }
   if( ncont != 0 )
      return;
   long long gcdinf = -1;
   if( possiblegcd )
   {
      gcdinf = (long long)(fabs(M.vals[last]) + P.feastol);
      for( k = last; k >= s && gcdinf >= 2; --k )
         if( inchk[k] )
            gcdinf = gcdLL(gcdinf, (long long)(fabs(M.vals[k]) + P.feastol));
   }
   else if( gcdisone )
      gcdinf = 1;
   if( gcdinf < 1 )
      return;

   // check-group values that the gcd part can complete into [lhs, rhs]: the first, second and
   // last of them
   const double step = (double)gcdinf;
   int nsols = 0;
   double minvalue = 0.0, maxvalue = 0.0;
   int64_t steps = 0;
   for( double value = epsCeil(P, minact - P.feastol); isLE(P, value, maxact); value += step )
   {
      if( ++steps > SAP_RANGED_STEPS )
         return;
      double value2 = value + g * epsCeil(P, (lhs - value) / g);
      if( !isGE(P, value2, lhs) )
         value2 += g;
      if( isLE(P, value2, rhs) )
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
      for( double value = epsFloor(P, maxact + P.feastol); isGE(P, value, minact); value -= step )
      {
         if( ++steps > 2 * (int64_t)SAP_RANGED_STEPS )
            return;
         double value2 = value + g * epsFloor(P, (rhs - value) / g);
         if( !isLE(P, value2, rhs) )
            value2 -= g;
         if( isGE(P, value2, lhs) )
         {
            maxvalue = value;
            break;
         }
      }
   }
   if( nsols == 0 )
   {
      rangedCut(C);
      return;
   }
   // the single unfixed column outside the check group, if there is exactly one
   int64_t w = -1;
   if( ninf == nunfixed - 1 )
      for( k = s; k < e; ++k )
         if( !inchk[k] && !isEQ(P, bnd[M.colidx[k]].x, bnd[M.colidx[k]].y) )
         {
            w = k;
            break;
         }
   if( ninf == 1 )
   {
      const double a = M.vals[first];
      const int32_t j = M.colidx[first];
      if( nsols == 1 )
      {
         submitLb(C, j, maxvalue / a, true);
         submitUb(C, j, maxvalue / a, true);
      }
      else
      {
         const double nlb = a < 0.0 ? maxvalue / a : minvalue / a;
         const double nub = a < 0.0 ? minvalue / a : maxvalue / a;
         if( nlb > bnd[j].x )
            submitLb(C, j, nlb, true);
         if( nub < bnd[j].y )
            submitUb(C, j, nub, true);
      }
   }
   else if( w >= 0 )
   {
      const double a = M.vals[w];
      const int32_t j = M.colidx[w];
      if( nsols == 1 )
      {
         const double fix = a < 0.0 ? epsFloor(P, (lhs - maxvalue) / a) : epsCeil(P, (lhs - maxvalue) / a);
         submitLb(C, j, fix, true);
         submitUb(C, j, fix, true);
      }
      else
      {
         const double nlb = a < 0.0 ? epsFloor(P, (rhs - minvalue) / a) : epsCeil(P, (lhs - maxvalue) / a);
         const double nub = a < 0.0 ? epsFloor(P, (lhs - maxvalue) / a) : epsCeil(P, (rhs - minvalue) / a);
         if( nlb > bnd[j].x )
            submitLb(C, j, nlb, true);
         if( nub < bnd[j].y )
            submitUb(C, j, nub, true);
      }
   }
}

// ranged-row step of a round (before the row sweep, on the same start-of-round bounds): the
// eligible rows marked for this round and not redundant
template <bool R>
__global__ __launch_bounds__(SAP_BLOCK) void k_ranged(Mat M, Vec V, Params P)
{
   Status* st = V.st;
   if( st->done )
      return;
   const int32_t round = st->round;
   if( V.dense )
   {
      const bool all = st->firstall && st->rlocal == 1;
      const size_t n = (size_t)V.nvec * (size_t)M.nranged;
      for( size_t i = blockIdx.x * (size_t)SAP_BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * SAP_BLOCK )
      {
         const int32_t v = (int32_t)(i / (size_t)M.nranged);
         const int32_t row = M.rlist[i - (size_t)v * (size_t)M.nranged];
         const size_t rflat = (size_t)v * (size_t)M.nrows + (size_t)row;
         if( (all || V.rowmark[rflat] == round) && !rowRedundant(V, rflat) && (R || !V.vcutoff[unitOfRow(M, v, row)]) )
            rangedRow<R>(M, V, P, v, row);
      }
   }
   else
   {
      for( int b = 0; b < SAP_NBINS; ++b )
      {
         const uint32_t n = st->wlcount[round & 1][b];
         const int2* wl = V.wl[round & 1][b];
         for( uint32_t i = blockIdx.x * SAP_BLOCK + threadIdx.x; i < n; i += gridDim.x * SAP_BLOCK )
         {
            const int2 it = wl[i];
            const size_t rflat = (size_t)it.y * (size_t)M.nrows + (size_t)it.x;
            if( (M.rowflags[it.x] & 2) && !rowRedundant(V, rflat) && (R || !V.vcutoff[unitOfRow(M, it.y, it.x)]) )
               rangedRow<R>(M, V, P, it.y, it.x);
         }
      }
   }
}

__global__ void k_reason_reset(Vec V, int64_t ncols)
{
   for( int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < ncols; j += (int64_t)gridDim.x * blockDim.x )
      V.rflag[j] = 0;
}

// ------------------------------------------------------------------------------------------
// round control and set-up kernels (kernels, not memsets, so a round can be graph-captured)
// ------------------------------------------------------------------------------------------
// one workgroup of SAP_STAT_SLOTS threads: folds the previous sweep's timestamps, then
// (thread 0) advances the round or declares the call done
// (thread 0) advance the round or declare the call done
__device__ __forceinline__ void roundAdvance(const Vec& V, int32_t maxrounds)
{
   Status* st = V.st;
   if( threadIdx.x != 0 || st->done )
      return;
   if( stopsNext(*st, V.dense != 0, maxrounds) )
   {
      st->done = 1;
      return;
   }
   const int32_t r = st->round + 1;
   st->round = r;
   st->rlocal += 1;
   st->clogn = 0;
   st->anymark[(r + 1) & 1] = 0;
   for( int b = 0; b < SAP_NBINS; ++b )
      st->wlcount[(r + 1) & 1][b] = 0;
   st->touched = 0;
}

__global__ __launch_bounds__(SAP_STAT_SLOTS) void k_round_begin(Vec V, int32_t maxrounds)
{
   foldStamps(V.st);
   roundAdvance(V, maxrounds);
}

// per-call reset of the vector state; status counters for the call.  gen (1..255) is the
// call's generation of the locally-redundant row flags (redGen); the call that starts
// generation 1 clears the first nclear flags (those of every vector) first.  With begin set
// (a launch of one workgroup, nothing enqueued between it and the first round) it also does
// the first round's k_round_begin, whose stamps were just cleared, saving that launch.
__global__ void k_call_reset(Vec V, size_t nclear, int32_t nvec, int32_t firstall, int32_t gen, int32_t begin,
   int32_t maxrounds)
{
   for( size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nclear; i += (size_t)gridDim.x * blockDim.x )
      V.redundant[i] = 0;
   const size_t gi = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
   const size_t nunits = (size_t)nvec * (size_t)V.nseg;
   for( size_t u = gi; u < nunits; u += (size_t)gridDim.x * blockDim.x )
   {
      V.vcutoff[u] = 0;
      V.vrounds[u] = 0;
   }
   if( gi < (size_t)nvec )
   {
      V.vstart[gi] = V.st->round + 1;
      V.vpend[gi] = firstall ? V.st->round + 1 : 0;
      V.vlogn[gi] = 0;
      if( V.ovln != NULL )
         V.ovln[gi] = 0;
   }
   if( gi == 0 )
   {
      Status* st = V.st;
      st->cont = 0;
      st->rlocal = 0;
      st->done = 0;
      st->cutoff = 0;
      st->firstall = firstall;
      st->rgenm1 = gen - 1;
      st->anymark[0] = 0;
      st->anymark[1] = 0;
      st->touched = 0;
      st->rows_swept = 0;
      st->nnz_swept = 0;
      st->nchg = 0;
      st->nrecomp = 0;
      st->cutrow = INT32_MAX;
      st->cutcol = INT32_MAX;
      st->clogn = 0;
      st->histn = 0;
      st->clogovf = 0;
      for( int b = 0; b < SAP_NBINS; ++b )
      {
         st->wlcount[0][b] = 0;
         st->wlcount[1][b] = 0;
      }
   }
   if( gi < SAP_STAT_SLOTS )
   {
      V.st->slot_nnz[gi] = 0;
      V.st->slot_rows[gi] = 0;
      V.st->slot_chg[gi] = 0;
      V.st->tslot0[gi] = ~0ULL;
      V.st->tslot1[gi] = 0;
   }
   if( gi == 0 )
   {
      V.st->sweep_ticks = 0;
      V.st->sweep_count = 0;
   }
   if( begin )
   {
      __syncthreads();
      roundAdvance(V, maxrounds);
   }
}

// sparse mode: initial worklists hold every row of every vector
__global__ void k_init_all(Mat M, Vec V, int32_t bin, int32_t nvec)
{
   Status* st = V.st;
   const int par = (st->round + 1) & 1;
   const int32_t base = M.binstart[bin];
   const size_t len = (size_t)(M.binstart[bin + 1] - base);
   const size_t tot = (size_t)nvec * len;
   int2* wl = V.wl[par][bin];
   for( size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x )
      wl[i] = make_int2(M.binrows[base + (int32_t)(i % len)], (int32_t)(i / len));
   for( size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < (size_t)nvec; v += (size_t)gridDim.x * blockDim.x )
      V.vpend[v] = st->round + 1;
   if( blockIdx.x == 0 && threadIdx.x == 0 )
      st->wlcount[par][bin] = (uint32_t)tot;
}

// initial marks / worklists from overlay columns (base bounds already a fixpoint)
__global__ void k_init_overlay(Mat M, Vec V, const int64_t* ovptr, const int32_t* ovcol, int32_t nvec)
{
   Status* st = V.st;
   const int32_t first = st->round + 1;
   const int par = first & 1;
   bool any = false;
   for( int32_t v = blockIdx.y; v < nvec; v += gridDim.y )
   {
      for( int64_t q = ovptr[v] + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < ovptr[v + 1];
           q += (int64_t)gridDim.x * blockDim.x )
      {
         const int32_t j = ovcol[q];
         for( int64_t r = M.colptr[j]; r < M.colptr[j + 1]; ++r )
         {
            const int32_t row = M.rowidx[r];
            const size_t rflat = (size_t)v * (size_t)M.nrows + (size_t)row;
            if( V.vpend[v] != first )
               V.vpend[v] = first;
            if( V.dense )
            {
               V.rowmark[rflat] = first;
               any = true;
            }
            else
            {
               const bool fresh = atomicMax(&V.rowmark[rflat], first) < first;
               waveAppendRow(st, V.wl[par], par, fresh, M.rowbin[row], make_int2(row, v));
            }
         }
      }
   }
   if( __ballot(any) != 0 && ((int)threadIdx.x & (SAP_WAVE - 1)) == 0 )
      st->anymark[par] = 1;
}

// ---- slots: admission, step control, collection (vector indices through a slot list) ----

// fresh state for admitted slots: redundancy flags, units, start round, change log length;
// the bounds of each slot are the base vector plus the slot's overlay
__global__ void k_slot_reset(Vec V, int64_t nrows, const int32_t* slots, int32_t n)
{
   const int32_t first = V.st->round + 1;
   for( int32_t i = blockIdx.y; i < n; i += gridDim.y )
   {
      const int32_t v = slots[i];
      uint8_t* red = V.redundant + (size_t)v * (size_t)nrows;
      for( int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * blockDim.x )
         red[r] = 0;
      if( blockIdx.x == 0 )
         for( int32_t q = (int32_t)threadIdx.x; q < V.nseg; q += (int32_t)blockDim.x )
         {
            V.vcutoff[(size_t)v * (size_t)V.nseg + (size_t)q] = 0;
            V.vrounds[(size_t)v * (size_t)V.nseg + (size_t)q] = 0;
         }
      if( blockIdx.x == 0 && threadIdx.x == 0 )
      {
         V.vstart[v] = first;
         V.vpend[v] = 0;
         V.vlogn[v] = 0;
         V.ovln[v] = 0;
      }
   }
}

__global__ void k_slot_base(double2* bnd, const double2* base, int64_t ncols, const int32_t* slots, int32_t n)
{
   for( int32_t i = blockIdx.y; i < n; i += gridDim.y )
   {
      double2* dst = bnd + (size_t)slots[i] * (size_t)ncols;
      for( int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < ncols; j += (int64_t)gridDim.x * blockDim.x )
         dst[j] = base[j];
   }
}

__global__ void k_slot_overlay(Vec V, int64_t ncols, const int32_t* slots, int32_t n, const int64_t* ovptr,
   const int32_t* ovcol, const double* ovlb, const double* ovub)
{
   for( int32_t i = blockIdx.y; i < n; i += gridDim.y )
   {
      const int32_t v = slots[i];
      double2* dst = V.bnd + (size_t)v * (size_t)ncols;
      int32_t* save = V.ovl + (size_t)v * (size_t)ncols;
      for( int64_t q = ovptr[i] + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < ovptr[i + 1];
           q += (int64_t)gridDim.x * blockDim.x )
      {
         dst[ovcol[q]] = make_double2(ovlb[q], ovub[q]);
         save[q - ovptr[i]] = ovcol[q];
      }
      if( blockIdx.x == 0 && threadIdx.x == 0 )
         V.ovln[v] = (int32_t)(ovptr[i + 1] - ovptr[i]);
   }
}

// back to the base vector: only the columns of the change log and of the overlay of the
// slot's previous start differ from it
__global__ void k_slot_restore(Vec V, const double2* base, int64_t ncols, const int32_t* slots, int32_t n)
{
   for( int32_t i = blockIdx.y; i < n; i += gridDim.y )
   {
      const int32_t v = slots[i];
      double2* dst = V.bnd + (size_t)v * (size_t)ncols;
      const int32_t* log = V.log + (size_t)v * (size_t)ncols;
      const int32_t* ovl = V.ovl + (size_t)v * (size_t)ncols;
      const int32_t nl = V.vlogn[v], no = V.ovln[v];
      for( int32_t q = blockIdx.x * (int32_t)blockDim.x + (int32_t)threadIdx.x; q < nl + no;
           q += (int32_t)(gridDim.x * blockDim.x) )
      {
         const int32_t j = q < nl ? log[q] : ovl[q - nl];
         dst[j] = base[j];
      }
   }
}

// one row of vector v marked for the round `first` (dense: round tag; sparse: worklist entry)
__device__ __forceinline__ void slotMarkRow(const Mat& M, const Vec& V, Status* st, int32_t v, int32_t row,
   int32_t first, int par, bool& any)
{
   const size_t rflat = (size_t)v * (size_t)M.nrows + (size_t)row;
   if( V.vpend[v] != first )
      V.vpend[v] = first;
   if( V.dense )
   {
      V.rowmark[rflat] = first;
      any = true;
   }
   else
   {
      const bool fresh = atomicMax(&V.rowmark[rflat], first) < first;
      waveAppendRow(st, V.wl[par], par, fresh, M.rowbin[row], make_int2(row, v));
   }
}

// marks for the first round of the admitted slots: the rows of the overlay columns (base is a
// fixpoint), or every row (grid: rows per slot in x, slots in y).  The overlay entries of all
// slots (ovptr[0] = 0 .. ovptr[n]) are taken SAP_BLOCK at a time over a one-dimensional grid,
// each thread finding its slot by binary search in ovptr, and the (column, row) pairs of a
// chunk are numbered across the workgroup (a scan of the column lengths in LDS).  So a
// wavefront appends up to 64 rows to the worklists at once, also when every slot has a single
// overlay column (probing): a workgroup or thread per slot or column would leave most lanes
// idle and make one append per row, and the appends of all slots contend on the few worklist
// counters.  Launched with SAP_BLOCK threads.
__global__ void __launch_bounds__(SAP_BLOCK) k_slot_mark(Mat M, Vec V, const int32_t* slots, int32_t n,
   const int64_t* ovptr, const int32_t* ovcol, int32_t allrows, int32_t chunk)
{
   __shared__ int64_t soff[SAP_BLOCK + 1];   // soff[k]: pairs of the chunk's columns before column k
   __shared__ int64_t sbeg[SAP_BLOCK];       // colptr of column k
   __shared__ int32_t svec[SAP_BLOCK];       // vector of column k
   __shared__ uint32_t scnt[SAP_NBINS];      // fresh rows of one pass per length bin
   __shared__ uint32_t sbase[SAP_NBINS];     // their first worklist position
   Status* st = V.st;
   const int32_t first = st->round + 1;
   const int par = first & 1;
   const int t = (int)threadIdx.x;
   bool any = false;
   if( allrows )
   {
      for( int32_t i = blockIdx.y; i < n; i += gridDim.y )
      {
         const int32_t v = slots[i];
         for( int64_t q = blockIdx.x * (int64_t)blockDim.x + t; q < M.nrows; q += (int64_t)gridDim.x * blockDim.x )
            slotMarkRow(M, V, st, v, (int32_t)q, first, par, any);
      }
   }
   else if( blockIdx.y == 0 && n > 0 )   // one-dimensional grid over the overlay entries
   {
      const int64_t nov = ovptr[n];
      // workgroup-uniform loop: c0 and nov depend on the block index and the arguments only
      for( int64_t c0 = (int64_t)blockIdx.x * chunk; c0 < nov; c0 += (int64_t)gridDim.x * chunk )
      {
         int64_t len = 0, beg = 0;
         int32_t v = 0;
         const int64_t e = c0 + t;
         if( t < chunk && e < nov )
         {
            // the slot of entry e: the largest i < n with ovptr[i] <= e (ovptr[0] = 0 <= e; such
            // an i has ovptr[i + 1] > e, as ovptr[n] = nov > e)
            int32_t lo = 0, hi = n - 1;
            while( lo < hi )
            {
               const int32_t mid = lo + (hi - lo + 1) / 2;
               if( ovptr[mid] <= e )
                  lo = mid;
               else
                  hi = mid - 1;
            }
            v = slots[lo];
            const int32_t j = ovcol[e];
            beg = M.colptr[j];
            len = M.colptr[j + 1] - beg;
         }
         if( t == 0 )
            soff[0] = 0;
         soff[t + 1] = len;
         sbeg[t] = beg;
         svec[t] = v;
         __syncthreads();
         for( int d = 1; d < SAP_BLOCK; d <<= 1 )
         {
            const int64_t add = t >= d ? soff[t + 1 - d] : 0;
            __syncthreads();
            soff[t + 1] += add;
            __syncthreads();
         }
         const int64_t total = soff[SAP_BLOCK];
         // passes of SAP_BLOCK pairs (workgroup-uniform: total is read from LDS); the fresh
         // rows of a pass are counted per length bin in LDS and appended with one global
         // atomic per bin and workgroup
         for( int64_t f0 = 0; f0 < total; f0 += SAP_BLOCK )
         {
            const int64_t f = f0 + t;
            bool fresh = false;
            int bin = 0;
            int2 item = make_int2(0, 0);
            if( f < total )
            {
               // the column of pair f: the largest k with soff[k] <= f (soff[0] = 0 <= f; such
               // a k has soff[k + 1] > f, so a column with rows)
               int lo = 0, hi = SAP_BLOCK - 1;
               while( lo < hi )
               {
                  const int mid = (lo + hi + 1) >> 1;
                  if( soff[mid] <= f )
                     lo = mid;
                  else
                     hi = mid - 1;
               }
               const int32_t rv = svec[lo];
               const int32_t row = M.rowidx[sbeg[lo] + (f - soff[lo])];
               const size_t rflat = (size_t)rv * (size_t)M.nrows + (size_t)row;
               if( V.vpend[rv] != first )
                  V.vpend[rv] = first;
               if( V.dense )
               {
                  V.rowmark[rflat] = first;
                  any = true;
               }
               else
               {
                  fresh = atomicMax(&V.rowmark[rflat], first) < first;
                  bin = M.rowbin[row];
                  item = make_int2(row, rv);
               }
            }
            blockAppendRows(st, V.wl[par], par, fresh, bin, item, scnt, sbase);
         }
         __syncthreads();   // soff, sbeg and svec are rewritten by the next chunk
      }
   }
   if( __ballot(any) != 0 && ((int)threadIdx.x & (SAP_WAVE - 1)) == 0 )
      st->anymark[par] = 1;
}

// start of a step of continuous batching: clear the stop flag and the step's statistics
__global__ void k_step_begin(Vec V)
{
   Status* st = V.st;
   const int t = (int)threadIdx.x;
   if( t < SAP_STAT_SLOTS )
   {
      st->slot_nnz[t] = 0;
      st->slot_rows[t] = 0;
      st->slot_chg[t] = 0;
      st->tslot0[t] = ~0ULL;
      st->tslot1[t] = 0;
   }
   if( t == 0 )
   {
      st->cont = 1;
      st->done = 0;
      st->cutoff = 0;
      st->rlocal = 0;
      st->firstall = 0;
      st->sweep_ticks = 0;
      st->sweep_count = 0;
   }
}

// batched call whose change lists come back with its last wait: after every round group, once
// the call has ended (its done flag, or the stop test the host applies after the group), the
// lists of vectors 0..n-1 are numbered into outoff (a cut-off vector's list is empty) and
// outoff[n] and hdr[0] (mapped host memory) receive their total; -1 while the call goes on.
// One workgroup.
__global__ __launch_bounds__(SAP_BLOCK) void k_list_offsets(Vec V, int32_t n, int32_t maxrounds, int64_t* outoff,
   int64_t* hdr)
{
   const Status* st = V.st;
   const bool ended = st->done || stopsNext(*st, V.dense != 0, maxrounds);
   __shared__ int64_t part[SAP_BLOCK];
   const int t = (int)threadIdx.x;
   int64_t carry = 0;
   for( int32_t c0 = 0; ended && c0 < n; c0 += SAP_BLOCK )
   {
      const int32_t v = c0 + t;
      const int64_t len = (v < n && V.vcutoff[v] == 0) ? (int64_t)V.vlogn[v] : 0;
      part[t] = len;
      __syncthreads();
      for( int d = 1; d < SAP_BLOCK; d <<= 1 )
      {
         const int64_t add = t >= d ? part[t - d] : 0;
         __syncthreads();
         part[t] += add;
         __syncthreads();
      }
      if( v < n )
         outoff[v] = carry + part[t] - len;
      carry += part[SAP_BLOCK - 1];
      __syncthreads();   // part is rewritten by the next chunk
   }
   if( t == 0 )
   {
      outoff[n] = ended ? carry : -1;
      hdr[0] = ended ? carry : -1;
   }
}

// the lists k_list_offsets numbered, written to mapped host memory when their total fits cap
// entries, and every logged column's bit cleared (a cut-off vector's too); nothing while the
// call goes on or when the lists do not fit (the logs stay for a collection with room)
__global__ void k_collect_ended(Vec V, int64_t ncols, int32_t n, const int64_t* outoff, int64_t cap, int32_t* outcol,
   double2* outbnd)
{
   const int64_t tot = outoff[n];
   if( tot < 0 || tot > cap )
      return;
   for( int32_t v = blockIdx.y; v < n; v += gridDim.y )
   {
      const int32_t cnt = V.vlogn[v];
      const bool keep = V.vcutoff[v] == 0;
      const int64_t o = outoff[v];
      const int32_t* log = V.log + (size_t)v * (size_t)ncols;
      for( int32_t q = blockIdx.x * (int32_t)blockDim.x + (int32_t)threadIdx.x; q < cnt; q += (int32_t)(gridDim.x * blockDim.x) )
      {
         const int32_t j = log[q];
         const size_t flat = (size_t)v * (size_t)ncols + (size_t)j;
         if( keep )
         {
            outcol[o + q] = j;
            outbnd[o + q] = V.bnd[flat];
         }
         atomicAnd(&V.logbits[flat >> 5], ~(1u << (flat & 31)));
      }
   }
}

// changed columns of the listed slots with their final bounds, at outoff[i] for slot i; the
// log bits of the collected entries are cleared for the slot's next admission
__global__ void k_collect(Vec V, int64_t ncols, const int32_t* slots, int32_t n, const int64_t* outoff,
   int32_t* outcol, double2* outbnd)
{
   for( int32_t i = blockIdx.y; i < n; i += gridDim.y )
   {
      const int32_t v = slots[i];
      const int32_t cnt = V.vlogn[v];
      const int32_t* log = V.log + (size_t)v * (size_t)ncols;
      for( int32_t q = blockIdx.x * (int32_t)blockDim.x + (int32_t)threadIdx.x; q < cnt; q += (int32_t)(gridDim.x * blockDim.x) )
      {
         const int32_t j = log[q];
         const size_t flat = (size_t)v * (size_t)ncols + (size_t)j;
         if( outcol != NULL )
         {
            outcol[outoff[i] + q] = j;
            outbnd[outoff[i] + q] = V.bnd[flat];
         }
         atomicAnd(&V.logbits[flat >> 5], ~(1u << (flat & 31)));
      }
   }
}

// probing: *bad receives the smallest index i whose column pcol[i] is not an integral column
// with base bounds exactly [0, 1]
__global__ void k_probe_check(Mat M, const double2* base, const int32_t* pcol, int32_t n, int32_t* bad)
{
   for( int32_t i = blockIdx.x * (int32_t)blockDim.x + (int32_t)threadIdx.x; i < n; i += (int32_t)(gridDim.x * blockDim.x) )
   {
      const int32_t j = pcol[i];
      const double2 bj = base[j];
      if( !((M.colflags[j] & 1) && bj.x == 0.0 && bj.y == 1.0) )
         atomicMin(bad, i);
   }
}

// probing: bounds of column cols[i] in vector vec[i] (vec == NULL: in src itself)
__global__ void k_gather_bounds(const double2* src, int64_t ncols, const int32_t* cols, const int32_t* vec, int32_t n,
   double2* out)
{
   for( int32_t i = blockIdx.x * (int32_t)blockDim.x + (int32_t)threadIdx.x; i < n; i += (int32_t)(gridDim.x * blockDim.x) )
      out[i] = src[(vec != NULL ? (size_t)vec[i] * (size_t)ncols : 0) + (size_t)cols[i]];
}

// probing: probe p owns vectors 2p (x = 0) and 2p+1 (x = 1).  When neither branch is cut off,
// the columns other than x found in both change logs are written at outoff[p] + (count[p]++)
// with their bounds in both branches.  The shorter log is scanned and the other branch's log
// bits are tested, so the work is the smaller change count.
__global__ void k_probe_pairs(Vec V, int64_t ncols, const int32_t* pcol, int32_t np, const int64_t* outoff,
   int32_t* count, int32_t* outcol, double2* out0, double2* out1)
{
   for( int32_t p = blockIdx.y; p < np; p += gridDim.y )
   {
      const int32_t v0 = 2 * p, v1 = 2 * p + 1;
      if( V.vcutoff[v0] != 0 || V.vcutoff[v1] != 0 )
         continue;
      const int32_t n0 = V.vlogn[v0], n1 = V.vlogn[v1];
      const int32_t vs = n0 <= n1 ? v0 : v1;
      const int32_t vo = n0 <= n1 ? v1 : v0;
      const int32_t ns = n0 <= n1 ? n0 : n1;
      const int32_t x = pcol[p];
      const int32_t* log = V.log + (size_t)vs * (size_t)ncols;
      for( int32_t q = blockIdx.x * (int32_t)blockDim.x + (int32_t)threadIdx.x; q < ns; q += (int32_t)(gridDim.x * blockDim.x) )
      {
         const int32_t j = log[q];
         const size_t flat = (size_t)vo * (size_t)ncols + (size_t)j;
         if( j == x || ((V.logbits[flat >> 5] >> (flat & 31)) & 1u) == 0 )
            continue;
         const int64_t at = outoff[p] + atomicAdd(&count[p], 1);
         outcol[at] = j;
         out0[at] = V.bnd[(size_t)v0 * (size_t)ncols + (size_t)j];
         out1[at] = V.bnd[(size_t)v1 * (size_t)ncols + (size_t)j];
      }
   }
}

// bounds of all vectors from a base vector plus overlays
__global__ void k_broadcast_base(double2* bnd, const double2* base, int64_t ncols, int32_t nvec)
{
   const size_t tot = (size_t)nvec * (size_t)ncols;
   for( size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x )
      bnd[i] = base[i % (size_t)ncols];
}

__global__ void k_apply_overlay(double2* bnd, int64_t ncols, const int64_t* ovptr, const int32_t* ovcol,
   const double* ovlb, const double* ovub, int32_t nvec)
{
   for( int32_t v = blockIdx.y; v < nvec; v += gridDim.y )
      for( int64_t q = ovptr[v] + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < ovptr[v + 1];
           q += (int64_t)gridDim.x * blockDim.x )
         bnd[(size_t)v * (size_t)ncols + (size_t)ovcol[q]] = make_double2(ovlb[q], ovub[q]);
}

__global__ void k_fill_i64(long long* p, long long val, size_t n)
{
   for( size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x )
      p[i] = val;
}

__global__ void k_fill_i32(int32_t* p, int32_t val, size_t n)
{
   for( size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x )
      p[i] = val;
}

// word copy between device memory and mapped pinned host memory.  The batched calls move
// their small lists this way, as a kernel on the call's own stream, instead of through
// hipMemcpyAsync, whose copies of pinned memory measured slower for these sizes.
__global__ void k_copy_words(uint32_t* dst, const uint32_t* src, size_t n)
{
   for( size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x )
      dst[i] = src[i];
}

// up to four word copies of one upload (piece y of the grid's y dimension)
struct CopyPieces
{
   uint32_t* dst[4];
   const uint32_t* src[4];
   size_t n[4];
};

__global__ void k_copy_pieces(CopyPieces c)
{
   const int y = (int)blockIdx.y;
   uint32_t* dst = c.dst[y];
   const uint32_t* src = c.src[y];
   const size_t n = c.n[y];
   for( size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x )
      dst[i] = src[i];
}

__global__ void k_iota_i32(int32_t* p, size_t n)
{
   for( size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x )
      p[i] = (int32_t)i;
}

// the per-vector outcome of a call on vectors 0..n-1 of a one-instance bound set, written
// packed into mapped host memory: change-log length, cutoff flag, last round and start round
__global__ void k_pack_meta(Vec V, int32_t n, int32_t* out)
{
   for( int32_t v = blockIdx.x * (int32_t)blockDim.x + (int32_t)threadIdx.x; v < n; v += (int32_t)(gridDim.x * blockDim.x) )
   {
      out[v] = V.vlogn[v];
      out[(size_t)n + v] = V.vcutoff[v];
      out[2 * (size_t)n + v] = V.vrounds[v];
      out[3 * (size_t)n + v] = V.vstart[v];
   }
}

__global__ void k_interleave(double2* dst, const double* lb, const double* ub, int64_t n)
{
   for( int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x )
      dst[i] = make_double2(lb[i], ub[i]);
}

__global__ void k_deinterleave(const double2* src, double* lb, double* ub, int64_t n)
{
   for( int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x )
   {
      double2 b = src[i];
      lb[i] = b.x;
      ub[i] = b.y;
   }
}

// the memory floor of the row sweep: the sweep's own streams and gathers without its
// arithmetic.  Lane i of the grid takes nonzeros 4i..4i+3 (the four positions of a lane of a
// width-4 tile), reads them as one vector of column indices and two of values, so that a
// wavefront's loads cover one contiguous range each (reading them one nonzero at a time per
// lane costs ~30 % more on the 1M-row benchmark instance), and gathers the four columns'
// (lb, ub); the sum is kept live by a store that only a sum of exactly -1 would make (never,
// for the bench's data).  The last nnz % 4 nonzeros are read one by one.
__global__ __launch_bounds__(SAP_BLOCK) void k_gather_floor(Mat M, const double2* bnd, int64_t nnz, double* sink)
{
   const int64_t nq = nnz / 4;
   double acc = 0.0;
   for( int64_t q = (int64_t)blockIdx.x * SAP_BLOCK + threadIdx.x; q < nq; q += (int64_t)gridDim.x * SAP_BLOCK )
   {
      const int4 c = *(const int4*)(M.colidx + 4 * q);
      const double2 v0 = *(const double2*)(M.vals + 4 * q);
      const double2 v1 = *(const double2*)(M.vals + 4 * q + 2);
      const double2 b0 = bnd[c.x], b1 = bnd[c.y], b2 = bnd[c.z], b3 = bnd[c.w];
      acc += v0.x * (b0.x + b0.y) + v0.y * (b1.x + b1.y) + v1.x * (b2.x + b2.y) + v1.y * (b3.x + b3.y);
   }
   if( blockIdx.x == 0 && threadIdx.x == 0 )
      for( int64_t k = 4 * nq; k < nnz; ++k )
      {
         const double2 b = bnd[M.colidx[k]];
         acc += M.vals[k] * (b.x + b.y);
      }
   if( acc == -1.0 )
      sink[0] = acc;
}

// row activities for diagnostics: one wavefront per row (chunked canonical tree)
__global__ __launch_bounds__(SAP_BLOCK) void k_row_activities(Mat M, Params P, const double2* bnd, const int32_t* perm,
   double* minact, double* maxact, int32_t* counts, double* mad)
{
   const int lane = (int)threadIdx.x & (SAP_WAVE - 1);
   const int64_t wid = (blockIdx.x * (int64_t)SAP_BLOCK + threadIdx.x) / SAP_WAVE;
   const int64_t nw = (int64_t)gridDim.x * (SAP_BLOCK / SAP_WAVE);
   for( int64_t row = wid; row < M.nrows; row += nw )
   {
      const int64_t s = M.rowptr[row];
      const int64_t n = M.rowptr[row + 1] - s;
      const int64_t nch = (n + SAP_WAVE - 1) / SAP_WAVE;
      WavePart w = waveSubtree(M, P, bnd, s, n, 0, nch, (M.rowflags[row] & 1) != 0);
      if( lane == 0 )
      {
         const int64_t orow = perm[row];
         minact[orow] = ddVal(w.minact);
         maxact[orow] = ddVal(w.maxact);
         for( int q = 0; q < 8; ++q )
            counts[8 * orow + q] = w.c[q];
         mad[orow] = w.inf ? P.infinity : w.mad;
      }
   }
}

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------
// ---- scratch of the no-allocation entry point (sap_propagate_device_noalloc) ----
// One dense single-vector bound set (bounds, candidate keys, round tags, row marks, redundancy
// flags, per-unit counters and the status block) for matrices of up to SAP_NOALLOC_MAXCOLS
// columns and SAP_NOALLOC_MAXROWS rows: 56 B per column, 5 B per row, 256-byte aligned pieces.
#define SAP_NOALLOC_MAXCOLS ((int64_t)1 << 20)
#define SAP_NOALLOC_MAXROWS ((int64_t)1 << 20)
#define SAP_NOALLOC_BYTES ((size_t)SAP_NOALLOC_MAXCOLS * 56 + (size_t)SAP_NOALLOC_MAXROWS * 5 + ((size_t)1 << 16))
static __device__ __align__(256) unsigned char g_noalloc_scratch[SAP_NOALLOC_BYTES];

struct sap_context
{
   int32_t device;
   int32_t wallclock_khz;   // rate of wall_clock64() on the device
   sap_tuning tuning;       // applied to matrices created on this context
   int64_t nmatrices;       // live matrices (a context outlives its matrices)
   hipStream_t stream;
   hipStream_t cstream;     // round graphs are captured here: never launched on, so a capture
                            // never begins on a stream with work of other bound sets in flight
   hipEvent_t ev[4];
   Status* hstatus;   // pinned, mapped and coherent: the status after a round group ...
   Status* hstatusd;  // ... written by a copy kernel through this device address
   unsigned char* scratch;   // this device's address of the no-allocation scratch array
   int32_t resident;         // sweep workgroups the device holds at once (0: unknown)
};

struct sap_matrix
{
   sap_context* ctx;
   int64_t nrows, ncols, nnz;
   int64_t device_bytes;
   int32_t hasforced;
   int64_t* rowptr;
   int32_t* colidx;
   double* vals;
   double2* sides;
   uint8_t* colflags;
   uint8_t* nzflags;
   uint8_t* rowflags;
   uint8_t* rowbin;
   int32_t* binrows;
   int64_t* colptr;
   int32_t* rowidx;
   int32_t* perm;        // device row position -> caller's row index
   int32_t* rlist;       // device rows eligible for ranged-row propagation
   int32_t nranged;
   int32_t binstart[SAP_NBINS + 1];
   int32_t ntiles;
   int64_t* tstart;
   unsigned long long* tmask;
   int32_t* trow0;
   uint8_t* tlev;
   unsigned long long* tpad;
   sap_tuning tuning;    // row binning and launch shape it was built with
   int32_t nseg;         // packed instances (1: a single instance)
   int32_t* rowseg;      // device, by device row position (NULL when nseg == 1)
   int32_t* colseg;      // device (NULL when nseg == 1)
   std::vector<int64_t>* rowoff;   // host: segment row / column ranges (caller order)
   std::vector<int64_t>* coloff;
   sap_bounds* single;   // cached one-vector state for sap_propagate
   sap_bounds* rsingle;  // cached one-vector state with reason recording (sap_propagate_reasons)
   int64_t nbounds;      // live caller-created bound sets (a matrix outlives them)
};

// rounds captured into one hipGraph: a short graph for the first launch of a call (many
// propagations end after one or two rounds) and a long one for the rest
#define SAP_GRAPH_SHORT 2
#define SAP_GRAPH_LONG 8
#define SAP_NGRAPHS 3

struct RoundGraph
{
   hipGraphExec_t exec;
   int32_t rounds;
   int32_t valid;
   hipEvent_t ev[2 * SAP_GRAPH_LONG];   // around each captured sweep launch
};

// everything a captured round bakes into its kernel arguments: the parameters the kernels
// read, the vector state (device pointers included: a buffer that is reallocated after the
// capture must not be reached through the old pointer), the matrix and the sweep layout.  The
// raw bytes are compared, so padding only ever causes a spurious recapture.
struct GraphKey
{
   Params P;
   Vec V;
   Mat M;
   Layout L;
};

struct sap_bounds
{
   sap_matrix* mat;
   int32_t nvec;
   Vec V;
   Status* dstatus;
   int64_t device_bytes;
   double* stage;   // device staging for host lb/ub (2*ncols)
   int32_t roundbase;
   int32_t graphbroken;   // capture failed once: use the host-driven loop
   int32_t redgen;        // generation of the redundant-row flags of the last call (0: none yet)
   RoundGraph graph[SAP_NGRAPHS];   // SAP_GRAPH_SHORT, SAP_GRAPH_LONG and 4 rounds
   GraphKey key;
   // persistent staging for batches and slots (grown on demand, never per call)
   int64_t* ovptr;       // nvec+1
   int32_t* ovcol;
   double* ovlb;
   double* ovub;
   int64_t ovcap;
   double2* base2;       // ncols
   int32_t* slotidx;     // nvec: slot list of an admission / collection
   int32_t* slotidx2;    // nvec: second list (restored slots)
   int64_t* outoff;      // nvec+1: output offsets of a collection
   int32_t* outcol;      // gathered changes of a collection
   double2* outbnd;
   int64_t outcap;
   int32_t* pcol;        // nvec: probe columns of a probing chunk
   int32_t* pcount;      // nvec: per-probe output counters
   double2* outbnd2;     // probing: bounds in the second branch, beside outbnd (capacity outcap2)
   int64_t outcap2;
   int32_t logready;     // change log allocated and cleared
   int32_t slotsmode;    // slots are in use (continuous batching)
   int32_t inflight;     // a step was launched and not yet waited for
   std::vector<int32_t>* active;   // host: 0 free, 1 running, 2 finished and not yet collected
   Status* hstat;        // pinned: status after the last step
   int32_t* hpend;       // pinned: vpend then vcutoff after the last step (2*nvec)
   hipEvent_t evs[2];    // around a step
   int64_t gen;          // bumped by every write to the vectors' bounds outside the change log
   int32_t internal;     // the matrix's own one-vector state (not counted in nbounds)
   // per vector: the base vector it was last started from; while that base is unchanged, only
   // the columns of the change log and of the overlay differ from it
   std::vector<const sap_bounds*>* srcb;
   std::vector<int32_t>* srcv;
   std::vector<int64_t>* srcgen;
   // host side of the batched calls' small transfers: pinned memory, so that no copy is a
   // synchronous pageable one (uploads are handed out in order, see upload(); downloads land
   // in hdn, read after the stream is synchronised)
   int32_t* ident;       // nvec: 0, 1, ..., nvec-1 (the slot list of a one-shot batch)
   char* hup;            // pinned, mapped and coherent: uploads are staged here ...
   char* hupd;           // ... and read by k_copy_words through this device address
   size_t hupcap, hupused;
   char* hdn;            // pinned landing area, written by k_copy_words through hdnd
   char* hdnd;
   size_t hdncap;
   // vectors 0..logmax-1 may hold change-log bits that no collection has cleared yet
   int32_t logmax;
   int32_t begun;   // the last beginCall also began the call's first round
   int32_t lastrounds;   // rounds of the last graph-driven call on b (predicts the next one's)
};

// the fields of the vector state that one call narrows (the vectors it works on, the
// single-vector shortcut, change logging); only these are put back after the call, so that a
// buffer allocated during the call (ranged-row flags) stays attached to the bound set
struct CallScope
{
   int32_t nvec, single, logchg;
   explicit CallScope(const Vec& V) : nvec(V.nvec), single(V.single), logchg(V.logchg) {}
   void restore(Vec& V) const
   {
      V.nvec = nvec;
      V.single = single;
      V.logchg = logchg;
   }
};

static int hipToRc(hipError_t e)
{
   if( e == hipSuccess )
      return SAP_OKAY;
   if( e == hipErrorOutOfMemory )
      return SAP_NOMEMORY;
   if( e == hipErrorInvalidValue || e == hipErrorInvalidDevice )
      return SAP_INVALIDDATA;
   return SAP_ERROR;
}

#define HIPCHK(x)                                                                                    \
   do                                                                                                \
   {                                                                                                 \
      hipError_t _e = (x);                                                                           \
      if( _e != hipSuccess )                                                                         \
      {                                                                                              \
         fprintf(stderr, "[scip_amd] %s failed: %s (%s:%d)\n", #x, hipGetErrorString(_e), __FILE__, \
            __LINE__);                                                                               \
         return hipToRc(_e);                                                                         \
      }                                                                                              \
   } while( 0 )

template <typename T>
static int devAlloc(T** p, size_t count, int64_t* acc)
{
   size_t bytes = sizeof(T) * (count > 0 ? count : 1);
   hipError_t e = hipMalloc((void**)p, bytes);
   if( e != hipSuccess )
   {
      *p = NULL;
      fprintf(stderr, "[scip_amd] hipMalloc(%zu) failed: %s\n", bytes, hipGetErrorString(e));
      return hipToRc(e);
   }
   if( acc )
      *acc += (int64_t)bytes;
   return SAP_OKAY;
}

#define RCCHK(x)                \
   do                           \
   {                            \
      int _rc = (x);            \
      if( _rc != SAP_OKAY )     \
         return _rc;            \
   } while( 0 )

static inline unsigned gridFor(size_t n, unsigned cap = 8192)
{
   size_t g = (n + SAP_BLOCK - 1) / SAP_BLOCK;
   if( g < 1 )
      g = 1;
   if( g > cap )
      g = cap;
   return (unsigned)g;
}

static Params toParams(const sap_params* p)
{
   Params q;
   q.infinity = p->infinity;
   q.epsilon = p->epsilon;
   q.sumepsilon = p->sumepsilon;
   q.feastol = p->feastol;
   q.boundstreps = p->boundstreps;
   q.hugeval = p->hugeval;
   q.recompfac = p->recompfac;
   q.maxeasy = p->maxeasyactivitydelta;
   q.maxrounds = p->maxrounds;
   q.flags = p->flags;
   if( q.flags & SAP_FLAG_NOTIGHTEN )
      q.flags &= ~SAP_FLAG_RANGEDROWS;   // ranged rows run only together with the tightening
   return q;
}

static Mat toMat(const sap_matrix* m)
{
   Mat M;
   M.nrows = m->nrows;
   M.ncols = m->ncols;
   M.rowptr = m->rowptr;
   M.colidx = m->colidx;
   M.vals = m->vals;
   M.sides = m->sides;
   M.colflags = m->colflags;
   M.nzflags = m->nzflags;
   M.rowflags = m->rowflags;
   M.rowbin = m->rowbin;
   M.binrows = m->binrows;
   M.colptr = m->colptr;
   M.rowidx = m->rowidx;
   for( int b = 0; b <= SAP_NBINS; ++b )
      M.binstart[b] = m->binstart[b];
   M.ntiles = m->ntiles;
   M.pad = 0;
   M.tstart = m->tstart;
   M.tmask = m->tmask;
   M.trow0 = m->trow0;
   M.tlev = m->tlev;
   M.tpad = m->tpad;
   M.twidth = m->tuning.tilewidth;
   M.pad3 = 0;
   M.nseg = m->nseg;
   M.rowbins = m->tuning.rowbins;
   M.rowseg = m->nseg > 1 ? m->rowseg : NULL;
   M.colseg = m->nseg > 1 ? m->colseg : NULL;
   M.perm = m->perm;
   M.rlist = m->rlist;
   M.nranged = m->nranged;
   M.pad2 = 0;
   return M;
}

// host threads for the upload's preprocessing
template <typename F>
static void parallelFor(int64_t n, int64_t grain, F fn)
{
   int nt = (int)std::thread::hardware_concurrency();
   nt = nt < 1 ? 1 : (nt > 16 ? 16 : nt);
   if( n <= grain || nt == 1 )
   {
      fn(0, n, 0);
      return;
   }
   const int64_t chunks = std::min<int64_t>(nt, (n + grain - 1) / grain);
   std::vector<std::thread> th;
   for( int64_t t = 0; t < chunks; ++t )
      th.emplace_back([&, t]() { fn(n * t / chunks, n * (t + 1) / chunks, (int)t); });
   for( auto& x : th )
      x.join();
}

// host array without value initialisation: the parallel passes write every element, and
// zeroing hundreds of MB first would be a serial pass of its own
template <typename T>
struct HostArray
{
   std::unique_ptr<T[]> p;
   explicit HostArray(size_t n) : p(new T[n > 0 ? n : 1]) {}
   T* data() { return p.get(); }
   T& operator[](size_t i) { return p[i]; }
   const T& operator[](size_t i) const { return p[i]; }
};

// column-major copy (colptr, rowidx) of the stored rows, without atomics: thread t owns the
// columns [ncols*t/T, ncols*(t+1)/T) and scans every nonzero twice, counting, then (after a
// prefix sum) writing the rows of its own columns, so each column's rows come out ascending
static void buildColumns(int64_t nrows, int64_t ncols, int64_t nnz, const int64_t* prowptr, const int32_t* pcol,
   int64_t* colptr, int32_t* rowidx)
{
   int nt = (int)std::thread::hardware_concurrency();
   nt = nt < 1 ? 1 : (nt > 16 ? 16 : nt);
   if( nnz < (1 << 18) || ncols < nt )
      nt = 1;
   for( int64_t j = 0; j <= ncols; ++j )
      colptr[j] = 0;
   auto owned = [&](int t, int64_t& lo, int64_t& hi) {
      lo = ncols * t / nt;
      hi = ncols * (t + 1) / nt;
   };
   std::vector<std::thread> th;
   for( int t = 0; t < nt; ++t )
      th.emplace_back([&, t]() {
         int64_t lo, hi;
         owned(t, lo, hi);
         for( int64_t k = 0; k < nnz; ++k )
         {
            const int64_t c = pcol[k];
            if( c >= lo && c < hi )
               colptr[c + 1]++;
         }
      });
   for( auto& x : th )
      x.join();
   th.clear();
   for( int64_t j = 0; j < ncols; ++j )
      colptr[j + 1] += colptr[j];
   for( int t = 0; t < nt; ++t )
      th.emplace_back([&, t]() {
         int64_t lo, hi;
         owned(t, lo, hi);
         std::vector<int64_t> cur(colptr + lo, colptr + hi);
         for( int64_t r = 0; r < nrows; ++r )
            for( int64_t k = prowptr[r]; k < prowptr[r + 1]; ++k )
            {
               const int64_t c = pcol[k];
               if( c >= lo && c < hi )
                  rowidx[cur[(size_t)(c - lo)]++] = (int32_t)r;
            }
      });
   for( auto& x : th )
      x.join();
}

static inline int lenBin(int64_t L, int32_t longrow)
{
   for( int q = 0; q < 6; ++q )
      if( L <= kBinMaxLen[q] )
         return q;
   if( L > longrow )
      return SAP_NBINS - 1;
   return L <= kBinMaxLen[6] ? 6 : 7;
}

extern "C" {

void sap_params_default(sap_params* p)
{
   if( p == NULL )
      return;
   p->infinity = 1e20;
   p->epsilon = 1e-9;
   p->sumepsilon = 1e-6;
   p->feastol = 1e-6;
   p->boundstreps = 0.05;
   p->hugeval = 1e15;
   p->recompfac = 1e6;
   p->maxeasyactivitydelta = 1e6;
   p->maxrounds = -1;
   p->flags = 0;
}

void sap_tuning_default(sap_tuning* t)
{
   if( t == NULL )
      return;
   t->tilebins = 6;
   t->longrow = 1024;
   t->capscale = 50;
   t->rowbins = 0;
   t->tilewidth = 4;
}

int sap_context_set_tuning(sap_context* c, const sap_tuning* t)
{
   if( c == NULL || t == NULL || t->tilebins < 0 || t->tilebins > 6 || t->longrow < 64 || t->longrow > (1 << 24)
      || t->capscale < 10 || t->capscale > 1000 || t->rowbins < 0 || t->rowbins > 3
      || (t->tilewidth != 1 && t->tilewidth != 2 && t->tilewidth != 4) || (t->tilebins == 6 && t->tilewidth == 1) )
      return SAP_INVALIDDATA;
   c->tuning = *t;
   return SAP_OKAY;
}

int sap_context_get_tuning(const sap_context* c, sap_tuning* t)
{
   if( c == NULL || t == NULL )
      return SAP_INVALIDDATA;
   *t = c->tuning;
   return SAP_OKAY;
}

const char* sap_error_string(int rc)
{
   switch( rc )
   {
   case SAP_OKAY: return "okay";
   case SAP_ERROR: return "unspecified error (HIP runtime)";
   case SAP_NOMEMORY: return "insufficient device or host memory";
   case SAP_INVALIDCALL: return "invalid call";
   case SAP_INVALIDDATA: return "invalid input data";
   case SAP_NOTIMPLEMENTED: return "not implemented";
   default: return "unknown return code";
   }
}

int sap_device_count(int32_t* n)
{
   int c = 0;
   if( n == NULL )
      return SAP_INVALIDDATA;
   if( hipGetDeviceCount(&c) != hipSuccess )
      c = 0;
   *n = c;
   return SAP_OKAY;
}

int sap_context_create(int32_t device, sap_context** out)
{
   if( out == NULL )
      return SAP_INVALIDDATA;
   *out = NULL;
   int nd = 0;
   if( hipGetDeviceCount(&nd) != hipSuccess || nd == 0 )
   {
      fprintf(stderr, "[scip_amd] no HIP device available\n");
      return SAP_INVALIDCALL;
   }
   if( device < 0 || device >= nd )
      return SAP_INVALIDDATA;
   HIPCHK(hipSetDevice(device));
   sap_context* c = (sap_context*)calloc(1, sizeof(sap_context));
   if( c == NULL )
      return SAP_NOMEMORY;
   c->device = device;
   sap_tuning_default(&c->tuning);
   if( hipDeviceGetAttribute(&c->wallclock_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess
      || c->wallclock_khz <= 0 )
      c->wallclock_khz = 100000;   // 100 MHz reference clock
   HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
   HIPCHK(hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking));
   for( int i = 0; i < 4; ++i )
      HIPCHK(hipEventCreate(&c->ev[i]));
   HIPCHK(hipHostMalloc((void**)&c->hstatus, sizeof(Status), hipHostMallocMapped | hipHostMallocCoherent));
   HIPCHK(hipHostGetDevicePointer((void**)&c->hstatusd, c->hstatus, 0));
   // the device address of the no-allocation entry point's scratch array on this device (each
   // device holds its own copy; the symbol's name is not a device address in host code)
   HIPCHK(hipGetSymbolAddress((void**)&c->scratch, HIP_SYMBOL(g_noalloc_scratch)));
   int ncu = 0, per = 0;
   if( hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess
      && hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_sweep<false>, SAP_BLOCK, 0) == hipSuccess && ncu > 0
      && per > 0 )
      c->resident = ncu * per;
   *out = c;
   return SAP_OKAY;
}

int sap_context_free(sap_context* c)
{
   if( c == NULL )
      return SAP_OKAY;
   const int64_t live = __atomic_load_n(&c->nmatrices, __ATOMIC_RELAXED);
   if( live > 0 )
   {
      fprintf(stderr, "[scip_amd] sap_context_free: %ld matrices still live; free them first\n", (long)live);
      return SAP_INVALIDCALL;
   }
   (void)hipSetDevice(c->device);
   (void)hipStreamSynchronize(c->stream);
   for( int i = 0; i < 4; ++i )
      (void)hipEventDestroy(c->ev[i]);
   (void)hipHostFree(c->hstatus);
   (void)hipStreamDestroy(c->cstream);
   (void)hipStreamDestroy(c->stream);
   free(c);
   return SAP_OKAY;
}

static void matrixRelease(sap_matrix* m)
{
   (void)hipFree(m->rowptr);
   (void)hipFree(m->colidx);
   (void)hipFree(m->vals);
   (void)hipFree(m->sides);
   (void)hipFree(m->colflags);
   (void)hipFree(m->nzflags);
   (void)hipFree(m->rowflags);
   (void)hipFree(m->rowbin);
   (void)hipFree(m->binrows);
   (void)hipFree(m->colptr);
   (void)hipFree(m->rowidx);
   (void)hipFree(m->perm);
   (void)hipFree(m->rlist);
   (void)hipFree(m->tstart);
   (void)hipFree(m->tmask);
   (void)hipFree(m->trow0);
   (void)hipFree(m->tlev);
   (void)hipFree(m->tpad);
   (void)hipFree(m->rowseg);
   (void)hipFree(m->colseg);
   delete m->rowoff;
   delete m->coloff;
}

// upload of one instance (nseg == 1) or of nseg instances packed block-diagonally, instance s
// owning rows [rowoff[s], rowoff[s+1]) and columns [coloff[s], coloff[s+1])
static int matrixCreate(sap_context* ctx, int64_t nrows, int64_t ncols, const int64_t* rowptr, const int32_t* colidx,
   const double* vals, const double* lhs, const double* rhs, const uint8_t* integral, const double* glb,
   const double* gub, int32_t nseg, const int64_t* rowoff, const int64_t* coloff, sap_matrix** out)
{
   if( ctx == NULL || out == NULL || rowptr == NULL || lhs == NULL || rhs == NULL || integral == NULL || glb == NULL
      || gub == NULL || nrows < 0 || ncols < 0 || nrows > INT32_MAX || ncols > INT32_MAX || nseg < 1 )
      return SAP_INVALIDDATA;
   *out = NULL;
   const sap_tuning tune = ctx->tuning;
   const int64_t nnz = rowptr[nrows];
   if( rowptr[0] != 0 || nnz < 0 || (nnz > 0 && (colidx == NULL || vals == NULL)) )
      return SAP_INVALIDDATA;
   for( int64_t i = 0; i < nrows; ++i )
      if( rowptr[i + 1] < rowptr[i] || rowptr[i + 1] > nnz )
         return SAP_INVALIDDATA;
   // segment of every row and column (instances must not share columns)
   std::vector<int32_t> oseg, cseg;
   if( nseg > 1 )
   {
      if( rowoff == NULL || coloff == NULL || rowoff[0] != 0 || coloff[0] != 0 || rowoff[nseg] != nrows
         || coloff[nseg] != ncols )
         return SAP_INVALIDDATA;
      for( int32_t q = 0; q < nseg; ++q )
         if( rowoff[q + 1] < rowoff[q] || coloff[q + 1] < coloff[q] )
            return SAP_INVALIDDATA;
      oseg.resize((size_t)nrows);
      cseg.resize((size_t)ncols);
      std::atomic<int> bad(0);
      parallelFor(nseg, 1, [&](int64_t q0, int64_t q1, int) {
         for( int64_t q = q0; q < q1; ++q )
         {
            for( int64_t i = rowoff[q]; i < rowoff[q + 1]; ++i )
            {
               oseg[(size_t)i] = (int32_t)q;
               for( int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k )
                  if( colidx[k] < coloff[q] || colidx[k] >= coloff[q + 1] )
                     bad = 1;
            }
            for( int64_t j = coloff[q]; j < coloff[q + 1]; ++j )
               cseg[(size_t)j] = (int32_t)q;
         }
      });
      if( bad )
         return SAP_INVALIDDATA;
   }
   // host-side validation and derived arrays; rows are stored grouped by length bin
   // (stable), so each bin is a contiguous range and sub-wavefront groups load contiguously
   std::vector<uint8_t> colflags((size_t)ncols);
   for( int64_t j = 0; j < ncols; ++j )
   {
      if( glb[j] != glb[j] || gub[j] != gub[j] )
         return SAP_INVALIDDATA;
      bool isint = integral[j] != 0;
      bool isbin = isint && glb[j] >= 0.0 && gub[j] <= 1.0;
      colflags[(size_t)j] = (uint8_t)((isint ? 1 : 0) | (isbin ? 2 : 0));
   }
   std::vector<uint8_t> obin((size_t)nrows);
   std::atomic<int> bad(0), forced(0);
   std::vector<std::array<int64_t, SAP_NBINS>> tcount(16);
   for( auto& c : tcount )
      c.fill(0);
   parallelFor(nrows, 1 << 16, [&](int64_t i0, int64_t i1, int t) {
      std::vector<int32_t> tmp;
      for( int64_t i = i0; i < i1; ++i )
      {
         const int64_t st0 = rowptr[i], e = rowptr[i + 1];
         if( lhs[i] != lhs[i] || rhs[i] != rhs[i] )
            bad = 1;
         bool ascending = true;
         for( int64_t k = st0; k < e; ++k )
         {
            const int32_t c = colidx[k];
            if( c < 0 || c >= ncols || vals[k] == 0.0 || vals[k] != vals[k] )
               bad = 1;
            if( k > st0 && c <= colidx[k - 1] )
               ascending = false;
         }
         // duplicate columns within a row are not allowed (SCIP merges them)
         if( !ascending )
         {
            tmp.assign(colidx + st0, colidx + e);
            std::sort(tmp.begin(), tmp.end());
            for( size_t q = 1; q < tmp.size(); ++q )
               if( tmp[q] == tmp[q - 1] )
                  bad = 1;
         }
         const int64_t L = e - st0;
         if( L == 1 )
            forced = 1;
         const int b = lenBin(L, tune.longrow);
         obin[(size_t)i] = (uint8_t)b;
         tcount[(size_t)t][(size_t)b]++;
      }
   });
   if( bad )
      return SAP_INVALIDDATA;
   const int32_t hasforced = forced.load();
   int64_t bincount[SAP_NBINS] = {};
   for( auto& c : tcount )
      for( int b = 0; b < SAP_NBINS; ++b )
         bincount[b] += c[(size_t)b];
   int32_t binstart[SAP_NBINS + 1];
   binstart[0] = 0;
   for( int b = 0; b < SAP_NBINS; ++b )
      binstart[b + 1] = binstart[b] + (int32_t)bincount[b];
   std::vector<int32_t> perm((size_t)nrows);
   {
      int32_t fill[SAP_NBINS];
      for( int b = 0; b < SAP_NBINS; ++b )
         fill[b] = binstart[b];
      for( int64_t i = 0; i < nrows; ++i )
         perm[(size_t)fill[obin[(size_t)i]]++] = (int32_t)i;
   }
   // the tiled bins (1..64 nonzeros per row, 1..128 with tile width 2) are re-ordered tile by
   // tile: first-fit decreasing on the row lengths, i.e. each tile is opened by the longest row
   // left and filled with the longest rows that still fit.  A tile's cost is nearly
   // independent of how many of its 64 lanes are used, so fewer, fuller tiles mean less work;
   // on log-normal row lengths this packs ~100 % of the lanes against ~82 % for consecutive
   // rows in bin order.  Rows of one length keep their bin order, so the layout is
   // deterministic.
   const int32_t tfirst = binstart[1 + std::min(tune.rowbins, tune.tilebins)];
   const int32_t tend = std::max(tfirst, binstart[1 + tune.tilebins]);
   // With tile width W > 1 a row takes its length rounded up to a multiple of W positions of
   // a tile of 64*W positions.
   const int tw = tune.tilewidth;
   std::vector<int32_t> tilerows;   // rows per tile, in tile order
   if( tend > tfirst )
   {
      const int cap = SAP_WAVE * tw;
      std::vector<std::vector<int32_t>> bysize((size_t)cap + 2);
      for( int32_t r = tfirst; r < tend; ++r )
      {
         const int32_t i = perm[(size_t)r];
         const int64_t L = rowptr[i + 1] - rowptr[i];
         bysize[(size_t)((L + tw - 1) / tw * tw)].push_back(i);
      }
      std::vector<size_t> head((size_t)cap + 2, 0);
      int32_t w = tfirst;
      for( int big = cap; big >= 1; --big )
         while( head[(size_t)big] < bysize[(size_t)big].size() )
         {
            perm[(size_t)w++] = bysize[(size_t)big][head[(size_t)big]++];
            int space = cap - big, nr = 1;
            for( int l = std::min(space, big); l >= 1 && space > 0; --l )
               while( space >= l && head[(size_t)l] < bysize[(size_t)l].size() )
               {
                  perm[(size_t)w++] = bysize[(size_t)l][head[(size_t)l]++];
                  space -= l;
                  ++nr;
               }
            tilerows.push_back(nr);
         }
   }
   std::vector<int64_t> prowptr((size_t)nrows + 1);
   prowptr[0] = 0;
   for( int64_t r = 0; r < nrows; ++r )
   {
      const int64_t i = perm[(size_t)r];
      prowptr[(size_t)r + 1] = prowptr[(size_t)r] + (rowptr[i + 1] - rowptr[i]);
   }
   HostArray<int32_t> pcol((size_t)nnz);
   HostArray<double> pval((size_t)nnz);
   HostArray<uint8_t> pflag((size_t)nnz);
   std::vector<double2> sides((size_t)nrows);
   std::vector<uint8_t> rowflags((size_t)nrows), rowbin((size_t)nrows);
   std::vector<int64_t> colptr((size_t)ncols + 1);
   HostArray<int32_t> rowidx((size_t)nnz);
   parallelFor(nrows, 1 << 16, [&](int64_t r0, int64_t r1, int) {
      for( int64_t r = r0; r < r1; ++r )
      {
         const int64_t i = perm[(size_t)r];
         const int64_t st0 = rowptr[i], L = rowptr[i + 1] - st0, o = prowptr[(size_t)r];
         bool allbin = true;
         for( int64_t k = 0; k < L; ++k )
         {
            const int32_t c = colidx[st0 + k];
            pcol[(size_t)(o + k)] = c;
            pval[(size_t)(o + k)] = vals[st0 + k];
            pflag[(size_t)(o + k)] = colflags[(size_t)c];
            if( !(colflags[(size_t)c] & 2) )
               allbin = false;
         }
         sides[(size_t)r] = make_double2(lhs[i], rhs[i]);
         // bit 1: eligible for ranged-row propagation (two sides of magnitude below SCIP's
         // default infinity, at least 3 nonzeros)
         const bool ranged = L >= 3 && fabs(lhs[i]) < 1e20 && fabs(rhs[i]) < 1e20;
         rowflags[(size_t)r] = (uint8_t)((allbin ? 1 : 0) | (ranged ? 2 : 0));
         rowbin[(size_t)r] = obin[(size_t)i];
      }
   });
   buildColumns(nrows, ncols, nnz, prowptr.data(), pcol.data(), colptr.data(), rowidx.data());

   // tiles over the packed rows of the tiled bins (consecutive rows, at most 64 lanes)
   std::vector<int64_t> tstart;
   std::vector<unsigned long long> tmask;
   std::vector<int32_t> trow0;
   std::vector<uint8_t> tlev;
   std::vector<unsigned long long> tpad;   // two masks per tile (wide tiles)
   {
      auto levels = [](int l) {
         int v = 0;
         while( (1 << v) < l )
            ++v;
         return (uint8_t)v;
      };
      int32_t r = tfirst;
      for( const int32_t nr : tilerows )
      {
         tstart.push_back(prowptr[(size_t)r]);
         trow0.push_back(r);
         unsigned long long mask = 0, pad0 = 0, pad1 = 0;
         int lanes = 0, maxl = 0;
         for( int32_t q = 0; q < nr; ++q, ++r )
         {
            const int L = (int)(prowptr[(size_t)r + 1] - prowptr[(size_t)r]);
            const int pad = (tw - L % tw) % tw;
            mask |= 1ULL << lanes;
            if( pad & 1 )
               pad0 |= 1ULL << lanes;
            if( pad & 2 )
               pad1 |= 1ULL << lanes;
            lanes += (L + tw - 1) / tw;
            maxl = L > maxl ? L : maxl;
         }
         tmask.push_back(mask);
         tpad.push_back(pad0);
         tpad.push_back(pad1);
         tlev.push_back(levels(maxl));
      }
      tstart.push_back(prowptr[(size_t)tend]);
   }
   // rows of each bin by device position (bin b: binrows[binstart[b] .. binstart[b+1])); the
   // identity outside the tiled bins, whose rows are in tile order
   std::vector<int32_t> binrows((size_t)nrows);
   {
      int32_t fill[SAP_NBINS];
      for( int b = 0; b < SAP_NBINS; ++b )
         fill[b] = binstart[b];
      for( int64_t r = 0; r < nrows; ++r )
         binrows[(size_t)fill[rowbin[(size_t)r]]++] = (int32_t)r;
   }

   HIPCHK(hipSetDevice(ctx->device));
   sap_matrix* m = (sap_matrix*)calloc(1, sizeof(sap_matrix));
   if( m == NULL )
      return SAP_NOMEMORY;
   m->ctx = ctx;
   m->tuning = tune;
   m->ntiles = (int32_t)tmask.size();
   m->nrows = nrows;
   m->ncols = ncols;
   m->nnz = nnz;
   std::vector<int32_t> rlist;
   for( int64_t r = 0; r < nrows; ++r )
      if( rowflags[(size_t)r] & 2 )
         rlist.push_back((int32_t)r);
   m->nranged = (int32_t)rlist.size();
   // ranged-row propagation submits forced candidates, so matrices with eligible rows keep them
   m->hasforced = hasforced || m->nranged > 0;
   for( int b = 0; b <= SAP_NBINS; ++b )
      m->binstart[b] = binstart[b];
   int rc = SAP_OKAY;
#define MALLOC_OR_FAIL(ptr, cnt)                           \
   do                                                      \
   {                                                       \
      rc = devAlloc(&(ptr), (size_t)(cnt), &m->device_bytes); \
      if( rc != SAP_OKAY )                                 \
      {                                                    \
         matrixRelease(m);                                 \
         free(m);                                          \
         return rc;                                        \
      }                                                    \
   } while( 0 )
   MALLOC_OR_FAIL(m->rowptr, nrows + 1);
   MALLOC_OR_FAIL(m->colidx, nnz);
   MALLOC_OR_FAIL(m->vals, nnz);
   MALLOC_OR_FAIL(m->sides, nrows);
   MALLOC_OR_FAIL(m->colflags, ncols);
   MALLOC_OR_FAIL(m->nzflags, nnz);
   MALLOC_OR_FAIL(m->rowflags, nrows);
   MALLOC_OR_FAIL(m->rowbin, nrows);
   MALLOC_OR_FAIL(m->binrows, nrows);
   MALLOC_OR_FAIL(m->colptr, ncols + 1);
   MALLOC_OR_FAIL(m->rowidx, nnz);
   MALLOC_OR_FAIL(m->perm, nrows);
   MALLOC_OR_FAIL(m->rlist, rlist.size() + 1);
   MALLOC_OR_FAIL(m->tstart, tstart.size());
   MALLOC_OR_FAIL(m->tmask, tmask.size());
   MALLOC_OR_FAIL(m->trow0, trow0.size());
   MALLOC_OR_FAIL(m->tlev, tlev.size());
   MALLOC_OR_FAIL(m->tpad, tpad.size());
   m->nseg = nseg;
   if( nseg > 1 )
   {
      MALLOC_OR_FAIL(m->rowseg, nrows);
      MALLOC_OR_FAIL(m->colseg, ncols);
      m->rowoff = new std::vector<int64_t>(rowoff, rowoff + nseg + 1);
      m->coloff = new std::vector<int64_t>(coloff, coloff + nseg + 1);
   }
#undef MALLOC_OR_FAIL
   hipStream_t s = ctx->stream;
   HIPCHK(hipMemcpyAsync(m->rowptr, prowptr.data(), sizeof(int64_t) * (size_t)(nrows + 1), hipMemcpyHostToDevice, s));
   if( nnz > 0 )
   {
      HIPCHK(hipMemcpyAsync(m->colidx, pcol.data(), sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(m->vals, pval.data(), sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(m->nzflags, pflag.data(), (size_t)nnz, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(m->rowidx, rowidx.data(), sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice, s));
   }
   if( nrows > 0 )
   {
      HIPCHK(hipMemcpyAsync(m->sides, sides.data(), sizeof(double2) * (size_t)nrows, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(m->rowflags, rowflags.data(), (size_t)nrows, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(m->rowbin, rowbin.data(), (size_t)nrows, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(m->binrows, binrows.data(), sizeof(int32_t) * (size_t)nrows, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(m->perm, perm.data(), sizeof(int32_t) * (size_t)nrows, hipMemcpyHostToDevice, s));
      if( !rlist.empty() )
         HIPCHK(hipMemcpyAsync(m->rlist, rlist.data(), sizeof(int32_t) * rlist.size(), hipMemcpyHostToDevice, s));
   }
   if( ncols > 0 )
      HIPCHK(hipMemcpyAsync(m->colflags, colflags.data(), (size_t)ncols, hipMemcpyHostToDevice, s));
   HIPCHK(hipMemcpyAsync(m->colptr, colptr.data(), sizeof(int64_t) * (size_t)(ncols + 1), hipMemcpyHostToDevice, s));
   HIPCHK(hipMemcpyAsync(m->tstart, tstart.data(), sizeof(int64_t) * tstart.size(), hipMemcpyHostToDevice, s));
   if( !tmask.empty() )
   {
      HIPCHK(hipMemcpyAsync(m->tmask, tmask.data(), sizeof(unsigned long long) * tmask.size(), hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(m->trow0, trow0.data(), sizeof(int32_t) * trow0.size(), hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(m->tlev, tlev.data(), tlev.size(), hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(m->tpad, tpad.data(), sizeof(unsigned long long) * tpad.size(), hipMemcpyHostToDevice,
         s));
   }
   if( nseg > 1 )
   {
      std::vector<int32_t> prseg((size_t)nrows);
      for( int64_t r = 0; r < nrows; ++r )
         prseg[(size_t)r] = oseg[(size_t)perm[(size_t)r]];
      if( nrows > 0 )
         HIPCHK(hipMemcpyAsync(m->rowseg, prseg.data(), sizeof(int32_t) * (size_t)nrows, hipMemcpyHostToDevice, s));
      if( ncols > 0 )
         HIPCHK(hipMemcpyAsync(m->colseg, cseg.data(), sizeof(int32_t) * (size_t)ncols, hipMemcpyHostToDevice, s));
      HIPCHK(hipStreamSynchronize(s));
   }
   HIPCHK(hipStreamSynchronize(s));
   __atomic_fetch_add(&ctx->nmatrices, 1, __ATOMIC_RELAXED);
   *out = m;
   return SAP_OKAY;
}

int sap_matrix_create(sap_context* ctx, int64_t nrows, int64_t ncols, const int64_t* rowptr, const int32_t* colidx,
   const double* vals, const double* lhs, const double* rhs, const uint8_t* integral, const double* glb,
   const double* gub, sap_matrix** out)
{
   return matrixCreate(ctx, nrows, ncols, rowptr, colidx, vals, lhs, rhs, integral, glb, gub, 1, NULL, NULL, out);
}

static int checkParams(const sap_params* p);

int sap_autotune(sap_context* ctx, int64_t nrows, int64_t ncols, const int64_t* rowptr, const int32_t* colidx,
   const double* vals, const double* lhs, const double* rhs, const uint8_t* integral, const double* lb,
   const double* ub, const sap_params* params, int64_t samplerows, uint64_t seed, sap_tuning* best, double* best_ms,
   double* default_ms)
{
   if( ctx == NULL || rowptr == NULL || lb == NULL || ub == NULL || best == NULL || samplerows < 1 || nrows < 0
      || lhs == NULL || rhs == NULL || integral == NULL || ncols < 0
      || (rowptr[nrows] > 0 && (colidx == NULL || vals == NULL)) )
      return SAP_INVALIDDATA;
   RCCHK(checkParams(params));
   // row sample (every row when there are few): splitmix64 scores, smallest samplerows kept
   std::vector<int64_t> rows;
   if( nrows <= samplerows )
   {
      rows.resize((size_t)nrows);
      for( int64_t i = 0; i < nrows; ++i )
         rows[(size_t)i] = i;
   }
   else
   {
      std::vector<std::pair<uint64_t, int64_t>> sc((size_t)nrows);
      for( int64_t i = 0; i < nrows; ++i )
      {
         uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(i + 1);
         z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
         z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
         sc[(size_t)i] = std::make_pair(z ^ (z >> 31), i);
      }
      std::nth_element(sc.begin(), sc.begin() + samplerows, sc.end());
      rows.resize((size_t)samplerows);
      for( int64_t q = 0; q < samplerows; ++q )
         rows[(size_t)q] = sc[(size_t)q].second;
      std::sort(rows.begin(), rows.end());
   }
   const int64_t sn = (int64_t)rows.size();
   std::vector<int64_t> srp((size_t)sn + 1, 0);
   std::vector<int32_t> sci;
   std::vector<double> sva, slh((size_t)sn), srh((size_t)sn);
   for( int64_t q = 0; q < sn; ++q )
   {
      const int64_t i = rows[(size_t)q];
      for( int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k )
      {
         sci.push_back(colidx[k]);
         sva.push_back(vals[k]);
      }
      srp[(size_t)q + 1] = (int64_t)sci.size();
      slh[(size_t)q] = lhs[i];
      srh[(size_t)q] = rhs[i];
   }
   // only the columns the sample uses, renumbered in their order (rows stay ascending): the
   // sample propagates the same on them alone, and a system with many more columns than the
   // sample touches (packed instances) does not copy whole bound vectors per candidate
   std::vector<int32_t> used(sci);
   std::sort(used.begin(), used.end());
   used.erase(std::unique(used.begin(), used.end()), used.end());
   if( !used.empty() && (used.front() < 0 || (int64_t)used.back() >= ncols) )
      return SAP_INVALIDDATA;
   for( int32_t& c : sci )
      c = (int32_t)(std::lower_bound(used.begin(), used.end(), c) - used.begin());
   const int64_t snc = (int64_t)used.size();
   std::vector<uint8_t> sint((size_t)snc);
   std::vector<double> slb((size_t)snc), sub((size_t)snc);
   for( int64_t u = 0; u < snc; ++u )
   {
      const int32_t j = used[(size_t)u];
      sint[(size_t)u] = integral[j];
      slb[(size_t)u] = lb[j];
      sub[(size_t)u] = ub[j];
   }
   const sap_tuning saved = ctx->tuning;
   sap_tuning def;
   sap_tuning_default(&def);
   static const int32_t rbs[] = { 0, 0, 0, 0, 2, 3 };
   static const int32_t tbs[] = { 3, 4, 5, 6, 6, 6 };
   static const int32_t tws[] = { 2, 4 };
   static const int32_t lrs[] = { 1024, 4096, 16384 };
   static const int32_t cps[] = { 25, 50, 100 };
   std::vector<double> wl((size_t)snc), wu((size_t)snc);
   double bestt = 1e300, deft = -1.0;
   sap_tuning bt = def;
   int rc = SAP_OKAY;
   for( int a = 0; a < 6 && rc == SAP_OKAY; ++a )
      for( int w = 0; w < 2 && rc == SAP_OKAY; ++w )
         for( int l = 0; l < 3 && rc == SAP_OKAY; ++l )
            for( int c = 0; c < 3 && rc == SAP_OKAY; ++c )
            {
               sap_tuning t = def;
               t.rowbins = rbs[a];
               t.tilebins = tbs[a];
               t.tilewidth = tws[w];
               t.longrow = lrs[l];
               t.capscale = cps[c];
               ctx->tuning = t;
               sap_matrix* m = NULL;
               rc = matrixCreate(ctx, sn, snc, srp.data(), sci.data(), sva.data(), slh.data(), srh.data(),
                  sint.data(), slb.data(), sub.data(), 1, NULL, NULL, &m);
               if( rc != SAP_OKAY )
                  break;
               // candidates differ only in the sweep's layout: they are ranked by the sweep time,
               // measured with HIP events around every sweep launch of the host-driven round loop
               // (no round graphs are captured for a matrix that lives for three calls)
               sap_params hp = *params;
               hp.flags |= SAP_FLAG_HOSTLOOP;
               double tmin = 1e300;
               for( int rep = 0; rep < 3 && rc == SAP_OKAY; ++rep )
               {
                  wl = slb;
                  wu = sub;
                  sap_stats st;
                  rc = sap_propagate(m, &hp, wl.data(), wu.data(), &st);
                  if( rep > 0 )   // the first call includes one-time set-up
                     tmin = std::min(tmin, st.sweep_ms);
               }
               sap_matrix_free(m);
               if( t.tilebins == def.tilebins && t.longrow == def.longrow && t.capscale == def.capscale
                  && t.rowbins == def.rowbins && t.tilewidth == def.tilewidth )
                  deft = tmin;
               if( tmin < bestt )
               {
                  bestt = tmin;
                  bt = t;
               }
            }
   ctx->tuning = saved;
   RCCHK(rc);
   *best = bt;
   if( best_ms )
      *best_ms = bestt;
   if( default_ms )
      *default_ms = deft;
   return SAP_OKAY;
}

int sap_matrix_create_packed(sap_context* ctx, int32_t ninst, const int64_t* rowoff, const int64_t* coloff,
   int64_t nrows, int64_t ncols, const int64_t* rowptr, const int32_t* colidx, const double* vals, const double* lhs,
   const double* rhs, const uint8_t* integral, const double* glb, const double* gub, sap_matrix** out)
{
   if( ninst < 1 )
      return SAP_INVALIDDATA;
   return matrixCreate(ctx, nrows, ncols, rowptr, colidx, vals, lhs, rhs, integral, glb, gub, ninst, rowoff, coloff,
      out);
}

int sap_matrix_free(sap_matrix* m)
{
   if( m == NULL )
      return SAP_OKAY;
   if( m->nbounds > 0 )
   {
      fprintf(stderr, "[scip_amd] sap_matrix_free: %ld bound sets still live; free them first\n", (long)m->nbounds);
      return SAP_INVALIDCALL;
   }
   (void)hipSetDevice(m->ctx->device);
   (void)hipStreamSynchronize(m->ctx->stream);
   if( m->single )
      sap_bounds_free(m->single);
   if( m->rsingle )
      sap_bounds_free(m->rsingle);
   __atomic_fetch_sub(&m->ctx->nmatrices, 1, __ATOMIC_RELAXED);
   matrixRelease(m);
   free(m);
   return SAP_OKAY;
}

int sap_matrix_set_context(sap_matrix* m, sap_context* ctx)
{
   if( m == NULL || ctx == NULL || ctx->device != m->ctx->device )
      return SAP_INVALIDDATA;
   if( ctx == m->ctx )
      return SAP_OKAY;
   HIPCHK(hipSetDevice(ctx->device));
   // the matrix's work on its old stream has ended (engine calls return with their work
   // done; a slots step must have been waited for)
   HIPCHK(hipStreamSynchronize(m->ctx->stream));
   __atomic_fetch_sub(&m->ctx->nmatrices, 1, __ATOMIC_RELAXED);
   __atomic_fetch_add(&ctx->nmatrices, 1, __ATOMIC_RELAXED);
   m->ctx = ctx;
   return SAP_OKAY;
}

int sap_matrix_info(const sap_matrix* m, int64_t* nrows, int64_t* ncols, int64_t* nnz, int64_t* device_bytes)
{
   if( m == NULL )
      return SAP_INVALIDDATA;
   if( nrows )
      *nrows = m->nrows;
   if( ncols )
      *ncols = m->ncols;
   if( nnz )
      *nnz = m->nnz;
   if( device_bytes )
      *device_bytes = m->device_bytes;
   return SAP_OKAY;
}

static void boundsRelease(sap_bounds* b)
{
   Vec& V = b->V;
   (void)hipFree(V.bnd);
   (void)hipFree(V.clb);
   (void)hipFree(V.cub);
   (void)hipFree(V.flb);
   (void)hipFree(V.fub);
   (void)hipFree(V.coltag);
   (void)hipFree(V.chgtag);
   (void)hipFree(V.rowmark);
   (void)hipFree(V.redundant);
   (void)hipFree(V.vcutoff);
   (void)hipFree(V.vrounds);
   (void)hipFree(V.vstart);
   (void)hipFree(V.vpend);
   (void)hipFree(V.vlogn);
   (void)hipFree(V.log);
   (void)hipFree(V.logbits);
   (void)hipFree(b->ovptr);
   (void)hipFree(b->ovcol);
   (void)hipFree(b->ovlb);
   (void)hipFree(b->ovub);
   (void)hipFree(b->base2);
   (void)hipFree(b->slotidx);
   (void)hipFree(b->slotidx2);
   (void)hipFree(b->outoff);
   (void)hipFree(b->outcol);
   (void)hipFree(b->outbnd);
   (void)hipFree(b->pcol);
   (void)hipFree(b->pcount);
   (void)hipFree(b->outbnd2);
   delete b->active;
   b->active = NULL;
   delete b->srcb;
   delete b->srcv;
   delete b->srcgen;
   b->srcb = NULL;
   b->srcv = NULL;
   b->srcgen = NULL;
   (void)hipFree(V.ovl);
   (void)hipFree(V.ovln);
   (void)hipFree(V.clog);
   (void)hipFree(V.rflag);
   (void)hipFree(V.rkey);
   (void)hipFree(V.rrow);
   (void)hipFree(V.hist);
   (void)hipFree(V.rgflag);
   if( b->hstat )
      (void)hipHostFree(b->hstat);
   if( b->hpend )
      (void)hipHostFree(b->hpend);
   b->hstat = NULL;
   b->hpend = NULL;
   (void)hipFree(b->ident);
   b->ident = NULL;
   if( b->hup )
      (void)hipHostFree(b->hup);
   if( b->hdn )
      (void)hipHostFree(b->hdn);
   b->hup = b->hupd = NULL;
   b->hdn = b->hdnd = NULL;
   b->hupcap = b->hupused = b->hdncap = 0;
   for( int e = 0; e < 2; ++e )
      if( b->evs[e] != NULL )
         (void)hipEventDestroy(b->evs[e]);
   for( int p = 0; p < 2; ++p )
      for( int q = 0; q < SAP_NBINS; ++q )
         (void)hipFree(V.wl[p][q]);
   (void)hipFree(V.touched);
   (void)hipFree(b->dstatus);
   (void)hipFree(b->stage);
   for( int g = 0; g < SAP_NGRAPHS; ++g )
   {
      if( b->graph[g].valid )
         (void)hipGraphExecDestroy(b->graph[g].exec);
      b->graph[g].valid = 0;
      for( int e = 0; e < 2 * SAP_GRAPH_LONG; ++e )
         if( b->graph[g].ev[e] != NULL )
            (void)hipEventDestroy(b->graph[g].ev[e]);
   }
}

// dense marks are scanned every round: worthwhile while the scan (4 B per row and per column
// and vector) stays small against the row sweep itself
#define SAP_DENSE_LIMIT ((size_t)1 << 24)

int sap_bounds_create(sap_matrix* m, int32_t nvec, sap_bounds** out)
{
   return sap_bounds_create_mode(m, nvec, SAP_MODE_AUTO, out);
}

int sap_bounds_create_mode(sap_matrix* m, int32_t nvec, int32_t mode, sap_bounds** out)
{
   if( m == NULL || out == NULL || nvec < 1 || nvec >= (1 << 30) || mode < SAP_MODE_AUTO || mode > SAP_MODE_SPARSE )
      return SAP_INVALIDDATA;
   *out = NULL;
   if( (int64_t)nvec * m->nrows > (int64_t)UINT32_MAX || (int64_t)nvec * m->ncols > (int64_t)UINT32_MAX )
      return SAP_INVALIDDATA;
   HIPCHK(hipSetDevice(m->ctx->device));
   sap_bounds* b = (sap_bounds*)calloc(1, sizeof(sap_bounds));
   if( b == NULL )
      return SAP_NOMEMORY;
   b->mat = m;
   b->nvec = nvec;
   Vec& V = b->V;
   V.nvec = nvec;
   V.nseg = m->nseg;
   V.single = (nvec == 1 && m->nseg == 1);
   V.hasforced = m->hasforced;
   const size_t NC = (size_t)nvec * (size_t)m->ncols;
   const size_t NR = (size_t)nvec * (size_t)m->nrows;
   if( mode == SAP_MODE_AUTO )
      V.dense = (nvec == 1 || NC + NR <= SAP_DENSE_LIMIT) ? 1 : 0;
   else
      V.dense = (mode == SAP_MODE_DENSE) ? 1 : 0;
   int rc = SAP_OKAY;
#define BALLOC(ptr, cnt)                                       \
   do                                                          \
   {                                                           \
      rc = devAlloc(&(ptr), (size_t)(cnt), &b->device_bytes);  \
      if( rc != SAP_OKAY )                                     \
      {                                                        \
         boundsRelease(b);                                     \
         free(b);                                              \
         return rc;                                            \
      }                                                        \
   } while( 0 )
   BALLOC(V.bnd, NC);
   BALLOC(V.clb, NC);
   BALLOC(V.cub, NC);
   if( m->hasforced )
   {
      BALLOC(V.flb, NC);
      BALLOC(V.fub, NC);
   }
   BALLOC(V.coltag, NC);
   BALLOC(V.chgtag, NC);
   BALLOC(V.rowmark, NR);
   BALLOC(V.redundant, NR);
   BALLOC(V.vcutoff, (size_t)nvec * (size_t)m->nseg);
   BALLOC(V.vrounds, (size_t)nvec * (size_t)m->nseg);
   BALLOC(V.vstart, nvec);
   BALLOC(V.vpend, nvec);
   BALLOC(V.vlogn, nvec);
   BALLOC(b->ovptr, (size_t)nvec + 1);
   BALLOC(b->base2, (size_t)(m->ncols > 0 ? m->ncols : 1));
   BALLOC(b->slotidx, nvec);
   BALLOC(b->slotidx2, nvec);
   BALLOC(b->outoff, (size_t)nvec + 1);
   BALLOC(b->ident, nvec);
   if( !V.dense )
   {
      for( int p = 0; p < 2; ++p )
         for( int q = 0; q < SAP_NBINS; ++q )
            BALLOC(V.wl[p][q], (size_t)nvec * (size_t)(m->binstart[q + 1] - m->binstart[q]));
      BALLOC(V.touched, NC);
   }
   BALLOC(b->dstatus, 1);
   BALLOC(b->stage, 2 * (size_t)m->ncols);
#undef BALLOC
   for( int e = 0; e < 2; ++e )
      if( hipEventCreate(&b->evs[e]) != hipSuccess )
      {
         boundsRelease(b);
         free(b);
         return SAP_ERROR;
      }
   V.st = b->dstatus;
   hipStream_t s = m->ctx->stream;
   k_fill_i64<<<gridFor(NC), SAP_BLOCK, 0, s>>>(V.clb, KEY_NEG_INF, NC);
   k_fill_i64<<<gridFor(NC), SAP_BLOCK, 0, s>>>(V.cub, KEY_POS_INF, NC);
   if( m->hasforced )
   {
      k_fill_i64<<<gridFor(NC), SAP_BLOCK, 0, s>>>(V.flb, KEY_NEG_INF, NC);
      k_fill_i64<<<gridFor(NC), SAP_BLOCK, 0, s>>>(V.fub, KEY_POS_INF, NC);
   }
   k_fill_i32<<<gridFor(NC), SAP_BLOCK, 0, s>>>(V.coltag, 0, NC);
   k_fill_i32<<<gridFor(NC), SAP_BLOCK, 0, s>>>(V.chgtag, 0, NC);
   k_fill_i32<<<gridFor(NR), SAP_BLOCK, 0, s>>>(V.rowmark, 0, NR);
   k_fill_i32<<<gridFor((size_t)nvec), SAP_BLOCK, 0, s>>>(V.vpend, 0, (size_t)nvec);
   k_fill_i32<<<gridFor((size_t)nvec), SAP_BLOCK, 0, s>>>(V.vstart, 0, (size_t)nvec);
   k_fill_i32<<<gridFor((size_t)nvec), SAP_BLOCK, 0, s>>>(V.vlogn, 0, (size_t)nvec);
   k_iota_i32<<<gridFor((size_t)nvec), SAP_BLOCK, 0, s>>>(b->ident, (size_t)nvec);
   HIPCHK(hipGetLastError());
   HIPCHK(hipMemsetAsync(b->dstatus, 0, sizeof(Status), s));
   HIPCHK(hipStreamSynchronize(s));
   m->nbounds++;
   *out = b;
   return SAP_OKAY;
}

int sap_bounds_free(sap_bounds* b)
{
   if( b == NULL )
      return SAP_OKAY;
   (void)hipSetDevice(b->mat->ctx->device);
   (void)hipStreamSynchronize(b->mat->ctx->stream);
   if( !b->internal )
      b->mat->nbounds--;
   boundsRelease(b);
   free(b);
   return SAP_OKAY;
}

// vector v (all for v < 0) of b was written outside the change log
static void touchedVectors(sap_bounds* b, int32_t v)
{
   b->gen++;
   if( b->srcb == NULL )
      return;
   const int32_t v0 = v < 0 ? 0 : v, v1 = v < 0 ? b->nvec : v + 1;
   for( int32_t q = v0; q < v1; ++q )
      (*b->srcb)[(size_t)q] = NULL;
}

int sap_bounds_upload(sap_bounds* b, int32_t v, const double* lb, const double* ub)
{
   if( b == NULL || lb == NULL || ub == NULL || v >= b->nvec )
      return SAP_INVALIDDATA;
   sap_matrix* m = b->mat;
   hipStream_t s = m->ctx->stream;
   HIPCHK(hipSetDevice(m->ctx->device));
   const int64_t nc = m->ncols;
   for( int64_t j = 0; j < nc; ++j )
      if( lb[j] != lb[j] || ub[j] != ub[j] )
         return SAP_INVALIDDATA;
   HIPCHK(hipMemcpyAsync(b->stage, lb, sizeof(double) * (size_t)nc, hipMemcpyHostToDevice, s));
   HIPCHK(hipMemcpyAsync(b->stage + nc, ub, sizeof(double) * (size_t)nc, hipMemcpyHostToDevice, s));
   int32_t v0 = v < 0 ? 0 : v, v1 = v < 0 ? b->nvec : v + 1;
   for( int32_t q = v0; q < v1; ++q )
      k_interleave<<<gridFor((size_t)nc), SAP_BLOCK, 0, s>>>(b->V.bnd + (size_t)q * (size_t)nc, b->stage,
         b->stage + nc, nc);
   HIPCHK(hipGetLastError());
   touchedVectors(b, v);
   HIPCHK(hipStreamSynchronize(s));
   return SAP_OKAY;
}

int sap_bounds_download(sap_bounds* b, int32_t v, double* lb, double* ub)
{
   if( b == NULL || lb == NULL || ub == NULL || v < 0 || v >= b->nvec )
      return SAP_INVALIDDATA;
   sap_matrix* m = b->mat;
   hipStream_t s = m->ctx->stream;
   HIPCHK(hipSetDevice(m->ctx->device));
   const int64_t nc = m->ncols;
   k_deinterleave<<<gridFor((size_t)nc), SAP_BLOCK, 0, s>>>(b->V.bnd + (size_t)v * (size_t)nc, b->stage,
      b->stage + nc, nc);
   HIPCHK(hipGetLastError());
   HIPCHK(hipMemcpyAsync(lb, b->stage, sizeof(double) * (size_t)nc, hipMemcpyDeviceToHost, s));
   HIPCHK(hipMemcpyAsync(ub, b->stage + nc, sizeof(double) * (size_t)nc, hipMemcpyDeviceToHost, s));
   HIPCHK(hipStreamSynchronize(s));
   return SAP_OKAY;
}

int sap_bounds_copy(sap_bounds* dst, int32_t dv, const sap_bounds* src, int32_t sv)
{
   if( dst == NULL || src == NULL || dst->mat != src->mat || dv < 0 || dv >= dst->nvec || sv < 0 || sv >= src->nvec )
      return SAP_INVALIDDATA;
   sap_matrix* m = dst->mat;
   HIPCHK(hipSetDevice(m->ctx->device));
   const size_t nc = (size_t)m->ncols;
   HIPCHK(hipMemcpyAsync(dst->V.bnd + (size_t)dv * nc, src->V.bnd + (size_t)sv * nc, sizeof(double2) * nc,
      hipMemcpyDeviceToDevice, m->ctx->stream));
   touchedVectors(dst, dv);
   return SAP_OKAY;
}

// grid partition of the sweep launch over the bins, by worklist capacity
static Layout makeLayout(const sap_bounds* b)
{
   Layout L;
   const sap_matrix* m = b->mat;
   const size_t nvec = (size_t)b->V.nvec;
   // items per workgroup and block caps per kind (kind 0 tiles, kind 1+b bin b); the caps
   // scale with the matrix's tuning
   static const int kPerBlock[SAP_NKINDS] = { SAP_BLOCK / SAP_WAVE, SAP_BLOCK / 4, SAP_BLOCK / 4, SAP_BLOCK / 8,
      SAP_BLOCK / 16, SAP_BLOCK / 32, SAP_BLOCK / 64, SAP_BLOCK / SAP_WAVE, SAP_BLOCK / SAP_WAVE, 1 };
   static const int kCap[SAP_NKINDS] = { 4096, 2048, 2048, 2048, 2048, 2048, 2048, 2048, 2048, 1024 };
   const int tb = m->tuning.tilebins, rb = m->tuning.rowbins;
   int acc = 0;
   for( int q = 0; q < SAP_NKINDS; ++q )
   {
      L.blockstart[q] = acc;
      const int bin = q - 1;
      size_t items;
      if( q == 0 )
         items = b->V.dense ? nvec * (size_t)(m->ntiles > 0 ? m->ntiles : (m->binstart[1] > m->binstart[0] ? 1 : 0)) : 0;
      else if( b->V.dense && (bin == 0 || (bin > rb && bin <= tb)) )
         items = 0;   // empty rows and the tiled bins go through the tiles in dense mode
      else
         items = nvec * (size_t)(m->binstart[q] - m->binstart[q - 1]);
      const int per = (bin >= 1 && bin <= rb) ? SAP_BLOCK : kPerBlock[q];   // one row per lane
      size_t blocks = (items + per - 1) / per;
      const size_t cap = std::max<size_t>(64, (size_t)kCap[q] * (size_t)m->tuning.capscale / 100);
      if( blocks > cap )
         blocks = cap;
      // the tiles' grid-stride loop runs in generations of resident workgroups: at most one
      if( q == 0 && m->ctx->resident > 0 && blocks > (size_t)m->ctx->resident )
         blocks = (size_t)m->ctx->resident;
      acc += (int)blocks;
   }
   L.blockstart[SAP_NKINDS] = acc;
   return L;
}

// host twin of foldStamps, on a status read back from the device
static void hostFoldStamps(Status* h)
{
   unsigned long long tmin = ~0ULL, tmax = 0;
   for( int q = 0; q < SAP_STAT_SLOTS; ++q )
   {
      tmin = h->tslot0[q] < tmin ? h->tslot0[q] : tmin;
      tmax = h->tslot1[q] > tmax ? h->tslot1[q] : tmax;
   }
   if( tmin != ~0ULL && tmax >= tmin )
   {
      h->sweep_ticks += tmax - tmin;
      h->sweep_count += 1;
   }
}

static void foldStats(sap_context* c)
{
   // fold the per-workgroup statistic slots into the host copy of the status
   Status* h = c->hstatus;
   unsigned long long a = 0, r = 0, g = 0;
   for( int q = 0; q < SAP_STAT_SLOTS; ++q )
   {
      a += h->slot_nnz[q];
      r += h->slot_rows[q];
      g += h->slot_chg[q];
   }
   h->nnz_swept = a;
   h->rows_swept = r;
   h->nchg = g;
}

static void fillStats(sap_context* c, double tot, double sweepms, int64_t launches, sap_stats* stats)
{
   if( stats == NULL )
      return;
   const Status* h = c->hstatus;
   stats->rounds = h->rlocal;
   stats->nchgbds = (int64_t)h->nchg;
   stats->rows_swept = (int64_t)h->rows_swept;
   stats->nnz_swept = (int64_t)h->nnz_swept;
   stats->device_ms = tot;
   stats->sweep_ms = sweepms;
   stats->sweep_launches = launches;
   stats->nrecompute = (int64_t)h->nrecomp;
}

// enqueue one round on stream s: control (unless the call's reset already began this round),
// sweep (between two events when given), column step
static void enqueueRound(sap_bounds* b, const Mat& M, const Params& P, const Layout& L, hipStream_t s, hipEvent_t e0,
   hipEvent_t e1, bool begun = false)
{
   sap_matrix* m = b->mat;
   const int nsweep = L.blockstart[SAP_NKINDS];
   const unsigned colgrid = gridFor((size_t)b->V.nvec * (size_t)m->ncols, 4096);
   if( !begun )
      k_round_begin<<<1, SAP_STAT_SLOTS, 0, s>>>(b->V, P.maxrounds);
   if( (P.flags & SAP_FLAG_RANGEDROWS) && m->nranged > 0 && b->V.reasons )
      k_ranged<true><<<gridFor((size_t)b->V.nvec * (size_t)m->nranged, 4096), SAP_BLOCK, 0, s>>>(M, b->V, P);
   else if( (P.flags & SAP_FLAG_RANGEDROWS) && m->nranged > 0 )
      k_ranged<false><<<gridFor((size_t)b->V.nvec * (size_t)m->nranged, 4096), SAP_BLOCK, 0, s>>>(M, b->V, P);
   if( e0 != NULL )
      (void)hipEventRecord(e0, s);
   if( nsweep > 0 && b->V.reasons )
      k_sweep<true><<<nsweep, SAP_BLOCK, 0, s>>>(M, b->V, P, L);
   else if( nsweep > 0 )
      k_sweep<false><<<nsweep, SAP_BLOCK, 0, s>>>(M, b->V, P, L);
   if( e1 != NULL )
      (void)hipEventRecord(e1, s);
   k_columns<<<colgrid, SAP_BLOCK, 0, s>>>(M, b->V, P);
   if( b->V.reasons )
   {
      k_reason<<<gridFor(2 * (size_t)m->nnz, 4096), SAP_BLOCK, 0, s>>>(b->V);
      k_history<<<gridFor((size_t)m->ncols, 4096), SAP_BLOCK, 0, s>>>(M, b->V);
   }
   k_mark<<<colgrid, SAP_BLOCK, 0, s>>>(M, b->V);
}

// ranged-row propagation keeps one check-group flag per (vector, nonzero); allocated on the
// first call that asks for it, before any round graph that uses it is captured
static int ensureRanged(sap_bounds* b, const Params& P)
{
   if( !(P.flags & SAP_FLAG_RANGEDROWS) || b->mat->nranged == 0 || b->V.rgflag != NULL )
      return SAP_OKAY;
   return devAlloc(&b->V.rgflag, (size_t)b->nvec * (size_t)b->mat->nnz + 1, &b->device_bytes);
}

// host-driven loop: one synchronisation per round
static int runRoundsHost(sap_bounds* b, const Params& P, sap_stats* stats)
{
   RCCHK(ensureRanged(b, P));
   sap_matrix* m = b->mat;
   sap_context* c = m->ctx;
   hipStream_t s = c->stream;
   const Mat M = toMat(m);
   const Layout L = makeLayout(b);
   double sweepms = 0.0;
   int64_t launches = 0;
   int32_t started = 0;
   HIPCHK(hipEventRecord(c->ev[0], s));
   for( int it = 0;; ++it )
   {
      // every round tightens some bound strictly (boundstreps, forced rows, sign fixing), so
      // the loop terminates; the cap guards against a defect turning into an endless loop
      if( it > SAP_ROUND_CAP )
      {
         fprintf(stderr, "[scip_amd] round cap %d exceeded\n", SAP_ROUND_CAP);
         return SAP_ERROR;
      }
      enqueueRound(b, M, P, L, s, c->ev[2], c->ev[3], it == 0 && b->begun);
      b->begun = 0;
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(c->hstatus, b->dstatus, sizeof(Status), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      // the round ran (its sweep was timed) when its control kernel advanced the round count;
      // the last one already reports done at its end
      if( c->hstatus->rlocal > started )
      {
         started = c->hstatus->rlocal;
         float ms = 0.0f;
         HIPCHK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
         sweepms += ms;
         ++launches;
      }
      if( c->hstatus->done )
         break;
   }
   HIPCHK(hipEventRecord(c->ev[1], s));
   HIPCHK(hipEventSynchronize(c->ev[1]));
   float tot = 0.0f;
   HIPCHK(hipEventElapsedTime(&tot, c->ev[0], c->ev[1]));
   foldStats(c);
   fillStats(c, tot, sweepms, launches, stats);
   return SAP_OKAY;
}

// capture `rounds` rounds into g on the context's capture stream, which nothing is ever
// launched on; on failure the stream is left out of capture mode
static bool captureGraph(sap_bounds* b, const Params& P, int rounds, RoundGraph* g)
{
   sap_matrix* m = b->mat;
   hipStream_t s = m->ctx->cstream;
   const Mat M = toMat(m);
   const Layout L = makeLayout(b);
   if( hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess )
      return false;
   for( int r = 0; r < rounds; ++r )
      enqueueRound(b, M, P, L, s, NULL, NULL);
   hipGraph_t graph = NULL;
   const hipError_t e1 = hipStreamEndCapture(s, &graph);
   if( e1 != hipSuccess || graph == NULL || hipGetLastError() != hipSuccess )
   {
      if( graph != NULL )
         (void)hipGraphDestroy(graph);
      return false;
   }
   const hipError_t e2 = hipGraphInstantiate(&g->exec, graph, NULL, NULL, 0);
   (void)hipGraphDestroy(graph);
   if( e2 != hipSuccess )
      return false;
   g->rounds = rounds;
   g->valid = 1;
   return true;
}

static void dropGraphs(sap_bounds* b);

static bool ensureGraphs(sap_bounds* b, const Params& P)
{
   GraphKey k;
   memset(&k, 0, sizeof(k));
   Params kp = P;
   kp.flags = P.flags & (SAP_FLAG_RANGEDROWS | SAP_FLAG_NOTIGHTEN);   // the other flags steer the host side only
   const Mat M = toMat(b->mat);
   const Layout L = makeLayout(b);
   memcpy(&k.P, &kp, sizeof(Params));
   memcpy(&k.V, &b->V, sizeof(Vec));
   memcpy(&k.M, &M, sizeof(Mat));
   memcpy(&k.L, &L, sizeof(Layout));
   bool valid = memcmp(&k, &b->key, sizeof(k)) == 0;
   for( int g = 0; g < SAP_NGRAPHS; ++g )
      valid = valid && b->graph[g].valid;
   if( valid )
      return true;
   dropGraphs(b);
   if( !captureGraph(b, P, SAP_GRAPH_SHORT, &b->graph[0]) || !captureGraph(b, P, SAP_GRAPH_LONG, &b->graph[1])
      || !captureGraph(b, P, 4, &b->graph[2]) )
   {
      (void)hipGetLastError();
      b->graphbroken = 1;
      fprintf(stderr, "[scip_amd] round graph capture unavailable; using the host-driven loop\n");
      return false;
   }
   b->key = k;
   return true;
}

// rounds until the device reports done: a first round enqueued directly, then graph launches
// of 2, 4, then 8 rounds each, one synchronisation per group (per round with
// SAP_FLAG_HOSTLOOP).  The last kernel of every
// round already decides whether another round would run, so a call ends right after the
// group of its last round; growing groups bound both the synchronisations (log2 of the rounds)
// and the rounds launched after the end (fewer than the rounds run).
// node calls: change lists of up to this many entries in all come back with the call's last
// wait, in a landing area of 20 B per entry (others are collected after it)
#define SAP_LISTS_ONEREAD ((int64_t)1 << 18)

// change lists of a batched call brought back by its last wait (k_list_offsets,
// k_collect_ended): device pointers into b's landing area
struct ListSpec
{
   int64_t cap;       // entries that fit
   int64_t* hdr;      // total, or -1 (call not ended)
   int32_t* col;      // cap columns
   double2* bnd;      // cap (lb, ub)
};

// metan > 0: the outcomes of vectors 0..metan-1 (k_pack_meta) are also in b's landing area,
// whose room the caller has made, when it returns, read by the call's last wait; with ls also
// their change lists
static int runRounds(sap_bounds* b, const Params& P, sap_stats* stats, int32_t metan = 0, const ListSpec* ls = NULL)
{
   RCCHK(ensureRanged(b, P));
   if( (P.flags & SAP_FLAG_HOSTLOOP) || b->graphbroken || !ensureGraphs(b, P) )
   {
      RCCHK(runRoundsHost(b, P, stats));
      if( metan > 0 )
      {
         hipStream_t s = b->mat->ctx->stream;
         k_pack_meta<<<gridFor((size_t)metan), SAP_BLOCK, 0, s>>>(b->V, metan, (int32_t*)b->hdnd);
         HIPCHK(hipGetLastError());
         HIPCHK(hipStreamSynchronize(s));
      }
      return SAP_OKAY;
   }
   sap_context* c = b->mat->ctx;
   hipStream_t s = c->stream;
   double sweepms = 0.0;
   int64_t launches = 0;
   int32_t prev = 0;
   HIPCHK(hipEventRecord(c->ev[0], s));
   for( int it = 0;; ++it )
   {
      if( it > SAP_ROUND_CAP / SAP_GRAPH_SHORT )
      {
         fprintf(stderr, "[scip_amd] round cap %d exceeded\n", SAP_ROUND_CAP);
         return SAP_ERROR;
      }
      // the first group's one round is enqueued directly (measured 0.5-1.5 % faster per call
      // than launching a one-round graph); then graphs of 2, 4 and 8 rounds.  A call on several
      // vectors rarely ends after its first round, nor does one on a bound set whose last call
      // took more: its two-round graph follows without a wait (rounds after the end cost only
      // their early exits)
      static const int kSchedule[3] = { 0, 2, 1 };   // 2, 4, 8 rounds
      if( it == 0 )
      {
         const Mat M0 = toMat(b->mat);
         const Layout L0 = makeLayout(b);
         enqueueRound(b, M0, P, L0, s, NULL, NULL, b->begun != 0);
         b->begun = 0;
         HIPCHK(hipGetLastError());
         if( b->V.nvec > 1 || b->lastrounds > 1 )
         {
            HIPCHK(hipGraphLaunch(b->graph[kSchedule[0]].exec, s));
            it = 1;
         }
      }
      else
         HIPCHK(hipGraphLaunch(b->graph[kSchedule[it < 3 ? it - 1 : 2]].exec, s));
      // the status goes to mapped host memory by a copy kernel on the stream
      k_copy_words<<<gridFor(sizeof(Status) / 4), SAP_BLOCK, 0, s>>>((uint32_t*)c->hstatusd, (const uint32_t*)b->dstatus,
         sizeof(Status) / 4);
      if( metan > 0 )
         k_pack_meta<<<gridFor((size_t)metan), SAP_BLOCK, 0, s>>>(b->V, metan, (int32_t*)b->hdnd);
      HIPCHK(hipGetLastError());
      // the end event of the call rides on every group's synchronisation, so that the last
      // one needs no further round trip
      HIPCHK(hipEventRecord(c->ev[1], s));
      if( ls != NULL && metan > 0 )
      {
         k_list_offsets<<<1, SAP_BLOCK, 0, s>>>(b->V, metan, P.maxrounds, b->outoff, ls->hdr);
         k_collect_ended<<<dim3(1, (unsigned)std::min(metan, 4096)), SAP_BLOCK, 0, s>>>(b->V, b->mat->ncols, metan,
            b->outoff, ls->cap, ls->col, ls->bnd);
         HIPCHK(hipGetLastError());
      }
      HIPCHK(hipStreamSynchronize(s));
      prev = c->hstatus->rlocal;
      if( c->hstatus->done )
         break;
      if( stopsNext(*c->hstatus, b->V.dense != 0, P.maxrounds) )
      {
         // the group's last round was the call's last: no further group; its sweep stamps,
         // which the next round's control kernel would fold, are folded here (the next call
         // resets them on the device)
         hostFoldStamps(c->hstatus);
         break;
      }
   }
   // sweep launch durations from the in-kernel wall clock (events recorded inside a graph do
   // not time on this runtime)
   sweepms = (double)c->hstatus->sweep_ticks / (double)c->wallclock_khz;
   launches = (int64_t)c->hstatus->sweep_count;
   b->lastrounds = c->hstatus->rlocal;
   (void)prev;
   float tot = 0.0f;
   HIPCHK(hipEventElapsedTime(&tot, c->ev[0], c->ev[1]));
   foldStats(c);
   fillStats(c, tot, sweepms, launches, stats);
   return SAP_OKAY;
}

// prepare a call on the first nvec vectors: reset per-call state; with firstall the first
// round sweeps every row (dense mode), else rows are queued explicitly.  maxrounds >= -1:
// the caller enqueues nothing before the first round, so when the reset is one workgroup it
// also begins that round (b->begun), and the first round is then enqueued without its
// control kernel
static int beginCall(sap_bounds* b, int32_t nvec, int32_t firstall, int32_t maxrounds = -2)
{
   sap_matrix* m = b->mat;
   hipStream_t s = m->ctx->stream;
   // the redundant-row flags are tagged with a generation per call instead of being cleared
   // per call (nvec * nrows bytes); they are cleared only when the generation starts again
   // at 1 (the first call on b, and every 255th call after it)
   const int32_t gen = (b->redgen >= 1 && b->redgen < 255) ? b->redgen + 1 : 1;
   const size_t nclear = gen == 1 ? (size_t)b->nvec * (size_t)m->nrows : 0;
   const size_t nunits = (size_t)nvec * (size_t)m->nseg;
   size_t work = nclear > nunits ? nclear : nunits;
   if( work < (size_t)nvec )
      work = (size_t)nvec;
   if( work < SAP_STAT_SLOTS )
      work = SAP_STAT_SLOTS;
   const unsigned grid = gridFor(work);
   const int32_t begin = (maxrounds >= -1 && grid == 1) ? 1 : 0;
   static_assert(SAP_BLOCK == SAP_STAT_SLOTS, "one workgroup of k_call_reset covers the stat slots");
   k_call_reset<<<grid, SAP_BLOCK, 0, s>>>(b->V, nclear, nvec, firstall, gen, begin, maxrounds);
   HIPCHK(hipGetLastError());
   b->redgen = gen;
   b->begun = begin;
   return SAP_OKAY;
}

// sparse mode: all rows of the first nvec vectors form the first round's worklists
static int queueAllRows(sap_bounds* b, int32_t nvec)
{
   sap_matrix* m = b->mat;
   hipStream_t s = m->ctx->stream;
   const Mat M = toMat(m);
   for( int q = 0; q < SAP_NBINS; ++q )
   {
      const int32_t len = m->binstart[q + 1] - m->binstart[q];
      if( len > 0 )
         k_init_all<<<gridFor((size_t)nvec * (size_t)len), SAP_BLOCK, 0, s>>>(M, b->V, q, nvec);
   }
   HIPCHK(hipGetLastError());
   return SAP_OKAY;
}

static int checkParams(const sap_params* p)
{
   if( p == NULL )
      return SAP_INVALIDDATA;
   if( !(p->infinity > 0.0) || !(p->epsilon >= 0.0) || !(p->sumepsilon >= 0.0) || !(p->feastol >= 0.0)
      || !(p->boundstreps >= 0.0) || !(p->hugeval > 0.0) || !(p->recompfac > 0.0) || p->maxrounds < -1 )
      return SAP_INVALIDDATA;
   return SAP_OKAY;
}

// the matrix's own one-vector state (for sap_propagate / packed / autotune), not counted
// among the caller's bound sets
static int createSingle(sap_matrix* m)
{
   RCCHK(sap_bounds_create(m, 1, &m->single));
   m->single->internal = 1;
   m->nbounds--;
   return SAP_OKAY;
}

int sap_propagate_resident(sap_bounds* b, const sap_params* params, sap_stats* stats)
{
   if( b == NULL )
      return SAP_INVALIDDATA;
   RCCHK(checkParams(params));
   sap_matrix* m = b->mat;
   sap_context* c = m->ctx;
   HIPCHK(hipSetDevice(c->device));
   const Params P = toParams(params);
   if( b->slotsmode )
      return SAP_INVALIDCALL;   // the vectors serve as slots of continuous batching
   const CallScope saved(b->V);
   b->V.nvec = 1;
   b->V.single = (m->nseg == 1) ? 1 : 0;   // packed instances: a cutoff stops only its instance
   int rc = beginCall(b, 1, b->V.dense, b->V.dense ? P.maxrounds : -2);
   if( rc == SAP_OKAY && !b->V.dense )
      rc = queueAllRows(b, 1);
   if( rc == SAP_OKAY )
      rc = runRounds(b, P, stats);
   saved.restore(b->V);
   touchedVectors(b, 0);
   RCCHK(rc);
   if( stats )
      stats->result = c->hstatus->cutoff ? SAP_CUTOFF : (c->hstatus->nchg > 0 ? SAP_REDUCEDDOM : SAP_DIDNOTFIND);
   return SAP_OKAY;
}

// destroy the round graphs of b; the stream they are launched on is drained first, so no
// launch of them can still be in flight
static void dropGraphs(sap_bounds* b)
{
   bool any = false;
   for( int g = 0; g < SAP_NGRAPHS; ++g )
      any = any || b->graph[g].valid;
   if( !any )
      return;
   (void)hipStreamSynchronize(b->mat->ctx->stream);
   for( int g = 0; g < SAP_NGRAPHS; ++g )
      if( b->graph[g].valid )
      {
         (void)hipGraphExecDestroy(b->graph[g].exec);
         b->graph[g].valid = 0;
      }
}

// the matrix's one-vector state with reason recording: candidate log of 2 entries per
// nonzero (a row is swept once per round and submits at most a lower and an upper bound
// candidate per nonzero), per-column outcome, winning keys and rows, and the history
static int createReasons(sap_matrix* m)
{
   if( m->nrows >= ((int64_t)1 << 28) || m->ncols >= ((int64_t)1 << 30) || 2 * m->nnz + 64 > (int64_t)UINT32_MAX )
      return SAP_NOTIMPLEMENTED;   // row * 8 and column * 2 must fit the record fields
   sap_bounds* b = NULL;
   RCCHK(sap_bounds_create_mode(m, 1, SAP_MODE_DENSE, &b));
   b->internal = 1;
   m->nbounds--;
   m->rsingle = b;
   Vec& V = b->V;
   V.clogcap = (uint32_t)(2 * m->nnz + 64);
   V.histcap = (uint32_t)std::max<int64_t>(4096, std::min<int64_t>(2 * m->ncols, (int64_t)UINT32_MAX));
   RCCHK(devAlloc(&V.clog, (size_t)V.clogcap, &b->device_bytes));
   RCCHK(devAlloc(&V.rflag, (size_t)m->ncols + 1, &b->device_bytes));
   RCCHK(devAlloc(&V.rkey, 2 * (size_t)m->ncols + 2, &b->device_bytes));
   RCCHK(devAlloc(&V.rrow, 2 * (size_t)m->ncols + 2, &b->device_bytes));
   RCCHK(devAlloc(&V.hist, (size_t)V.histcap, &b->device_bytes));
   V.reasons = 1;
   return SAP_OKAY;
}

// history entry order: round, column, lower before upper
static bool histLess(const HistRec& a, const HistRec& b)
{
   return a.round != b.round ? a.round < b.round : a.colside < b.colside;
}

int sap_propagate_reasons(sap_matrix* m, const sap_params* params, double* lb, double* ub, int64_t histcap,
   int64_t* nhist, int32_t* hround, int32_t* hcol, int32_t* hbound, double* hval, int32_t* hrow, int32_t* hrule,
   sap_conflict* conflict, sap_stats* stats)
{
   if( m == NULL || lb == NULL || ub == NULL || nhist == NULL || histcap < 0 || m->nseg != 1
      || (histcap > 0 && (hround == NULL || hcol == NULL || hbound == NULL || hval == NULL || hrow == NULL)) )
      return SAP_INVALIDDATA;
   RCCHK(checkParams(params));
   sap_context* c = m->ctx;
   HIPCHK(hipSetDevice(c->device));
   hipStream_t s = c->stream;
   if( m->rsingle == NULL )
      RCCHK(createReasons(m));
   sap_bounds* b = m->rsingle;
   const Params P = toParams(params);
   const int64_t nc = m->ncols;
   std::vector<double> lb0(lb, lb + nc), ub0(ub, ub + nc);
   sap_stats local;
   for( int attempt = 0;; ++attempt )
   {
      RCCHK(sap_bounds_upload(b, 0, lb0.data(), ub0.data()));
      k_reason_reset<<<gridFor((size_t)nc), SAP_BLOCK, 0, s>>>(b->V, nc);
      HIPCHK(hipGetLastError());
      memset(&local, 0, sizeof(local));
      b->V.nvec = 1;
      b->V.single = 1;
      RCCHK(beginCall(b, 1, 1));
      RCCHK(runRounds(b, P, &local));
      const Status* h = c->hstatus;
      if( h->clogovf )
         return SAP_ERROR;   // cannot happen: the log holds two candidates per nonzero
      if( h->histn <= b->V.histcap || attempt > 0 )
         break;
      // more accepted changes than the history holds: again with room for all of them
      const uint32_t need = h->histn;
      dropGraphs(b);
      (void)hipFree(b->V.hist);
      b->V.hist = NULL;
      RCCHK(devAlloc(&b->V.hist, (size_t)need, &b->device_bytes));
      b->V.histcap = need;
   }
   const Status h = *c->hstatus;
   const bool cut = h.cutoff != 0;
   local.result = cut ? SAP_CUTOFF : (h.nchg > 0 ? SAP_REDUCEDDOM : SAP_DIDNOTFIND);
   if( stats )
      *stats = local;
   // the history of the rounds before the last (a round ending infeasible adds nothing)
   const int64_t nh = (int64_t)std::min<uint32_t>(h.histn, b->V.histcap);
   *nhist = nh;
   if( conflict )
   {
      memset(conflict, 0, sizeof(*conflict));
      conflict->round = cut ? h.rlocal : 0;
      conflict->row = conflict->col = conflict->rowlb = conflict->rowub = -1;
      if( cut && h.cutrow != INT32_MAX )
      {
         conflict->kind = SAP_CONFLICT_ROW;
         const int code = h.cutrow & 7;
         conflict->row = h.cutrow >> 3;
         conflict->side = code == 2 ? 2 : (code == 0 || code == 3) ? 0 : 1;
      }
      else if( cut && h.cutcol != INT32_MAX )
      {
         uint8_t f = 0;
         int32_t rr[2];
         HIPCHK(hipMemcpyAsync(&f, b->V.rflag + h.cutcol, 1, hipMemcpyDeviceToHost, s));
         HIPCHK(hipMemcpyAsync(rr, b->V.rrow + 2 * (size_t)h.cutcol, sizeof(rr), hipMemcpyDeviceToHost, s));
         HIPCHK(hipStreamSynchronize(s));
         conflict->kind = SAP_CONFLICT_COLUMN;
         conflict->col = h.cutcol;
         conflict->rowlb = (f & RF_LB) ? rr[0] >> 1 : -1;
         conflict->rowub = (f & RF_UB) ? rr[1] >> 1 : -1;
         conflict->rules = ((f & RF_LB) ? (rr[0] & 1) : 0) | ((f & RF_UB) ? (rr[1] & 1) << 1 : 0);
      }
   }
   if( !cut )
      RCCHK(sap_bounds_download(b, 0, lb, ub));
   if( nh > histcap )
      return SAP_NOMEMORY;
   if( nh > 0 )
   {
      std::vector<HistRec> hh((size_t)nh);
      HIPCHK(hipMemcpyAsync(hh.data(), b->V.hist, sizeof(HistRec) * (size_t)nh, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      std::sort(hh.begin(), hh.end(), histLess);
      for( int64_t q = 0; q < nh; ++q )
      {
         const HistRec& r = hh[(size_t)q];
         hround[q] = r.round;
         hcol[q] = r.colside >> 1;
         hbound[q] = r.colside & 1;
         hval[q] = r.val;
         hrow[q] = r.row >> 1;
         if( hrule )
            hrule[q] = r.row & 1;
      }
   }
   return SAP_OKAY;
}

int sap_propagate(sap_matrix* m, const sap_params* params, double* lb, double* ub, sap_stats* stats)
{
   if( m == NULL || lb == NULL || ub == NULL )
      return SAP_INVALIDDATA;
   RCCHK(checkParams(params));
   if( m->single == NULL )
      RCCHK(createSingle(m));
   RCCHK(sap_bounds_upload(m->single, 0, lb, ub));
   sap_stats local;
   memset(&local, 0, sizeof(local));
   RCCHK(sap_propagate_resident(m->single, params, &local));
   if( local.result != SAP_CUTOFF )
      RCCHK(sap_bounds_download(m->single, 0, lb, ub));
   if( stats )
      *stats = local;
   return SAP_OKAY;
}


}   // extern "C"

// ---- persistent staging, change log and slot bookkeeping ----

template <typename T>
static int growDev(T** p, int64_t have, int64_t need)
{
   if( need <= have && *p != NULL )
      return SAP_OKAY;
   (void)hipFree(*p);
   *p = NULL;
   return devAlloc(p, (size_t)need, NULL);
}

static int ensureOverlay(sap_bounds* b, int64_t nov)
{
   if( nov <= b->ovcap && b->ovcol != NULL )
      return SAP_OKAY;
   int64_t cap = nov > 2 * b->ovcap ? nov : 2 * b->ovcap;
   if( cap < 1024 )
      cap = 1024;
   RCCHK(growDev(&b->ovcol, 0, cap));
   RCCHK(growDev(&b->ovlb, 0, cap));
   RCCHK(growDev(&b->ovub, 0, cap));
   b->ovcap = cap;
   return SAP_OKAY;
}

static int ensureOut(sap_bounds* b, int64_t n)
{
   if( n <= b->outcap && b->outcol != NULL )
      return SAP_OKAY;
   int64_t cap = n > 2 * b->outcap ? n : 2 * b->outcap;
   if( cap < 4096 )
      cap = 4096;
   RCCHK(growDev(&b->outcol, 0, cap));
   RCCHK(growDev(&b->outbnd, 0, cap));
   b->outcap = cap;
   return SAP_OKAY;
}

// the change log (nvec*ncols entries, one bit per column and vector) is allocated on first use
static int ensureLog(sap_bounds* b)
{
   if( b->logready )
      return SAP_OKAY;
   sap_matrix* m = b->mat;
   const size_t NC = (size_t)b->nvec * (size_t)m->ncols;
   RCCHK(devAlloc(&b->V.log, NC, &b->device_bytes));
   RCCHK(devAlloc(&b->V.logbits, NC / 32 + 1, &b->device_bytes));
   RCCHK(devAlloc(&b->V.ovl, NC, &b->device_bytes));
   RCCHK(devAlloc(&b->V.ovln, (size_t)b->nvec, &b->device_bytes));
   k_fill_i32<<<gridFor(NC / 32 + 1), SAP_BLOCK, 0, m->ctx->stream>>>((int32_t*)b->V.logbits, 0, NC / 32 + 1);
   k_fill_i32<<<gridFor((size_t)b->nvec), SAP_BLOCK, 0, m->ctx->stream>>>(b->V.ovln, 0, (size_t)b->nvec);
   HIPCHK(hipGetLastError());
   b->srcb = new std::vector<const sap_bounds*>((size_t)b->nvec, (const sap_bounds*)NULL);
   b->srcv = new std::vector<int32_t>((size_t)b->nvec, 0);
   b->srcgen = new std::vector<int64_t>((size_t)b->nvec, 0);
   b->logready = 1;
   return SAP_OKAY;
}

// validate overlays and copy them to the persistent device staging
// pinned host memory that the device reads and writes directly (mapped, coherent), with its
// device address
static int pinnedArea(size_t bytes, char** h, char** d)
{
   *h = NULL;
   *d = NULL;
   HIPCHK(hipHostMalloc((void**)h, bytes, hipHostMallocMapped | hipHostMallocCoherent));
   const hipError_t e = hipHostGetDevicePointer((void**)d, *h, 0);
   if( e != hipSuccess || *d == NULL )
   {
      (void)hipHostFree(*h);
      *h = NULL;
      *d = NULL;
      fprintf(stderr, "[scip_amd] hipHostGetDevicePointer failed: %s\n", hipGetErrorString(e));
      return SAP_ERROR;
   }
   return SAP_OKAY;
}

static int copyWords(hipStream_t s, void* dst, const void* src, size_t bytes)
{
   if( bytes % 4 != 0 )
      return SAP_ERROR;
   const size_t n = bytes / 4;
   k_copy_words<<<gridFor(n), SAP_BLOCK, 0, s>>>((uint32_t*)dst, (const uint32_t*)src, n);
   HIPCHK(hipGetLastError());
   return SAP_OKAY;
}

// Host-to-device copies for the batched calls (at most four at once): the bytes go to the bound
// set's pinned upload area first, and one copy kernel on the call's stream moves them, so
// nothing waits (a pageable source makes every copy synchronous) and the caller's arrays may
// be reused at once.  Pieces of the area are handed out in order; when it is full, the stream
// is drained, after which every earlier piece has been read, and the area starts over (grown
// if the pieces need more).
struct Upload
{
   void* dst;
   const void* src;
   size_t bytes;
};

static int uploadMany(sap_bounds* b, const Upload* u, int count)
{
   if( count < 1 || count > 4 )
      return SAP_ERROR;
   hipStream_t s = b->mat->ctx->stream;
   size_t need = 0;
   for( int i = 0; i < count; ++i )
   {
      if( u[i].bytes % 4 != 0 )
         return SAP_ERROR;
      need += (u[i].bytes + 255) & ~(size_t)255;
   }
   if( need == 0 )
      return SAP_OKAY;
   if( b->hupused + need > b->hupcap )
   {
      HIPCHK(hipStreamSynchronize(s));
      b->hupused = 0;
      if( need > b->hupcap )
      {
         const size_t cap = std::max(std::max(need, 2 * b->hupcap), (size_t)1 << 16);
         if( b->hup != NULL )
            (void)hipHostFree(b->hup);
         b->hup = NULL;
         b->hupd = NULL;
         b->hupcap = 0;
         RCCHK(pinnedArea(cap, &b->hup, &b->hupd));
         b->hupcap = cap;
      }
   }
   CopyPieces c;
   memset(&c, 0, sizeof(c));
   size_t most = 0;
   for( int i = 0; i < count; ++i )
   {
      if( u[i].bytes > 0 )
         memcpy(b->hup + b->hupused, u[i].src, u[i].bytes);
      c.dst[i] = (uint32_t*)u[i].dst;
      c.src[i] = (const uint32_t*)(b->hupd + b->hupused);
      c.n[i] = u[i].bytes / 4;
      most = std::max(most, c.n[i]);
      b->hupused += (u[i].bytes + 255) & ~(size_t)255;
   }
   k_copy_pieces<<<dim3(gridFor(most), (unsigned)count), SAP_BLOCK, 0, s>>>(c);
   HIPCHK(hipGetLastError());
   return SAP_OKAY;
}

static int upload(sap_bounds* b, void* ddst, const void* hsrc, size_t bytes)
{
   const Upload u = { ddst, hsrc, bytes };
   return uploadMany(b, &u, 1);
}

// pinned landing area of at least `bytes` for device-to-host copies of the batched calls
// (download()); the caller synchronises the stream before reading it and has read it before
// the next call here
static int downloadArea(sap_bounds* b, size_t bytes, char** out)
{
   if( bytes > b->hdncap )
   {
      HIPCHK(hipStreamSynchronize(b->mat->ctx->stream));
      const size_t cap = std::max(std::max(bytes, 2 * b->hdncap), (size_t)1 << 16);
      if( b->hdn != NULL )
         (void)hipHostFree(b->hdn);
      b->hdn = NULL;
      b->hdnd = NULL;
      b->hdncap = 0;
      RCCHK(pinnedArea(cap, &b->hdn, &b->hdnd));
      b->hdncap = cap;
   }
   *out = b->hdn;
   return SAP_OKAY;
}

// device-to-host copy into the landing area at h (inside b->hdn), by a copy kernel
static int download(sap_bounds* b, char* h, const void* dsrc, size_t bytes)
{
   if( bytes == 0 )
      return SAP_OKAY;
   if( h < b->hdn || h + bytes > b->hdn + b->hdncap )
      return SAP_ERROR;
   return copyWords(b->mat->ctx->stream, b->hdnd + (h - b->hdn), dsrc, bytes);
}

// change-log length, cutoff flag, last round and start round of vectors 0..n-1 (one instance),
// in one download
// the outcomes k_pack_meta left in the landing area (vectors 0..n-1)
static void readMeta(const sap_bounds* b, int32_t n, std::vector<int32_t>& logn, std::vector<int32_t>& cut,
   std::vector<int32_t>* vr, std::vector<int32_t>* vs)
{
   const int32_t* p = (const int32_t*)b->hdn;
   logn.assign(p, p + n);
   cut.assign(p + n, p + 2 * (size_t)n);
   if( vr != NULL )
      vr->assign(p + 2 * (size_t)n, p + 3 * (size_t)n);
   if( vs != NULL )
      vs->assign(p + 3 * (size_t)n, p + 4 * (size_t)n);
}

static int fetchMeta(sap_bounds* b, int32_t n, std::vector<int32_t>& logn, std::vector<int32_t>& cut,
   std::vector<int32_t>* vr, std::vector<int32_t>* vs)
{
   hipStream_t s = b->mat->ctx->stream;
   char* h = NULL;
   RCCHK(downloadArea(b, 16 * (size_t)n, &h));
   k_pack_meta<<<gridFor((size_t)n), SAP_BLOCK, 0, s>>>(b->V, n, (int32_t*)b->hdnd);
   HIPCHK(hipGetLastError());
   HIPCHK(hipStreamSynchronize(s));
   readMeta(b, n, logn, cut, vr, vs);
   return SAP_OKAY;
}

static int stageOverlay(sap_bounds* b, int32_t n, const int64_t* ovptr, const int32_t* ovcol, const double* ovlb,
   const double* ovub, int64_t* novout)
{
   const int64_t nc = b->mat->ncols;
   const int64_t nov = ovptr ? ovptr[n] : 0;
   if( ovptr != NULL && (ovptr[0] != 0 || nov < 0 || (nov > 0 && (ovcol == NULL || ovlb == NULL || ovub == NULL))) )
      return SAP_INVALIDDATA;
   for( int32_t v = 0; ovptr && v < n; ++v )
      if( ovptr[v + 1] < ovptr[v] || ovptr[v + 1] - ovptr[v] > nc )
         return SAP_INVALIDDATA;
   // without early exits, so that the check vectorises
   int bad = 0;
   for( int64_t q = 0; q < nov; ++q )
      bad |= (ovcol[q] < 0) | (ovcol[q] >= nc) | (ovlb[q] != ovlb[q]) | (ovub[q] != ovub[q]);
   if( bad )
      return SAP_INVALIDDATA;
   RCCHK(ensureOverlay(b, nov));
   hipStream_t s = b->mat->ctx->stream;
   if( ovptr == NULL )
   {
      k_fill_i64<<<gridFor((size_t)n + 1), SAP_BLOCK, 0, s>>>((long long*)b->ovptr, 0, (size_t)n + 1);
      HIPCHK(hipGetLastError());
   }
   Upload u[4];
   int nu = 0;
   if( ovptr != NULL )
      u[nu++] = Upload{ b->ovptr, ovptr, sizeof(int64_t) * ((size_t)n + 1) };
   if( nov > 0 )
   {
      u[nu++] = Upload{ b->ovcol, ovcol, sizeof(int32_t) * (size_t)nov };
      u[nu++] = Upload{ b->ovlb, ovlb, sizeof(double) * (size_t)nov };
      u[nu++] = Upload{ b->ovub, ovub, sizeof(double) * (size_t)nov };
   }
   if( nu > 0 )
      RCCHK(uploadMany(b, u, nu));
   *novout = nov;
   return SAP_OKAY;
}

// overlay entries per workgroup chunk of k_slot_mark (1..SAP_BLOCK) and its grid: about 256
// workgroups whatever the number of entries, so few entries (one per slot when probing) still
// spread over the chip and many entries make few worklist appends
static int32_t slotMarkChunk(int64_t nov)
{
   const int64_t c = (nov + 255) / 256;
   return (int32_t)(c < 1 ? 1 : (c > SAP_BLOCK ? SAP_BLOCK : c));
}

static unsigned slotMarkBlocks(int64_t nov)
{
   const int64_t c = slotMarkChunk(nov);
   const int64_t g = (nov + c - 1) / c;
   return (unsigned)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

static dim3 slotGrid(size_t perslot, int32_t n)
{
   return dim3(gridFor(perslot, 64), (unsigned)std::min<int32_t>(n > 0 ? n : 1, 65535));
}

// bounds of slots hslots[0..n) from a base vector: a slot last started from the same, unchanged
// base vector (B, bv) gets back only the columns of its change log and overlay; any other
// slot gets a full copy (B == NULL: dbase is staged host data, always copied)
static int initFromBase(sap_bounds* b, int32_t n, const int32_t* hslots, const sap_bounds* B, int32_t bv,
   const double2* dbase)
{
   sap_matrix* m = b->mat;
   hipStream_t s = m->ctx->stream;
   std::vector<int32_t> full, rest;
   for( int32_t i = 0; i < n; ++i )
   {
      const size_t v = (size_t)hslots[i];
      const bool same = B != NULL && B != b && (*b->srcb)[v] == B && (*b->srcv)[v] == bv && (*b->srcgen)[v] == B->gen;
      (same ? rest : full).push_back(hslots[i]);
      (*b->srcb)[v] = B;
      (*b->srcv)[v] = bv;
      (*b->srcgen)[v] = B != NULL ? B->gen : 0;
   }
   // both lists use slotidx2: the second upload is queued behind the first list's kernel on
   // the same stream, so it cannot overwrite the list that kernel reads.  A list that is
   // vectors 0..k-1 in order is b->ident's first k entries and needs no upload.
   auto leading = [](const std::vector<int32_t>& l) {
      for( size_t i = 0; i < l.size(); ++i )
         if( l[i] != (int32_t)i )
            return false;
      return true;
   };
   if( !rest.empty() )
   {
      const int32_t* d = b->ident;
      if( !leading(rest) )
      {
         RCCHK(upload(b, b->slotidx2, rest.data(), sizeof(int32_t) * rest.size()));
         d = b->slotidx2;
      }
      k_slot_restore<<<slotGrid(1024, (int32_t)rest.size()), SAP_BLOCK, 0, s>>>(b->V, dbase, m->ncols, d,
         (int32_t)rest.size());
      HIPCHK(hipGetLastError());
   }
   if( !full.empty() )
   {
      const int32_t* d = b->ident;
      if( !leading(full) )
      {
         RCCHK(upload(b, b->slotidx2, full.data(), sizeof(int32_t) * full.size()));
         d = b->slotidx2;
      }
      k_slot_base<<<slotGrid((size_t)m->ncols, (int32_t)full.size()), SAP_BLOCK, 0, s>>>(b->V.bnd, dbase, m->ncols,
         d, (int32_t)full.size());
      HIPCHK(hipGetLastError());
   }
   return SAP_OKAY;
}

// set up slots hslots[0..n) (device copy dslots) from the base vector and the staged overlays
static int admitSlots(sap_bounds* b, int32_t n, const int32_t* hslots, const int32_t* dslots, const sap_bounds* B,
   int32_t bv, const double2* dbase, int64_t nov, int32_t allrows)
{
   sap_matrix* m = b->mat;
   hipStream_t s = m->ctx->stream;
   const Mat M = toMat(m);
   RCCHK(initFromBase(b, n, hslots, B, bv, dbase));
   k_slot_reset<<<slotGrid((size_t)m->nrows, n), SAP_BLOCK, 0, s>>>(b->V, m->nrows, dslots, n);
   if( nov > 0 )
      k_slot_overlay<<<slotGrid((size_t)(nov / n + 1), n), SAP_BLOCK, 0, s>>>(b->V, m->ncols, dslots, n, b->ovptr,
         b->ovcol, b->ovlb, b->ovub);
   if( allrows || nov > 0 )
      k_slot_mark<<<allrows ? slotGrid((size_t)m->nrows, n) : dim3(slotMarkBlocks(nov)), SAP_BLOCK, 0, s>>>(M, b->V,
         dslots, n, b->ovptr, b->ovcol, allrows, slotMarkChunk(nov));
   HIPCHK(hipGetLastError());
   return SAP_OKAY;
}

// per-slot verdicts and sorted change lists of slots hslots[0..n); clears their change logs.
// With chgcol == NULL only verdicts (and chgptr counts, when given) are produced.
// change lists [off[i], off[i+1]) as gathered (hc, hb) into the caller's arrays; the log order
// depends on atomics, so each list is sorted by column (short lists, the common case, by
// insertion into the output)
static void emitLists(int32_t n, const int64_t* off, const int32_t* hc, const double2* hb, int32_t* chgcol,
   double* chglb, double* chgub)
{
   std::vector<int64_t> idx;
   for( int32_t i = 0; i < n; ++i )
   {
      const int64_t a = off[i], e = off[i + 1];
      if( e - a <= 16 )
      {
         for( int64_t q = a; q < e; ++q )
         {
            const int32_t c = hc[(size_t)q];
            const double2 bq = hb[(size_t)q];
            int64_t w = q;
            while( w > a && chgcol[w - 1] > c )
            {
               chgcol[w] = chgcol[w - 1];
               if( chglb )
                  chglb[w] = chglb[w - 1];
               if( chgub )
                  chgub[w] = chgub[w - 1];
               --w;
            }
            chgcol[w] = c;
            if( chglb )
               chglb[w] = bq.x;
            if( chgub )
               chgub[w] = bq.y;
         }
         continue;
      }
      idx.resize((size_t)(e - a));
      for( int64_t q = a; q < e; ++q )
         idx[(size_t)(q - a)] = q;
      std::sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) { return hc[(size_t)x] < hc[(size_t)y]; });
      for( int64_t q = a; q < e; ++q )
      {
         const int64_t src = idx[(size_t)(q - a)];
         chgcol[q] = hc[(size_t)src];
         if( chglb )
            chglb[q] = hb[(size_t)src].x;
         if( chgub )
            chgub[q] = hb[(size_t)src].y;
      }
   }
}

// premeta > 0: the outcomes of vectors 0..premeta-1, which hold every listed slot, are already
// in the landing area (batchCall with meta)
static int collectSlots(sap_bounds* b, int32_t n, const int32_t* hslots, int32_t* result, int32_t* rounds,
   int64_t chgcap, int64_t* chgptr, int32_t* chgcol, double* chglb, double* chgub, int32_t premeta = 0)
{
   sap_matrix* m = b->mat;
   hipStream_t s = m->ctx->stream;
   std::vector<int32_t> logn, cut, vr, vs;
   bool have = premeta > 0;
   for( int32_t i = 0; have && i < n; ++i )
      have = hslots[i] >= 0 && hslots[i] < premeta;
   if( have )
      readMeta(b, premeta, logn, cut, &vr, &vs);
   else
      RCCHK(fetchMeta(b, b->nvec, logn, cut, &vr, &vs));
   // output offsets: a cut-off slot has an empty list
   std::vector<int64_t> off((size_t)n + 1, 0);
   int32_t maxlog = 0;
   for( int32_t i = 0; i < n; ++i )
   {
      const size_t v = (size_t)hslots[i];
      off[(size_t)i + 1] = off[(size_t)i] + (cut[v] ? 0 : logn[v]);
      maxlog = std::max(maxlog, logn[v]);
      if( result )
         result[i] = cut[v] ? SAP_CUTOFF : (logn[v] > 0 ? SAP_REDUCEDDOM : SAP_DIDNOTFIND);
      if( rounds )
         rounds[i] = vr[v] > 0 ? vr[v] - vs[v] + 1 : 0;
   }
   const int64_t tot = off[(size_t)n];
   if( chgptr )
      memcpy(chgptr, off.data(), sizeof(int64_t) * ((size_t)n + 1));
   const bool want = chgcol != NULL;
   if( want && tot > chgcap )
      return SAP_NOMEMORY;   // nothing cleared: the caller may collect again with room
   // from here on every listed slot's log bits are cleared; when the slots cover every vector
   // that may hold bits, no later call needs to clear them first
   int32_t covered = 0;
   {
      std::vector<char> seen((size_t)std::max(b->logmax, 0), 0);
      for( int32_t i = 0; i < n; ++i )
         if( hslots[i] < b->logmax && !seen[(size_t)hslots[i]] )
         {
            seen[(size_t)hslots[i]] = 1;
            ++covered;
         }
   }
   const bool allclear = covered == b->logmax;
   if( maxlog == 0 )
   {
      if( allclear )
         b->logmax = 0;
      return SAP_OKAY;
   }
   // slots whose logs are only cleared, and slots whose changes are gathered
   std::vector<int32_t> clr, keep;
   std::vector<int64_t> koff;
   for( int32_t i = 0; i < n; ++i )
   {
      const int32_t v = hslots[i];
      if( logn[(size_t)v] == 0 )
         continue;
      if( want && !cut[(size_t)v] )
      {
         keep.push_back(v);
         koff.push_back(off[(size_t)i]);
      }
      else
         clr.push_back(v);
   }
   if( !clr.empty() )
   {
      RCCHK(upload(b, b->slotidx2, clr.data(), sizeof(int32_t) * clr.size()));
      k_collect<<<slotGrid((size_t)maxlog, (int32_t)clr.size()), SAP_BLOCK, 0, s>>>(b->V, m->ncols, b->slotidx2,
         (int32_t)clr.size(), b->outoff, NULL, NULL);
      HIPCHK(hipGetLastError());
   }
   if( keep.empty() )
   {
      HIPCHK(hipStreamSynchronize(s));
      if( allclear )
         b->logmax = 0;
      return SAP_OKAY;
   }
   RCCHK(ensureOut(b, tot));
   const Upload lists[2] = { { b->slotidx, keep.data(), sizeof(int32_t) * keep.size() },
      { b->outoff, koff.data(), sizeof(int64_t) * koff.size() } };
   RCCHK(uploadMany(b, lists, 2));
   k_collect<<<slotGrid((size_t)maxlog, (int32_t)keep.size()), SAP_BLOCK, 0, s>>>(b->V, m->ncols, b->slotidx,
      (int32_t)keep.size(), b->outoff, b->outcol, b->outbnd);
   HIPCHK(hipGetLastError());
   const size_t cbytes = ((sizeof(int32_t) * (size_t)tot) + 255) & ~(size_t)255;
   char* h = NULL;
   RCCHK(downloadArea(b, cbytes + sizeof(double2) * (size_t)tot, &h));
   const int32_t* hc = (const int32_t*)h;
   const double2* hb = (const double2*)(h + cbytes);
   RCCHK(download(b, h, b->outcol, sizeof(int32_t) * (size_t)tot));
   RCCHK(download(b, h + cbytes, b->outbnd, sizeof(double2) * (size_t)tot));
   HIPCHK(hipStreamSynchronize(s));
   if( allclear )
      b->logmax = 0;
   emitLists(n, off.data(), hc, hb, chgcol, chglb, chgub);
   return SAP_OKAY;
}

// clear the change logs of all vectors (before a one-shot call reuses them)
static int clearLogs(sap_bounds* b)
{
   if( b->logmax == 0 )
      return SAP_OKAY;   // every log bit written so far has been collected
   sap_matrix* m = b->mat;
   hipStream_t s = m->ctx->stream;
   std::vector<int32_t> logn, cut;
   RCCHK(fetchMeta(b, b->nvec, logn, cut, NULL, NULL));
   int32_t maxlog = 0;
   for( int32_t v = 0; v < b->nvec; ++v )
      maxlog = std::max(maxlog, logn[(size_t)v]);
   if( maxlog > 0 )
   {
      k_collect<<<slotGrid((size_t)maxlog, b->nvec), SAP_BLOCK, 0, s>>>(b->V, m->ncols, b->ident, b->nvec, b->outoff,
         NULL, NULL);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(s));
   }
   b->logmax = 0;
   return SAP_OKAY;
}

// one-shot propagation of vectors 0..nvec-1 from a device base vector plus the staged overlays
static int batchCall(sap_bounds* b, const sap_params* params, int32_t nvec, const sap_bounds* B, int32_t bv,
   const double2* dbase, int64_t nov, sap_stats* stats, bool meta = false, const ListSpec* ls = NULL)
{
   sap_matrix* m = b->mat;
   hipStream_t s = m->ctx->stream;
   if( b->slotsmode )
      return SAP_INVALIDCALL;   // the vectors serve as slots of continuous batching
   const Params P = toParams(params);
   const bool fromfix = (params->flags & SAP_FLAG_BASE_IS_FIXPOINT) != 0;
   RCCHK(ensureLog(b));
   RCCHK(clearLogs(b));
   std::vector<int32_t> ident((size_t)nvec);
   for( int32_t v = 0; v < nvec; ++v )
      ident[(size_t)v] = v;
   RCCHK(initFromBase(b, nvec, ident.data(), B, bv, dbase));
   b->gen++;
   // the call works on the first nvec vectors (slot list b->ident); the single-vector cutoff
   // shortcut only applies when the batch is one vector
   const CallScope saved(b->V);
   b->V.nvec = nvec;
   b->V.single = (nvec == 1 && m->nseg == 1) ? 1 : 0;
   b->V.logchg = 1;
   b->logmax = std::max(b->logmax, nvec);
   int rc = beginCall(b, nvec, (!fromfix && b->V.dense) ? 1 : 0);
   if( rc == SAP_OKAY )
   {
      if( nov > 0 )
         k_slot_overlay<<<slotGrid((size_t)(nov / nvec + 1), nvec), SAP_BLOCK, 0, s>>>(b->V, m->ncols, b->ident,
            nvec, b->ovptr, b->ovcol, b->ovlb, b->ovub);
      if( fromfix && nov > 0 )
         k_slot_mark<<<dim3(slotMarkBlocks(nov)), SAP_BLOCK, 0, s>>>(toMat(m), b->V, b->ident, nvec,
            b->ovptr, b->ovcol, 0, slotMarkChunk(nov));
      rc = hipToRc(hipGetLastError());
   }
   if( rc == SAP_OKAY && !fromfix && !b->V.dense )
      rc = queueAllRows(b, nvec);
   // meta: the vectors' outcomes come back with the last round group (readMeta), saving the
   // caller a wait; ls: their change lists too (the caller has made the landing area's room)
   char* h = NULL;
   if( rc == SAP_OKAY && meta && ls == NULL )
      rc = downloadArea(b, 16 * (size_t)nvec, &h);
   if( rc == SAP_OKAY )
      rc = runRounds(b, P, stats, meta ? nvec : 0, meta ? ls : NULL);
   saved.restore(b->V);
   return rc;
}

extern "C" {

int sap_propagate_batch(sap_bounds* b, const sap_params* params, int32_t nvec, const double* base_lb,
   const double* base_ub, const int64_t* ovptr, const int32_t* ovcol, const double* ovlb, const double* ovub,
   int32_t* result, int32_t* rounds, double* out_lb, double* out_ub, sap_stats* stats)
{
   if( b == NULL || base_lb == NULL || base_ub == NULL || nvec < 1 || nvec > b->nvec || b->mat->nseg != 1 )
      return SAP_INVALIDDATA;
   RCCHK(checkParams(params));
   sap_matrix* m = b->mat;
   sap_context* c = m->ctx;
   HIPCHK(hipSetDevice(c->device));
   hipStream_t s = c->stream;
   const int64_t nc = m->ncols;
   int64_t nov = 0;
   RCCHK(stageOverlay(b, nvec, ovptr, ovcol, ovlb, ovub, &nov));
   HIPCHK(hipMemcpyAsync(b->stage, base_lb, sizeof(double) * (size_t)nc, hipMemcpyHostToDevice, s));
   HIPCHK(hipMemcpyAsync(b->stage + nc, base_ub, sizeof(double) * (size_t)nc, hipMemcpyHostToDevice, s));
   k_interleave<<<gridFor((size_t)nc), SAP_BLOCK, 0, s>>>(b->base2, b->stage, b->stage + nc, nc);
   HIPCHK(hipGetLastError());
   RCCHK(batchCall(b, params, nvec, NULL, 0, b->base2, nov, stats, true));
   std::vector<int32_t> ident((size_t)nvec), res((size_t)nvec);
   for( int32_t v = 0; v < nvec; ++v )
      ident[(size_t)v] = v;
   RCCHK(collectSlots(b, nvec, ident.data(), res.data(), rounds, 0, NULL, NULL, NULL, NULL, nvec));
   if( result )
      memcpy(result, res.data(), sizeof(int32_t) * (size_t)nvec);
   if( out_lb != NULL && out_ub != NULL )
      for( int32_t v = 0; v < nvec; ++v )
         RCCHK(sap_bounds_download(b, v, out_lb + (size_t)v * (size_t)nc, out_ub + (size_t)v * (size_t)nc));
   if( stats )
   {
      int32_t anycut = 0;
      for( int32_t v = 0; v < nvec; ++v )
         anycut |= res[(size_t)v] == SAP_CUTOFF;
      stats->result = anycut ? SAP_CUTOFF : (stats->nchgbds > 0 ? SAP_REDUCEDDOM : SAP_DIDNOTFIND);
   }
   return SAP_OKAY;
}

int sap_propagate_nodes(sap_bounds* b, const sap_params* params, int32_t nvec, const sap_bounds* base,
   int32_t basevec, const int64_t* ovptr, const int32_t* ovcol, const double* ovlb, const double* ovub,
   int32_t* result, int32_t* rounds, int64_t chgcap, int64_t* chgptr, int32_t* chgcol, double* chglb,
   double* chgub, sap_stats* stats)
{
   if( b == NULL || base == NULL || base->mat != b->mat || basevec < 0 || basevec >= base->nvec || nvec < 1
      || nvec > b->nvec || b->mat->nseg != 1 || (chgcol != NULL && chgptr == NULL) || chgcap < 0 )
      return SAP_INVALIDDATA;
   RCCHK(checkParams(params));
   sap_matrix* m = b->mat;
   HIPCHK(hipSetDevice(m->ctx->device));
   int64_t nov = 0;
   RCCHK(stageOverlay(b, nvec, ovptr, ovcol, ovlb, ovub, &nov));
   const double2* dbase = base->V.bnd + (size_t)basevec * (size_t)m->ncols;
   if( base == b && basevec < nvec )
   {
      // the base is one of the vectors being overwritten: keep a copy
      HIPCHK(hipMemcpyAsync(b->base2, dbase, sizeof(double2) * (size_t)m->ncols, hipMemcpyDeviceToDevice,
         m->ctx->stream));
      dbase = b->base2;
   }
   // change lists that fit the caller's room (up to SAP_LISTS_ONEREAD entries) come back with
   // the call's last wait: landing area = outcomes of every vector, total, columns, bounds
   const size_t mb = (16 * (size_t)b->nvec + 255) & ~(size_t)255;
   const size_t cb = (sizeof(int32_t) * (size_t)chgcap + 255) & ~(size_t)255;
   const bool oneread = chgcol != NULL && chgcap > 0 && chgcap <= SAP_LISTS_ONEREAD;
   ListSpec ls;
   char* h = NULL;
   if( oneread )
   {
      RCCHK(downloadArea(b, mb + 256 + cb + sizeof(double2) * (size_t)chgcap, &h));
      ls.cap = chgcap;
      ls.hdr = (int64_t*)(b->hdnd + mb);
      ls.col = (int32_t*)(b->hdnd + mb + 256);
      ls.bnd = (double2*)(b->hdnd + mb + 256 + cb);
      *(volatile int64_t*)(h + mb) = -1;   // stays -1 unless the device numbered the lists
   }
   RCCHK(batchCall(b, params, nvec, base == b ? NULL : base, basevec, dbase, nov, stats, true,
      oneread ? &ls : NULL));
   std::vector<int32_t> res((size_t)nvec);
   const int64_t tot = oneread ? *(volatile int64_t*)(b->hdn + mb) : -1;
   if( oneread && tot >= 0 && tot <= chgcap )
   {
      // the lists are in the landing area and the logs cleared: outcomes and offsets as a
      // collection gives them
      std::vector<int32_t> logn, cut, vr, vs;
      readMeta(b, nvec, logn, cut, &vr, &vs);
      std::vector<int64_t> off((size_t)nvec + 1, 0);
      for( int32_t v = 0; v < nvec; ++v )
      {
         off[(size_t)v + 1] = off[(size_t)v] + (cut[(size_t)v] ? 0 : logn[(size_t)v]);
         res[(size_t)v] = cut[(size_t)v] ? SAP_CUTOFF : (logn[(size_t)v] > 0 ? SAP_REDUCEDDOM : SAP_DIDNOTFIND);
         if( rounds )
            rounds[v] = vr[(size_t)v] > 0 ? vr[(size_t)v] - vs[(size_t)v] + 1 : 0;
      }
      if( off[(size_t)nvec] != tot )
         return SAP_ERROR;
      memcpy(chgptr, off.data(), sizeof(int64_t) * ((size_t)nvec + 1));
      if( b->logmax <= nvec )
         b->logmax = 0;
      emitLists(nvec, off.data(), (const int32_t*)(b->hdn + mb + 256), (const double2*)(b->hdn + mb + 256 + cb),
         chgcol, chglb, chgub);
   }
   else
   {
      std::vector<int32_t> ident((size_t)nvec);
      for( int32_t v = 0; v < nvec; ++v )
         ident[(size_t)v] = v;
      RCCHK(collectSlots(b, nvec, ident.data(), res.data(), rounds, chgcap, chgptr, chgcol, chglb, chgub, nvec));
   }
   if( result )
      memcpy(result, res.data(), sizeof(int32_t) * (size_t)nvec);
   if( stats )
   {
      int32_t anycut = 0;
      for( int32_t v = 0; v < nvec; ++v )
         anycut |= res[(size_t)v] == SAP_CUTOFF;
      stats->result = anycut ? SAP_CUTOFF : (stats->nchgbds > 0 ? SAP_REDUCEDDOM : SAP_DIDNOTFIND);
   }
   return SAP_OKAY;
}

// probing: pair lists up to this many entries are read back together with their counts
#define SAP_PROBE_ONEREAD 4096

// probing buffers: probe columns and counters (nvec each), pair outputs for n entries
static int ensureProbe(sap_bounds* b, int64_t n)
{
   if( b->pcol == NULL )
   {
      RCCHK(devAlloc(&b->pcol, (size_t)b->nvec, NULL));
      RCCHK(devAlloc(&b->pcount, (size_t)b->nvec, NULL));
   }
   RCCHK(ensureOut(b, n));
   if( b->outbnd2 == NULL || b->outcap2 < b->outcap )
   {
      RCCHK(growDev(&b->outbnd2, 0, b->outcap));
      b->outcap2 = b->outcap;
   }
   return SAP_OKAY;
}

// probing of binaries (downub == NULL: branches x = 0 / x = 1) or of split branches
// x <= downub[p] / x >= uplb[p]; in the split form the probe column itself is listed with its
// bounds in both branches
static int probeCore(sap_bounds* b, const sap_params* params, int32_t nprobe, const int32_t* probecol,
   const double* downub, const double* uplb, const sap_bounds* base, int32_t basevec, int32_t* verdict,
   int32_t* rounds, int64_t dedcap, int64_t* dedptr, int32_t* dedcol, double* dlb0, double* dub0, double* dlb1,
   double* dub1, sap_stats* stats)
{
   if( b == NULL || base == NULL || base->mat != b->mat || basevec < 0 || basevec >= base->nvec || b->nvec < 2
      || nprobe < 0 || (nprobe > 0 && (probecol == NULL || verdict == NULL)) || b->mat->nseg != 1 || dedcap < 0
      || (dedcol != NULL && (dedptr == NULL || dlb0 == NULL || dub0 == NULL || dlb1 == NULL || dub1 == NULL))
      || (downub != NULL && nprobe > 0 && uplb == NULL) )
      return SAP_INVALIDDATA;
   const bool split = downub != NULL;
   RCCHK(checkParams(params));
   sap_matrix* m = b->mat;
   const int64_t nc = m->ncols;
   for( int32_t p = 0; p < nprobe; ++p )
      if( probecol[p] < 0 || probecol[p] >= nc )
         return SAP_INVALIDDATA;
   if( b->slotsmode )
      return SAP_INVALIDCALL;
   HIPCHK(hipSetDevice(m->ctx->device));
   hipStream_t s = m->ctx->stream;
   const double2* dbase = base->V.bnd + (size_t)basevec * (size_t)nc;
   if( base == b )
   {
      // every vector of b is overwritten by the branches: keep a copy of the base
      HIPCHK(hipMemcpyAsync(b->base2, dbase, sizeof(double2) * (size_t)nc, hipMemcpyDeviceToDevice, s));
      dbase = b->base2;
   }
   RCCHK(ensureProbe(b, b->nvec));
   const int32_t per = b->nvec / 2;
   // binary probing: every probe column must be an unfixed binary of the base; split probing:
   // base lb <= downub < uplb <= base ub (the base bounds of the probe columns are kept)
   std::vector<double2> pbase(split ? (size_t)nprobe : 0);
   for( int32_t p0 = 0; p0 < nprobe; p0 += b->nvec )
   {
      const int32_t n = std::min<int32_t>(b->nvec, nprobe - p0);
      RCCHK(upload(b, b->pcol, probecol + p0, sizeof(int32_t) * (size_t)n));
      char* h = NULL;
      if( split )
      {
         RCCHK(downloadArea(b, sizeof(double2) * (size_t)n, &h));
         k_gather_bounds<<<gridFor((size_t)n), SAP_BLOCK, 0, s>>>(dbase, nc, b->pcol, NULL, n, b->outbnd);
         HIPCHK(hipGetLastError());
         RCCHK(download(b, h, b->outbnd, sizeof(double2) * (size_t)n));
         HIPCHK(hipStreamSynchronize(s));
         memcpy(pbase.data() + p0, h, sizeof(double2) * (size_t)n);
         for( int32_t q = p0; q < p0 + n; ++q )
            if( !(downub[q] >= pbase[(size_t)q].x && downub[q] < uplb[q] && uplb[q] <= pbase[(size_t)q].y) )
               return SAP_INVALIDDATA;
         continue;
      }
      RCCHK(downloadArea(b, sizeof(int32_t), &h));
      k_fill_i32<<<1, 64, 0, s>>>(b->pcount, INT32_MAX, 1);
      k_probe_check<<<gridFor((size_t)n), SAP_BLOCK, 0, s>>>(toMat(m), dbase, b->pcol, n, b->pcount);
      HIPCHK(hipGetLastError());
      RCCHK(download(b, h, b->pcount, sizeof(int32_t)));
      HIPCHK(hipStreamSynchronize(s));
      int32_t bad = INT32_MAX;
      memcpy(&bad, h, sizeof(int32_t));
      if( bad != INT32_MAX )
         return SAP_INVALIDDATA;
   }
   sap_params prm = *params;
   std::vector<int64_t> cnt((size_t)nprobe + 1, 0);
   std::vector<int32_t> hcol;
   std::vector<double2> hb0, hb1;
   sap_stats acc;
   memset(&acc, 0, sizeof(acc));
   int32_t anyboth = 0, anyinfo = 0;
   std::vector<int64_t> ovptr, off;
   std::vector<int32_t> ovcol, logn, cut, vr, vs, pc;
   std::vector<double> ovlb, ovub;
   for( int32_t p0 = 0; p0 < nprobe; p0 += per )
   {
      const int32_t np = std::min<int32_t>(per, nprobe - p0);
      const int32_t nv = 2 * np;
      ovptr.resize((size_t)nv + 1);
      ovcol.resize((size_t)nv);
      ovlb.resize((size_t)nv);
      ovub.resize((size_t)nv);
      for( int32_t v = 0; v <= nv; ++v )
         ovptr[(size_t)v] = v;
      for( int32_t p = 0; p < np; ++p )
      {
         const int32_t q = p0 + p;
         const int32_t x = probecol[q];
         ovcol[(size_t)(2 * p)] = x;
         ovlb[(size_t)(2 * p)] = split ? pbase[(size_t)q].x : 0.0;
         ovub[(size_t)(2 * p)] = split ? downub[q] : 0.0;
         ovcol[(size_t)(2 * p + 1)] = x;
         ovlb[(size_t)(2 * p + 1)] = split ? uplb[q] : 1.0;
         ovub[(size_t)(2 * p + 1)] = split ? pbase[(size_t)q].y : 1.0;
      }
      int64_t nov = 0;
      RCCHK(stageOverlay(b, nv, ovptr.data(), ovcol.data(), ovlb.data(), ovub.data(), &nov));
      sap_stats cs;
      memset(&cs, 0, sizeof(cs));
      RCCHK(batchCall(b, &prm, nv, base == b ? NULL : base, basevec, dbase, nov, &cs, true));
      acc.rounds = std::max(acc.rounds, cs.rounds);
      acc.nchgbds += cs.nchgbds;
      acc.rows_swept += cs.rows_swept;
      acc.nnz_swept += cs.nnz_swept;
      acc.device_ms += cs.device_ms;
      acc.sweep_ms += cs.sweep_ms;
      acc.sweep_launches += cs.sweep_launches;
      // room for the columns changed in both branches: at most the shorter log per probe; the
      // same outcomes, read back with the last round group, give the verdicts and rounds
      readMeta(b, nv, logn, cut, &vr, &vs);
      off.assign((size_t)np + 1, 0);
      int32_t maxs = 0;
      for( int32_t p = 0; p < np; ++p )
      {
         const size_t v0 = (size_t)(2 * p), v1 = v0 + 1;
         const int32_t ns = (cut[v0] || cut[v1]) ? 0 : std::min(logn[v0], logn[v1]);
         off[(size_t)p + 1] = off[(size_t)p] + ns;
         maxs = std::max(maxs, ns);
      }
      const int64_t room = off[(size_t)np];
      pc.assign((size_t)np, 0);
      std::vector<int32_t> tc((size_t)room);
      std::vector<double2> t0((size_t)room), t1((size_t)room);
      if( room > 0 )
      {
         RCCHK(ensureProbe(b, room));
         // pair lists of at most SAP_PROBE_ONEREAD entries come back with the counts in one
         // read; longer ones only once the counts show that some pair was found
         const bool one = room <= SAP_PROBE_ONEREAD;
         const size_t pb = (sizeof(int32_t) * (size_t)np + 255) & ~(size_t)255;
         const size_t cb = (sizeof(int32_t) * (size_t)room + 255) & ~(size_t)255;
         const size_t bb = (sizeof(double2) * (size_t)room + 255) & ~(size_t)255;
         char* h = NULL;
         RCCHK(downloadArea(b, pb + (one ? cb + 2 * bb : 0), &h));
         const Upload lists[2] = { { b->pcol, probecol + p0, sizeof(int32_t) * (size_t)np },
            { b->outoff, off.data(), sizeof(int64_t) * (size_t)np } };
         RCCHK(uploadMany(b, lists, 2));
         k_fill_i32<<<gridFor((size_t)np), SAP_BLOCK, 0, s>>>(b->pcount, 0, (size_t)np);
         k_probe_pairs<<<slotGrid((size_t)maxs, np), SAP_BLOCK, 0, s>>>(b->V, nc, b->pcol, np, b->outoff, b->pcount,
            b->outcol, b->outbnd, b->outbnd2);
         HIPCHK(hipGetLastError());
         RCCHK(download(b, h, b->pcount, sizeof(int32_t) * (size_t)np));
         if( one )
         {
            RCCHK(download(b, h + pb, b->outcol, sizeof(int32_t) * (size_t)room));
            RCCHK(download(b, h + pb + cb, b->outbnd, sizeof(double2) * (size_t)room));
            RCCHK(download(b, h + pb + cb + bb, b->outbnd2, sizeof(double2) * (size_t)room));
         }
         HIPCHK(hipStreamSynchronize(s));
         memcpy(pc.data(), h, sizeof(int32_t) * (size_t)np);
         int64_t got = 0;
         for( int32_t p = 0; p < np; ++p )
            got += pc[(size_t)p];
         if( got > 0 && !one )
         {
            RCCHK(downloadArea(b, cb + 2 * bb, &h));
            RCCHK(download(b, h, b->outcol, sizeof(int32_t) * (size_t)room));
            RCCHK(download(b, h + cb, b->outbnd, sizeof(double2) * (size_t)room));
            RCCHK(download(b, h + cb + bb, b->outbnd2, sizeof(double2) * (size_t)room));
            HIPCHK(hipStreamSynchronize(s));
            memcpy(tc.data(), h, sizeof(int32_t) * (size_t)room);
            memcpy(t0.data(), h + cb, sizeof(double2) * (size_t)room);
            memcpy(t1.data(), h + cb + bb, sizeof(double2) * (size_t)room);
         }
         else if( got > 0 )
         {
            memcpy(tc.data(), h + pb, sizeof(int32_t) * (size_t)room);
            memcpy(t0.data(), h + pb + cb, sizeof(double2) * (size_t)room);
            memcpy(t1.data(), h + pb + cb + bb, sizeof(double2) * (size_t)room);
         }
      }
      // split probing: the probe columns' own bounds in both branches
      std::vector<double2> x0, x1;
      if( split )
      {
         std::vector<int32_t> xv((size_t)nv), xc((size_t)nv);
         for( int32_t v = 0; v < nv; ++v )
         {
            xv[(size_t)v] = v;
            xc[(size_t)v] = probecol[p0 + v / 2];
         }
         x0.resize((size_t)nv);
         char* h = NULL;
         RCCHK(downloadArea(b, sizeof(double2) * (size_t)nv, &h));
         const Upload lists[2] = { { b->pcol, xc.data(), sizeof(int32_t) * (size_t)nv },
            { b->pcount, xv.data(), sizeof(int32_t) * (size_t)nv } };
         RCCHK(uploadMany(b, lists, 2));
         k_gather_bounds<<<gridFor((size_t)nv), SAP_BLOCK, 0, s>>>(b->V.bnd, nc, b->pcol, b->pcount, nv, b->outbnd);
         HIPCHK(hipGetLastError());
         RCCHK(download(b, h, b->outbnd, sizeof(double2) * (size_t)nv));
         HIPCHK(hipStreamSynchronize(s));
         memcpy(x0.data(), h, sizeof(double2) * (size_t)nv);
      }
      // the change logs of the chunk's vectors are cleared for the next chunk (after the pair
      // kernel, in stream order; nothing to wait for)
      int32_t maxlog = 0;
      for( int32_t v = 0; v < nv; ++v )
         maxlog = std::max(maxlog, logn[(size_t)v]);
      if( maxlog > 0 )
      {
         k_collect<<<slotGrid((size_t)maxlog, nv), SAP_BLOCK, 0, s>>>(b->V, nc, b->ident, nv, b->outoff, NULL, NULL);
         HIPCHK(hipGetLastError());
      }
      if( b->logmax <= nv )
         b->logmax = 0;
      // verdicts (a cut-off branch is infeasible) and rounds, as a collection would report them
      std::vector<int64_t> idx;
      for( int32_t p = 0; p < np; ++p )
      {
         const int32_t q = p0 + p;
         const size_t v0 = (size_t)(2 * p), v1 = v0 + 1;
         verdict[q] = (cut[v0] ? SAP_PROBE_DOWN_INFEASIBLE : 0) | (cut[v1] ? SAP_PROBE_UP_INFEASIBLE : 0);
         anyboth |= verdict[q] == (SAP_PROBE_DOWN_INFEASIBLE | SAP_PROBE_UP_INFEASIBLE);
         anyinfo |= verdict[q] != 0 || pc[(size_t)p] > 0;
         if( rounds )
         {
            rounds[2 * (size_t)q] = vr[v0] > 0 ? vr[v0] - vs[v0] + 1 : 0;
            rounds[2 * (size_t)q + 1] = vr[v1] > 0 ? vr[v1] - vs[v1] + 1 : 0;
         }
         // the order within a probe depends on atomics: sort by column
         const int64_t a = off[(size_t)p];
         idx.resize((size_t)pc[(size_t)p]);
         for( int32_t k = 0; k < pc[(size_t)p]; ++k )
            idx[(size_t)k] = a + k;
         std::sort(idx.begin(), idx.end(), [&](int64_t u, int64_t w) { return tc[(size_t)u] < tc[(size_t)w]; });
         const int32_t x = probecol[q];
         const bool addx = split && verdict[q] == 0;
         bool xdone = !addx;
         for( int64_t k : idx )
         {
            if( !xdone && x < tc[(size_t)k] )
            {
               hcol.push_back(x);
               hb0.push_back(x0[(size_t)(2 * p)]);
               hb1.push_back(x0[(size_t)(2 * p + 1)]);
               xdone = true;
            }
            hcol.push_back(tc[(size_t)k]);
            hb0.push_back(t0[(size_t)k]);
            hb1.push_back(t1[(size_t)k]);
         }
         if( !xdone )
         {
            hcol.push_back(x);
            hb0.push_back(x0[(size_t)(2 * p)]);
            hb1.push_back(x0[(size_t)(2 * p + 1)]);
         }
         cnt[(size_t)q + 1] = pc[(size_t)p] + (addx ? 1 : 0);
      }
   }
   for( int32_t p = 0; p < nprobe; ++p )
      cnt[(size_t)p + 1] += cnt[(size_t)p];
   if( dedptr )
      memcpy(dedptr, cnt.data(), sizeof(int64_t) * ((size_t)nprobe + 1));
   if( stats )
   {
      *stats = acc;
      stats->result = anyboth ? SAP_CUTOFF : (anyinfo ? SAP_REDUCEDDOM : SAP_DIDNOTFIND);
   }
   if( dedcol == NULL )
      return SAP_OKAY;
   if( cnt[(size_t)nprobe] > dedcap )
      return SAP_NOMEMORY;
   for( size_t k = 0; k < hcol.size(); ++k )
   {
      dedcol[k] = hcol[k];
      dlb0[k] = hb0[k].x;
      dub0[k] = hb0[k].y;
      dlb1[k] = hb1[k].x;
      dub1[k] = hb1[k].y;
   }
   return SAP_OKAY;
}

int sap_probe(sap_bounds* b, const sap_params* params, int32_t nprobe, const int32_t* probecol,
   const sap_bounds* base, int32_t basevec, int32_t* verdict, int32_t* rounds, int64_t dedcap, int64_t* dedptr,
   int32_t* dedcol, double* dlb0, double* dub0, double* dlb1, double* dub1, sap_stats* stats)
{
   return probeCore(b, params, nprobe, probecol, NULL, NULL, base, basevec, verdict, rounds, dedcap, dedptr, dedcol,
      dlb0, dub0, dlb1, dub1, stats);
}

int sap_probe_split(sap_bounds* b, const sap_params* params, int32_t nprobe, const int32_t* probecol,
   const double* downub, const double* uplb, const sap_bounds* base, int32_t basevec, int32_t* verdict,
   int32_t* rounds, int64_t dedcap, int64_t* dedptr, int32_t* dedcol, double* dlb0, double* dub0, double* dlb1,
   double* dub1, sap_stats* stats)
{
   if( nprobe > 0 && (downub == NULL || uplb == NULL) )
      return SAP_INVALIDDATA;
   static const double none = 0.0;
   return probeCore(b, params, nprobe, probecol, downub != NULL ? downub : &none, uplb != NULL ? uplb : &none, base,
      basevec, verdict, rounds, dedcap, dedptr, dedcol, dlb0, dub0, dlb1, dub1, stats);
}

int sap_propagate_packed(sap_matrix* m, const sap_params* params, double* lb, double* ub, int32_t* result,
   int32_t* rounds, sap_stats* stats)
{
   if( m == NULL || lb == NULL || ub == NULL )
      return SAP_INVALIDDATA;
   RCCHK(checkParams(params));
   HIPCHK(hipSetDevice(m->ctx->device));
   if( m->single == NULL )
      RCCHK(createSingle(m));
   sap_bounds* b = m->single;
   hipStream_t s = m->ctx->stream;
   const int32_t ns = m->nseg;
   std::vector<double> in_lb(lb, lb + m->ncols), in_ub(ub, ub + m->ncols);
   RCCHK(sap_bounds_upload(b, 0, lb, ub));
   sap_stats local;
   memset(&local, 0, sizeof(local));
   RCCHK(sap_propagate_resident(b, params, &local));
   std::vector<int32_t> cut((size_t)ns), vr((size_t)ns);
   int32_t vs = 0;
   HIPCHK(hipMemcpyAsync(cut.data(), b->V.vcutoff, sizeof(int32_t) * (size_t)ns, hipMemcpyDeviceToHost, s));
   HIPCHK(hipMemcpyAsync(vr.data(), b->V.vrounds, sizeof(int32_t) * (size_t)ns, hipMemcpyDeviceToHost, s));
   HIPCHK(hipMemcpyAsync(&vs, b->V.vstart, sizeof(int32_t), hipMemcpyDeviceToHost, s));
   HIPCHK(hipStreamSynchronize(s));
   RCCHK(sap_bounds_download(b, 0, lb, ub));
   int32_t anycut = 0;
   for( int32_t q = 0; q < ns; ++q )
   {
      const int64_t c0 = ns > 1 ? (*m->coloff)[(size_t)q] : 0, c1 = ns > 1 ? (*m->coloff)[(size_t)q + 1] : m->ncols;
      bool changed = false;
      for( int64_t j = c0; j < c1 && !changed; ++j )
         changed = lb[j] != in_lb[(size_t)j] || ub[j] != in_ub[(size_t)j];
      const int32_t r = cut[(size_t)q] ? SAP_CUTOFF : (changed ? SAP_REDUCEDDOM : SAP_DIDNOTFIND);
      anycut |= cut[(size_t)q];
      if( result )
         result[q] = r;
      if( rounds )
         rounds[q] = ns > 1 ? (vr[(size_t)q] > 0 ? vr[(size_t)q] - vs + 1 : 0) : local.rounds;
   }
   if( ns == 1 && result )
      result[0] = local.result;
   if( stats )
   {
      *stats = local;
      if( ns > 1 )
         stats->result = anycut ? SAP_CUTOFF : (local.nchgbds > 0 ? SAP_REDUCEDDOM : SAP_DIDNOTFIND);
   }
   return SAP_OKAY;
}

// ---- continuous batching ----

static int slotsInit(sap_bounds* b)
{
   if( b->slotsmode )
      return SAP_OKAY;
   if( b->mat->nseg != 1 )
      return SAP_INVALIDDATA;
   RCCHK(ensureLog(b));
   RCCHK(clearLogs(b));
   b->logmax = b->nvec;   // any slot may hold log bits from here on
   b->active = new std::vector<int32_t>((size_t)b->nvec, 0);
   if( b->hstat == NULL )
   {
      HIPCHK(hipHostMalloc((void**)&b->hstat, sizeof(Status), hipHostMallocDefault));
      HIPCHK(hipHostMalloc((void**)&b->hpend, sizeof(int32_t) * 2 * (size_t)b->nvec, hipHostMallocDefault));
   }
   b->V.single = 0;
   b->V.logchg = 1;
   b->slotsmode = 1;
   return SAP_OKAY;
}

int sap_slots_admit(sap_bounds* b, int32_t n, const int32_t* slots, const sap_bounds* base, int32_t basevec,
   const int64_t* ovptr, const int32_t* ovcol, const double* ovlb, const double* ovub, int32_t flags)
{
   if( b == NULL || base == NULL || base->mat != b->mat || basevec < 0 || basevec >= base->nvec || n < 0
      || n > b->nvec || (n > 0 && slots == NULL) )
      return SAP_INVALIDDATA;
   if( base == b )
      return SAP_INVALIDDATA;   // the base must not be a slot
   sap_matrix* m = b->mat;
   HIPCHK(hipSetDevice(m->ctx->device));
   RCCHK(slotsInit(b));
   if( n == 0 )
      return SAP_OKAY;
   std::vector<char> seen((size_t)b->nvec, 0);
   for( int32_t i = 0; i < n; ++i )
   {
      const int32_t v = slots[i];
      if( v < 0 || v >= b->nvec || seen[(size_t)v] || (*b->active)[(size_t)v] != 0 )
         return SAP_INVALIDDATA;
      seen[(size_t)v] = 1;
   }
   int64_t nov = 0;
   RCCHK(stageOverlay(b, n, ovptr, ovcol, ovlb, ovub, &nov));
   RCCHK(upload(b, b->slotidx, slots, sizeof(int32_t) * (size_t)n));
   RCCHK(admitSlots(b, n, slots, b->slotidx, base, basevec, base->V.bnd + (size_t)basevec * (size_t)m->ncols, nov,
      (flags & SAP_FLAG_BASE_IS_FIXPOINT) ? 0 : 1));
   b->gen++;
   for( int32_t i = 0; i < n; ++i )
      (*b->active)[(size_t)slots[i]] = 1;
   // the staging buffers are reused by the next call: wait for the admission kernels
   HIPCHK(hipStreamSynchronize(m->ctx->stream));
   return SAP_OKAY;
}

int sap_slots_launch(sap_bounds* b, const sap_params* params, int32_t rounds)
{
   if( b == NULL || !b->slotsmode || (rounds != SAP_STEP_SHORT && rounds != SAP_STEP_LONG) )
      return b != NULL && !b->slotsmode ? SAP_INVALIDCALL : SAP_INVALIDDATA;
   RCCHK(checkParams(params));
   if( b->inflight )
      return SAP_INVALIDCALL;   // the previous step has not been waited for
   sap_matrix* m = b->mat;
   sap_context* c = m->ctx;
   HIPCHK(hipSetDevice(c->device));
   hipStream_t s = c->stream;
   const Params P = toParams(params);
   RCCHK(ensureRanged(b, P));
   HIPCHK(hipEventRecord(b->evs[0], s));
   k_step_begin<<<1, SAP_STAT_SLOTS, 0, s>>>(b->V);
   HIPCHK(hipGetLastError());
   // the step's rounds are enqueued directly, not launched as a captured graph: the server
   // steps many pools on several streams, and graphs saved only ~2 % of its saturated
   // capacity (37.0-37.2k against 37.8-38.3k requests/s on the serving workload)
   const Mat M = toMat(m);
   const Layout L = makeLayout(b);
   for( int r = 0; r < rounds; ++r )
      enqueueRound(b, M, P, L, s, NULL, NULL);
   HIPCHK(hipGetLastError());
   HIPCHK(hipMemcpyAsync(b->hstat, b->dstatus, sizeof(Status), hipMemcpyDeviceToHost, s));
   HIPCHK(hipMemcpyAsync(b->hpend, b->V.vpend, sizeof(int32_t) * (size_t)b->nvec, hipMemcpyDeviceToHost, s));
   HIPCHK(hipMemcpyAsync(b->hpend + b->nvec, b->V.vcutoff, sizeof(int32_t) * (size_t)b->nvec, hipMemcpyDeviceToHost,
      s));
   HIPCHK(hipEventRecord(b->evs[1], s));
   b->inflight = 1;
   return SAP_OKAY;
}

int sap_slots_query(sap_bounds* b, int32_t* ready)
{
   if( b == NULL || ready == NULL )
      return SAP_INVALIDDATA;
   if( !b->inflight )
   {
      *ready = 1;
      return SAP_OKAY;
   }
   const hipError_t e = hipEventQuery(b->evs[1]);
   if( e == hipErrorNotReady )
   {
      *ready = 0;
      return SAP_OKAY;
   }
   HIPCHK(e);
   *ready = 1;
   return SAP_OKAY;
}

int sap_slots_wait(sap_bounds* b, int32_t* nfinished, int32_t* finished, sap_stats* stats)
{
   if( b == NULL || !b->slotsmode || nfinished == NULL )
      return SAP_INVALIDDATA;
   *nfinished = 0;
   if( !b->inflight )
      return SAP_INVALIDCALL;
   HIPCHK(hipSetDevice(b->mat->ctx->device));
   HIPCHK(hipEventSynchronize(b->evs[1]));
   b->inflight = 0;
   const Status* h = b->hstat;
   int32_t nf = 0;
   for( int32_t v = 0; v < b->nvec; ++v )
   {
      if( (*b->active)[(size_t)v] != 1 )
         continue;
      const bool cut = b->hpend[b->nvec + v] != 0;
      const bool pending = b->hpend[v] > h->round;
      if( cut || !pending )
      {
         (*b->active)[(size_t)v] = 2;
         if( finished )
            finished[nf] = v;
         ++nf;
      }
   }
   *nfinished = nf;
   if( stats )
   {
      unsigned long long a = 0, r = 0, g = 0;
      for( int q = 0; q < SAP_STAT_SLOTS; ++q )
      {
         a += h->slot_nnz[q];
         r += h->slot_rows[q];
         g += h->slot_chg[q];
      }
      float ms = 0.0f;
      HIPCHK(hipEventElapsedTime(&ms, b->evs[0], b->evs[1]));
      memset(stats, 0, sizeof(*stats));
      stats->result = SAP_DIDNOTFIND;
      stats->rounds = h->rlocal;
      stats->nchgbds = (int64_t)g;
      stats->rows_swept = (int64_t)r;
      stats->nnz_swept = (int64_t)a;
      stats->device_ms = ms;
      stats->sweep_ms = (double)h->sweep_ticks / (double)b->mat->ctx->wallclock_khz;
      stats->sweep_launches = (int64_t)h->sweep_count;
   }
   return SAP_OKAY;
}

int sap_slots_step(sap_bounds* b, const sap_params* params, int32_t rounds, int32_t* nfinished, int32_t* finished,
   sap_stats* stats)
{
   RCCHK(sap_slots_launch(b, params, rounds));
   return sap_slots_wait(b, nfinished, finished, stats);
}

int sap_slots_collect(sap_bounds* b, int32_t n, const int32_t* slots, int32_t* result, int32_t* rounds,
   int64_t chgcap, int64_t* chgptr, int32_t* chgcol, double* chglb, double* chgub)
{
   if( b == NULL || !b->slotsmode || n < 0 || (n > 0 && slots == NULL) || (chgcol != NULL && chgptr == NULL)
      || chgcap < 0 )
      return SAP_INVALIDDATA;
   if( b->inflight )
      return SAP_INVALIDCALL;
   for( int32_t i = 0; i < n; ++i )
      if( slots[i] < 0 || slots[i] >= b->nvec || (*b->active)[(size_t)slots[i]] != 2 )
         return SAP_INVALIDDATA;
   if( n == 0 )
      return SAP_OKAY;
   HIPCHK(hipSetDevice(b->mat->ctx->device));
   RCCHK(collectSlots(b, n, slots, result, rounds, chgcap, chgptr, chgcol, chglb, chgub));
   for( int32_t i = 0; i < n; ++i )
      (*b->active)[(size_t)slots[i]] = 0;
   return SAP_OKAY;
}

int sap_propagate_device_noalloc(sap_matrix* m, const sap_params* params, double* dlb, double* dub, sap_stats* stats)
{
   if( m == NULL || dlb == NULL || dub == NULL )
      return SAP_INVALIDDATA;
   RCCHK(checkParams(params));
   if( m->nseg != 1 || m->ncols > SAP_NOALLOC_MAXCOLS || m->nrows > SAP_NOALLOC_MAXROWS )
   {
      fprintf(stderr, "[scip_amd] sap_propagate_device_noalloc: at most %lld columns and %lld rows of one instance "
         "(got %lld x %lld, %d instances)\n", (long long)SAP_NOALLOC_MAXCOLS, (long long)SAP_NOALLOC_MAXROWS,
         (long long)m->nrows, (long long)m->ncols, (int)m->nseg);
      return SAP_INVALIDDATA;
   }
   if( (params->flags & SAP_FLAG_RANGEDROWS) && !(params->flags & SAP_FLAG_NOTIGHTEN) && m->nranged > 0 )
   {
      fprintf(stderr, "[scip_amd] sap_propagate_device_noalloc: ranged-row propagation needs per-nonzero "
         "flags and is not available without allocation\n");
      return SAP_INVALIDDATA;
   }
   sap_context* c = m->ctx;
   HIPCHK(hipSetDevice(c->device));
   hipStream_t s = c->stream;
   const size_t nc = (size_t)m->ncols, nr = (size_t)m->nrows;
   // a host-side bound set whose device arrays are pieces of this device's scratch array
   sap_bounds b;
   memset(&b, 0, sizeof(b));
   b.mat = m;
   b.nvec = 1;
   b.graphbroken = 1;   // rounds from the host loop: capturing a graph would allocate
   b.internal = 1;
   Vec& V = b.V;
   V.nvec = 1;
   V.nseg = 1;
   V.single = 1;
   V.dense = 1;
   V.hasforced = m->hasforced;
   unsigned char* next = c->scratch;
   auto take = [&next](size_t bytes) {
      unsigned char* r = next;
      next += (bytes + 255) & ~(size_t)255;
      return (void*)r;
   };
   V.bnd = (double2*)take(16 * nc);
   V.clb = (long long*)take(8 * nc);
   V.cub = (long long*)take(8 * nc);
   if( m->hasforced )
   {
      V.flb = (long long*)take(8 * nc);
      V.fub = (long long*)take(8 * nc);
   }
   V.coltag = (int32_t*)take(4 * nc);
   V.chgtag = (int32_t*)take(4 * nc);
   V.rowmark = (int32_t*)take(4 * nr);
   V.redundant = (uint8_t*)take(nr);
   V.vcutoff = (int32_t*)take(4);
   V.vrounds = (int32_t*)take(4);
   V.vstart = (int32_t*)take(4);
   V.vpend = (int32_t*)take(4);
   V.vlogn = (int32_t*)take(4);
   b.dstatus = (Status*)take(sizeof(Status));
   V.st = b.dstatus;
   if( (size_t)(next - c->scratch) > SAP_NOALLOC_BYTES )
      return SAP_ERROR;   // cannot happen: the size covers the largest admitted matrix
   // the array keeps the last call's contents: everything a call reads before writing is
   // cleared with fill kernels (no memset nodes, so a caller may capture the call)
   k_fill_i64<<<gridFor(nc), SAP_BLOCK, 0, s>>>(V.clb, KEY_NEG_INF, nc);
   k_fill_i64<<<gridFor(nc), SAP_BLOCK, 0, s>>>(V.cub, KEY_POS_INF, nc);
   if( m->hasforced )
   {
      k_fill_i64<<<gridFor(nc), SAP_BLOCK, 0, s>>>(V.flb, KEY_NEG_INF, nc);
      k_fill_i64<<<gridFor(nc), SAP_BLOCK, 0, s>>>(V.fub, KEY_POS_INF, nc);
   }
   k_fill_i32<<<gridFor(nc), SAP_BLOCK, 0, s>>>(V.coltag, 0, nc);
   k_fill_i32<<<gridFor(nc), SAP_BLOCK, 0, s>>>(V.chgtag, 0, nc);
   k_fill_i32<<<gridFor(nr), SAP_BLOCK, 0, s>>>(V.rowmark, 0, nr);
   k_fill_i32<<<1, SAP_BLOCK, 0, s>>>(V.vpend, 0, 1);
   k_fill_i32<<<1, SAP_BLOCK, 0, s>>>(V.vstart, 0, 1);
   k_fill_i32<<<1, SAP_BLOCK, 0, s>>>(V.vlogn, 0, 1);
   k_fill_i32<<<gridFor(sizeof(Status) / 4), SAP_BLOCK, 0, s>>>((int32_t*)b.dstatus, 0, sizeof(Status) / 4);
   k_interleave<<<gridFor(nc), SAP_BLOCK, 0, s>>>(V.bnd, dlb, dub, (int64_t)nc);
   HIPCHK(hipGetLastError());
   sap_stats local;
   memset(&local, 0, sizeof(local));
   const Params P = toParams(params);
   RCCHK(beginCall(&b, 1, 1));   // generation 1: the redundancy flags are cleared too
   RCCHK(runRoundsHost(&b, P, &local));
   local.result = c->hstatus->cutoff ? SAP_CUTOFF : (c->hstatus->nchg > 0 ? SAP_REDUCEDDOM : SAP_DIDNOTFIND);
   if( local.result != SAP_CUTOFF )
   {
      k_deinterleave<<<gridFor(nc), SAP_BLOCK, 0, s>>>(V.bnd, dlb, dub, (int64_t)nc);
      HIPCHK(hipGetLastError());
   }
   HIPCHK(hipStreamSynchronize(s));
   if( stats )
      *stats = local;
   return SAP_OKAY;
}

int sap_bench_gather(sap_bounds* b, int32_t reps, int32_t blocks, double* ms_per_pass)
{
   if( b == NULL || reps < 1 || blocks < 0 || ms_per_pass == NULL )
      return SAP_INVALIDDATA;
   sap_matrix* m = b->mat;
   sap_context* c = m->ctx;
   HIPCHK(hipSetDevice(c->device));
   hipStream_t s = c->stream;
   double* sink = NULL;
   RCCHK(devAlloc(&sink, 1, NULL));
   const Mat M = toMat(m);
   // blocks 0: one lane per four nonzeros, no grid-stride loop (the most loads in flight)
   const size_t lanes = (size_t)((m->nnz + 3) / 4);
   const unsigned grid = blocks > 0 ? (unsigned)blocks : (unsigned)std::max<size_t>(1, (lanes + SAP_BLOCK - 1) / SAP_BLOCK);
   k_gather_floor<<<grid, SAP_BLOCK, 0, s>>>(M, b->V.bnd, m->nnz, sink);   // warm-up
   HIPCHK(hipEventRecord(c->ev[0], s));
   for( int32_t r = 0; r < reps; ++r )
      k_gather_floor<<<grid, SAP_BLOCK, 0, s>>>(M, b->V.bnd, m->nnz, sink);
   HIPCHK(hipEventRecord(c->ev[1], s));
   HIPCHK(hipGetLastError());
   HIPCHK(hipEventSynchronize(c->ev[1]));
   float ms = 0.0f;
   HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
   (void)hipFree(sink);
   *ms_per_pass = (double)ms / reps;
   return SAP_OKAY;
}

int sap_row_activities(sap_matrix* m, const sap_params* params, const double* lb, const double* ub, double* minact,
   double* maxact, int32_t* counts, double* maxactdelta)
{
   if( m == NULL || lb == NULL || ub == NULL )
      return SAP_INVALIDDATA;
   RCCHK(checkParams(params));
   sap_context* c = m->ctx;
   HIPCHK(hipSetDevice(c->device));
   hipStream_t s = c->stream;
   if( m->single == NULL )
      RCCHK(createSingle(m));
   RCCHK(sap_bounds_upload(m->single, 0, lb, ub));
   const int64_t nr = m->nrows;
   double *dmin = NULL, *dmax = NULL, *dmad = NULL;
   int32_t* dcnt = NULL;
   RCCHK(devAlloc(&dmin, (size_t)nr, NULL));
   RCCHK(devAlloc(&dmax, (size_t)nr, NULL));
   RCCHK(devAlloc(&dmad, (size_t)nr, NULL));
   RCCHK(devAlloc(&dcnt, 8 * (size_t)nr, NULL));
   const Params P = toParams(params);
   if( nr > 0 )
      k_row_activities<<<gridFor((size_t)nr * SAP_WAVE, 8192), SAP_BLOCK, 0, s>>>(toMat(m), P, m->single->V.bnd, m->perm,
         dmin, dmax, dcnt, dmad);
   HIPCHK(hipGetLastError());
   if( minact && nr > 0 )
      HIPCHK(hipMemcpyAsync(minact, dmin, sizeof(double) * (size_t)nr, hipMemcpyDeviceToHost, s));
   if( maxact && nr > 0 )
      HIPCHK(hipMemcpyAsync(maxact, dmax, sizeof(double) * (size_t)nr, hipMemcpyDeviceToHost, s));
   if( maxactdelta && nr > 0 )
      HIPCHK(hipMemcpyAsync(maxactdelta, dmad, sizeof(double) * (size_t)nr, hipMemcpyDeviceToHost, s));
   if( counts && nr > 0 )
      HIPCHK(hipMemcpyAsync(counts, dcnt, sizeof(int32_t) * 8 * (size_t)nr, hipMemcpyDeviceToHost, s));
   HIPCHK(hipStreamSynchronize(s));
   (void)hipFree(dmin);
   (void)hipFree(dmax);
   (void)hipFree(dmad);
   (void)hipFree(dcnt);
   return SAP_OKAY;
}

}  // extern "C"
