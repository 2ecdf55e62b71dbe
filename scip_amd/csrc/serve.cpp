// serve.cpp -- request serving on top of the propagation engine: resident problems, admission
// by HBM footprint with LRU eviction, continuous batching of node requests in per-problem slot
// pools over several HIP streams, and a real-time replay of an arrival trace (see
// include/scip_amd_serve.h).
//
// Uploads run on loader threads, each with a context (HIP stream) of its own, so that building
// a problem's matrix, root fixpoint and slot pool never stalls the scheduling loop: the loop
// reserves the problem's footprint (evicting idle problems), hands the problem to the loaders,
// keeps stepping the resident pools, and moves the finished matrix to its serving context
// (sap_matrix_set_context) when it collects it.  Several loaders build different problems at
// once: a build is partly serial host work (validation, packing, copies to the device).
//
// Host code only: every device operation goes through the engine's C ABI.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "scip_amd_prop.h"
#include "scip_amd_serve.h"

namespace {

struct Problem
{
   int64_t nrows = 0, ncols = 0;
   std::vector<int64_t> rowptr;
   std::vector<int32_t> colidx;
   std::vector<double> vals, lhs, rhs, lb, ub;
   std::vector<uint8_t> integral;
   int32_t stream = 0;        // context (HIP stream) the problem lives on
   sap_matrix* mat = NULL;    // NULL: not resident
   sap_bounds* root = NULL;   // root fixpoint (base of every request)
   sap_bounds* pool = NULL;   // slots
   int32_t slots = 0;         // slots of the pool
   int32_t rootcut = 0;       // the root is infeasible: every request is a cutoff
   int64_t bytes = 0;         // device bytes while resident
   uint64_t lastuse = 0;
   std::deque<int32_t> queue;         // waiting requests
   std::vector<int32_t> slotreq;      // slot -> request (-1: free)
   std::vector<int32_t> freeslots;
   int32_t running = 0;
   bool inflight = false;
   // upload by a loader thread: reserved and handed over, not yet collected
   bool loading = false;
   int loadrc = SAP_OKAY;
   sap_params lparams;            // parameters of the root propagation
   sap_matrix* lmat = NULL;       // what the loader built (on its own context)
   sap_bounds* lroot = NULL;
   sap_bounds* lpool = NULL;
   int32_t lrootcut = 0;
};

struct Request
{
   double arrival = 0.0;
   int32_t prob = 0;
   int32_t result = 0, rounds = 0, nchg = 0;
   double latency_ms = 0.0;
   std::vector<int32_t> col;
   std::vector<double> clb, cub;
   bool done = false;
   bool cold = false;   // its problem was uploaded while it waited
};

// an evicted problem's device objects, freed by a loader thread (freeing waits for the
// device, which the scheduling loop must not)
struct Dead
{
   sap_matrix* mat;
   sap_bounds* root;
   sap_bounds* pool;
};

}   // namespace

struct sap_server
{
   int32_t device = 0;
   std::vector<sap_context*> ctx;
   int64_t budget = 0;
   int32_t slots = 0;
   std::vector<Problem*> probs;
   int64_t resident = 0;
   uint64_t tick = 0;
   // per replay
   sap_serve_stats st;
   int32_t ndone = 0;
   std::vector<uint8_t> cold;   // per request of the last replay: waited for an upload
   // the loader threads and their contexts; jobs and finished uploads are guarded by mu
   std::vector<sap_context*> lctx;
   std::vector<std::thread> loaders;
   std::mutex mu;
   std::condition_variable cv;
   std::deque<Problem*> jobs;
   std::deque<Problem*> loaded;
   std::deque<Dead> frees;   // evicted problems' objects, freed by the loaders
   bool stop = false;
   int32_t nloading = 0;   // submitted and not yet collected (scheduling thread only)
};

namespace {

// loader threads: builds of different problems overlap (measured on the serving workload's
// 512 instances, a build of 3.9M nonzeros takes ~40 ms, of 16.8M ~130 ms, partly serial)
constexpr int SAP_SERVER_LOADERS = 3;

// device bytes of a resident problem (upper bound, before it is uploaded): matrix with CSC and
// tiles, the root vector and a slot pool of `slots` vectors with change logs
int64_t footprint(const Problem* p, int32_t slots)
{
   const int64_t nnz = p->rowptr.empty() ? 0 : p->rowptr.back();
   const int64_t mat = nnz * (8 + 4 + 4 + 1) + nnz / 4 + p->nrows * (8 + 16 + 1 + 1 + 1 + 4 + 4 + 8) + p->ncols * (1 + 8 + 4)
      + 4096;
   const int64_t vec1 = p->ncols * (16 + 16 + 16 + 4 + 4 + 32) + p->nrows * (4 + 1 + 16);
   const int64_t pool = (int64_t)slots * (p->ncols * (16 + 16 + 16 + 4 + 4 + 8 + 4 + 4 + 1) + p->nrows * (4 + 1 + 16))
      + p->ncols * 32 + 65536;
   return mat + vec1 + pool;
}

double nowSeconds()
{
   using namespace std::chrono;
   return duration<double>(steady_clock::now().time_since_epoch()).count();
}

void freeDead(const Dead& d)
{
   if( d.pool )
      sap_bounds_free(d.pool);
   if( d.root )
      sap_bounds_free(d.root);
   if( d.mat )
      sap_matrix_free(d.mat);
}

// releases p's device objects: they go to the loader threads, which free them
void evict(sap_server* s, Problem* p)
{
   Dead d = { p->mat, p->root, p->pool };
   p->pool = NULL;
   p->root = NULL;
   p->mat = NULL;
   s->resident -= p->bytes;
   p->bytes = 0;
   s->st.evictions++;
   {
      std::lock_guard<std::mutex> lock(s->mu);
      s->frees.push_back(d);
   }
   s->cv.notify_one();
}

// the least recently used resident problem without running, queued or in-flight work
Problem* lruVictim(sap_server* s)
{
   Problem* victim = NULL;
   for( Problem* q : s->probs )
      if( q->mat != NULL && q->running == 0 && q->queue.empty() && !q->inflight
         && (victim == NULL || q->lastuse < victim->lastuse) )
         victim = q;
   return victim;
}

// builds p's matrix, root fixpoint and slot pool on context c (a loader's); on failure
// nothing is left allocated
int buildProblem(Problem* p, sap_context* c)
{
   p->lmat = NULL;
   p->lroot = NULL;
   p->lpool = NULL;
   int rc = sap_matrix_create(c, p->nrows, p->ncols, p->rowptr.data(), p->colidx.data(), p->vals.data(),
      p->lhs.data(), p->rhs.data(), p->integral.data(), p->lb.data(), p->ub.data(), &p->lmat);
   if( rc != SAP_OKAY )
   {
      p->lmat = NULL;
      return rc;
   }
   rc = sap_bounds_create_mode(p->lmat, 1, SAP_MODE_DENSE, &p->lroot);
   if( rc == SAP_OKAY )
      rc = sap_bounds_upload(p->lroot, 0, p->lb.data(), p->ub.data());
   sap_stats rst;
   memset(&rst, 0, sizeof(rst));
   if( rc == SAP_OKAY )
   {
      sap_params rp = p->lparams;
      rp.flags = 0;
      rc = sap_propagate_resident(p->lroot, &rp, &rst);
   }
   if( rc == SAP_OKAY )
      rc = sap_bounds_create(p->lmat, p->slots, &p->lpool);
   if( rc != SAP_OKAY )
   {
      if( p->lpool )
         sap_bounds_free(p->lpool);
      if( p->lroot )
         sap_bounds_free(p->lroot);
      sap_matrix_free(p->lmat);
      p->lpool = NULL;
      p->lroot = NULL;
      p->lmat = NULL;
      return rc;
   }
   p->lrootcut = rst.result == SAP_CUTOFF ? 1 : 0;
   return SAP_OKAY;
}

// loader thread k: builds the problems handed to it, one after another, on context lctx[k]
void loaderMain(sap_server* s, int k)
{
   for( ;; )
   {
      Problem* p = NULL;
      std::deque<Dead> dead;
      {
         std::unique_lock<std::mutex> lock(s->mu);
         s->cv.wait(lock, [s] { return s->stop || !s->jobs.empty() || !s->frees.empty(); });
         dead.swap(s->frees);
         if( s->jobs.empty() && dead.empty() )
            return;   // stop requested and nothing left
         if( !s->jobs.empty() )
         {
            p = s->jobs.front();
            s->jobs.pop_front();
         }
      }
      // evicted problems first: their memory makes room for the next upload
      for( const Dead& d : dead )
         freeDead(d);
      if( p == NULL )
         continue;
      const int rc = buildProblem(p, s->lctx[(size_t)k]);
      {
         std::lock_guard<std::mutex> lock(s->mu);
         p->loadrc = rc;
         s->loaded.push_back(p);
      }
   }
}

// makes room for need bytes: 1 fits (reserved), 0 not now (the resident problems are busy)
int reserve(sap_server* s, int64_t need)
{
   while( s->resident + need > s->budget )
   {
      Problem* victim = lruVictim(s);
      if( victim == NULL )
         return 0;
      evict(s, victim);
   }
   s->resident += need;
   s->st.peak_bytes = std::max(s->st.peak_bytes, s->resident);
   return 1;
}

// the built problem takes its place: its matrix moves to the problem's serving context
int handOver(sap_server* s, Problem* p)
{
   p->loading = false;
   int rc = p->loadrc;
   if( rc == SAP_OKAY )
      rc = sap_matrix_set_context(p->lmat, s->ctx[(size_t)p->stream]);
   if( rc != SAP_OKAY )
   {
      if( p->lpool )
         sap_bounds_free(p->lpool);
      if( p->lroot )
         sap_bounds_free(p->lroot);
      if( p->lmat )
         sap_matrix_free(p->lmat);
      p->lpool = NULL;
      p->lroot = NULL;
      p->lmat = NULL;
      s->resident -= p->bytes;
      p->bytes = 0;
      return rc;
   }
   p->mat = p->lmat;
   p->root = p->lroot;
   p->pool = p->lpool;
   p->rootcut = p->lrootcut;
   p->lmat = NULL;
   p->lroot = NULL;
   p->lpool = NULL;
   p->slotreq.assign((size_t)p->slots, -1);
   p->freeslots.clear();
   for( int32_t k = p->slots - 1; k >= 0; --k )
      p->freeslots.push_back(k);
   p->running = 0;
   s->st.uploads++;
   return SAP_OKAY;
}

// starts an upload of p on a loader thread: 0 submitted or waiting for room, < 0 error
int submitLoad(sap_server* s, Problem* p, const sap_params* params)
{
   const int64_t need = footprint(p, p->slots);
   if( need > s->budget )
      return SAP_NOMEMORY;
   if( !reserve(s, need) )
      return 0;
   p->bytes = need;
   p->lparams = *params;
   p->loading = true;
   p->loadrc = SAP_OKAY;
   s->nloading++;
   {
      std::lock_guard<std::mutex> lock(s->mu);
      s->jobs.push_back(p);
   }
   s->cv.notify_one();
   return 0;
}

// hands over every upload the loaders have finished; the first error is returned
int collectLoads(sap_server* s)
{
   std::deque<Problem*> fin;
   {
      std::lock_guard<std::mutex> lock(s->mu);
      fin.swap(s->loaded);
   }
   int rc = SAP_OKAY;
   for( Problem* p : fin )
   {
      s->nloading--;
      const int r = handOver(s, p);
      if( r != SAP_OKAY && rc == SAP_OKAY )
         rc = r;
   }
   return rc;
}

// waits until the loaders have finished every submitted upload and hands them over
int drainLoads(sap_server* s)
{
   int rc = SAP_OKAY;
   while( s->nloading > 0 )
   {
      const int r = collectLoads(s);
      if( r != SAP_OKAY && rc == SAP_OKAY )
         rc = r;
      if( s->nloading > 0 )
         std::this_thread::sleep_for(std::chrono::microseconds(50));
   }
   return rc;
}

// 1: resident (uploaded now if it was not); 0: does not fit now; < 0 SAP error code.  The
// upload runs on this thread, on the first loader's context (the loaders have no work at the
// times this is called).
int makeResident(sap_server* s, Problem* p, const sap_params* params)
{
   if( p->mat != NULL )
      return 1;
   const int64_t need = footprint(p, p->slots);
   if( need > s->budget )
      return SAP_NOMEMORY;
   if( !reserve(s, need) )
      return 0;
   p->bytes = need;
   p->lparams = *params;
   p->loadrc = buildProblem(p, s->lctx[0]);
   const int rc = handOver(s, p);
   return rc == SAP_OKAY ? 1 : rc;
}

void finish(sap_server* s, Request& r, double t0, int32_t result, int32_t rounds)
{
   r.result = result;
   r.rounds = rounds;
   r.nchg = (int32_t)r.col.size();
   r.latency_ms = 1000.0 * (nowSeconds() - t0 - r.arrival);
   r.done = true;
   s->ndone++;
}

// admit up to n queued requests of p into free slots
int admit(sap_server* s, Problem* p, std::vector<Request>& req, const int64_t* ovptr, const int32_t* ovcol,
   const double* ovlb, const double* ovub, int32_t maxn, double t0)
{
   if( p->rootcut )
   {
      while( !p->queue.empty() )
      {
         finish(s, req[(size_t)p->queue.front()], t0, SAP_CUTOFF, 0);
         p->queue.pop_front();
      }
      return SAP_OKAY;
   }
   std::vector<int32_t> slots;
   std::vector<int64_t> ptr(1, 0);
   std::vector<int32_t> col;
   std::vector<double> lo, hi;
   while( !p->queue.empty() && !p->freeslots.empty() && (int32_t)slots.size() < maxn )
   {
      const int32_t i = p->queue.front();
      p->queue.pop_front();
      const int32_t k = p->freeslots.back();
      p->freeslots.pop_back();
      p->slotreq[(size_t)k] = i;
      slots.push_back(k);
      for( int64_t q = ovptr[i]; q < ovptr[i + 1]; ++q )
      {
         col.push_back(ovcol[q]);
         lo.push_back(ovlb[q]);
         hi.push_back(ovub[q]);
      }
      ptr.push_back((int64_t)col.size());
   }
   if( slots.empty() )
      return SAP_OKAY;
   const int rc = sap_slots_admit(p->pool, (int32_t)slots.size(), slots.data(), p->root, 0, ptr.data(),
      col.empty() ? NULL : col.data(), lo.empty() ? NULL : lo.data(), hi.empty() ? NULL : hi.data(),
      SAP_FLAG_BASE_IS_FIXPOINT);
   if( rc != SAP_OKAY )
      return rc;
   p->running += (int32_t)slots.size();
   p->lastuse = ++s->tick;
   s->st.admissions++;
   return SAP_OKAY;
}

// wait for p's step and collect its finished slots
int harvest(sap_server* s, Problem* p, std::vector<Request>& req, bool wantchg, double t0)
{
   int32_t nf = 0;
   std::vector<int32_t> fin((size_t)p->slots);
   sap_stats st;
   int rc = sap_slots_wait(p->pool, &nf, fin.data(), &st);
   p->inflight = false;
   if( rc != SAP_OKAY )
      return rc;
   s->st.nnz_swept += st.nnz_swept;
   s->st.rows_swept += st.rows_swept;
   if( nf == 0 )
      return SAP_OKAY;
   std::vector<int32_t> res((size_t)nf), rnd((size_t)nf);
   std::vector<int64_t> cptr((size_t)nf + 1);
   int64_t cap = 64 * (int64_t)nf + 1024;
   std::vector<int32_t> ccol;
   std::vector<double> clb, cub;
   for( ;; )
   {
      ccol.resize((size_t)cap);
      clb.resize(wantchg ? (size_t)cap : 1);
      cub.resize(wantchg ? (size_t)cap : 1);
      rc = sap_slots_collect(p->pool, nf, fin.data(), res.data(), rnd.data(), cap, cptr.data(), ccol.data(),
         wantchg ? clb.data() : NULL, wantchg ? cub.data() : NULL);
      if( rc == SAP_NOMEMORY && cptr[(size_t)nf] > cap )
      {
         cap = cptr[(size_t)nf];
         continue;
      }
      break;
   }
   if( rc != SAP_OKAY )
      return rc;
   for( int32_t q = 0; q < nf; ++q )
   {
      const int32_t k = fin[(size_t)q];
      Request& r = req[(size_t)p->slotreq[(size_t)k]];
      if( wantchg )
      {
         r.col.assign(ccol.begin() + cptr[(size_t)q], ccol.begin() + cptr[(size_t)q + 1]);
         r.clb.assign(clb.begin() + cptr[(size_t)q], clb.begin() + cptr[(size_t)q + 1]);
         r.cub.assign(cub.begin() + cptr[(size_t)q], cub.begin() + cptr[(size_t)q + 1]);
      }
      else
         r.col.resize((size_t)(cptr[(size_t)q + 1] - cptr[(size_t)q]));
      finish(s, r, t0, res[(size_t)q], rnd[(size_t)q]);
      p->slotreq[(size_t)k] = -1;
      p->freeslots.push_back(k);
   }
   p->running -= nf;
   p->lastuse = ++s->tick;
   return SAP_OKAY;
}

void waitUntil(double t)
{
   for( ;; )
   {
      const double d = t - nowSeconds();
      if( d <= 0.0 )
         return;
      if( d > 3e-4 )
         std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(1e6 * d) - 200));
   }
}

}   // namespace

extern "C" {

int sap_server_create(int32_t device, int32_t nstreams, int64_t hbm_budget, int32_t slots_per_pool, sap_server** out)
{
   if( out == NULL || nstreams < 1 || nstreams > 16 || hbm_budget <= 0 || slots_per_pool < 1 )
      return SAP_INVALIDDATA;
   *out = NULL;
   sap_server* s = new sap_server();
   s->device = device;
   s->budget = hbm_budget;
   s->slots = slots_per_pool;
   for( int32_t k = 0; k < nstreams + SAP_SERVER_LOADERS; ++k )
   {
      sap_context* c = NULL;
      const int rc = sap_context_create(device, &c);
      if( rc != SAP_OKAY )
      {
         for( sap_context* d : s->ctx )
            sap_context_free(d);
         for( sap_context* d : s->lctx )
            sap_context_free(d);
         delete s;
         return rc;
      }
      if( k < nstreams )
         s->ctx.push_back(c);
      else
         s->lctx.push_back(c);   // a loader's own context
   }
   for( int k = 0; k < SAP_SERVER_LOADERS; ++k )
      s->loaders.emplace_back(loaderMain, s, k);
   *out = s;
   return SAP_OKAY;
}

int sap_server_free(sap_server* s)
{
   if( s == NULL )
      return SAP_OKAY;
   (void)drainLoads(s);
   {
      std::lock_guard<std::mutex> lock(s->mu);
      s->stop = true;
   }
   s->cv.notify_all();
   for( std::thread& t : s->loaders )
      if( t.joinable() )
         t.join();   // each leaves once no job and no evicted problem is left
   for( Problem* p : s->probs )
   {
      if( p->mat )
         freeDead(Dead{ p->mat, p->root, p->pool });
      delete p;
   }
   for( sap_context* c : s->ctx )
      sap_context_free(c);
   for( sap_context* c : s->lctx )
      sap_context_free(c);
   delete s;
   return SAP_OKAY;
}

int sap_server_add_problem(sap_server* s, int64_t nrows, int64_t ncols, const int64_t* rowptr, const int32_t* colidx,
   const double* vals, const double* lhs, const double* rhs, const uint8_t* integral, const double* lb,
   const double* ub, int32_t* id)
{
   if( s == NULL || id == NULL || rowptr == NULL || nrows < 0 || ncols < 0 || lhs == NULL || rhs == NULL
      || integral == NULL || lb == NULL || ub == NULL || rowptr[0] != 0 || rowptr[nrows] < 0 )
      return SAP_INVALIDDATA;
   const int64_t nnz = rowptr[nrows];
   if( nnz > 0 && (colidx == NULL || vals == NULL) )
      return SAP_INVALIDDATA;
   Problem* p = new Problem();
   p->nrows = nrows;
   p->ncols = ncols;
   p->rowptr.assign(rowptr, rowptr + nrows + 1);
   p->colidx.assign(colidx, colidx + nnz);
   p->vals.assign(vals, vals + nnz);
   p->lhs.assign(lhs, lhs + nrows);
   p->rhs.assign(rhs, rhs + nrows);
   p->integral.assign(integral, integral + ncols);
   p->lb.assign(lb, lb + ncols);
   p->ub.assign(ub, ub + ncols);
   p->stream = (int32_t)(s->probs.size() % s->ctx.size());
   p->slots = s->slots;
   *id = (int32_t)s->probs.size();
   s->probs.push_back(p);
   return SAP_OKAY;
}

int sap_server_set_slots(sap_server* s, int32_t id, int32_t slots)
{
   if( s == NULL || id < 0 || id >= (int32_t)s->probs.size() || slots < 1 )
      return SAP_INVALIDDATA;
   Problem* p = s->probs[(size_t)id];
   if( p->running > 0 || !p->queue.empty() || p->inflight || p->loading )
      return SAP_INVALIDCALL;
   if( p->mat != NULL && p->slots != slots )
   {
      // the pool is sized at upload: release the problem, the next use uploads it again
      evict(s, p);
      s->st.evictions--;
   }
   p->slots = slots;
   return SAP_OKAY;
}

int sap_server_cold(sap_server* s, int32_t nreq, uint8_t* cold)
{
   if( s == NULL || cold == NULL || nreq != (int32_t)s->cold.size() )
      return SAP_INVALIDDATA;
   memcpy(cold, s->cold.data(), (size_t)nreq);
   return SAP_OKAY;
}

int sap_server_set_budget(sap_server* s, int64_t hbm_budget)
{
   if( s == NULL || hbm_budget <= 0 )
      return SAP_INVALIDDATA;
   s->budget = hbm_budget;
   while( s->resident > s->budget )
   {
      Problem* victim = lruVictim(s);
      if( victim == NULL )
         return SAP_NOMEMORY;
      evict(s, victim);
   }
   return SAP_OKAY;
}

int sap_server_footprint(sap_server* s, int32_t id, int64_t* bytes)
{
   if( s == NULL || bytes == NULL || id < 0 || id >= (int32_t)s->probs.size() )
      return SAP_INVALIDDATA;
   const Problem* p = s->probs[(size_t)id];
   *bytes = footprint(p, p->slots);
   return SAP_OKAY;
}

int sap_server_warm(sap_server* s, int32_t id)
{
   if( s == NULL || id < 0 || id >= (int32_t)s->probs.size() )
      return SAP_INVALIDDATA;
   sap_params prm;
   sap_params_default(&prm);
   Problem* p = s->probs[(size_t)id];
   const int r = makeResident(s, p, &prm);
   if( r == 1 )
      p->lastuse = ++s->tick;   // warming counts as a use for the LRU order
   return r == 1 ? SAP_OKAY : (r == 0 ? SAP_NOMEMORY : r);
}

int sap_server_replay(sap_server* s, const sap_params* params, int32_t mode, int32_t steprounds, int32_t nreq,
   const double* arrival_s, const int32_t* prob, const int64_t* ovptr, const int32_t* ovcol, const double* ovlb,
   const double* ovub, int32_t* result, int32_t* rounds, int32_t* nchg, double* latency_ms, int64_t chgcap,
   int64_t* chgptr, int32_t* chgcol, double* chglb, double* chgub, sap_serve_stats* stats)
{
   if( s == NULL || params == NULL || nreq < 0 || (nreq > 0 && (arrival_s == NULL || prob == NULL || ovptr == NULL))
      || (mode != SAP_SERVE_CONTINUOUS && mode != SAP_SERVE_PLAIN)
      || (steprounds != SAP_STEP_SHORT && steprounds != SAP_STEP_LONG) || (chgptr != NULL && chgcol == NULL) )
      return SAP_INVALIDDATA;
   for( int32_t i = 0; i < nreq; ++i )
      if( prob[i] < 0 || prob[i] >= (int32_t)s->probs.size() || (i > 0 && arrival_s[i] < arrival_s[i - 1])
         || ovptr[i + 1] < ovptr[i] )
         return SAP_INVALIDDATA;
   memset(&s->st, 0, sizeof(s->st));
   s->st.requests = nreq;
   s->st.peak_bytes = s->resident;
   s->ndone = 0;
   std::vector<Request> req((size_t)nreq);
   for( int32_t i = 0; i < nreq; ++i )
   {
      req[(size_t)i].arrival = arrival_s[i];
      req[(size_t)i].prob = prob[i];
   }
   const bool wantchg = chgptr != NULL;
   int rc = SAP_OKAY;
   const double t0 = nowSeconds();
   int32_t next = 0;
   if( mode == SAP_SERVE_PLAIN )
   {
      // one request at a time, in arrival order
      for( int32_t i = 0; i < nreq && rc == SAP_OKAY; ++i )
      {
         waitUntil(t0 + arrival_s[i]);
         Problem* p = s->probs[(size_t)prob[i]];
         req[(size_t)i].cold = p->mat == NULL;
         int r = makeResident(s, p, params);
         if( r != 1 )
         {
            rc = r == 0 ? SAP_NOMEMORY : r;
            break;
         }
         p->queue.push_back(i);
         rc = admit(s, p, req, ovptr, ovcol, ovlb, ovub, 1, t0);
         s->st.max_running = std::max<int64_t>(s->st.max_running, p->running);
         while( rc == SAP_OKAY && p->running > 0 )
         {
            rc = sap_slots_launch(p->pool, params, steprounds);
            if( rc == SAP_OKAY )
            {
               p->inflight = true;
               s->st.steps++;
               rc = harvest(s, p, req, wantchg, t0);
            }
         }
      }
   }
   else
   {
      // event-driven: a pool is collected as soon as its step has ended and, with requests
      // still running or waiting, stepped again at once; arrivals and admissions are handled
      // between these checks, so no pool waits for the steps of the others
      while( s->ndone < nreq && rc == SAP_OKAY )
      {
         // problems the loaders have finished join their serving contexts
         rc = collectLoads(s);
         if( rc != SAP_OKAY )
            break;
         const double now = nowSeconds() - t0;
         while( next < nreq && arrival_s[next] <= now )
         {
            Problem* p = s->probs[(size_t)prob[next]];
            req[(size_t)next].cold = p->mat == NULL;   // it waits for an upload
            p->queue.push_back(next);
            ++next;
         }
         // admissions into pools between steps (problems with waiting requests); a problem
         // that is not resident is handed to the loader, and its requests wait for it
         for( Problem* p : s->probs )
         {
            if( p->queue.empty() || p->inflight || p->loading )
               continue;
            if( p->mat == NULL )
            {
               rc = submitLoad(s, p, params);
               if( rc != SAP_OKAY && rc != 0 )
                  break;
               rc = SAP_OKAY;
               continue;
            }
            rc = admit(s, p, req, ovptr, ovcol, ovlb, ovub, p->slots, t0);
            if( rc != SAP_OKAY )
               break;
         }
         if( rc != SAP_OKAY )
            break;
         // a step for every pool with running slots and none in flight
         int64_t running = 0;
         for( Problem* p : s->probs )
         {
            running += p->running;
            if( p->running > 0 && !p->inflight )
            {
               rc = sap_slots_launch(p->pool, params, steprounds);
               if( rc != SAP_OKAY )
                  break;
               p->inflight = true;
               s->st.steps++;
            }
         }
         if( rc != SAP_OKAY )
            break;
         s->st.max_running = std::max(s->st.max_running, running);
         // collect the pools whose step has ended
         bool busy = false;
         for( Problem* p : s->probs )
         {
            if( !p->inflight )
               continue;
            int32_t ready = 0;
            rc = sap_slots_query(p->pool, &ready);
            if( rc != SAP_OKAY )
               break;
            if( ready )
               rc = harvest(s, p, req, wantchg, t0);
            else
               busy = true;
            if( rc != SAP_OKAY )
               break;
         }
         if( rc != SAP_OKAY )
            break;
         if( !busy && next < nreq )
         {
            bool waiting = false;
            for( Problem* p : s->probs )
               waiting = waiting || !p->queue.empty() || p->running > 0;
            if( !waiting )
               waitUntil(t0 + arrival_s[next]);
         }
      }
   }
   // uploads still running (an error ended the loop early) are waited for and kept
   const int drc = drainLoads(s);
   if( rc == SAP_OKAY )
      rc = drc;
   const double t1 = nowSeconds();
   s->st.wall_s = t1 - t0 - (nreq > 0 ? arrival_s[0] : 0.0);
   s->cold.assign((size_t)nreq, 0);
   for( int32_t i = 0; i < nreq; ++i )
      s->cold[(size_t)i] = req[(size_t)i].cold ? 1 : 0;
   std::vector<double> lat;
   int64_t tot = 0;
   for( int32_t i = 0; i < nreq; ++i )
   {
      const Request& r = req[(size_t)i];
      if( result )
         result[i] = r.result;
      if( rounds )
         rounds[i] = r.rounds;
      if( nchg )
         nchg[i] = r.nchg;
      if( latency_ms )
         latency_ms[i] = r.latency_ms;
      lat.push_back(r.latency_ms);
      if( chgptr )
      {
         chgptr[i] = tot;
         tot += (int64_t)r.col.size();
      }
   }
   if( chgptr )
   {
      chgptr[nreq] = tot;
      if( tot > chgcap )
      {
         if( rc == SAP_OKAY )
            rc = SAP_NOMEMORY;
      }
      else
         for( int32_t i = 0; i < nreq; ++i )
         {
            const Request& r = req[(size_t)i];
            for( size_t q = 0; q < r.col.size(); ++q )
            {
               chgcol[chgptr[i] + (int64_t)q] = r.col[q];
               if( chglb )
                  chglb[chgptr[i] + (int64_t)q] = r.clb[q];
               if( chgub )
                  chgub[chgptr[i] + (int64_t)q] = r.cub[q];
            }
         }
   }
   if( !lat.empty() )
   {
      std::sort(lat.begin(), lat.end());
      auto pct = [&](double f) { return lat[std::min(lat.size() - 1, (size_t)(f * (double)(lat.size() - 1) + 0.5))]; };
      s->st.p50_ms = pct(0.50);
      s->st.p99_ms = pct(0.99);
      s->st.max_ms = lat.back();
   }
   if( stats )
      *stats = s->st;
   return rc;
}

}   // extern "C"
