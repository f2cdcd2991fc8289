// serve.hip -- request server of the trj_hip library (trj_srv_*): continuous
// batching of independent RDF / S(Q) requests.  Each slot is a context (own
// HIP stream) driven by its own worker thread, so the host preparation of one
// request (q-set plan, staging) overlaps the device work of the others.  A
// request is admitted in FIFO order when a worker is free and the device
// memory it may need fits the HBM budget next to what every slot holds: a
// running slot holds its context's retained buffers plus the request's
// footprint, an idle slot its retained buffers (contexts keep what they grew).
// When the next request does not fit, the idle slots give their buffers back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/trj_hip.h"
#include "kernels.h"
#include "trj_internal.h"

namespace {

struct Req {
  trj_request r;
  int64_t footprint = 0;
  int state = 0;  // 0 queued, 1 running, 2 done
  int rc = TRJ_OK;
  std::string err;
  std::vector<int64_t> counts, nq;
  std::vector<double> sqf, sqfp;
};

}  // namespace

struct trj_srv {
  int device = 0;
  int64_t budget = 0;
  int64_t reserved = 0;       // sum of held[]
  int64_t peak = 0;           // largest `reserved` so far
  int64_t misses = 0;         // requests whose context grew past what was held for them
  int running = 0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int64_t> queue;
  std::unordered_map<int64_t, Req> reqs;
  std::vector<trj_ctx*> ctxs;
  std::vector<int64_t> retained;  // device bytes each slot's context holds between requests
  std::vector<int64_t> held;      // what each slot counts against the budget now
  std::vector<char> busy;
  std::vector<std::thread> workers;
  bool stop = false;
  int64_t next = 1;
  std::string err;
};

namespace {

int sfail(trj_srv* s, int code, const std::string& m) {
  if (s) s->err = m;
  return code;
}

// One request on one context: init for its parameters, the frame, results.
void execute(trj_ctx* ctx, Req& q) {
  const trj_request& r = q.r;
  int rc = TRJ_OK;
  int32_t ls = 0, nit = 0;
  trj_sq_info info{};
  // Every request initialises its slot's context afresh: no first-frame
  // autotuning of the pair kernels (it would time candidates per request).
  trj_rdf_params rp = r.rdf;
  rp.flags |= TRJ_RDF_NO_AUTOTUNE;
  trj_sq_params sp = r.sq;
  sp.flags |= TRJ_SQ_NO_AUTOTUNE;
  if (r.what & TRJ_WANT_RDF) rc = trj_rdf_init(ctx, &rp, r.box, &ls, &nit);
  if (rc == TRJ_OK && (r.what & TRJ_WANT_SQ)) rc = trj_sq_init(ctx, &sp, r.box, &info);
  if (rc == TRJ_OK) rc = trj_frame(ctx, r.pos, r.dtype, TRJ_AOS, r.itype, r.n, r.box, r.what);
  if (rc == TRJ_OK && (r.what & TRJ_WANT_RDF)) {
    q.counts.assign(static_cast<size_t>(ls) * nit, 0);
    rc = trj_rdf_counts(ctx, q.counts.data());
  }
  if (rc == TRJ_OK && (r.what & TRJ_WANT_SQ)) {
    const size_t ns = static_cast<size_t>(info.nqmax) + 1;
    q.sqf.assign(ns, 0.0);
    q.sqfp.assign(ns * r.sq.nsp, 0.0);
    q.nq.assign(ns, 0);
    rc = trj_sq_shells(ctx, q.sqf.data(), q.sqfp.data(), q.nq.data());
  }
  q.rc = rc;
  if (rc != TRJ_OK) q.err = trj_last_error(ctx);
}

void set_held(trj_srv* s, int k, int64_t v) {
  s->reserved += v - s->held[k];
  s->held[k] = v;
  s->peak = std::max(s->peak, s->reserved);
}

// Slot k may take a request of footprint fp when its context's retained
// buffers plus fp fit next to what the other slots hold.
bool fits(const trj_srv* s, int k, int64_t fp) {
  return s->reserved - s->held[k] + s->retained[k] + fp <= s->budget;
}

// Gives the buffers of every idle slot back to the device (called with the
// lock held: an idle slot's worker is waiting on the lock, so its context is
// not in use).
void release_idle(trj_srv* s) {
  for (size_t j = 0; j < s->ctxs.size(); ++j)
    if (!s->busy[j] && s->retained[j] > 0 && trj::ctx_release(s->ctxs[j]) == TRJ_OK) {
      s->retained[j] = 0;
      set_held(s, static_cast<int>(j), 0);
    }
}

void worker(trj_srv* s, int k) {
  (void)hipSetDevice(s->device);
  for (;;) {
    int64_t t;
    Req* q;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stop) return;
        if (!s->queue.empty()) {
          const int64_t fp = s->reqs[s->queue.front()].footprint;
          if (!fits(s, k, fp)) release_idle(s);
          // Every request's footprint is at most the budget (trj_srv_submit),
          // so with all idle slots released the request fits once the
          // running ones are done.
          if (fits(s, k, fp)) break;
        }
        s->cv.wait(lk);
      }
      t = s->queue.front();
      s->queue.pop_front();
      q = &s->reqs[t];  // element references are stable in unordered_map
      q->state = 1;
      s->busy[k] = 1;
      set_held(s, k, s->retained[k] + q->footprint);
      s->running += 1;
    }
    execute(s->ctxs[k], *q);
    const int64_t after = trj_device_bytes(s->ctxs[k]);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      q->state = 2;
      if (after > s->held[k]) s->misses += 1;
      s->retained[k] = after;
      s->busy[k] = 0;
      set_held(s, k, after);
      s->running -= 1;
    }
    s->cv.notify_all();
  }
}

}  // namespace

extern "C" {

int64_t trj_request_footprint(const trj_request* r) {
  // An upper bound on the device bytes a new context allocates for the
  // request (every buffer run_frame and the inits size for it, each with the
  // 1/8 headroom of the context's allocator).
  if (!r || r->n < 1) return 0;
  const int64_t n = r->n;
  const int nsp = std::max(1, std::max(r->rdf.nsp, r->sq.nsp));
  const int64_t nblk = (n + 255) / 256;
  int64_t b = n * 3 * 8 + n * 4;             // staged input, types
  b += n * 24 + 64;                          // wrapped FP64 positions, flags
  b += 2 * int64_t(4) * nsp * nblk + 4 * (nsp + 3);  // species sort
  if (r->what & TRJ_WANT_RDF) {
    const int nit = r->rdf.nsp * (r->rdf.nsp + 1) / 2;
    int32_t ls = 0;
    double rc = 0.0;
    if (trj_rdf_plan(&r->rdf, r->box, &ls, &rc) == TRJ_OK && rc > 0.0) {
      int64_t ncell = 1;
      for (int k = 0; k < r->rdf.ndim; ++k)
        ncell *= std::max<int64_t>(1, std::min<int64_t>(static_cast<int64_t>(r->box[k] / rc), 2048));
      b += int64_t(2048) * nit * ls * 4;     // per-workgroup histogram slices
      b += int64_t(nit) * ls * 8 + int64_t(ls + 2) * 8;  // totals, bin edges
      b += 12 * (ncell + 1);                 // cell counts, starts, cursors
      b += std::max<int64_t>(4 * n, ncell) * 4;  // per-workgroup cell counts of the LDS cell sort
    }
    b += n * (4 + 24 + 4);                   // cell ids, cell-sorted positions and types
  }
  if (r->what & TRJ_WANT_SQ) {
    trj::SqPlan S;
    std::string err;
    if (trj::sq_plan(r->sq, r->box, &S, &err) == TRJ_OK) {
      const int64_t tiles = static_cast<int64_t>(S.tiles.size());
      const int64_t nax = static_cast<int64_t>(S.axis_shell.size());
      // the largest split count a fresh context may ask for: the FP64
      // autotuner's shortest split length unless it is switched off
      int ks = tiles > 0 ? trj::contraction_splits(static_cast<int>(tiles), n, r->sq.precision, r->sq.nsp) : 0;
      if (tiles > 0 && r->sq.precision == TRJ_PREC_FP64 && !(r->sq.flags & TRJ_SQ_NO_AUTOTUNE))
        ks = std::max(ks, trj::contraction_splits(static_cast<int>(tiles), n, r->sq.precision, r->sq.nsp, 16384));
      const int64_t rho_entry = static_cast<int64_t>(trj::rho_entry_bytes(r->sq.precision));
      b += int64_t(ks) * r->sq.nsp * tiles * S.tk.vec() * rho_entry;  // rho partials
      b += n * 96;                                             // turn planes
      b += trj::sq_anchor_bytes(S, n);                         // FP64 anchor table
      b += tiles * (16 + 2 * S.tk.vec()) + 32;                 // tiles, shells of tile vectors
      b += int64_t(S.nqmax + 1) * 8 * (1 + r->sq.nsp);         // shell sums
      b += nax * 16 + static_cast<int64_t>(trj::sq_list_scratch_bytes(static_cast<int>(nax), r->sq.nsp));
    }
  }
  return b + b / 8 + (int64_t(1) << 20);
}

int trj_srv_create(trj_srv** out, int device, int32_t nslots, int64_t budget) {
  if (!out || nslots < 1 || nslots > 64) return TRJ_EINVAL;
  *out = nullptr;
  trj_srv* s = new trj_srv();
  s->device = device;
  s->budget = budget > 0 ? budget : (int64_t(64) << 30);
  s->ctxs.assign(nslots, nullptr);
  s->retained.assign(nslots, 0);
  s->held.assign(nslots, 0);
  s->busy.assign(nslots, 0);
  for (int k = 0; k < nslots; ++k) {
    const int rc = trj_create(&s->ctxs[k], device);
    if (rc != TRJ_OK) {
      trj_srv_destroy(s);
      return rc;
    }
  }
  for (int k = 0; k < nslots; ++k) s->workers.emplace_back(worker, s, k);
  *out = s;
  return TRJ_OK;
}

void trj_srv_destroy(trj_srv* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  for (std::thread& w : s->workers) w.join();
  for (trj_ctx* c : s->ctxs)
    if (c) trj_destroy(c);
  delete s;
}

const char* trj_srv_last_error(const trj_srv* s) { return s ? s->err.c_str() : "null server"; }

int trj_srv_submit(trj_srv* s, const trj_request* r, int64_t* ticket) {
  if (!s || !r || !ticket) return TRJ_EINVAL;
  if (r->n < 1 || !r->pos || !r->itype || !(r->what & (TRJ_WANT_RDF | TRJ_WANT_SQ)))
    return sfail(s, TRJ_EINVAL, "request: empty, NULL input or nothing asked");
  Req q;
  q.r = *r;
  q.footprint = trj_request_footprint(r);
  if (q.footprint > s->budget)
    return sfail(s, TRJ_ELIMIT, "request needs up to " + std::to_string(q.footprint) +
                                    " device bytes, above the server's HBM budget of " + std::to_string(s->budget));
  {
    std::lock_guard<std::mutex> lk(s->mu);
    const int64_t t = s->next++;
    s->reqs.emplace(t, std::move(q));
    s->queue.push_back(t);
    *ticket = t;
  }
  s->cv.notify_all();
  return TRJ_OK;
}

int trj_srv_poll(trj_srv* s, int64_t ticket, int32_t* done) {
  if (!s || !done) return TRJ_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return sfail(s, TRJ_EINVAL, "unknown ticket");
  *done = it->second.state == 2 ? 1 : 0;
  return TRJ_OK;
}

int trj_srv_result(trj_srv* s, int64_t ticket, int64_t* counts, double* sqf, double* sqfp, int64_t* nq) {
  if (!s) return TRJ_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return sfail(s, TRJ_EINVAL, "unknown ticket");
  Req& q = it->second;
  if (q.state != 2) return sfail(s, TRJ_ESTATE, "request not finished");
  const int rc = q.rc;
  if (rc == TRJ_OK) {
    if (counts) std::copy(q.counts.begin(), q.counts.end(), counts);
    if (sqf) std::copy(q.sqf.begin(), q.sqf.end(), sqf);
    if (sqfp) std::copy(q.sqfp.begin(), q.sqfp.end(), sqfp);
    if (nq) std::copy(q.nq.begin(), q.nq.end(), nq);
  } else {
    s->err = q.err;
  }
  s->reqs.erase(it);
  return rc;
}

int trj_srv_stats(const trj_srv* s_in, int32_t* queued, int32_t* running, int64_t* reserved) {
  if (!s_in) return TRJ_EINVAL;
  trj_srv* s = const_cast<trj_srv*>(s_in);
  std::lock_guard<std::mutex> lk(s->mu);
  if (queued) *queued = static_cast<int32_t>(s->queue.size());
  if (running) *running = s->running;
  if (reserved) *reserved = s->reserved;
  return TRJ_OK;
}

int trj_srv_memory(const trj_srv* s_in, int64_t* reserved, int64_t* peak, int64_t* misses) {
  if (!s_in) return TRJ_EINVAL;
  trj_srv* s = const_cast<trj_srv*>(s_in);
  std::lock_guard<std::mutex> lk(s->mu);
  if (reserved) *reserved = s->reserved;
  if (peak) *peak = s->peak;
  if (misses) *misses = s->misses;
  return TRJ_OK;
}

}  // extern "C"
