// serve.hip -- request server of the trj_hip library (trj_srv_*): continuous
// batching of independent RDF / S(Q) requests.  Each slot is a context (own
// HIP stream) driven by its own worker thread, so the host preparation of one
// request (q-set plan, staging) overlaps the device work of the others.  A
// request is admitted in FIFO order when a worker is free and its estimated
// device footprint fits next to the running requests.
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
  int64_t budget = 0, reserved = 0;
  int running = 0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int64_t> queue;
  std::unordered_map<int64_t, Req> reqs;
  std::vector<trj_ctx*> ctxs;
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
  if (r.what & TRJ_WANT_RDF) rc = trj_rdf_init(ctx, &r.rdf, r.box, &ls, &nit);
  if (rc == TRJ_OK && (r.what & TRJ_WANT_SQ)) rc = trj_sq_init(ctx, &r.sq, r.box, &info);
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

void worker(trj_srv* s, int k) {
  (void)hipSetDevice(s->device);
  for (;;) {
    int64_t t;
    Req* q;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] {
        if (s->stop) return true;
        if (s->queue.empty()) return false;
        const Req& h = s->reqs[s->queue.front()];
        return s->running == 0 || s->reserved + h.footprint <= s->budget;
      });
      if (s->stop) return;
      t = s->queue.front();
      s->queue.pop_front();
      q = &s->reqs[t];  // element references are stable in unordered_map
      q->state = 1;
      s->reserved += q->footprint;
      s->running += 1;
    }
    execute(s->ctxs[k], *q);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      q->state = 2;
      s->reserved -= q->footprint;
      s->running -= 1;
    }
    s->cv.notify_all();
  }
}

}  // namespace

extern "C" {

int64_t trj_request_footprint(const trj_request* r) {
  if (!r || r->n < 1) return 0;
  const int64_t n = r->n;
  int64_t b = n * 3 * 8 + n * 4;  // staged input
  b += n * 24 + n * 12;           // wrapped positions, species sort
  if (r->what & TRJ_WANT_RDF) {
    const int nit = r->rdf.nsp * (r->rdf.nsp + 1) / 2;
    int32_t ls = 0;
    double rc = 0.0;
    if (trj_rdf_plan(&r->rdf, r->box, &ls, &rc) == TRJ_OK) b += int64_t(2048) * nit * ls * 4 + nit * ls * 8;
    b += n * 40;  // cell list
  }
  if (r->what & TRJ_WANT_SQ) {
    trj_sq_info info;
    if (trj_sq_plan(&r->sq, r->box, &info, nullptr) == TRJ_OK) {
      const trj::TileShape tk = trj::tile_shape(r->sq.precision);
      const int64_t tiles = (info.nvec_grid + tk.vec() - 1) / tk.vec() + 1;
      int dev = 0, slots = 0;  // the splits depend on the CU count (16-bit)
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&slots, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        slots = 0;
      const int ks = trj::contraction_splits(static_cast<int>(tiles), n, r->sq.precision, r->sq.nsp, slots);
      b += int64_t(ks) * r->sq.nsp * tiles * tk.vec() * 16 + n * 96;
    }
  }
  return b;
}

int trj_srv_create(trj_srv** out, int device, int32_t nslots, int64_t budget) {
  if (!out || nslots < 1 || nslots > 64) return TRJ_EINVAL;
  *out = nullptr;
  trj_srv* s = new trj_srv();
  s->device = device;
  s->budget = budget > 0 ? budget : (int64_t(64) << 30);
  s->ctxs.assign(nslots, nullptr);
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

}  // extern "C"
