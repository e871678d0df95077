// server.hip -- the serving loop: continuous batching of create/repair
// requests over one device.
//
// The reference handles one PAR2 set per Par2Creator / Par2Repairer call
// (src/creator.rs:261-531, src/repair.rs:736-1125); a server in front of it
// would queue whole calls.  Here requests are admitted in arrival order while
// their HBM footprint fits the budget, and admitted requests are run as
// batches (par2hip_batch_run_dev) by a few lane threads, each with its own
// context (streams, parameter block, scratch): a request that arrives while
// other batches are on the device joins the next batch a free lane forms, and
// batches of different lanes overlap on the device.
//
// Outputs (recovery or rebuilt slices, MD5/CRC32 arrays) come from a device
// pool owned by the server (best-fit free list, 256-byte granules) and stay
// resident until the client releases the request.  The input slices are
// device-resident and owned by the client; their bytes are charged to the
// budget for the request's lifetime, as is an estimate of the batch working
// memory (coefficients, inverses and nibble tables).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "pool.hpp"

namespace par2hip {
namespace {

enum class State { kPending, kAdmitted, kRunning, kDone };

struct Request {
  uint64_t ticket = 0;
  par2hip_job job{};  // output pointers filled at admission
  std::vector<uint32_t> exps, missing;
  int32_t hashes = 0;
  uint64_t footprint = 0, out_bytes = 0, hash_bytes = 0, work_bytes = 0;
  uint64_t out_off = UINT64_MAX, hash_off = UINT64_MAX;
  State state = State::kPending;
  int32_t status = PAR2HIP_OK;
  int32_t batch_size = 0;
  double t_arrival = 0, t_submit = 0, t_admit = 0, t_start = 0, t_done = 0;
};

// HBM bytes a request holds while it is in the server: its input slices, its
// outputs and its share of the batch working memory (coefficient matrix,
// inverse blocks, nibble tables; repairs with gapped exponents take the
// single-set path, whose working memory is bounded by the same terms).
uint64_t footprint_of(const par2hip_job &j, int hashes, uint64_t *out_bytes, uint64_t *hash_bytes,
                      uint64_t *work_bytes) {
  const uint64_t n = j.n, s = j.slice_size, c = j.count;
  *out_bytes = c * s;
  *hash_bytes = hashes ? n * 20 : 0;
  uint64_t work = batch_job_arena_bytes(j);
  if (work == 0 && j.kind == PAR2HIP_JOB_REPAIR && c)
    work = 4 * c * c + 2 * c * (n + 8) * 2 + (uint64_t)gf_tab_entries((int)c, (int)n) * 16 * 2;
  *work_bytes = work;
  return n * s + *out_bytes + *hash_bytes + work;
}

}  // namespace
}  // namespace par2hip

using par2hip::Request;
using par2hip::State;

struct par2hip_server {
  int32_t device = 0;
  uint64_t budget = 0, charged = 0;
  size_t max_batch = 0;
  void *pool_base = nullptr;
  par2hip::RangePool pool;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> pending, admitted;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::vector<par2hip_ctx *> ctxs;
  std::vector<std::thread> lanes;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  par2hip_server_stats stats{};

  uint8_t *ptr(uint64_t off) const { return static_cast<uint8_t *>(pool_base) + off; }

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }

  // Admit pending requests in arrival order while the budget and the pool
  // allow (the first one that does not fit blocks the ones behind it, so
  // large requests are not starved).  Caller holds mu.
  void try_admit() {
    bool any = false;
    while (!pending.empty()) {
      Request *r = pending.front();
      if (charged + r->footprint > budget) break;
      uint64_t oo = UINT64_MAX, ho = UINT64_MAX;
      if (r->out_bytes && (oo = pool.alloc(r->out_bytes)) == UINT64_MAX) break;
      if (r->hash_bytes && (ho = pool.alloc(r->hash_bytes)) == UINT64_MAX) {
        if (oo != UINT64_MAX) pool.release(oo);
        break;
      }
      r->out_off = oo;
      r->hash_off = ho;
      if (r->job.kind == PAR2HIP_JOB_ENCODE)
        r->job.d_recovery = oo == UINT64_MAX ? nullptr : ptr(oo);
      else
        r->job.d_out = oo == UINT64_MAX ? nullptr : ptr(oo);
      if (ho != UINT64_MAX) {
        r->job.d_md5 = ptr(ho);
        r->job.d_crc = ptr(ho + r->job.n * 16);
      }
      charged += r->footprint;
      stats.peak_charged = std::max(stats.peak_charged, charged);
      r->state = State::kAdmitted;
      r->t_admit = now_ms();
      pending.pop_front();
      admitted.push_back(r);
      any = true;
    }
    if (any) cv_work.notify_all();
  }

  // One lane: take everything admitted so far (up to max_batch) as one
  // batch, run it on this lane's context, publish the results.
  void lane_loop(par2hip_ctx *ctx) {
    (void)hipSetDevice(device);
    std::vector<Request *> batch;
    std::vector<par2hip_job> jobs;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !admitted.empty(); });
        if (admitted.empty()) return;  // stop requested and nothing left
        batch.clear();
        while (!admitted.empty() && batch.size() < max_batch) {
          batch.push_back(admitted.front());
          admitted.pop_front();
        }
        const double t = now_ms();
        for (Request *r : batch) {
          r->state = State::kRunning;
          r->t_start = t;
          r->batch_size = (int32_t)batch.size();
        }
        stats.batches++;
        stats.max_batch_seen = std::max<uint64_t>(stats.max_batch_seen, batch.size());
      }
      jobs.clear();
      uint64_t work = par2hip::kBatchArenaExtra;
      for (Request *r : batch) {
        jobs.push_back(r->job);
        work += r->work_bytes;
      }
      // the batch working memory comes from the pool (its bytes are already
      // charged to the requests); the lane's own scratch only when the pool
      // is too fragmented to give it in one piece
      uint64_t arena_off;
      {
        std::lock_guard<std::mutex> lk(mu);
        arena_off = pool.alloc(work);
        if (arena_off == UINT64_MAX) stats.arena_fallbacks++;
      }
      int32_t rc;
      {
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        (void)hipSetDevice(device);
        rc = par2hip::batch_run(ctx, jobs.data(), jobs.size(), arena_off == UINT64_MAX ? nullptr : ptr(arena_off),
                                arena_off == UINT64_MAX ? 0 : work);
      }
      if (rc == PAR2HIP_OK) rc = par2hip_sync(ctx);
      if (rc != PAR2HIP_OK) (void)par2hip_sync(ctx);  // nothing of this batch may still run
      if (arena_off != UINT64_MAX) {
        std::lock_guard<std::mutex> lk(mu);
        pool.release(arena_off);
      }
      std::vector<int32_t> status(batch.size(), rc);
      if (rc != PAR2HIP_OK && batch.size() > 1) {
        // isolate the failing request(s): run each one alone
        for (size_t i = 0; i < batch.size(); i++) {
          status[i] = par2hip_batch_run_dev(ctx, &jobs[i], 1);
          if (status[i] == PAR2HIP_OK) status[i] = par2hip_sync(ctx);
        }
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        const double t = now_ms();
        for (size_t i = 0; i < batch.size(); i++) {
          batch[i]->status = status[i];
          batch[i]->t_done = t;
          batch[i]->state = State::kDone;
          stats.completed++;
          if (status[i] != PAR2HIP_OK) stats.failed++;
        }
      }
      cv_done.notify_all();
    }
  }
};

extern "C" {

int32_t par2hip_server_create(int32_t device, uint64_t hbm_budget, int32_t lanes, int32_t max_batch,
                              par2hip_server **out) {
  if (!out || hbm_budget == 0 || lanes < 1 || lanes > 8 || max_batch < 1) return PAR2HIP_EINVAL;
  *out = nullptr;
  if (par2hip_available() <= 0) return PAR2HIP_ENODEV;
  auto srv = std::make_unique<par2hip_server>();
  srv->device = device;
  srv->budget = hbm_budget;
  srv->max_batch = (size_t)max_batch;
  (void)hipSetDevice(device);
  const uint64_t pool_bytes = (hbm_budget + par2hip::kGranule - 1) & ~(par2hip::kGranule - 1);
  if (hipMalloc(&srv->pool_base, pool_bytes) != hipSuccess) return PAR2HIP_ENOMEM;
  srv->pool.init(pool_bytes);
  int32_t rc;
  for (int i = 0; i < lanes; i++) {
    par2hip_ctx *c = nullptr;
    if ((rc = par2hip::ctx_create(device, false, &c))) {  // lanes run batches only
      for (par2hip_ctx *x : srv->ctxs) par2hip_ctx_destroy(x);
      (void)hipFree(srv->pool_base);
      return rc;
    }
    srv->ctxs.push_back(c);
  }
  par2hip_server *s = srv.release();
  for (par2hip_ctx *c : s->ctxs) s->lanes.emplace_back([s, c] { s->lane_loop(c); });
  *out = s;
  return PAR2HIP_OK;
}

int32_t par2hip_server_destroy(par2hip_server *srv) {
  if (!srv) return PAR2HIP_EINVAL;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
    srv->pending.clear();  // never admitted: nothing on the device
  }
  srv->cv_work.notify_all();
  for (std::thread &t : srv->lanes) t.join();
  for (par2hip_ctx *c : srv->ctxs) par2hip_ctx_destroy(c);
  (void)hipFree(srv->pool_base);
  delete srv;
  return PAR2HIP_OK;
}

double par2hip_server_now_ms(par2hip_server *srv) { return srv ? srv->now_ms() : 0.0; }

int32_t par2hip_server_submit(par2hip_server *srv, const par2hip_request *req, uint64_t *ticket) {
  if (!srv || !req || !ticket) return PAR2HIP_EINVAL;
  const par2hip_job &j = req->job;
  if (!j.d_slices || j.n == 0 || j.n > 32768 || j.slice_size == 0 || (j.slice_size & 1)) return PAR2HIP_EINVAL;
  if (j.kind != PAR2HIP_JOB_ENCODE && j.kind != PAR2HIP_JOB_REPAIR) return PAR2HIP_EINVAL;
  if (j.count && !j.exps) return PAR2HIP_EINVAL;
  if (j.kind == PAR2HIP_JOB_ENCODE && j.count > 65535) return PAR2HIP_EINVAL;
  if (req->hashes && j.slice_size >= par2hip::kMd5MaxSliceBytes) return PAR2HIP_EINVAL;
  if (j.kind == PAR2HIP_JOB_REPAIR) {
    if (j.count > j.n || (j.count && (!j.missing || !j.d_recovery))) return PAR2HIP_EINVAL;
    std::vector<char> seen(j.n, 0);
    for (uint64_t k = 0; k < j.count; k++) {
      if (j.missing[k] >= j.n || seen[j.missing[k]]) return PAR2HIP_EINVAL;
      seen[j.missing[k]] = 1;
    }
  }
  auto r = std::make_unique<Request>();
  r->job = j;
  r->hashes = req->hashes ? 1 : 0;
  r->exps.assign(j.exps ? j.exps : nullptr, j.exps ? j.exps + j.count : nullptr);
  r->job.exps = r->exps.data();
  if (j.kind == PAR2HIP_JOB_REPAIR) {
    r->missing.assign(j.missing ? j.missing : nullptr, j.missing ? j.missing + j.count : nullptr);
    r->job.missing = r->missing.data();
    r->job.d_out = nullptr;
  } else {
    r->job.d_recovery = nullptr;
    r->job.missing = nullptr;
    r->job.d_out = nullptr;
  }
  r->job.d_md5 = nullptr;
  r->job.d_crc = nullptr;
  r->footprint = par2hip::footprint_of(j, r->hashes, &r->out_bytes, &r->hash_bytes, &r->work_bytes);
  std::lock_guard<std::mutex> lk(srv->mu);
  if (r->footprint > srv->budget || r->out_bytes + r->hash_bytes + 2 * par2hip::kGranule > srv->pool.size())
    return PAR2HIP_ENOMEM;  // could never be admitted
  r->ticket = srv->next_ticket++;
  r->t_submit = srv->now_ms();
  r->t_arrival = req->arrival_ms > 0 ? req->arrival_ms : r->t_submit;
  *ticket = r->ticket;
  Request *raw = r.get();
  srv->reqs.emplace(raw->ticket, std::move(r));
  srv->pending.push_back(raw);
  srv->stats.submitted++;
  srv->try_admit();
  return PAR2HIP_OK;
}

int32_t par2hip_server_wait_timeout(par2hip_server *srv, uint64_t ticket, double timeout_ms, par2hip_result *res) {
  if (!srv || !res || !(timeout_ms >= 0)) return PAR2HIP_EINVAL;
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) return PAR2HIP_EINVAL;
  Request *r = it->second.get();
  auto done = [&] { return r->state == State::kDone; };
  if (std::isinf(timeout_ms)) {
    srv->cv_done.wait(lk, done);
  } else {
    const auto limit = std::chrono::steady_clock::now() +
                       std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                           std::chrono::duration<double, std::milli>(std::min(timeout_ms, 1e12)));
    if (!srv->cv_done.wait_until(lk, limit, done)) return PAR2HIP_ETIMEDOUT;
  }
  res->status = r->status;
  res->batch_size = r->batch_size;
  res->d_out = r->out_off == UINT64_MAX ? nullptr : srv->ptr(r->out_off);
  res->d_md5 = r->job.d_md5;
  res->d_crc = r->job.d_crc;
  res->t_arrival_ms = r->t_arrival;
  res->t_submit_ms = r->t_submit;
  res->t_admit_ms = r->t_admit;
  res->t_start_ms = r->t_start;
  res->t_done_ms = r->t_done;
  return PAR2HIP_OK;
}

int32_t par2hip_server_wait(par2hip_server *srv, uint64_t ticket, par2hip_result *res) {
  return par2hip_server_wait_timeout(srv, ticket, HUGE_VAL, res);
}

int32_t par2hip_server_release(par2hip_server *srv, uint64_t ticket) {
  if (!srv) return PAR2HIP_EINVAL;
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) return PAR2HIP_EINVAL;
  Request *r = it->second.get();
  if (r->state != State::kDone) return PAR2HIP_EINVAL;
  if (r->out_off != UINT64_MAX) srv->pool.release(r->out_off);
  if (r->hash_off != UINT64_MAX) srv->pool.release(r->hash_off);
  srv->charged -= r->footprint;
  srv->reqs.erase(it);
  srv->try_admit();
  return PAR2HIP_OK;
}

int32_t par2hip_server_get_stats(par2hip_server *srv, par2hip_server_stats *st) {
  if (!srv || !st) return PAR2HIP_EINVAL;
  std::lock_guard<std::mutex> lk(srv->mu);
  *st = srv->stats;
  st->charged = srv->charged;
  st->pending = srv->pending.size();
  st->admitted = srv->admitted.size();
  return PAR2HIP_OK;
}

}  // extern "C"
