// server.cpp -- render service with continuous batching (C-ABI mr_server_*).
//
// One worker thread owns `slots` execution slots (scratch + stream).  It
// admits queued requests in arrival order into the next free slot while the
// HBM budget allows, launches the batch (all stages of the pipeline once for
// the whole batch) and the read-back of its outputs into pinned staging, and
// returns to admitting: requests arriving while batches run join the next
// batch, and batches on different slots overlap on the device.  When a
// slot's completion event fires, outputs are copied to the callers' buffers
// and waiters are released.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "minirender_amd.h"
#include "context.h"

using namespace mr;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t ticket = 0;
  mr_frame frame{};
  std::vector<mr_draw> draws;
  mr_outputs out{};
  uint64_t mem = 0;
  double submit_us = 0, launch_us = 0, done_us = 0;
  int64_t batch_id = -1;
  int32_t batch_size = 0;
  int32_t status = MR_OK;
  bool done = false;
};

struct Flight {
  Slot slot;
  bool busy = false;
  BatchPlan plan;
  std::vector<Request*> reqs;
  std::vector<mr_outputs> outs;  // device staging per request
  std::vector<size_t> offs;
  std::vector<char> staged;  // per request x output kind: read back via staging
  uint64_t mem = 0;
  int64_t batch_id = -1;
};

// Device memory a request holds while in flight: outputs, winner ids, edge
// checkpoints, per-triangle and per-tile scratch (items 32 B, bins ~12 B,
// regions; tile counters 2 x 64 B + 16 B), a conservative estimate.
uint64_t request_bytes(mr_ctx* c, const mr_frame& f, const mr_outputs& o) {
  const uint64_t npix = (uint64_t)f.width * f.height;
  uint64_t b = npix * 20;  // winner ids, edge checkpoints
  if (o.rgb8) b += npix * 3;
  if (o.depth) b += npix * 4;
  if (o.winner) b += npix * 4;
  if (o.rgb) b += npix * 12;
  if (o.normals) b += npix * 12;
  if (o.points) b += npix * 12;
  uint64_t tris = 0;
  for (int d = 0; d < f.ndraws; d++) {
    const int h = f.draws[d].mesh;
    if (h >= 0 && h < (int)c->meshes.size()) tris += (uint64_t)c->meshes[h].ntri;
  }
  const uint64_t tiles = ((uint64_t)(f.width + c->tile_w - 1) / c->tile_w) *
                         ((uint64_t)(f.height + c->tile_h - 1) / c->tile_h);
  return b + tris * 64 + tiles * 160;  // + 16 B per triangle: checkpoint bases
}

}  // namespace

struct mr_server {
  mr_ctx* ctx = nullptr;
  mr_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> live;
  std::vector<std::unique_ptr<Flight>> flights;
  uint64_t next_ticket = 1;
  uint64_t mem_in_flight = 0;
  int64_t next_batch = 0;
  bool stop = false;
  std::thread worker;
  Clock::time_point t0;

  double now_us() const {
    return std::chrono::duration<double, std::micro>(Clock::now() - t0).count();
  }
  void run();
  bool launch(Flight& f, std::vector<Request*>& reqs);
  int queue_readback(Flight& f);
  void complete(Flight& f);
};

// True for host memory the HIP runtime has pinned (hipHostMalloc, torch's
// pin_memory): the device can copy straight into it.
static bool is_pinned(const void* p) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeHost;
}

static const size_t kOutBytes[kOutKinds] = {3, 4, 4, 12, 12, 12};

// Queue the read-back of every host output of the batch on the slot's
// stream: pinned destinations directly, the rest into pinned staging (copied
// out by complete()).  Marks which outputs need the staging copy.
int mr_server::queue_readback(Flight& f) {
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  f.staged.assign(f.reqs.size() * kOutKinds, 0);
  size_t need = 0;
  for (size_t j = 0; j < f.reqs.size(); j++) {
    const mr_outputs& o = f.reqs[j]->out;
    if (o.on_device) continue;
    const size_t npix = (size_t)f.plan.jobs[j].W * f.plan.jobs[j].H;
    const void* dst[kOutKinds] = {o.rgb8, o.depth, o.winner, o.rgb, o.normals, o.points};
    for (int k = 0; k < kOutKinds; k++)
      if (dst[k] && !is_pinned(dst[k])) {
        f.staged[j * kOutKinds + k] = 1;
        need = std::max(need, f.offs[j * kOutKinds + k] + npix * kOutBytes[k]);
      }
  }
  if (need && host_alloc(f.slot.sc.h_outputs, need)) return MR_ERR_NOMEM;
  for (size_t j = 0; j < f.reqs.size(); j++) {
    const mr_outputs& o = f.reqs[j]->out;
    if (o.on_device) continue;
    const size_t npix = (size_t)f.plan.jobs[j].W * f.plan.jobs[j].H;
    void* dst[kOutKinds] = {o.rgb8, o.depth, o.winner, o.rgb, o.normals, o.points};
    for (int k = 0; k < kOutKinds; k++) {
      if (!dst[k]) continue;
      const size_t off = f.offs[j * kOutKinds + k];
      void* to = f.staged[j * kOutKinds + k] ? (char*)f.slot.sc.h_outputs.p + off : dst[k];
      if (hipMemcpyAsync(to, (const char*)f.slot.sc.outputs.p + off, npix * kOutBytes[k],
                         hipMemcpyDeviceToHost, f.slot.stream) != hipSuccess) {
        (void)hipGetLastError();
        return MR_ERR_HIP;
      }
    }
  }
  return MR_OK;
}

bool mr_server::launch(Flight& f, std::vector<Request*>& reqs) {
  std::vector<mr_frame> frames;
  frames.reserve(reqs.size());
  f.outs.assign(reqs.size(), mr_outputs{});
  for (size_t i = 0; i < reqs.size(); i++) {
    frames.push_back(reqs[i]->frame);
    f.outs[i] = reqs[i]->out;
  }
  f.reqs = reqs;
  int rc;
  {
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    hipSetDevice(ctx->device);
    rc = build_batch(ctx, frames.data(), (int)frames.size(), f.plan);
    if (!rc) rc = bind_outputs(ctx, f.slot, f.plan, f.outs.data(), f.offs);
    if (!rc) rc = enqueue_batch(ctx, f.slot, f.plan, false);
    if (!rc) rc = queue_readback(f);
    if (!rc && hipEventRecord(f.slot.done, f.slot.stream) != hipSuccess) rc = MR_ERR_HIP;
  }
  const double t = now_us();
  f.batch_id = next_batch++;
  for (Request* r : reqs) {
    r->launch_us = t;
    r->batch_id = f.batch_id;
    r->batch_size = (int32_t)reqs.size();
  }
  if (rc) {  // fail the whole batch
    if (f.slot.stream) hipStreamSynchronize(f.slot.stream);
    std::lock_guard<std::mutex> lk(mu);
    for (Request* r : reqs) {
      r->status = rc;
      r->done = true;
      r->done_us = now_us();
    }
    mem_in_flight -= f.mem;
    f.mem = 0;
    f.reqs.clear();
    cv_done.notify_all();
    return false;
  }
  f.busy = true;
  return true;
}

void mr_server::complete(Flight& f) {
  int rc;
  {
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    hipSetDevice(ctx->device);
    rc = finish_batch(ctx, f.slot, f.plan, false);
    if (!rc && f.slot.retries) {
      // re-run after an overflow: read the outputs back again
      rc = queue_readback(f);
      if (!rc && hipStreamSynchronize(f.slot.stream) != hipSuccess) rc = MR_ERR_HIP;
    }
  }
  for (size_t j = 0; j < f.reqs.size() && rc == MR_OK; j++) {
    const mr_outputs& o = f.reqs[j]->out;
    if (o.on_device) continue;
    const size_t npix = (size_t)f.plan.jobs[j].W * f.plan.jobs[j].H;
    void* dst[kOutKinds] = {o.rgb8, o.depth, o.winner, o.rgb, o.normals, o.points};
    for (int k = 0; k < kOutKinds; k++)
      if (dst[k] && f.staged[j * kOutKinds + k])
        memcpy(dst[k], (const char*)f.slot.sc.h_outputs.p + f.offs[j * kOutKinds + k],
               npix * kOutBytes[k]);
  }
  std::lock_guard<std::mutex> lk(mu);
  const double t = now_us();
  for (Request* r : f.reqs) {
    r->status = rc;
    r->done_us = t;
    r->done = true;
  }
  f.reqs.clear();
  f.busy = false;
  mem_in_flight -= f.mem;
  f.mem = 0;
  cv_done.notify_all();
}

void mr_server::run() {
  hipSetDevice(ctx->device);
  const uint64_t budget = cfg.hbm_budget;
  for (;;) {
    bool any_busy = false;
    for (auto& fp : flights) {
      if (!fp->busy) continue;
      const hipError_t q = hipEventQuery(fp->slot.done);
      if (q == hipSuccess) complete(*fp);
      else if (q != hipErrorNotReady) {
        (void)hipGetLastError();
        complete(*fp);  // finish_batch reports the error
      } else {
        any_busy = true;
      }
    }
    Flight* free_slot = nullptr;
    for (auto& fp : flights)
      if (!fp->busy) { free_slot = fp.get(); break; }
    std::vector<Request*> take;
    {
      std::unique_lock<std::mutex> lk(mu);
      if (stop && queue.empty() && !any_busy) break;
      if (free_slot && !queue.empty()) {
        // Wait briefly for more requests when the queue is short.
        if ((int)queue.size() < cfg.max_batch && cfg.max_wait_us > 0 && any_busy) {
          const double oldest = queue.front()->submit_us;
          if (now_us() - oldest < cfg.max_wait_us) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            continue;
          }
        }
        uint64_t mem = 0;
        while (!queue.empty() && (int)take.size() < cfg.max_batch) {
          Request* r = queue.front();
          if (budget && mem_in_flight + mem + r->mem > budget && !take.empty()) break;
          if (budget && mem_in_flight + mem + r->mem > budget && any_busy) break;
          queue.pop_front();
          take.push_back(r);
          mem += r->mem;
        }
        mem_in_flight += mem;
        if (!take.empty()) free_slot->mem = mem;
      } else if (!any_busy) {
        cv_work.wait_for(lk, std::chrono::milliseconds(5));
        continue;
      }
    }
    if (!take.empty()) {
      launch(*free_slot, take);
      continue;
    }
    if (any_busy) std::this_thread::sleep_for(std::chrono::microseconds(10));
  }
}

extern "C" {

int mr_server_create(mr_ctx* ctx, const mr_server_config* cfg, mr_server** out) {
  if (!ctx || !cfg || !out || cfg->max_batch < 1 || cfg->slots < 1 || cfg->slots > 16 ||
      cfg->max_wait_us < 0)
    return MR_ERR_ARG;
  auto s = std::make_unique<mr_server>();
  s->ctx = ctx;
  s->cfg = *cfg;
  s->t0 = Clock::now();
  for (int i = 0; i < cfg->slots; i++) {
    auto f = std::make_unique<Flight>();
    if (slot_init(f->slot, nullptr) != MR_OK) {
      for (auto& g : s->flights) slot_free(ctx, g->slot);
      slot_free(ctx, f->slot);
      return MR_ERR_HIP;
    }
    s->flights.push_back(std::move(f));
  }
  mr_server* raw = s.release();
  raw->worker = std::thread([raw] { raw->run(); });
  *out = raw;
  return MR_OK;
}

int mr_server_destroy(mr_server* s) {
  if (!s) return MR_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  {
    std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
    for (auto& f : s->flights) slot_free(s->ctx, f->slot);
  }
  delete s;
  return MR_OK;
}

int mr_server_submit(mr_server* s, const mr_frame* frame, const mr_outputs* out,
                     uint64_t* ticket) {
  if (!s || !frame || !out || !ticket || frame->ndraws < 0 ||
      (frame->ndraws > 0 && !frame->draws))
    return MR_ERR_ARG;
  auto r = std::make_unique<Request>();
  r->frame = *frame;
  r->draws.assign(frame->draws, frame->draws + frame->ndraws);
  r->frame.draws = r->draws.data();
  r->out = *out;
  {
    std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
    r->mem = request_bytes(s->ctx, r->frame, r->out);
  }
  if (s->cfg.hbm_budget && r->mem > s->cfg.hbm_budget) return MR_ERR_BUDGET;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return MR_ERR_CLOSED;
  r->ticket = s->next_ticket++;
  r->submit_us = s->now_us();
  *ticket = r->ticket;
  s->queue.push_back(r.get());
  s->live[r->ticket] = std::move(r);
  s->cv_work.notify_one();
  return MR_OK;
}

int mr_server_poll(mr_server* s, uint64_t ticket, int32_t* done) {
  if (!s || !done) return MR_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->live.find(ticket);
  if (it == s->live.end()) return MR_ERR_HANDLE;
  *done = it->second->done ? 1 : 0;
  return MR_OK;
}

int mr_server_wait(mr_server* s, uint64_t ticket, int32_t timeout_ms, mr_request_info* info) {
  if (!s) return MR_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->live.find(ticket);
  if (it == s->live.end()) return MR_ERR_HANDLE;
  Request* r = it->second.get();
  auto pred = [r] { return r->done; };
  if (timeout_ms < 0) s->cv_done.wait(lk, pred);
  else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred))
    return MR_ERR_TIMEOUT;
  if (info) {
    info->submit_us = r->submit_us;
    info->launch_us = r->launch_us;
    info->done_us = r->done_us;
    info->batch_id = r->batch_id;
    info->batch_size = r->batch_size;
    info->status = r->status;
  }
  const int st = r->status;
  s->live.erase(it);
  return st;
}

int mr_server_run_trace(mr_server* s, const mr_frame* frames, int32_t n,
                        const double* arrival_s, const mr_outputs* outs,
                        double* latency_ms, double* makespan_s) {
  if (!s || !frames || !arrival_s || !outs || n < 0) return MR_ERR_ARG;
  std::vector<uint64_t> tickets((size_t)n, 0);
  std::vector<double> sched((size_t)n, 0);
  const double base = s->now_us();
  int rc_all = MR_OK;
  for (int i = 0; i < n; i++) {
    const double target = base + arrival_s[i] * 1e6;
    while (s->now_us() < target) {
      const double left = target - s->now_us();
      if (left > 200) std::this_thread::sleep_for(std::chrono::microseconds((int)(left - 100)));
      else std::this_thread::yield();
    }
    sched[i] = target;
    const int rc = mr_server_submit(s, &frames[i], &outs[i], &tickets[i]);
    if (rc) {
      rc_all = rc;
      tickets[i] = 0;
    }
  }
  double last = base;
  for (int i = 0; i < n; i++) {
    if (!tickets[i]) {
      if (latency_ms) latency_ms[i] = -1;
      continue;
    }
    mr_request_info info{};
    const int rc = mr_server_wait(s, tickets[i], -1, &info);
    if (rc && !rc_all) rc_all = rc;
    if (latency_ms) latency_ms[i] = (info.done_us - sched[i]) / 1e3;
    last = std::max(last, info.done_us);
  }
  if (makespan_s) makespan_s[0] = (last - (base + (n ? arrival_s[0] * 1e6 : 0))) / 1e6;
  return rc_all;
}

}  // extern "C"
