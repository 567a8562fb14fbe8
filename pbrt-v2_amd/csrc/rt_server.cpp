// rt_server.cpp -- serving form of the rayhierarchy accelerator (rt_server_* in
// include/pbrt_amd_rt.h).
//
// One loop thread per server owns a HIP stream and a batch.  Each tick it admits queued
// requests while their device memory fits the budget, generates the camera rays of the newly
// admitted ones on the device, and runs ONE batch holding the primary rays of those requests
// and the shadow rays of the requests whose primaries finished in the previous tick.  After
// the batch it builds the shadow rays of the primaries just traced and completes the
// requests whose shadow rays were traced (copying their outputs back when asked to).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/pbrt_amd_rt.h"
#include "rt_internal.h"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// A request in flight: its device block holds the camera rays, the hits, the shadow rays and
// the occlusion flags.
struct Req {
  uint64_t ticket = 0;
  rt_request q{};
  size_t n = 0;      // rays of the camera image
  size_t block = 0;  // bytes of the device block
  size_t charge = 0; // bytes counted against the budget (block + traversal scratch)
  char* dmem = nullptr;
  int phase = 0;  // 0 queued, 1 primary rays in the next batch, 2 shadow rays, 3 done
  rt_request_stats st{};
  float* ray[8] = {};
  float* sray[8] = {};
  uint8_t* svalid = nullptr;
  int32_t* id = nullptr;
  float *t = nullptr, *u = nullptr, *v = nullptr;
  uint8_t* occ = nullptr;

  rt_rays primary() const {
    rt_rays r{ray[0], ray[1], ray[2], ray[3], ray[4], ray[5], ray[6], ray[7], nullptr,
              q.width * q.spp, q.height};
    return r;
  }
  rt_rays shadow() const {
    rt_rays r{sray[0], sray[1], sray[2], sray[3], sray[4], sray[5], sray[6], sray[7], svalid,
              q.width * q.spp, q.height};
    return r;
  }
};

}  // namespace

struct rt_server {
  rt_ctx* ctx = nullptr;
  std::vector<rt_scene*> scenes;
  uint64_t budget = 0, used = 0;
  int max_active = 1;
  rt_batch* batch = nullptr;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req*> pending;
  std::vector<Req*> active;
  std::unordered_map<uint64_t, Req*> all;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread th;
  std::multimap<size_t, char*> pool;  // freed device blocks by size (loop thread only)
  uint64_t ticks = 0;
  int max_tick = 0;
};

namespace {

int to_status(hipError_t e) {
  if (e == hipSuccess) return RT_OK;
  return e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;
}

// Takes a pooled block of at least `bytes` (at most twice as large) or allocates one.
int take_block(rt_server* S, size_t bytes, char** out) {
  auto it = S->pool.lower_bound(bytes);
  if (it != S->pool.end() && it->first <= 2 * bytes) {
    *out = it->second;
    S->pool.erase(it);
    return RT_OK;
  }
  hipError_t e = hipMalloc((void**)out, bytes);
  if (e == hipErrorOutOfMemory && !S->pool.empty()) {
    for (auto& kv : S->pool) (void)hipFree(kv.second);
    S->pool.clear();
    e = hipMalloc((void**)out, bytes);
  }
  return to_status(e);
}

int setup_request(rt_server* S, Req* r) {
  const size_t n = r->n;
  char* p = nullptr;
  int st = take_block(S, r->block, &p);
  if (st) return st;
  r->dmem = p;
  size_t off = 0;
  for (int c = 0; c < 8; ++c, off += al(4 * n)) r->ray[c] = (float*)(p + off);
  for (int c = 0; c < 8; ++c, off += al(4 * n)) r->sray[c] = (float*)(p + off);
  r->svalid = (uint8_t*)(p + off);
  off += al(n);
  r->id = (int32_t*)(p + off);
  off += al(4 * n);
  r->t = (float*)(p + off);
  off += al(4 * n);
  r->u = (float*)(p + off);
  off += al(4 * n);
  r->v = (float*)(p + off);
  off += al(4 * n);
  r->occ = (uint8_t*)(p + off);
  // the same double-precision steps as the workloads' camera (degrees -> radians, halve)
  const double tan_half = std::tan((double)r->q.fov_deg * (3.14159265358979323846 / 180.0) / 2.0);
  const float aspect = (float)((double)r->q.width / (double)r->q.height);
  return to_status(rti::launch_camera_rays(r->q.width, r->q.height, r->q.spp, r->q.origin,
                                           (float)tan_half, aspect, r->q.seed, r->ray[0],
                                           r->ray[1], r->ray[2], r->ray[3], r->ray[4], r->ray[5],
                                           r->ray[6], r->ray[7], S->stream));
}

void finish(rt_server* S, Req* r, int status) {
  r->st.status = status;
  r->st.done_s = now_s();
  r->phase = 3;
  if (r->dmem) S->pool.emplace(r->block, r->dmem);
  r->dmem = nullptr;
  S->used -= r->charge;
}

void loop(rt_server* S) {
  (void)hipSetDevice(S->ctx->device);
  std::vector<Req*> admitted, tick;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] { return S->stop || !S->pending.empty() || !S->active.empty(); });
      if (S->pending.empty() && S->active.empty()) {
        if (S->stop) break;
        continue;
      }
      admitted.clear();
      while (!S->pending.empty() && (int)S->active.size() < S->max_active) {
        Req* r = S->pending.front();
        // a request larger than the budget still runs when nothing else is in flight
        if (S->used + r->charge > S->budget && !S->active.empty()) break;
        S->pending.pop_front();
        S->used += r->charge;
        S->active.push_back(r);
        admitted.push_back(r);
      }
      tick = S->active;
    }
    for (Req* r : admitted) {
      r->st.admit_s = now_s();
      const int st = setup_request(S, r);
      if (st) r->st.status = st;
      else r->phase = 1;
    }
    // the tick's batch: primaries of the admitted, shadow rays of last tick's primaries
    int st = rt_batch_clear(S->batch);
    int njobs = 0;
    for (Req* r : tick) {
      if (r->st.status) continue;
      if (r->phase == 1) {
        const rt_rays rays = r->primary();
        const rt_hits hits{r->id, r->t, r->u, r->v, nullptr, nullptr};
        if (!st) st = rt_batch_add_closest(S->batch, S->scenes[r->q.scene], &rays,
                                           RT_CLOSEST_PRIMARY, &hits);
        ++njobs;
      } else if (r->phase == 2) {
        const rt_rays rays = r->shadow();
        if (!st) st = rt_batch_add_occluded(S->batch, S->scenes[r->q.scene], &rays, r->occ);
        ++njobs;
      }
    }
    if (!st) st = rt_batch_run(S->batch, S->stream);
    for (Req* r : tick) {
      if (st || r->st.status) continue;
      if (r->phase == 1) {
        const rt_rays rays = r->primary();
        st = rt_make_shadow_rays(&rays, r->id, r->t, r->q.light, r->sray[0], r->sray[1],
                                 r->sray[2], r->sray[3], r->sray[4], r->sray[5], r->sray[6],
                                 r->sray[7], r->svalid, S->stream);
        r->phase = 2;
      } else if (r->phase == 2) {
        const size_t n = r->n;
        hipError_t e = hipSuccess;
        if (r->q.host_id)
          e = hipMemcpyAsync(r->q.host_id, r->id, 4 * n, hipMemcpyDeviceToHost, S->stream);
        if (e == hipSuccess && r->q.host_t)
          e = hipMemcpyAsync(r->q.host_t, r->t, 4 * n, hipMemcpyDeviceToHost, S->stream);
        if (e == hipSuccess && r->q.host_occluded)
          e = hipMemcpyAsync(r->q.host_occluded, r->occ, n, hipMemcpyDeviceToHost, S->stream);
        if (e != hipSuccess) st = to_status(e);
        r->phase = 4;  // completes after the stream drains
      }
    }
    const hipError_t se = hipStreamSynchronize(S->stream);
    if (!st && se != hipSuccess) st = to_status(se);
    {
      std::lock_guard<std::mutex> lk(S->mu);
      ++S->ticks;
      S->max_tick = std::max(S->max_tick, njobs);
      std::vector<Req*> keep;
      for (Req* r : S->active) {
        const bool in_tick = std::find(tick.begin(), tick.end(), r) != tick.end();
        if (r->st.status) {
          finish(S, r, r->st.status);
        } else if (in_tick && st) {
          finish(S, r, st);  // the tick failed: its requests fail with its status
        } else if (r->phase == 4) {
          finish(S, r, RT_OK);
        } else {
          keep.push_back(r);
        }
      }
      S->active.swap(keep);
    }
    S->cv_done.notify_all();
  }
}

}  // namespace

extern "C" {

int rt_server_create(rt_ctx* ctx, rt_scene* const* scenes, int32_t nscenes, uint64_t mem_budget,
                     int32_t max_active, rt_server** out) {
  if (!ctx || !scenes || nscenes <= 0 || !out || max_active <= 0) return RT_ERR_INVALID;
  *out = nullptr;
  if (hipSetDevice(ctx->device) != hipSuccess) return RT_ERR_HIP;
  rt_server* S = new (std::nothrow) rt_server();
  if (!S) return RT_ERR_OOM;
  S->ctx = ctx;
  S->scenes.assign(scenes, scenes + nscenes);
  for (rt_scene* s : S->scenes)
    if (!s) {
      delete S;
      return RT_ERR_INVALID;
    }
  S->max_active = max_active;
  int st = rt_batch_create(ctx, &S->batch);
  if (st) {
    delete S;
    return st;
  }
  if (mem_budget == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
    mem_budget = free_b / 2;
  }
  S->budget = mem_budget;
  if (hipStreamCreate(&S->stream) != hipSuccess) {
    rt_batch_destroy(S->batch);
    delete S;
    return RT_ERR_HIP;
  }
  S->th = std::thread(loop, S);
  *out = S;
  return RT_OK;
}

int rt_server_submit(rt_server* S, const rt_request* q, uint64_t* ticket) {
  if (!S || !q || !ticket) return RT_ERR_INVALID;
  if (q->scene < 0 || q->scene >= (int32_t)S->scenes.size()) return RT_ERR_INVALID;
  if (q->width <= 0 || q->height <= 0 || q->spp <= 0 || q->spp > 64) return RT_ERR_INVALID;
  if (!(q->fov_deg > 0.f && q->fov_deg < 180.f)) return RT_ERR_INVALID;
  const uint64_t n = (uint64_t)q->width * q->spp * q->height;
  if (n >= (1ull << 30)) return RT_ERR_UNSUPPORTED;
  Req* r = new (std::nothrow) Req();
  if (!r) return RT_ERR_OOM;
  r->q = *q;
  r->n = (size_t)n;
  r->block = 16 * al(4 * r->n) + 2 * al(r->n) + 4 * al(4 * r->n);
  // traversal scratch in the batch: keys and repair keys (12 B/ray), hierarchy nodes
  // (< 64 B per 12 rays), passing-cluster lists (4 B per block and cluster), work units
  const uint64_t blocks = (n + 255) / 256;
  const uint64_t ncl = rti::scene_nclusters(S->scenes[q->scene]);
  r->charge = r->block + 12 * n + 6 * n + blocks * ncl * 4 + blocks * (ncl / 2 + 1) * 16;
  r->st.primary_rays = n;
  r->st.submit_s = now_s();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) {
      delete r;
      return RT_ERR_STATE;
    }
    r->ticket = S->next_ticket++;
    S->all.emplace(r->ticket, r);
    S->pending.push_back(r);
    *ticket = r->ticket;
  }
  S->cv_work.notify_one();
  return RT_OK;
}

int rt_server_wait(rt_server* S, uint64_t ticket, double timeout_s, rt_request_stats* stats) {
  if (!S) return RT_ERR_INVALID;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->all.find(ticket);
  if (it == S->all.end()) return RT_ERR_INVALID;
  Req* r = it->second;
  auto done = [&] { return r->phase == 3; };
  if (timeout_s < 0) {
    S->cv_done.wait(lk, done);
  } else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
    return RT_ERR_STATE;
  }
  if (stats) *stats = r->st;
  const int status = r->st.status;
  S->all.erase(it);
  delete r;
  return status;
}

int rt_server_counters(rt_server* S, uint64_t* ticks, int32_t* max_tick_requests) {
  if (!S || !ticks || !max_tick_requests) return RT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  *ticks = S->ticks;
  *max_tick_requests = S->max_tick;
  return RT_OK;
}

int rt_server_destroy(rt_server* S) {
  if (!S) return RT_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->th.joinable()) S->th.join();
  for (auto& kv : S->pool) (void)hipFree(kv.second);
  for (auto& kv : S->all) delete kv.second;
  rt_batch_destroy(S->batch);
  (void)hipStreamDestroy(S->stream);
  delete S;
  return RT_OK;
}

}  // extern "C"
