// rt_server.cpp -- serving form of the ray-space accelerators (rt_server_* in
// include/pbrt_amd_rt.h).
//
// One loop thread per server owns a HIP stream, two batches and a device pool.  Each tick it
// admits queued requests while their memory fits the budget, takes their device blocks from
// the pool (no allocation on the loop), generates their camera rays, and launches ONE batch
// holding the primary rays of the newly admitted requests and the shadow rays of the
// requests whose primaries ran in the previous tick; then it queues the shadow rays of the
// primaries just launched.  The loop does not wait for that tick: it waits for the previous
// one, completes the requests whose shadow rays it traced, and goes on to prepare the next
// tick while the current one runs, so the host work of a tick overlaps the device work of the
// one before it.  The two batches alternate, so a batch's tables are never rewritten while
// the device may still read them.
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

// First-fit suballocator over one device region (offsets aligned to 256 bytes), with
// coalescing of freed neighbours.
class Pool {
 public:
  void reset(size_t cap) {
    free_.clear();
    if (cap) free_[0] = cap;
  }
  bool take(size_t bytes, size_t* off) {
    bytes = al(std::max<size_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      return true;
    }
    return false;
  }
  void give(size_t off, size_t bytes) {
    bytes = al(std::max<size_t>(bytes, 1));
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<size_t, size_t> free_;  // offset -> size
};

// A request in flight: its device block holds the camera rays, the hits, the shadow rays and
// the occlusion flags.
struct Req {
  uint64_t ticket = 0;
  rt_request q{};
  size_t n = 0;      // rays of the camera image
  size_t block = 0;  // bytes of the device block
  size_t charge = 0; // bytes counted against the budget (block + traversal scratch)
  size_t off = 0;    // offset of the block in the pool
  bool has_block = false;
  // 0 queued, 1 primary rays in the next tick, 2 shadow rays in the next tick, 4 shadow
  // rays launched (done with tick `done_tick`), 3 complete
  int phase = 0;
  uint64_t done_tick = 0;
  rt_request_stats st{};
  float* ray[8] = {};
  float* sray[8] = {};
  uint8_t* svalid = nullptr;
  int32_t* id = nullptr;
  float *t = nullptr, *u = nullptr, *v = nullptr;
  uint8_t* occ = nullptr;
  uint32_t* nshadow = nullptr;  // valid shadow rays, counted on the device

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
  rt_batch* batch[2] = {nullptr, nullptr};
  hipStream_t stream = nullptr;  // ticks
  hipStream_t copy = nullptr;    // copies of results to the caller (non-blocking stream)
  hipEvent_t ev[2] = {nullptr, nullptr};
  char* pool_mem = nullptr;
  size_t pool_cap = 0;
  Pool pool;  // loop thread only
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req*> pending;
  std::vector<Req*> active;
  std::unordered_map<uint64_t, Req*> all;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread th;
  uint64_t ticks = 0;
  int max_tick = 0;
};

namespace {

int to_status(hipError_t e) {
  if (e == hipSuccess) return RT_OK;
  return e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;
}

size_t block_bytes(size_t n) { return 16 * al(4 * n) + 2 * al(n) + 4 * al(4 * n) + 256; }

int setup_request(rt_server* S, Req* r) {
  const size_t n = r->n;
  size_t off = 0;
  if (!S->pool.take(r->block, &off)) return RT_ERR_OOM;
  r->off = off;
  r->has_block = true;
  char* p = S->pool_mem + off;
  size_t o = 0;
  for (int c = 0; c < 8; ++c, o += al(4 * n)) r->ray[c] = (float*)(p + o);
  for (int c = 0; c < 8; ++c, o += al(4 * n)) r->sray[c] = (float*)(p + o);
  r->svalid = (uint8_t*)(p + o);
  o += al(n);
  r->id = (int32_t*)(p + o);
  o += al(4 * n);
  r->t = (float*)(p + o);
  o += al(4 * n);
  r->u = (float*)(p + o);
  o += al(4 * n);
  r->v = (float*)(p + o);
  o += al(4 * n);
  r->occ = (uint8_t*)(p + o);
  o += al(n);
  r->nshadow = (uint32_t*)(p + o);
  // the same double-precision steps as the workloads' camera (degrees -> radians, halve)
  const double tan_half = std::tan((double)r->q.fov_deg * (3.14159265358979323846 / 180.0) / 2.0);
  const float aspect = (float)((double)r->q.width / (double)r->q.height);
  return to_status(rti::launch_camera_rays(r->q.width, r->q.height, r->q.spp, r->q.origin,
                                           (float)tan_half, aspect, r->q.seed, r->ray[0],
                                           r->ray[1], r->ray[2], r->ray[3], r->ray[4], r->ray[5],
                                           r->ray[6], r->ray[7], S->stream));
}

// Copies the finished results to the caller's host arrays (on the copy stream, after the
// tick that produced them has completed) and returns the request's block to the pool.
int deliver(rt_server* S, Req* r) {
  const size_t n = r->n;
  hipError_t e = hipSuccess;
  if (r->st.status == RT_OK && r->has_block) {
    uint32_t ns = 0;
    e = hipMemcpyAsync(&ns, r->nshadow, 4, hipMemcpyDeviceToHost, S->copy);
    if (e == hipSuccess && r->q.host_id)
      e = hipMemcpyAsync(r->q.host_id, r->id, 4 * n, hipMemcpyDeviceToHost, S->copy);
    if (e == hipSuccess && r->q.host_t)
      e = hipMemcpyAsync(r->q.host_t, r->t, 4 * n, hipMemcpyDeviceToHost, S->copy);
    if (e == hipSuccess && r->q.host_occluded)
      e = hipMemcpyAsync(r->q.host_occluded, r->occ, n, hipMemcpyDeviceToHost, S->copy);
    if (e == hipSuccess) e = hipStreamSynchronize(S->copy);
    r->st.shadow_rays = ns;
  }
  return to_status(e);
}

// Marks a request complete (mutex held).
void finish(rt_server* S, Req* r, int status) {
  if (!r->st.status) r->st.status = status;
  r->st.done_s = now_s();
  r->phase = 3;
  if (r->has_block) S->pool.give(r->off, r->block);
  r->has_block = false;
  S->used -= r->charge;
}

// Waits for tick `k` and completes the requests it finished (or, when it failed, every
// request it carried).
void complete_tick(rt_server* S, uint64_t k, int tick_status, const std::vector<Req*>& carried) {
  hipError_t e = hipEventSynchronize(S->ev[k & 1]);
  int st = tick_status ? tick_status : to_status(e);
  std::vector<Req*> done;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    for (Req* r : S->active)
      if (r->phase == 4 && r->done_tick == k) done.push_back(r);
  }
  std::vector<int> dst(done.size(), RT_OK);
  for (size_t i = 0; i < done.size(); ++i) dst[i] = st ? st : deliver(S, done[i]);
  std::lock_guard<std::mutex> lk(S->mu);
  for (size_t i = 0; i < done.size(); ++i) finish(S, done[i], dst[i]);
  if (st)
    for (Req* r : carried)
      if (r->phase != 3) finish(S, r, st);
  // requests that could not start (no batch carried them) end with their own status
  for (Req* r : S->active)
    if (r->phase != 3 && r->phase != 4 && r->st.status) finish(S, r, r->st.status);
  std::vector<Req*> keep;
  for (Req* r : S->active)
    if (r->phase != 3) keep.push_back(r);
  S->active.swap(keep);
  S->cv_done.notify_all();
}

void loop(rt_server* S) {
  (void)hipSetDevice(S->ctx->device);
  std::vector<Req*> admitted, tick;
  std::vector<Req*> carried[2];  // requests of the tick in flight on each batch
  int status[2] = {RT_OK, RT_OK};
  uint64_t k = 0;             // next tick
  bool inflight = false;      // tick k - 1 has not been completed
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] {
        return S->stop || !S->pending.empty() || !S->active.empty() || inflight;
      });
      if (S->pending.empty() && S->active.empty() && !inflight) {
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
      tick.clear();
      for (Req* r : S->active)
        if (r->phase <= 2) tick.push_back(r);
    }
    for (Req* r : admitted) {
      r->st.admit_s = now_s();
      int st = setup_request(S, r);
      if (st == RT_ERR_OOM && inflight) {
        // the pool is fragmented by requests of the tick in flight: complete it and retry
        complete_tick(S, k - 1, status[(k - 1) & 1], carried[(k - 1) & 1]);
        inflight = false;
        st = setup_request(S, r);
      }
      if (st) r->st.status = st;
      else r->phase = 1;
    }
    // the tick's jobs: primaries of the admitted, shadow rays of last tick's primaries
    rt_batch* B = S->batch[k & 1];
    int st = rt_batch_clear(B);
    int njobs = 0;
    std::vector<Req*>& car = carried[k & 1];
    car.clear();
    for (Req* r : tick) {
      if (r->st.status) continue;
      if (r->phase == 1) {
        const rt_rays rays = r->primary();
        const rt_hits hits{r->id, r->t, r->u, r->v, nullptr, nullptr};
        if (!st) st = rt_batch_add_closest(B, S->scenes[r->q.scene], &rays, RT_CLOSEST_PRIMARY,
                                           &hits);
      } else if (r->phase == 2) {
        const rt_rays rays = r->shadow();
        if (!st) st = rt_batch_add_occluded(B, S->scenes[r->q.scene], &rays, r->occ);
      } else {
        continue;
      }
      car.push_back(r);
      ++njobs;
    }
    if (njobs == 0 && st == RT_OK) {
      // nothing to launch: finish the tick in flight, then fail what could not start
      if (inflight) complete_tick(S, k - 1, status[(k - 1) & 1], carried[(k - 1) & 1]);
      inflight = false;
      std::lock_guard<std::mutex> lk(S->mu);
      std::vector<Req*> keep;
      for (Req* r : S->active) {
        if (r->st.status) finish(S, r, r->st.status);
        if (r->phase != 3) keep.push_back(r);
      }
      S->active.swap(keep);
      S->cv_done.notify_all();
      continue;
    }
    if (!st) st = rt_batch_run(B, S->stream);
    for (Req* r : car) {
      if (st) break;
      if (r->phase == 1) {
        const rt_rays rays = r->primary();
        const rti::RaysDev pr{rays.ox,   rays.oy,   rays.oz, rays.dx,   rays.dy,
                              rays.dz,   rays.tmin, rays.tmax, nullptr, rays.width,
                              rays.height};
        st = to_status(rti::launch_fill_u8((uint8_t*)r->nshadow, 0, 4, S->stream));
        if (!st)
          st = to_status(rti::launch_shadow_rays(pr, r->id, r->t, r->q.light[0], r->q.light[1],
                                                 r->q.light[2], r->sray[0], r->sray[1],
                                                 r->sray[2], r->sray[3], r->sray[4], r->sray[5],
                                                 r->sray[6], r->sray[7], r->svalid, S->stream,
                                                 r->nshadow));
        r->phase = 2;
      } else if (r->phase == 2) {
        r->phase = 4;
        r->done_tick = k;
      }
    }
    const hipError_t re = hipEventRecord(S->ev[k & 1], S->stream);
    if (!st && re != hipSuccess) st = to_status(re);
    status[k & 1] = st;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      ++S->ticks;
      S->max_tick = std::max(S->max_tick, njobs);
    }
    // the device now holds tick k (and maybe k - 1): complete k - 1 while k runs
    if (inflight) complete_tick(S, k - 1, status[(k - 1) & 1], carried[(k - 1) & 1]);
    inflight = true;
    ++k;
    if (st) {  // a failed tick: settle it now
      complete_tick(S, k - 1, st, carried[(k - 1) & 1]);
      inflight = false;
    }
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
  for (rt_scene* s : S->scenes) {
    const int st = rti::scene_check_image(s, 16, 16);
    if (st) {
      delete S;
      return st;
    }
  }
  S->max_active = max_active;
  if (mem_budget == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
    mem_budget = std::min<uint64_t>(free_b / 2, (uint64_t)32 << 30);
  }
  S->budget = mem_budget;
  // the request blocks of every admitted request fit in the budget, so one region of that
  // size holds them; traversal scratch stays with the batches
  S->pool_cap = al((size_t)mem_budget);
  int st = RT_OK;
  if (hipMalloc((void**)&S->pool_mem, S->pool_cap) != hipSuccess) st = RT_ERR_OOM;
  if (!st) st = rt_batch_create(ctx, &S->batch[0]);
  if (!st) st = rt_batch_create(ctx, &S->batch[1]);
  if (!st && hipStreamCreate(&S->stream) != hipSuccess) st = RT_ERR_HIP;
  if (!st && hipStreamCreateWithFlags(&S->copy, hipStreamNonBlocking) != hipSuccess)
    st = RT_ERR_HIP;
  for (int i = 0; i < 2 && !st; ++i)
    if (hipEventCreateWithFlags(&S->ev[i], hipEventDisableTiming) != hipSuccess) st = RT_ERR_HIP;
  if (st) {
    for (int i = 0; i < 2; ++i) {
      if (S->batch[i]) rt_batch_destroy(S->batch[i]);
      if (S->ev[i]) (void)hipEventDestroy(S->ev[i]);
    }
    if (S->stream) (void)hipStreamDestroy(S->stream);
    if (S->copy) (void)hipStreamDestroy(S->copy);
    (void)hipFree(S->pool_mem);
    delete S;
    return st;
  }
  S->pool.reset(S->pool_cap);
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
  if (n >= (1ull << 30) || (uint64_t)q->width * q->spp >= (1ull << 31)) return RT_ERR_UNSUPPORTED;
  // the image must trace on its scene: checked here so that a bad request fails alone
  const int chk = rti::scene_check_image(S->scenes[q->scene], q->width * q->spp, q->height);
  if (chk) return chk;
  if (block_bytes((size_t)n) > S->pool_cap) return RT_ERR_OOM;
  Req* r = new (std::nothrow) Req();
  if (!r) return RT_ERR_OOM;
  r->q = *q;
  r->n = (size_t)n;
  r->block = block_bytes(r->n);
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
  for (auto& kv : S->all) delete kv.second;
  (void)hipStreamSynchronize(S->stream);
  for (int i = 0; i < 2; ++i) {
    rt_batch_destroy(S->batch[i]);
    (void)hipEventDestroy(S->ev[i]);
  }
  (void)hipStreamDestroy(S->stream);
  (void)hipStreamDestroy(S->copy);
  (void)hipFree(S->pool_mem);
  delete S;
  return RT_OK;
}

}  // extern "C"
