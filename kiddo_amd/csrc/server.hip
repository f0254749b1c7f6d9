// server.hip — continuous-batching request loop over resident trees.
//
// Requests (host query arrays and host result buffers) are submitted from any thread and
// cut into parts of at most `max_part` queries. A single loop thread
//   * admits pending parts whenever a stream slot is free and the device result arena has
//     room for the part's queries and results (admission by memory), taking first the oldest
//     part of a request with the fewest parts on the device, so that a large request cannot
//     hold every slot while smaller ones wait;
//   * coalesces the admitted part with other pending parts of the same kind (same tree,
//     family, k, radius and flags), at most one part per request, into one launch of up to
//     `max_coalesce` queries;
//   * hands each launch to its slot's worker thread.
// A slot's worker runs the launch on the slot's stream (host -> pinned staging -> HBM, the
// query kernel through the C ABI, HBM -> pinned staging), waits for it, copies every part's
// slice of the results into its request's host buffers, frees the arena range and the slot,
// and stamps a request's latency when its last part is answered. The workers' staging
// copies, launches and result copies run side by side (one loop thread doing all of them
// kept the device idle about half of the time at saturation). Launches on different slots
// overlap on the device; a request's results do not depend on what it was batched with,
// because every query is answered independently.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "tree.h"

namespace kd {
namespace {

using Clock = std::chrono::steady_clock;

struct Req {
  uint64_t ticket;
  kd_request r;
  Clock::time_point submitted;
  double latency_ms = 0;
  bool done = false;
  int status = KD_OK;
  uint32_t parts_left = 0;             // parts not answered yet (under the server's mutex)
  std::atomic<uint32_t> in_flight{0};  // parts on the device (read by the loop thread)
  uint64_t launch_mark = 0;            // loop thread only: the launch being assembled holds one of its parts
};

// Queries [first, first + m) of a request: the unit of admission.
struct Part {
  Req* req;
  uint64_t first, m;
  bool admitted = false;
};

// First-fit allocator over one device arena (and its pinned mirror); host bookkeeping only.
struct Arena {
  uint64_t size = 0;
  std::map<uint64_t, uint64_t> free_;  // offset -> length
  void init(uint64_t n) {
    size = n;
    free_.clear();
    free_[0] = n;
  }
  bool alloc(uint64_t n, uint64_t& off) {
    n = (n + 255) & ~uint64_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= n) {
        off = it->first;
        const uint64_t rest = it->second - n;
        free_.erase(it);
        if (rest) free_[off + n] = rest;
        return true;
      }
    return false;
  }
  void release(uint64_t off, uint64_t n) {
    n = (n + 255) & ~uint64_t(255);
    auto it = free_.emplace(off, n).first;
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
  uint64_t in_use() const {
    uint64_t f = 0;
    for (auto& kv : free_) f += kv.second;
    return size - f;
  }
};

struct Launch {
  int slot;
  std::vector<Part*> parts;
  std::vector<uint64_t> q_begin;  // per part: first query index in the launch
  uint64_t m = 0, off = 0, bytes = 0;
  uint64_t q_bytes = 0, items_off = 0, dists_off = 0, counts_off = 0;
  uint32_t k = 0;  // result row width (1 for nearest_one)
  size_t o_size = 0;
  hipEvent_t done_ev = nullptr;
};

// The worker thread of one stream slot and the launch handed to it.
struct SlotWorker {
  std::thread th;
  std::mutex mu;
  std::condition_variable cv;
  bool has = false, stop = false;
  Launch L;
};

}  // namespace
}  // namespace kd

struct kd_server {
  std::vector<kd_tree*> trees;
  uint32_t slots = 0;
  uint64_t max_coalesce = 0;
  std::atomic<uint64_t> max_part{16384};
  uint8_t* dev = nullptr;
  uint8_t* pinned = nullptr;
  kd::Arena arena;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  std::vector<bool> slot_busy;
  std::vector<std::unique_ptr<kd::SlotWorker>> workers;
  bool freed = false;  // a slot or arena range was released since the loop last looked
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<kd::Part*> submitted;
  std::map<uint64_t, kd::Req*> reqs;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread loop;
  // statistics
  std::atomic<uint64_t> st_done{0}, st_launches{0}, st_queries{0}, st_coalesced{0}, st_mem_waits{0};
  uint64_t peak_in_use = 0;
  int device = 0;
};

namespace kd {
namespace {

size_t out_size(const kd_tree* t) { return (t->dtype == KD_F32) ? 4 : 8; }
size_t a_size(const kd_tree* t) { return dtype_size(t->dtype); }

bool same_kind(const kd_request& a, const kd_request& b) {
  return a.tree == b.tree && a.family == b.family && a.k == b.k && a.flags == b.flags &&
         std::memcmp(&a.radius, &b.radius, sizeof(double)) == 0;
}

uint64_t launch_bytes(const kd_tree* t, const kd_request& r, uint64_t m) {
  const uint64_t w = r.family == KD_FAMILY_NEAREST_ONE || r.family == KD_FAMILY_APPROX ? 1 : r.k;
  auto al = [](uint64_t x) { return (x + 255) & ~uint64_t(255); };
  return al(m * t->dims * a_size(t)) + al(m * w * 4) + al(m * w * out_size(t)) + al(m * 4);
}

int run_launch(kd_server* S, Launch& L) {
  const kd_request& r0 = L.parts[0]->req->r;
  kd_tree* t = S->trees[r0.tree];
  hipStream_t s = S->streams[L.slot];
  auto al = [](uint64_t x) { return (x + 255) & ~uint64_t(255); };
  L.k = (r0.family == KD_FAMILY_NEAREST_ONE || r0.family == KD_FAMILY_APPROX) ? 1 : r0.k;
  L.o_size = out_size(t);
  L.q_bytes = L.m * t->dims * a_size(t);
  L.items_off = L.off + al(L.q_bytes);
  L.dists_off = L.items_off + al(L.m * L.k * 4);
  L.counts_off = L.dists_off + al(L.m * L.k * L.o_size);
  // gather the parts' queries into pinned staging, then one copy to HBM
  const size_t row = t->dims * a_size(t);
  for (size_t i = 0; i < L.parts.size(); ++i) {
    const Part* P = L.parts[i];
    std::memcpy(S->pinned + L.off + L.q_begin[i] * row,
                static_cast<const uint8_t*>(P->req->r.queries) + P->first * row, P->m * row);
  }
  KD_HIP(hipMemcpyAsync(S->dev + L.off, S->pinned + L.off, L.q_bytes, hipMemcpyHostToDevice, s));
  void* q = S->dev + L.off;
  uint32_t* items = reinterpret_cast<uint32_t*>(S->dev + L.items_off);
  void* dists = S->dev + L.dists_off;
  uint32_t* counts = reinterpret_cast<uint32_t*>(S->dev + L.counts_off);
  int rc;
  switch (r0.family) {
    case KD_FAMILY_NEAREST_ONE: rc = kd_nearest_one(t, q, L.m, r0.flags, items, dists, s); break;
    case KD_FAMILY_APPROX: rc = kd_approx_nearest_one(t, q, L.m, r0.flags, items, dists, s); break;
    case KD_FAMILY_NEAREST_N: rc = kd_nearest_n(t, q, L.m, r0.k, r0.radius, r0.flags, items, dists, counts, s); break;
    case KD_FAMILY_BEST_N_WITHIN:
      rc = kd_best_n_within(t, q, L.m, r0.radius, r0.k, r0.flags, items, dists, counts, s);
      break;
    default: rc = set_error(KD_ERR_UNSUPPORTED, "family %d is not served", r0.family);
  }
  if (rc) return rc;
  const uint64_t out_bytes = L.counts_off + L.m * 4 - L.items_off;
  KD_HIP(hipMemcpyAsync(S->pinned + L.items_off, S->dev + L.items_off, out_bytes, hipMemcpyDeviceToHost, s));
  KD_HIP(hipEventRecord(S->events[L.slot], s));
  L.done_ev = S->events[L.slot];
  return KD_OK;
}

// Runs on the slot's worker thread. Parts of one request may finish on several workers at
// once: each copies its own slice, then the request's counters change under the mutex, and
// the worker whose part was the last one marks the request done (its waiter may delete it
// as soon as the mutex is released).
void finish_launch(kd_server* S, Launch& L, int status) {
  if (status == KD_OK) {
    for (size_t i = 0; i < L.parts.size(); ++i) {
      const Part* P = L.parts[i];
      const kd_request& r = P->req->r;
      const uint64_t b = L.q_begin[i], m = P->m, f = P->first;
      if (r.items) std::memcpy(r.items + f * L.k, S->pinned + L.items_off + b * L.k * 4, m * L.k * 4);
      if (r.dists)
        std::memcpy(static_cast<uint8_t*>(r.dists) + f * L.k * L.o_size,
                    S->pinned + L.dists_off + b * L.k * L.o_size, m * L.k * L.o_size);
      if (r.counts && L.k > 1 && r.family != KD_FAMILY_NEAREST_ONE)
        std::memcpy(r.counts + f, S->pinned + L.counts_off + b * 4, m * 4);
    }
  }
  const auto now = Clock::now();
  std::lock_guard<std::mutex> g(S->mu);
  uint64_t finished = 0;
  for (Part* P : L.parts) {
    Req* R = P->req;
    if (status != KD_OK && R->status == KD_OK) R->status = status;
    --R->in_flight;
    if (--R->parts_left == 0) {
      R->latency_ms = std::chrono::duration<double, std::milli>(now - R->submitted).count();
      R->done = true;
      ++finished;
    }
    delete P;
  }
  L.parts.clear();
  S->arena.release(L.off, L.bytes);
  S->slot_busy[L.slot] = false;
  S->freed = true;
  S->st_done += finished;
  S->cv_done.notify_all();
  S->cv_work.notify_one();
}

// One stream slot: runs the launches the loop thread hands it, one at a time.
void slot_worker(kd_server* S, uint32_t j) {
  (void)hipSetDevice(S->device);
  SlotWorker& W = *S->workers[j];
  for (;;) {
    Launch L;
    {
      std::unique_lock<std::mutex> g(W.mu);
      W.cv.wait(g, [&] { return W.has || W.stop; });
      if (!W.has) return;
      L = std::move(W.L);
      W.has = false;
    }
    int rc = run_launch(S, L);
    if (rc == KD_OK && hipEventSynchronize(L.done_ev) != hipSuccess)
      rc = set_error(KD_ERR_DEVICE, "a served launch failed on the device");
    // after a failure, the stream's queued copies must end before the arena range is reused
    if (rc != KD_OK) (void)hipStreamSynchronize(S->streams[L.slot]);
    finish_launch(S, L, rc);
  }
}

// The loop thread: admission and coalescing; launches run on the slot workers.
void serve(kd_server* S) {
  (void)hipSetDevice(S->device);
  std::deque<Part*> pending;
  uint64_t serial = 0;
  bool progressed = true;
  for (;;) {
    {
      std::unique_lock<std::mutex> g(S->mu);
      // sleep until a request arrives or a worker frees a slot or memory (or 200 us pass)
      if (!progressed)
        S->cv_work.wait_for(g, std::chrono::microseconds(200),
                            [&] { return S->stop || !S->submitted.empty() || S->freed; });
      S->freed = false;
      const bool idle = std::none_of(S->slot_busy.begin(), S->slot_busy.end(), [](bool b) { return b; });
      if (S->stop && pending.empty() && S->submitted.empty() && idle) return;
      while (!S->submitted.empty()) {
        pending.push_back(S->submitted.front());
        S->submitted.pop_front();
      }
    }
    // admit in arrival order while a slot and memory are free
    progressed = false;
    while (!pending.empty()) {
      int slot = -1;
      {
        std::lock_guard<std::mutex> g(S->mu);
        for (uint32_t j = 0; j < S->slots; ++j)
          if (!S->slot_busy[j]) {
            slot = int(j);
            break;
          }
      }
      if (slot < 0) break;
      // the head: the oldest part among the requests with the fewest parts on the device
      auto hit = pending.begin();
      for (auto it = pending.begin(); it != pending.end(); ++it)
        if ((*it)->req->in_flight < (*hit)->req->in_flight) hit = it;
      Part* head = *hit;
      const kd_request& hr = head->req->r;
      kd_tree* t = S->trees[hr.tree];
      Launch L;
      L.slot = slot;
      L.parts.push_back(head);
      L.q_begin.push_back(0);
      L.m = head->m;
      head->req->launch_mark = ++serial;
      // other parts of the same kind, one per request, in arrival order
      for (auto it = pending.begin(); it != pending.end(); ++it) {
        Part* P = *it;
        if (P->req->launch_mark == serial || !same_kind(P->req->r, hr)) continue;
        if (L.m + P->m > S->max_coalesce) continue;
        P->req->launch_mark = serial;
        L.q_begin.push_back(L.m);
        L.parts.push_back(P);
        L.m += P->m;
      }
      // shrink the group until it fits the arena; the head alone must eventually fit
      bool ok = false;
      {
        std::lock_guard<std::mutex> g(S->mu);
        while (!L.parts.empty()) {
          L.bytes = launch_bytes(t, hr, L.m);
          if (S->arena.alloc(L.bytes, L.off)) {
            ok = true;
            break;
          }
          if (L.parts.size() == 1) break;
          L.m -= L.parts.back()->m;
          L.parts.pop_back();
          L.q_begin.pop_back();
        }
        if (ok) {
          S->slot_busy[slot] = true;
          S->peak_in_use = std::max(S->peak_in_use, S->arena.in_use());
        }
      }
      if (!ok) {
        ++S->st_mem_waits;
        break;  // wait for memory to free up
      }
      for (Part* P : L.parts) {
        P->admitted = true;
        ++P->req->in_flight;
      }
      pending.erase(std::remove_if(pending.begin(), pending.end(), [](const Part* P) { return P->admitted; }),
                    pending.end());
      ++S->st_launches;
      S->st_queries += L.m;
      S->st_coalesced += L.parts.size() - 1;
      {
        SlotWorker& W = *S->workers[slot];
        std::lock_guard<std::mutex> g(W.mu);
        W.L = std::move(L);
        W.has = true;
      }
      S->workers[slot]->cv.notify_one();
      progressed = true;
    }
  }
}

}  // namespace
}  // namespace kd

extern "C" {

int kd_server_create(kd_tree* const* trees, uint32_t ntrees, uint64_t arena_bytes, uint32_t slots,
                     uint64_t max_coalesce, kd_server** out) {
  if (!trees || ntrees == 0 || !out) return kd::set_error(KD_ERR_INVALID, "null argument");
  if (slots == 0 || slots > 8) return kd::set_error(KD_ERR_INVALID, "slots must be in [1, 8]");
  if (arena_bytes < (1u << 20)) return kd::set_error(KD_ERR_INVALID, "arena below 1 MiB");
  auto* S = new kd_server;
  S->trees.assign(trees, trees + ntrees);
  for (kd_tree* t : S->trees)
    if (!t) {
      delete S;
      return kd::set_error(KD_ERR_INVALID, "null tree");
    }
  S->slots = slots;
  S->max_coalesce = max_coalesce ? max_coalesce : (uint64_t(1) << 18);
  (void)hipGetDevice(&S->device);
  S->arena.init(arena_bytes);
  int rc = KD_OK;
  if (hipMalloc(&S->dev, arena_bytes) != hipSuccess) rc = kd::set_error(KD_ERR_NO_MEMORY, "device arena");
  if (!rc && hipHostMalloc(&S->pinned, arena_bytes, hipHostMallocDefault) != hipSuccess)
    rc = kd::set_error(KD_ERR_NO_MEMORY, "pinned staging");
  S->streams.resize(slots, nullptr);
  S->events.resize(slots, nullptr);
  S->slot_busy.assign(slots, false);
  for (uint32_t j = 0; j < slots && !rc; ++j) {
    if (hipStreamCreateWithFlags(&S->streams[j], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&S->events[j], hipEventDisableTiming) != hipSuccess)
      rc = kd::set_error(KD_ERR_DEVICE, "stream/event creation failed");
  }
  if (rc) {
    for (auto s : S->streams)
      if (s) (void)hipStreamDestroy(s);
    for (auto e : S->events)
      if (e) (void)hipEventDestroy(e);
    if (S->dev) (void)hipFree(S->dev);
    if (S->pinned) (void)hipHostFree(S->pinned);
    delete S;
    return rc;
  }
  for (uint32_t j = 0; j < slots; ++j) S->workers.push_back(std::make_unique<kd::SlotWorker>());
  for (uint32_t j = 0; j < slots; ++j) S->workers[j]->th = std::thread(kd::slot_worker, S, j);
  S->loop = std::thread(kd::serve, S);
  *out = S;
  return KD_OK;
}

int kd_server_submit(kd_server* S, const kd_request* r, uint64_t* ticket) {
  if (!S || !r || !ticket) return kd::set_error(KD_ERR_INVALID, "null argument");
  if (r->tree < 0 || uint32_t(r->tree) >= S->trees.size()) return kd::set_error(KD_ERR_INVALID, "bad tree index");
  if (r->m == 0 || !r->queries) return kd::set_error(KD_ERR_INVALID, "empty request");
  if ((r->family == KD_FAMILY_NEAREST_N || r->family == KD_FAMILY_BEST_N_WITHIN) &&
      r->k == 0)
    return kd::set_error(KD_ERR_INVALID, "k must be >= 1");
  if (r->family < KD_FAMILY_NEAREST_ONE || r->family > KD_FAMILY_APPROX)
    return kd::set_error(KD_ERR_UNSUPPORTED, "family %d is not served", r->family);
  const uint64_t part = S->max_part ? std::min<uint64_t>(S->max_part, r->m) : r->m;
  if (kd::launch_bytes(S->trees[r->tree], *r, part) > S->arena.size)
    return kd::set_error(KD_ERR_CAPACITY, "a part of the request needs more memory than the arena holds");
  auto* R = new kd::Req;
  R->r = *r;
  R->submitted = kd::Clock::now();
  std::vector<kd::Part*> parts;
  for (uint64_t f = 0; f < r->m; f += part) parts.push_back(new kd::Part{R, f, std::min(part, r->m - f)});
  R->parts_left = uint32_t(parts.size());
  {
    std::lock_guard<std::mutex> g(S->mu);
    R->ticket = S->next_ticket++;
    S->reqs[R->ticket] = R;
    for (kd::Part* P : parts) S->submitted.push_back(P);
  }
  S->cv_work.notify_one();
  *ticket = R->ticket;
  return KD_OK;
}

int kd_server_wait(kd_server* S, uint64_t ticket, double* latency_ms) {
  if (!S) return kd::set_error(KD_ERR_INVALID, "server is null");
  std::unique_lock<std::mutex> g(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return kd::set_error(KD_ERR_INVALID, "unknown ticket");
  kd::Req* R = it->second;
  S->cv_done.wait(g, [&] { return R->done; });
  const int status = R->status;
  if (latency_ms) *latency_ms = R->latency_ms;
  S->reqs.erase(it);
  delete R;
  return status;
}

int kd_server_set_part(kd_server* S, uint64_t max_part) {
  if (!S) return kd::set_error(KD_ERR_INVALID, "server is null");
  if (max_part > 0xFFFFFFFFull) return kd::set_error(KD_ERR_INVALID, "parts hold < 2^32 queries");
  S->max_part = max_part;
  return KD_OK;
}

int kd_server_stats(kd_server* S, uint64_t* out6) {
  if (!S || !out6) return kd::set_error(KD_ERR_INVALID, "null argument");
  out6[0] = S->st_done;
  out6[1] = S->st_launches;
  out6[2] = S->st_queries;
  out6[3] = S->st_coalesced;
  out6[4] = S->st_mem_waits;
  std::lock_guard<std::mutex> g(S->mu);
  out6[5] = S->peak_in_use;
  return KD_OK;
}

int kd_server_destroy(kd_server* S) {
  if (!S) return KD_OK;
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->loop.joinable()) S->loop.join();  // returns once every admitted launch has finished
  for (auto& W : S->workers) {
    {
      std::lock_guard<std::mutex> g(W->mu);
      W->stop = true;
    }
    W->cv.notify_one();
    if (W->th.joinable()) W->th.join();
  }
  for (auto s : S->streams) (void)hipStreamDestroy(s);
  for (auto e : S->events) (void)hipEventDestroy(e);
  (void)hipFree(S->dev);
  (void)hipHostFree(S->pinned);
  for (auto& kv : S->reqs) delete kv.second;
  delete S;
  return KD_OK;
}

}  // extern "C"
