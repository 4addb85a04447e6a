// pa_server.hip -- a continuous-batching request server over the library's
// sort / select / histogram entry points, on one GPU.
//
// No reference counterpart: the reference is a blocking call-per-array
// library.  Its closest ideas are bundling small work items
// (BundleSmallWorkItems.cpp:36-144) and choosing an algorithm by free memory
// (MemoryUsage.cpp:20-86, RadixSortLSD.h:799-834).
//
// Jobs name device-resident arrays (keys sorted in place; select and
// histogram results written to device memory), so the server never blocks on
// PCIe.  A dispatcher thread takes jobs in arrival order:
//   - small u32 sorts waiting together are coalesced into one segmented batch
//     (pa_segmented_sort_ptrs_u32: 7 launches for the whole batch);
//   - every other job is launched on its own;
//   - admission is priced in the library's own scratch: every stream holds
//     the scratch block the library allocated for it (it grows to the largest
//     call run there and is kept for reuse), and a unit is placed only where
//     the sum over the server's streams after its call stays within the
//     budget.  The call's size comes from the same layout functions the
//     entry points allocate with (pa_job_scratch_bytes), so the prediction is
//     the allocation.  When no stream fits, idle streams give their blocks
//     back, and if that is not enough the dispatcher waits for units in
//     flight to finish.  A unit larger than the whole budget runs alone,
//     after every other block has been released (counted as oversize).
//   - among the streams that fit, the one whose held block grows least wins,
//     then the one with the least work in flight, so new work overlaps work
//     already on the GPU.
// A completer thread polls the units' events and stamps each job's
// completion time; latency = completion - submission (host steady clock).
// Each job's result is what the same call returns alone: units never share
// scratch (the library keeps scratch per stream and serialises each stream).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pa_internal.h"

namespace pa {

using Clock = std::chrono::steady_clock;

struct SJob {
  pa_job d;
  uint64_t ticket;
  Clock::time_point submit, done;
  int status = PA_OK;
  std::string error;  // the call's message when status is not PA_OK
  bool finished = false;
};

struct SUnit {
  std::vector<SJob*> jobs;
  int stream = 0;
  hipEvent_t ev = nullptr;
  size_t bytes = 0;  // scratch the unit's call asks the library for
  bool batch = false;
};

// Units in flight per stream before the dispatcher looks for another stream.
constexpr int kUnitsPerStream = 2;

struct Server {
  size_t budget = 0, peak = 0;
  size_t coalesce_max = 1 << 20;     // largest sort that joins a batch
  size_t batch_max_keys = 64 << 20;  // keys per batch
  std::vector<hipStream_t> streams;
  std::vector<int> busy;       // units in flight per stream
  std::vector<size_t> held;    // scratch the library holds for each stream
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<SJob*> pending;
  std::unordered_map<uint64_t, SJob*> jobs;
  std::vector<SUnit*> inflight;
  // Units taken off `pending` whose launch is not yet in `inflight` (the
  // dispatcher launches with the mutex released): drain and the completer's
  // exit wait for them too.
  int dispatching = 0;
  std::thread dispatcher, completer;
  bool stop = false;
  int device = 0;
  uint64_t next_ticket = 1;
  pa_server_stats st{};
};

// Scratch the library's call for one job asks for (device-resident keys and
// outputs, as the server requires).
static size_t call_bytes(const pa_job& j) {
  const bool u64 = j.key_type == PA_KEY_U64;
  switch (j.op) {
    case PA_OP_SORT: return sort_call_scratch_bytes(u64, j.n);
    case PA_OP_SELECT: return select_call_scratch_bytes(u64, j.n);
    case PA_OP_HIST: return hist_call_scratch_bytes(j.bits);
    default: return 0;
  }
}

// Scratch of a coalesced batch: the segmented sort's plan for these arrays.
static size_t batch_bytes(const SUnit& u) {
  std::vector<uint64_t> lens;
  for (SJob* j : u.jobs) lens.push_back(j->d.n);
  size_t b = 0;
  if (segmented_lens_scratch_bytes(lens.data(), lens.size(), &b) != PA_OK) {
    // The launch reports the error; price the batch as its keys plus a block.
    b = 0;
    for (SJob* j : u.jobs) b += j->d.n * 8;
  }
  return b;
}

// Block a stream holds after a call asking for `need` bytes: the library
// keeps a block that is large enough and otherwise replaces it.
static size_t held_after(size_t held, size_t need) {
  return held >= need ? held : scratch_alloc_bytes(need);
}

static size_t total_held(const Server& S) {
  size_t t = 0;
  for (size_t h : S.held) t += h;
  return t;
}

static bool coalescable(const Server& S, const SJob* j) {
  return j->d.op == PA_OP_SORT && j->d.key_type == PA_KEY_U32 && j->d.n <= S.coalesce_max;
}

static int launch_one(const pa_job& j, hipStream_t s) {
  switch (j.op) {
    case PA_OP_SORT:
      if (j.key_type == PA_KEY_U64) return pa_sort_u64((uint64_t*)j.keys, nullptr, j.n, s);
      return pa_sort_u32((uint32_t*)j.keys, nullptr, j.n, s);
    case PA_OP_SELECT:
      if (j.key_type == PA_KEY_U64) return pa_select_u64((const uint64_t*)j.keys, j.n, j.k, (uint64_t*)j.out, s);
      return pa_select_u32((const uint32_t*)j.keys, j.n, j.k, (uint32_t*)j.out, s);
    case PA_OP_HIST:
      if (j.key_type == PA_KEY_U64) return pa_hist_u64((const uint64_t*)j.keys, j.n, j.shift, j.bits, (uint64_t*)j.out, s);
      return pa_hist_u32((const uint32_t*)j.keys, j.n, j.shift, j.bits, (uint64_t*)j.out, s);
    default:
      return arg_error("server: unknown op");
  }
}

// Picks the stream for a unit asking for `need` bytes, under S->mu.  Among
// the streams with room for another unit whose held block after the call
// keeps the total within the budget, it prefers one that does not grow a
// block under a unit in flight (growing syncs the stream), then the least
// busy one, then the least total.  Returns -1 when none fits.
static int place(const Server& S, size_t need) {
  const size_t total = total_held(S);
  int best = -1;
  bool best_grow = false;
  size_t best_after = 0;
  for (int i = 0; i < (int)S.streams.size(); ++i) {
    if (S.busy[i] >= kUnitsPerStream) continue;
    const size_t a = total - S.held[i] + held_after(S.held[i], need);
    if (a > S.budget) continue;
    const bool grow = S.held[i] < need && S.busy[i] > 0;
    const bool better = best < 0 || grow < best_grow ||
                        (grow == best_grow && (S.busy[i] < S.busy[best] ||
                                               (S.busy[i] == S.busy[best] && a < best_after)));
    if (better) {
      best = i;
      best_grow = grow;
      best_after = a;
    }
  }
  return best;
}

// Admission of one unit, entered and left with `lk` held; returns the
// stream.  Every wait here ends: it waits only while units are in flight,
// and the completer retires them.
static int admit(Server* S, std::unique_lock<std::mutex>& lk, size_t need) {
  bool waited = false;
  std::vector<char> stuck(S->streams.size(), 0);  // streams whose release failed
  for (;;) {
    const int s = place(*S, need);
    if (s >= 0) return s;
    int in_flight = 0, any_room = -1, largest_idle = -1;
    size_t busy_held = 0;
    for (int i = 0; i < (int)S->streams.size(); ++i) {
      in_flight += S->busy[i];
      if (S->busy[i] > 0) busy_held += S->held[i];
      if (S->busy[i] < kUnitsPerStream && any_room < 0) any_room = i;
      if (S->busy[i] == 0 && S->held[i] > 0 && !stuck[i] &&
          (largest_idle < 0 || S->held[i] > S->held[largest_idle]))
        largest_idle = i;
    }
    // Would the unit fit once every idle stream's block is given back?
    bool fits_after_release = false;
    for (int i = 0; i < (int)S->streams.size(); ++i) {
      if (S->busy[i] >= kUnitsPerStream) continue;
      const size_t a = S->busy[i] ? busy_held - S->held[i] + held_after(S->held[i], need)
                                  : busy_held + scratch_alloc_bytes(need);
      if (a <= S->budget) fits_after_release = true;
    }
    if (largest_idle >= 0 && (fits_after_release || in_flight == 0)) {
      // Give back the largest idle block and look again.  Only the
      // dispatcher launches on the server's streams, so an idle stream
      // stays idle while the lock is released.
      hipStream_t st = S->streams[largest_idle];
      lk.unlock();
      if (release_stream_scratch(st) != PA_OK) (void)hipGetLastError();
      const size_t h = stream_scratch_held(st);
      lk.lock();
      if (h == S->held[largest_idle]) stuck[largest_idle] = 1;
      S->held[largest_idle] = h;
      S->st.scratch_releases++;
      continue;
    }
    if (in_flight > 0) {
      // Wait for units in flight to finish (or for a stream's quota to free).
      if (!waited && any_room >= 0) S->st.admission_waits++;
      waited = waited || any_room >= 0;
      S->cv_done.wait(lk);
      continue;
    }
    // Nothing in flight and no idle block left to give back: the unit is
    // larger than the budget on its own.  Run it alone.
    S->st.oversize_units++;
    return any_room >= 0 ? any_room : 0;
  }
}

static void dispatch_loop(Server* S) {
  (void)hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    S->cv_work.wait(lk, [&] {
      if (S->stop) return true;
      if (S->pending.empty()) return false;
      return *std::min_element(S->busy.begin(), S->busy.end()) < kUnitsPerStream;
    });
    if (S->stop && S->pending.empty()) return;
    if (S->pending.empty()) continue;

    // Form the next unit in arrival order.
    SUnit* u = new SUnit();
    SJob* head = S->pending.front();
    if (coalescable(*S, head)) {
      // A batch's scratch is about 4 bytes a key: keep it within a quarter
      // of the budget so batches never crowd out the other jobs.
      const size_t max_keys = std::max(S->coalesce_max, std::min(S->batch_max_keys, S->budget / 16));
      size_t keys = 0;
      for (auto it = S->pending.begin(); it != S->pending.end();) {
        if (coalescable(*S, *it) && keys + (*it)->d.n <= max_keys) {
          keys += (*it)->d.n;
          u->jobs.push_back(*it);
          it = S->pending.erase(it);
        } else {
          ++it;
        }
      }
      u->batch = u->jobs.size() > 1;
    } else {
      u->jobs.push_back(head);
      S->pending.pop_front();
    }
    S->dispatching++;
    lk.unlock();
    u->bytes = u->batch ? batch_bytes(*u) : call_bytes(u->jobs[0]->d);
    lk.lock();

    u->stream = admit(S, lk, u->bytes);
    S->busy[u->stream]++;
    hipStream_t s = S->streams[u->stream];
    lk.unlock();

    int st = PA_OK;
    if (u->batch) {
      std::vector<uint32_t*> ptrs;
      std::vector<uint64_t> lens;
      for (SJob* j : u->jobs) {
        ptrs.push_back((uint32_t*)j->d.keys);
        lens.push_back(j->d.n);
      }
      st = pa_segmented_sort_ptrs_u32(ptrs.data(), lens.data(), ptrs.size(), s);
    } else {
      st = launch_one(u->jobs[0]->d, s);
    }
    if (hipEventCreateWithFlags(&u->ev, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(u->ev, s) != hipSuccess) {
      (void)hipGetLastError();
      if (st == PA_OK) st = PA_ERR_HIP;
    }

    const std::string msg = st == PA_OK ? std::string() : std::string(pa_last_error_string());
    const size_t now_held = stream_scratch_held(s);
    lk.lock();
    S->held[u->stream] = now_held;
    S->peak = std::max(S->peak, total_held(*S));
    for (SJob* j : u->jobs) {
      j->status = st;
      j->error = msg;
    }
    S->st.launches++;
    if (u->batch) {
      S->st.coalesced_batches++;
      S->st.coalesced_jobs += u->jobs.size();
    }
    S->inflight.push_back(u);
    S->dispatching--;
    S->cv_done.notify_all();  // wakes the completer
  }
}

static void complete_loop(Server* S) {
  (void)hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    S->cv_done.wait(lk, [&] { return S->stop || !S->inflight.empty(); });
    if (S->stop && S->inflight.empty() && S->pending.empty() && S->dispatching == 0) return;
    std::vector<SUnit*> units = S->inflight;
    lk.unlock();
    std::vector<SUnit*> finished;
    for (SUnit* u : units) {
      hipError_t e = u->ev ? hipEventQuery(u->ev) : hipSuccess;
      if (e == hipSuccess) {
        finished.push_back(u);
      } else if (e != hipErrorNotReady) {
        (void)hipGetLastError();
        for (SJob* j : u->jobs) {
          j->status = PA_ERR_HIP;
          j->error = std::string("server: ") + hipGetErrorString(e);
        }
        finished.push_back(u);
      }
    }
    (void)hipGetLastError();
    const auto now = Clock::now();
    if (finished.empty()) std::this_thread::sleep_for(std::chrono::microseconds(20));
    lk.lock();
    for (SUnit* u : finished) {
      for (SJob* j : u->jobs) {
        j->done = now;
        j->finished = true;
        if (j->status == PA_OK)
          S->st.completed++;
        else
          S->st.failed++;
      }
      S->busy[u->stream]--;
      S->inflight.erase(std::find(S->inflight.begin(), S->inflight.end(), u));
      if (u->ev) (void)hipEventDestroy(u->ev);
      delete u;
    }
    if (!finished.empty()) {
      S->cv_done.notify_all();
      S->cv_work.notify_all();
    }
  }
}

}  // namespace pa

using namespace pa;

extern "C" {

pa_server* pa_server_create(size_t hbm_budget_bytes, int num_streams, size_t coalesce_max_keys) {
  if (num_streams < 1) num_streams = 1;
  Server* S = new Server();
  S->budget = hbm_budget_bytes ? hbm_budget_bytes : (size_t)64 << 30;
  if (coalesce_max_keys) S->coalesce_max = coalesce_max_keys;
  (void)hipGetDevice(&S->device);
  for (int i = 0; i < num_streams; ++i) {
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      for (hipStream_t t : S->streams) (void)hipStreamDestroy(t);
      delete S;
      set_last_error("pa_server_create: hipStreamCreateWithFlags failed");
      return nullptr;
    }
    S->streams.push_back(s);
    S->busy.push_back(0);
    S->held.push_back(0);
  }
  S->dispatcher = std::thread(dispatch_loop, S);
  S->completer = std::thread(complete_loop, S);
  return reinterpret_cast<pa_server*>(S);
}

int pa_server_submit(pa_server* srv, const pa_job* job, uint64_t* ticket) {
  Server* S = reinterpret_cast<Server*>(srv);
  if (!S || !job || !ticket) return arg_error("pa_server_submit: null argument");
  if (job->n == 0 || !job->keys) return arg_error("pa_server_submit: empty job");
  if (job->op != PA_OP_SORT && !job->out) return arg_error("pa_server_submit: select/histogram need an output");
  if (!is_device_ptr(job->keys) || (job->out && !is_device_ptr(job->out)))
    return arg_error("pa_server_submit: keys and outputs must be device memory");
  SJob* j = new SJob();
  j->d = *job;
  j->submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    j->ticket = S->next_ticket++;
    S->jobs[j->ticket] = j;
    S->pending.push_back(j);
    S->st.submitted++;
  }
  *ticket = j->ticket;
  S->cv_work.notify_all();
  return PA_OK;
}

int pa_server_wait(pa_server* srv, uint64_t ticket, double* latency_ms) {
  Server* S = reinterpret_cast<Server*>(srv);
  if (!S) return arg_error("pa_server_wait: null server");
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->jobs.find(ticket);
  if (it == S->jobs.end()) return arg_error("pa_server_wait: unknown ticket");
  SJob* j = it->second;
  S->cv_done.wait(lk, [&] { return j->finished; });
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(j->done - j->submit).count();
  const int st = j->status;
  // The message was recorded on the dispatcher's thread; hand it to this one.
  if (st != PA_OK) set_last_error(j->error.empty() ? std::string("pa_server_wait: job failed") : j->error);
  S->jobs.erase(it);
  delete j;
  return st;
}

int pa_server_drain(pa_server* srv) {
  Server* S = reinterpret_cast<Server*>(srv);
  if (!S) return arg_error("pa_server_drain: null server");
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv_done.wait(lk, [&] { return S->pending.empty() && S->inflight.empty() && S->dispatching == 0; });
  return PA_OK;
}

int pa_server_get_stats(pa_server* srv, pa_server_stats* out) {
  Server* S = reinterpret_cast<Server*>(srv);
  if (!S || !out) return arg_error("pa_server_get_stats: null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->st;
  out->peak_bytes_in_flight = S->peak;
  out->pending = S->pending.size();
  out->bytes_held = total_held(*S);
  return PA_OK;
}

int pa_job_scratch_bytes(const pa_job* job, size_t* bytes) {
  if (!job || !bytes) return arg_error("pa_job_scratch_bytes: null argument");
  if (job->op != PA_OP_SORT && job->op != PA_OP_SELECT && job->op != PA_OP_HIST)
    return arg_error("pa_job_scratch_bytes: unknown op");
  *bytes = scratch_alloc_bytes(call_bytes(*job));
  return PA_OK;
}

void pa_server_destroy(pa_server* srv) {
  Server* S = reinterpret_cast<Server*>(srv);
  if (!S) return;
  (void)pa_server_drain(srv);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  S->cv_done.notify_all();
  S->dispatcher.join();
  S->completer.join();
  for (hipStream_t s : S->streams) {
    // The library's block for this stream would otherwise outlive it.
    if (release_stream_scratch(s) != PA_OK) (void)hipGetLastError();
    (void)hipStreamDestroy(s);
  }
  for (auto& kv : S->jobs) delete kv.second;
  delete S;
}

}  // extern "C"
