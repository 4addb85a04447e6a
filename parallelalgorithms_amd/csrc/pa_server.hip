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
//   - a unit is admitted only while the estimated device scratch of the units
//     in flight stays within the budget, and goes to the stream with the least
//     work in flight, so new work overlaps work already on the GPU.
// A completer thread polls the units' events and stamps each job's
// completion time; latency = completion - submission (host steady clock).
// Each job's result is what the same call returns alone: units never share
// scratch (the library keeps scratch per stream and serialises each stream).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
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
  bool finished = false;
};

struct SUnit {
  std::vector<SJob*> jobs;
  int stream = 0;
  hipEvent_t ev = nullptr;
  size_t bytes = 0;
  bool batch = false;
};

struct Server {
  size_t budget = 0, in_use = 0, peak = 0;
  size_t coalesce_max = 1 << 20;     // largest sort that joins a batch
  size_t batch_max_keys = 64 << 20;  // keys per batch
  std::vector<hipStream_t> streams;
  std::vector<int> busy;  // units in flight per stream
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

static size_t job_bytes(const pa_job& j) {
  const size_t kb = j.key_type == PA_KEY_U64 ? 8 : 4;
  switch (j.op) {
    case PA_OP_SORT: return j.n * kb + j.n / 8192 * 8 * 1024 + (1 << 20);
    case PA_OP_SELECT: return j.n / 2 * kb + (1 << 20);
    default: return 1 << 20;
  }
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

static void dispatch_loop(Server* S) {
  (void)hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    S->cv_work.wait(lk, [&] {
      if (S->stop) return true;
      if (S->pending.empty()) return false;
      return *std::min_element(S->busy.begin(), S->busy.end()) < 2;
    });
    if (S->stop && S->pending.empty()) return;
    if (S->pending.empty()) continue;

    // Form the next unit in arrival order.
    SUnit* u = new SUnit();
    SJob* head = S->pending.front();
    if (coalescable(*S, head)) {
      size_t keys = 0;
      for (auto it = S->pending.begin(); it != S->pending.end();) {
        if (coalescable(*S, *it) && keys + (*it)->d.n <= S->batch_max_keys) {
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
    for (SJob* j : u->jobs) u->bytes += job_bytes(j->d);

    // Admission against the device budget (a unit larger than the whole
    // budget runs alone once nothing else is in flight).
    S->cv_done.wait(lk, [&] { return S->in_use == 0 || S->in_use + u->bytes <= S->budget || S->stop; });
    S->in_use += u->bytes;
    S->peak = std::max(S->peak, S->in_use);
    u->stream = (int)(std::min_element(S->busy.begin(), S->busy.end()) - S->busy.begin());
    S->busy[u->stream]++;
    S->dispatching++;
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

    lk.lock();
    for (SJob* j : u->jobs) j->status = st;
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
        for (SJob* j : u->jobs) j->status = PA_ERR_HIP;
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
      S->in_use -= u->bytes;
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
  for (hipStream_t s : S->streams) (void)hipStreamDestroy(s);
  for (auto& kv : S->jobs) delete kv.second;
  delete S;
}

}  // extern "C"
