// Request loop over one resident scene (fclx_server_*, include/fclx.h).
//
// Requests queue on the host.  A worker thread owns num_streams slots, each
// with its own HIP stream, device buffers (query mesh ids, the scene pose
// filled once per query slot, link poses, counts) and pinned staging.
// Whenever a slot is free it takes waiting requests, smallest first (oldest
// first once one has waited 1.5 ms), as long as they fit the batch size and
// the device-memory budget, moves their poses to the device (straight from
// the caller's memory when it is pinned, else through the slot's pinned
// staging, packed by the worker and its helper threads), sends a small
// table (each request's first query, link count and link mesh ids) from
// which a kernel writes every query's mesh ids, launches one first-contact
// collide batch (CollisionRequest(1, false) per link and waypoint), has a
// kernel write the counts into the slot's pinned host buffer (read-back
// copies queued behind the batch's kernels left the worker blocked in
// later copy calls more often) and records an event; when the event has
// fired it writes each request's results and reports its completion.
// Slots run concurrently on the device, so later requests overlap earlier
// ones.  Every query of a batch is
// independent of the others, so a request's result is the same alone and in
// any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../../include/fclx.h"

namespace fclx {
int context_device(const fclx_context* ctx);
int collide_batch_status(fclx_context* ctx, const fclx_collide_batch_args* a, void* stream,
                         int32_t* d_status);
int status_code(int32_t flags);
bool take_injected_failure(fclx_context* ctx);
}

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// Host memory from hipHostMalloc (or torch's pinned allocator): the device
// reads it directly, at the address returned.  Pageable memory: nullptr.
const double* pinned_device_address(const double* p) {
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return a.type == hipMemoryTypeHost ? static_cast<const double*>(a.devicePointer) : nullptr;
}

constexpr int64_t kQueryBytes = 8 + 96 + 96 + 4;  // ids, two poses, count
constexpr auto kAgeLimit = std::chrono::microseconds(1500);  // then a request goes first
constexpr int32_t kBatchDefault = 1 << 16;                 // queries per batch at most
constexpr int64_t kSlotMin = 1 << 18;  // queries a slot holds without growing
constexpr int kHelpers = 3;             // threads that pack poses beside the worker
constexpr int64_t kPackChunk = 1 << 18;  // bytes of pose copy per packing job

// A batch's counts and status flags into the slot's pinned host buffers,
// written by the device (no copy command that waits on the batch's kernel)
__global__ void publish_counts(const int32_t* counts, int32_t* h_counts, int64_t n,
                               const int32_t* status, int32_t* h_status) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) h_counts[i] = counts[i];
  if (i == 0) *h_status = *status;
}

// n 16-B words from pinned host memory into device memory, read by the
// device over the bus.  The loop moves every batch's input this way rather
// than with copy commands: those blocked the worker thread for 6-7 ms a few
// times a second under load.
__global__ void gather_words(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

// Query mesh ids of a batch from its request table: requests r = 0..R-1
// with first[r] .. first[r+1] their queries (waypoint-major), nlink[r]
// links whose mesh ids start at links[off[r]].
__global__ void fill_ids(int32_t* qm, const int64_t* first, const int32_t* nlink,
                         const int32_t* off, const int32_t* links, int nreq, int32_t scene,
                         int64_t nq, int32_t* status) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k == 0) *status = 0;  // the batch's status flags, before its collide kernels
  if (k >= nq) return;
  int lo = 0, hi = nreq - 1;  // the last r with first[r] <= k
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= k) lo = mid; else hi = mid - 1;
  }
  const int64_t j = k - first[lo];
  qm[2 * k] = scene;
  qm[2 * k + 1] = links[off[lo] + static_cast<int>(j % nlink[lo])];
}

__global__ void fill_pose(double* out, int n, fclx_server_config cfg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int k = 0; k < 12; ++k) out[12 * static_cast<size_t>(i) + k] = cfg.scene_pose[k];
}

struct Pending {
  fclx_request req;
  Clock::time_point t_submit;
  int64_t nq;
  bool pinned;                // link_poses is pinned host memory: read by the device directly
  const double* dev_poses;    // its device address
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  int64_t cap = 0;  // queries
  int32_t* d_qm = nullptr;
  double* d_tf1 = nullptr;
  double* d_tf2 = nullptr;
  int32_t* d_counts = nullptr;
  int64_t* d_num = nullptr;
  int32_t* d_status = nullptr;  // the batch's device status flags
  char* d_table = nullptr;      // the batch's request table (fill_ids)
  char* h_table = nullptr;
  double* h_tf2 = nullptr;
  int32_t* h_counts = nullptr;
  int32_t* h_status = nullptr;
  bool busy = false;
  int status = FCLX_OK;
  std::vector<Pending> batch;
  std::vector<int64_t> first;  // first query of each request of the batch
  int64_t queries = 0;
  Clock::time_point t_launch;
};

void free_slot_buffers(Slot& s) {
  (void)hipFree(s.d_qm);
  (void)hipFree(s.d_tf1);
  (void)hipFree(s.d_tf2);
  (void)hipFree(s.d_counts);
  (void)hipFree(s.d_num);
  (void)hipFree(s.d_status);
  (void)hipFree(s.d_table);
  if (s.h_table) (void)hipHostFree(s.h_table);
  s.d_table = s.h_table = nullptr;
  if (s.h_tf2) (void)hipHostFree(s.h_tf2);
  if (s.h_counts) (void)hipHostFree(s.h_counts);
  if (s.h_status) (void)hipHostFree(s.h_status);
  s.d_qm = s.d_counts = s.h_counts = s.d_status = s.h_status = nullptr;
  s.d_tf1 = s.d_tf2 = s.h_tf2 = nullptr;
  s.d_num = nullptr;
  s.cap = 0;
}

}  // namespace

struct fclx_server {
  fclx_context* ctx = nullptr;
  fclx_server_config cfg{};
  int device = 0;
  std::mutex mu;
  std::condition_variable cv;       // submissions and stop
  std::condition_variable cv_done;  // completions
  std::deque<Pending> queue;
  std::deque<fclx_completion> done;
  std::vector<Slot> slots;
  std::thread worker;
  bool stop = false;
  int64_t in_flight_bytes = 0;
  int64_t queued_queries = 0;  // of the waiting requests
  int64_t batches = 0, queries = 0;
  // pose packing: the worker posts copy jobs and takes a share of them
  // itself; helpers take the rest
  struct Job {
    void* dst;
    const void* src;
    size_t bytes;
  };
  std::vector<std::thread> helpers;
  std::mutex pack_mu;
  std::condition_variable pack_cv, pack_done_cv;
  std::vector<Job> jobs;
  size_t job_next = 0, jobs_done = 0;
  bool pack_stop = false;
};

namespace {

// Bytes of a batch's request table for up to n queries: the firsts (8 B,
// one more than the requests), link counts and offsets (4 B each), link
// mesh ids (4 B each; every request has at least as many queries as
// links, and at least one query).
size_t table_bytes(int64_t n) { return static_cast<size_t>(8 + 24 * n + 64); }

// Slot buffers for at least n queries (the slot is idle).
hipError_t slot_reserve(fclx_server* S, Slot& s, int64_t n) {
  if (n <= s.cap) return hipSuccess;
  free_slot_buffers(s);
  const int64_t cap = std::max<int64_t>(n, 1 << 16);
  hipError_t e;
  if ((e = hipMalloc(&s.d_qm, 8 * cap)) != hipSuccess ||
      (e = hipMalloc(&s.d_tf1, 96 * cap)) != hipSuccess ||
      (e = hipMalloc(&s.d_tf2, 96 * cap)) != hipSuccess ||
      (e = hipMalloc(&s.d_counts, 4 * cap)) != hipSuccess ||
      (e = hipMalloc(&s.d_num, 8)) != hipSuccess ||
      (e = hipMalloc(&s.d_status, 4)) != hipSuccess ||
      (e = hipHostMalloc(&s.h_status, 4)) != hipSuccess ||
      (e = hipMalloc(&s.d_table, table_bytes(cap))) != hipSuccess ||
      (e = hipHostMalloc(&s.h_table, table_bytes(cap))) != hipSuccess ||
      (e = hipHostMalloc(&s.h_tf2, 96 * cap)) != hipSuccess ||
      (e = hipHostMalloc(&s.h_counts, 4 * cap)) != hipSuccess) {
    free_slot_buffers(s);
    return e;
  }
  s.cap = cap;
  fill_pose<<<static_cast<unsigned>((cap + 255) / 256), 256, 0, s.stream>>>(
      s.d_tf1, static_cast<int>(cap), S->cfg);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // the stream once through the transfers of a batch (poses read from
  // pinned host memory, counts and status written to it, by kernels and by
  // a copy command, the fallback for unaligned poses), so that nothing is
  // set up inside the first requests' latency
  const size_t warm = static_cast<size_t>(std::min<int64_t>(cap, 1 << 16));
  std::memset(s.h_tf2, 0, 96 * warm);
  gather_words<<<static_cast<unsigned>(std::min<size_t>((6 * warm + 255) / 256, 1024)), 256, 0,
                 s.stream>>>(reinterpret_cast<const uint4*>(s.h_tf2),
                             reinterpret_cast<uint4*>(s.d_tf2), static_cast<int64_t>(6 * warm));
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(s.d_tf2, s.h_tf2, 96 * warm, hipMemcpyHostToDevice, s.stream)) !=
          hipSuccess)
    return e;
  publish_counts<<<static_cast<unsigned>((warm + 255) / 256), 256, 0, s.stream>>>(
      s.d_counts, s.h_counts, static_cast<int64_t>(warm), s.d_status, s.h_status);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return hipStreamSynchronize(s.stream);
}

void helper_main(fclx_server* S) {
  std::unique_lock<std::mutex> lock(S->pack_mu);
  while (true) {
    S->pack_cv.wait(lock, [S] { return S->pack_stop || S->job_next < S->jobs.size(); });
    if (S->pack_stop) return;
    const fclx_server::Job j = S->jobs[S->job_next++];
    lock.unlock();
    std::memcpy(j.dst, j.src, j.bytes);
    lock.lock();
    if (++S->jobs_done == S->jobs.size()) S->pack_done_cv.notify_all();
  }
}

// Runs the copy jobs on the worker and the helpers; returns when all are done.
void run_jobs(fclx_server* S, std::vector<fclx_server::Job>& jobs) {
  if (jobs.empty()) return;
  std::unique_lock<std::mutex> lock(S->pack_mu);
  S->jobs.swap(jobs);
  S->job_next = 0;
  S->jobs_done = 0;
  S->pack_cv.notify_all();
  while (S->job_next < S->jobs.size()) {
    const fclx_server::Job j = S->jobs[S->job_next++];
    lock.unlock();
    std::memcpy(j.dst, j.src, j.bytes);
    lock.lock();
    ++S->jobs_done;
  }
  S->pack_done_cv.wait(lock, [S] { return S->jobs_done == S->jobs.size(); });
  S->jobs.clear();
}

// Copies, launch and read-back of the slot's batch (filled by the caller).
int slot_launch(fclx_server* S, Slot& s) {
  int64_t k = 0;
  s.first.clear();
  // the request table: firsts, link counts, link offsets, link mesh ids
  const int R = static_cast<int>(s.batch.size());
  int64_t* first = reinterpret_cast<int64_t*>(s.h_table);
  int32_t* nlink = reinterpret_cast<int32_t*>(first + R + 1);
  int32_t* off = nlink + R;
  int32_t* links = off + R;
  int32_t nl = 0;
  std::vector<fclx_server::Job> jobs;
  for (int r = 0; r < R; ++r) {
    const Pending& p = s.batch[r];
    s.first.push_back(k);
    first[r] = k;
    nlink[r] = p.req.num_links;
    off[r] = nl;
    std::memcpy(links + nl, p.req.link_meshes, 4 * static_cast<size_t>(p.req.num_links));
    nl += p.req.num_links;
    if (!p.pinned) {
      const char* src = reinterpret_cast<const char*>(p.req.link_poses);
      char* dst = reinterpret_cast<char*>(s.h_tf2 + 12 * k);
      const size_t bytes = 96 * static_cast<size_t>(p.nq);
      for (size_t o = 0; o < bytes; o += kPackChunk)
        jobs.push_back({dst + o, src + o, std::min<size_t>(kPackChunk, bytes - o)});
    }
    k += p.nq;
  }
  first[R] = k;
  s.queries = k;
  const size_t tbytes = 8 * static_cast<size_t>(R + 1) + 4 * (2 * static_cast<size_t>(R) + nl);
  run_jobs(S, jobs);
  if (fclx::take_injected_failure(S->ctx)) return FCLX_ERR_DEVICE;
  // pinned host memory to the device: by the gather kernel from its device
  // address when 16-B aligned (the table is rounded up to whole words: its
  // buffers have room), else by a copy command from its host address
  auto to_device = [&](void* dst, const void* dev_src, const void* host_src, size_t bytes) {
    const int64_t n16 = static_cast<int64_t>((bytes + 15) / 16);
    if ((reinterpret_cast<uintptr_t>(dev_src) & 15) == 0 &&
        (bytes % 16 == 0 || host_src == s.h_table)) {
      const unsigned grid = static_cast<unsigned>(std::min<int64_t>((n16 + 255) / 256, 1024));
      gather_words<<<grid, 256, 0, s.stream>>>(static_cast<const uint4*>(dev_src),
                                               static_cast<uint4*>(dst), n16);
      return hipGetLastError() == hipSuccess;
    }
    return hipMemcpyAsync(dst, host_src, bytes, hipMemcpyHostToDevice, s.stream) == hipSuccess;
  };
  // the slot's own pinned buffers (hipHostMalloc) have one address for
  // host and device
  if (!to_device(s.d_table, s.h_table, s.h_table, tbytes)) return FCLX_ERR_DEVICE;
  // poses: pinned requests straight from the caller's memory, the others
  // from the staging in runs of consecutive requests
  int64_t run_first = -1;
  for (int r = 0; r <= R; ++r) {
    const bool staged = r < R && !s.batch[r].pinned;
    if (staged && run_first < 0) run_first = s.first[r];
    if (!staged && run_first >= 0) {
      const int64_t end = r < R ? s.first[r] : k;
      if (!to_device(s.d_tf2 + 12 * run_first, s.h_tf2 + 12 * run_first,
                     s.h_tf2 + 12 * run_first, 96 * static_cast<size_t>(end - run_first)))
        return FCLX_ERR_DEVICE;
      run_first = -1;
    }
    if (r < R && s.batch[r].pinned &&
        !to_device(s.d_tf2 + 12 * s.first[r], s.batch[r].dev_poses, s.batch[r].req.link_poses,
                   96 * static_cast<size_t>(s.batch[r].nq)))
      return FCLX_ERR_DEVICE;
  }
  {
    const int64_t* d_first = reinterpret_cast<const int64_t*>(s.d_table);
    const int32_t* d_nlink = reinterpret_cast<const int32_t*>(d_first + R + 1);
    const int32_t* d_off = d_nlink + R;
    const int32_t* d_links = d_off + R;
    fill_ids<<<static_cast<unsigned>((k + 255) / 256), 256, 0, s.stream>>>(
        s.d_qm, d_first, d_nlink, d_off, d_links, R, S->cfg.scene_mesh, k, s.d_status);
    if (hipGetLastError() != hipSuccess) return FCLX_ERR_DEVICE;
  }
  fclx_collide_batch_args a{};
  a.num_queries = static_cast<int32_t>(k);
  a.query_meshes = s.d_qm;
  a.tf1 = s.d_tf1;
  a.tf2 = s.d_tf2;
  a.num_max_contacts = 1;
  a.engine = S->cfg.engine;
  a.counts = s.d_counts;
  a.pairs = nullptr;
  a.pair_capacity = 0;
  a.num_pairs = s.d_num;
  int rc = fclx::collide_batch_status(S->ctx, &a, s.stream, s.d_status);
  if (rc != FCLX_OK) return rc;
  publish_counts<<<static_cast<unsigned>((k + 255) / 256), 256, 0, s.stream>>>(
      s.d_counts, s.h_counts, k, s.d_status, s.h_status);
  if (hipGetLastError() != hipSuccess || hipEventRecord(s.done, s.stream) != hipSuccess)
    return FCLX_ERR_DEVICE;
  return FCLX_OK;
}

// Results of a finished slot to its requests (called with mu not held).
void slot_retire(fclx_server* S, Slot& s, std::vector<fclx_completion>& out) {
  for (size_t i = 0; i < s.batch.size(); ++i) {
    const Pending& p = s.batch[i];
    if (s.status == FCLX_OK) {
      const int32_t* c = s.h_counts + s.first[i];
      for (int64_t j = 0; j < p.nq; ++j) p.req.result[j] = c[j] > 0 ? 1 : 0;
    }
    fclx_completion cp{};
    cp.tag = p.req.tag;
    cp.status = s.status;
    cp.num_queries = static_cast<int32_t>(p.nq);
    cp.queue_ms = std::chrono::duration<double, std::milli>(s.t_launch - p.t_submit).count();
    cp.latency_ms = ms_since(p.t_submit);
    out.push_back(cp);
  }
}

void worker_main(fclx_server* S) {
  (void)hipSetDevice(S->device);
  std::vector<fclx_completion> finished;
  std::unique_lock<std::mutex> lock(S->mu);
  while (true) {
    bool progress = false;
    // retire finished slots
    for (Slot& s : S->slots) {
      if (!s.busy) continue;
      lock.unlock();
      // any answer but "not ready" ends the batch: after a device error the
      // query keeps returning that error, and the slot must not wait for ever
      bool ready = s.status != FCLX_OK;
      if (!ready) {
        const hipError_t q = hipEventQuery(s.done);
        ready = q != hipErrorNotReady;
        if (q != hipSuccess && ready) {
          (void)hipGetLastError();
          s.status = FCLX_ERR_DEVICE;
        } else if (ready) {
          s.status = fclx::status_code(*s.h_status);
        }
      }
      if (ready) {
        finished.clear();
        slot_retire(S, s, finished);
      }
      lock.lock();
      if (!ready) continue;
      for (const Pending& p : s.batch) S->in_flight_bytes -= p.nq * kQueryBytes;
      s.batch.clear();
      s.busy = false;
      for (const auto& c : finished) S->done.push_back(c);
      S->cv_done.notify_all();
      progress = true;
    }
    // launch into free slots
    for (Slot& s : S->slots) {
      if (s.busy || S->queue.empty()) continue;
      int64_t nq = 0, bytes = 0;
      s.batch.clear();
      // a backlog goes out in larger batches (a share of it per stream, as
      // far as the slot holds without growing), which fill the device better
      // than many small ones
      const int64_t batch_cap = std::max<int64_t>(
          S->cfg.max_batch_queries,
          std::min<int64_t>(s.cap, S->queued_queries / static_cast<int64_t>(S->slots.size())));
      while (!S->queue.empty()) {
        // the smallest waiting request first, so that short requests do not
        // wait behind long ones; but the oldest first once it has waited
        // kAgeLimit, so that none waits for ever
        size_t pick = 0;
        if (Clock::now() - S->queue.front().t_submit < kAgeLimit) {
          for (size_t i = 1; i < S->queue.size(); ++i)
            if (S->queue[i].nq < S->queue[pick].nq) pick = i;
        }
        const Pending& p = S->queue[pick];
        const int64_t b = p.nq * kQueryBytes;
        const bool first = s.batch.empty();
        // a request larger than a batch or than the budget runs alone once
        // nothing else is in flight
        if (!first && nq + p.nq > batch_cap) break;
        if (S->in_flight_bytes + bytes + b > S->cfg.hbm_budget &&
            !(first && S->in_flight_bytes == 0))
          break;
        nq += p.nq;
        bytes += b;
        s.batch.push_back(p);
        S->queued_queries -= p.nq;
        S->queue.erase(S->queue.begin() + static_cast<std::ptrdiff_t>(pick));
      }
      if (s.batch.empty()) break;  // waits for memory to free up
      S->in_flight_bytes += bytes;
      s.busy = true;
      s.t_launch = Clock::now();
      ++S->batches;
      S->queries += nq;
      lock.unlock();
      int rc = slot_reserve(S, s, nq) == hipSuccess ? FCLX_OK : FCLX_ERR_OUT_OF_MEMORY;
      if (rc == FCLX_OK) rc = slot_launch(S, s);
      if (rc != FCLX_OK) (void)hipStreamSynchronize(s.stream);  // nothing left in flight
      lock.lock();
      s.status = rc;
      progress = true;
    }
    bool any_busy = false;
    for (const Slot& s : S->slots) any_busy |= s.busy;
    if (S->stop && S->queue.empty() && !any_busy) break;
    if (!progress) {
      if (any_busy)
        S->cv.wait_for(lock, std::chrono::microseconds(20));
      else
        S->cv.wait_for(lock, std::chrono::milliseconds(5));
    }
  }
}

}  // namespace

extern "C" {

int fclx_server_create(fclx_context* ctx, const fclx_server_config* cfg, fclx_server** out) {
  if (!ctx || !cfg || !out) return FCLX_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (cfg->num_streams < 1 || cfg->num_streams > 8 || cfg->hbm_budget < 0 ||
      cfg->max_batch_queries < 0 ||
      (cfg->engine != FCLX_ENGINE_FAST && cfg->engine != FCLX_ENGINE_PLAIN &&
       cfg->engine != FCLX_ENGINE_BOOL))
    return FCLX_ERR_INVALID_ARGUMENT;
  fclx_mesh_info info{};
  if (fclx_mesh_get_info(ctx, cfg->scene_mesh, &info) != FCLX_OK ||
      info.bv_type != FCLX_BV_OBBRSS)
    return FCLX_ERR_INVALID_ARGUMENT;
  auto* S = new fclx_server();
  S->ctx = ctx;
  S->cfg = *cfg;
  if (S->cfg.hbm_budget == 0) S->cfg.hbm_budget = int64_t(2) << 30;
  if (S->cfg.max_batch_queries == 0) S->cfg.max_batch_queries = kBatchDefault;
  S->device = fclx::context_device(ctx);
  (void)hipSetDevice(S->device);
  S->slots.resize(cfg->num_streams);
  for (Slot& s : S->slots) {
    if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
      for (Slot& t : S->slots) {
        if (t.stream) (void)hipStreamDestroy(t.stream);
        if (t.done) (void)hipEventDestroy(t.done);
      }
      delete S;
      return FCLX_ERR_DEVICE;
    }
  }
  // every slot at its full size now, so that no allocation (pinned host
  // memory above all) falls inside a request's latency
  const int64_t share = S->cfg.hbm_budget / (kQueryBytes * cfg->num_streams);
  const int64_t cap =
      std::max<int64_t>(1 << 16, std::min<int64_t>(
                                     std::max<int64_t>(S->cfg.max_batch_queries, kSlotMin), share));
  for (Slot& s : S->slots) {
    if (slot_reserve(S, s, cap) != hipSuccess) {
      for (Slot& t : S->slots) {
        free_slot_buffers(t);
        if (t.stream) (void)hipStreamDestroy(t.stream);
        if (t.done) (void)hipEventDestroy(t.done);
      }
      delete S;
      return FCLX_ERR_OUT_OF_MEMORY;
    }
  }
  // One batch of each launch shape the loop uses (4 query slots per
  // wavefront below 65,536 queries, 8 from there): the first launch of a
  // kernel costs milliseconds, which would otherwise fall inside the first
  // large request's latency.  The queries are the scene against itself at
  // its own pose, so each ends at its first leaf pair.
  {
    Slot& s = S->slots[0];
    const int n = static_cast<int>(std::min<int64_t>(s.cap, 1 << 16));
    std::vector<int32_t> ids(2 * static_cast<size_t>(n), cfg->scene_mesh);
    bool ok = hipMemcpy(s.d_qm, ids.data(), 8 * static_cast<size_t>(n), hipMemcpyHostToDevice) ==
                  hipSuccess &&
              hipMemcpy(s.d_tf2, s.d_tf1, 96 * static_cast<size_t>(n),
                        hipMemcpyDeviceToDevice) == hipSuccess;
    for (const int m : {n, std::min(n, 1024)}) {
      if (!ok) break;
      fclx_collide_batch_args a{};
      a.num_queries = m;
      a.query_meshes = s.d_qm;
      a.tf1 = s.d_tf1;
      a.tf2 = s.d_tf2;
      a.num_max_contacts = 1;
      a.engine = S->cfg.engine;
      a.counts = s.d_counts;
      a.num_pairs = s.d_num;
      ok = fclx::collide_batch_status(ctx, &a, s.stream, s.d_status) == FCLX_OK &&
           hipStreamSynchronize(s.stream) == hipSuccess;
    }
    (void)hipGetLastError();
  }
  for (int i = 0; i < kHelpers; ++i) S->helpers.emplace_back(helper_main, S);
  S->worker = std::thread(worker_main, S);
  *out = S;
  return FCLX_OK;
}

int fclx_server_submit(fclx_server* S, const fclx_request* req) {
  if (!S || !req || req->num_links < 1 || req->num_waypoints < 1 || !req->link_meshes ||
      !req->link_poses || !req->result)
    return FCLX_ERR_INVALID_ARGUMENT;
  const int64_t nq = static_cast<int64_t>(req->num_links) * req->num_waypoints;
  if (nq > (int64_t(1) << 30)) return FCLX_ERR_INVALID_ARGUMENT;
  for (int32_t l = 0; l < req->num_links; ++l) {
    fclx_mesh_info info{};
    if (fclx_mesh_get_info(S->ctx, req->link_meshes[l], &info) != FCLX_OK ||
        info.bv_type != FCLX_BV_OBBRSS)
      return FCLX_ERR_INVALID_ARGUMENT;
  }
  Pending p;
  p.req = *req;
  p.t_submit = Clock::now();
  p.nq = nq;
  p.dev_poses = pinned_device_address(req->link_poses);
  p.pinned = p.dev_poses != nullptr;
  {
    std::lock_guard<std::mutex> lock(S->mu);
    if (S->stop) return FCLX_ERR_INVALID_ARGUMENT;
    S->queue.push_back(p);
    S->queued_queries += p.nq;
  }
  S->cv.notify_one();
  return FCLX_OK;
}

int fclx_server_poll(fclx_server* S, fclx_completion* out, int32_t max, int32_t wait_ms,
                     int32_t* n) {
  if (!S || !n || max < 0 || (max > 0 && !out)) return FCLX_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lock(S->mu);
  if (S->done.empty() && wait_ms > 0)
    S->cv_done.wait_for(lock, std::chrono::milliseconds(wait_ms),
                        [S] { return !S->done.empty(); });
  int32_t k = 0;
  while (k < max && !S->done.empty()) {
    out[k++] = S->done.front();
    S->done.pop_front();
  }
  *n = k;
  return FCLX_OK;
}

int fclx_server_replay(fclx_server* S, const fclx_request* requests, const double* arrival_ms,
                       int32_t n, double* latency_ms, double* elapsed_ms, double* queue_ms) {
  if (!S || n < 0 || (n > 0 && (!requests || !arrival_ms || !latency_ms)))
    return FCLX_ERR_INVALID_ARGUMENT;
  const Clock::time_point t0 = Clock::now();
  std::unordered_map<int64_t, int32_t> index;  // tag -> request
  std::vector<fclx_completion> buf(256);
  int32_t got = 0, next = 0;
  int status = FCLX_OK;
  auto collect = [&](int32_t wait) {
    int32_t k = 0;
    fclx_server_poll(S, buf.data(), static_cast<int32_t>(buf.size()), wait, &k);
    for (int32_t i = 0; i < k; ++i) {
      auto it = index.find(buf[i].tag);
      if (it == index.end()) continue;
      latency_ms[it->second] = buf[i].latency_ms;
      if (queue_ms) queue_ms[it->second] = buf[i].queue_ms;
      if (buf[i].status != FCLX_OK) status = buf[i].status;
      ++got;
    }
  };
  while (next < n) {
    const auto due = t0 + std::chrono::duration_cast<Clock::duration>(
                              std::chrono::duration<double, std::milli>(arrival_ms[next]));
    if (Clock::now() < due) {
      collect(0);
      std::this_thread::sleep_until(due);
      continue;
    }
    fclx_request r = requests[next];
    r.tag = next;
    index[next] = next;
    const int rc = fclx_server_submit(S, &r);
    if (rc != FCLX_OK) return rc;
    ++next;
  }
  while (got < n) collect(100);
  if (elapsed_ms) *elapsed_ms = ms_since(t0);
  return status;
}

int fclx_server_stats(fclx_server* S, int64_t* batches, int64_t* queries) {
  if (!S) return FCLX_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(S->mu);
  if (batches) *batches = S->batches;
  if (queries) *queries = S->queries;
  return FCLX_OK;
}

int fclx_server_destroy(fclx_server* S) {
  if (!S) return FCLX_ERR_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> lock(S->mu);
    S->stop = true;
  }
  S->cv.notify_all();
  if (S->worker.joinable()) S->worker.join();
  {
    std::lock_guard<std::mutex> lock(S->pack_mu);
    S->pack_stop = true;
  }
  S->pack_cv.notify_all();
  for (auto& h : S->helpers) h.join();
  (void)hipSetDevice(S->device);
  for (Slot& s : S->slots) {
    free_slot_buffers(s);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    if (s.done) (void)hipEventDestroy(s.done);
  }
  delete S;
  return FCLX_OK;
}

}  // extern "C"
