// Serving loop (include/hpc_ode.h, hpo_server_*): continuous batching of chemistry-substep
// requests over device-resident cell streams.
//
// Device memory: every stream holds its rows y [n][K], pressures P [n] and statuses [n] from
// open to close (admitted against the HBM budget) in one block from the server's pool. The
// pool keeps freed blocks in power-of-two size classes and hands them out again, so opening
// and closing streams under load neither allocates nor frees device memory once the classes
// in use are warm (a hipFree waits for the whole device, which here is kept busy by the
// launches). Every launch slot owns a HIP stream, a work-queue counter, per-wave scratch
// sized for a whole-device launch of any of the server's mechanisms, and packing buffers of
// max_batch_cells rows (device; the rows twice with the FP64 fallback, which restarts failed
// cells of an in-place launch from a copy of its input) plus pinned host staging.
//
// Opening a stream does no device work: its rows are copied into pinned host memory (a pool
// like the device one), and the stream's first launch uploads them on its slot's stream, just
// before packing. An upload on a stream of its own shared the device's few hardware queues
// with the slots' persistent launches and waited behind them: at 70 % load opens took 8.5 s
// (median) and the load generator fell seconds behind its arrival times. Downloads of a
// stream's state (hpo_server_stream_get) run on a dedicated non-blocking stream, never on the
// null stream.
//
// A launch packs its requests' rows into the slot buffers (device-to-device), orders them by
// descending initial temperature on the device (opts.sort_cells = 1, the default), integrates
// them in place with the persistent BDF kernel, scatters them back to their streams and stages the
// rows callers asked for in pinned memory; an event marks the end. The scheduler thread polls
// the slots' events, completes finished batches (copies staged rows to caller buffers, wakes
// waiters) and fills free slots from the pending queue in arrival order. It holds the server
// lock only to move requests between queues: the HIP calls of a launch and the host copies of
// a completion run unlocked (their streams are in flight, so no client call touches them).
//
// Lifetime: a request holds a reference to its stream, so a stream closed between the
// completion of its substep and the wait for it stays valid until that wait returns.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "capi_internal.hpp"

using namespace hpo::capi;
using Clock = std::chrono::steady_clock;

namespace {

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// Device blocks (or pinned host blocks) in power-of-two size classes (at least 64 KiB), reused
// after release and freed only when the server is destroyed.
struct BlockPool {
  bool host = false;
  std::mutex mu;
  std::map<int64_t, std::vector<void*>> free_blocks;   // class bytes -> blocks
  std::vector<void*> all;
  static int64_t class_of(int64_t bytes) {
    int64_t c = 64 << 10;
    while (c < bytes) c <<= 1;
    return c;
  }
  void* get(int64_t cls) {
    {
      std::lock_guard<std::mutex> g(mu);
      auto& v = free_blocks[cls];
      if (!v.empty()) {
        void* p = v.back();
        v.pop_back();
        return p;
      }
    }
    void* p = nullptr;
    if (host)
      HPO_HIP(hipHostMalloc(&p, static_cast<size_t>(cls), hipHostMallocDefault));
    else
      p = dalloc<uint8_t>(static_cast<size_t>(cls));
    std::lock_guard<std::mutex> g(mu);
    all.push_back(p);
    return p;
  }
  void put(void* p, int64_t cls) {
    std::lock_guard<std::mutex> g(mu);
    free_blocks[cls].push_back(p);
  }
  void release() {
    for (void* p : all) (void)(host ? hipHostFree(p) : hipFree(p));
    all.clear();
    free_blocks.clear();
  }
};

struct SrvStream {
  int64_t id = 0;
  int mech = 0;
  int64_t n = 0, ld = 0, bytes = 0;   // bytes: the pool class charged to the budget
  void* block = nullptr;
  double *d_y = nullptr, *d_P = nullptr;
  int32_t* d_status = nullptr;
  bool in_flight = false;   // a substep of this stream is pending or running
  BlockPool* pool = nullptr;
  // Pinned copy of the opening rows and pressures, laid out as the device block's first
  // upload_bytes; the stream's first launch uploads it on its slot's stream.
  void* h_stage = nullptr;
  int64_t h_cls = 0, upload_bytes = 0;
  BlockPool* hpool = nullptr;
  ~SrvStream() {
    if (block && pool) pool->put(block, bytes);
    if (h_stage && hpool) hpool->put(h_stage, h_cls);
  }
};

struct SrvReq {
  int64_t id = 0;
  std::shared_ptr<SrvStream> st;
  int64_t ncells = 0;
  double dt = 0;
  double* y_out = nullptr;
  int64_t ld_out = 0;
  int32_t* status_out = nullptr;
  Clock::time_point t_submit, t_launch, t_done;
  void* h_stage = nullptr;     // the stream's opening rows, uploaded by this request's launch
  int64_t h_cls = 0;
  int64_t offset = 0;          // first row in its batch
  int64_t batch_cells = 0;
  int32_t batch_requests = 0;
  int32_t failed = 0;
  bool done = false;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  ExecRes res;
  int64_t cap = 0, ldmax = 0;
  double *d_y = nullptr, *d_P = nullptr, *d_dt = nullptr;
  int32_t* d_status = nullptr;
  int32_t* d_order = nullptr;     // work-queue order of the packed cells (sort_cells = 1)
  uint32_t* d_owork = nullptr;    // its bucket counters
  double *h_y = nullptr, *h_dt = nullptr;   // pinned
  int32_t* h_status = nullptr;              // pinned
  std::vector<SrvReq*> batch;
  bool busy = false;
};

}  // namespace

struct hpo_server {
  std::vector<hpo_mech*> mechs;
  hpo_server_opts opts{};
  int device = 0;
  std::vector<Slot> slots;
  BlockPool pool;
  BlockPool hpool;            // pinned staging of opening rows
  hipStream_t io = nullptr;   // downloads of stream state (hpo_server_stream_get)
  std::mutex io_mu;
  std::map<int64_t, std::shared_ptr<SrvStream>> streams;
  std::map<int64_t, std::unique_ptr<SrvReq>> reqs;
  std::deque<SrvReq*> pending;
  int64_t next_stream = 1, next_req = 1;
  int64_t used = 0, budget = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  std::thread th;
  std::string thread_error;

  void release() {
    for (auto& sl : slots) {
      for (void* p : {static_cast<void*>(sl.d_y), static_cast<void*>(sl.d_P), static_cast<void*>(sl.d_dt),
                      static_cast<void*>(sl.d_status), static_cast<void*>(sl.res.d_queue),
                      static_cast<void*>(sl.res.d_scratch), static_cast<void*>(sl.res.d_backup),
                      static_cast<void*>(sl.d_order), static_cast<void*>(sl.d_owork)})
        if (p) (void)hipFree(p);
      for (void* p : {static_cast<void*>(sl.h_y), static_cast<void*>(sl.h_dt), static_cast<void*>(sl.h_status)})
        if (p) (void)hipHostFree(p);
      if (sl.done) (void)hipEventDestroy(sl.done);
      if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    slots.clear();
    reqs.clear();
    pending.clear();
    streams.clear();
    if (io) (void)hipStreamDestroy(io);
    io = nullptr;
    pool.release();
    hpool.release();
  }

  // Packs a batch into slot k and launches it. Called by the scheduler thread WITHOUT the
  // lock: the batch's requests are out of the pending queue and their streams are in flight
  // (close and stream_get refuse them), so nothing here is shared with client calls.
  void launch(int k, const std::vector<SrvReq*>& batch) {
    Slot& sl = slots[k];
    hpo_mech* m = mechs[batch.front()->st->mech];
    const int64_t ld = m->host.K();
    int64_t total = 0;
    for (SrvReq* r : batch) total += r->st->n;
    const auto now = Clock::now();
    int64_t off = 0;
    for (SrvReq* r : batch) {
      SrvStream& st = *r->st;
      if (st.h_stage) {
        // first substep of the stream: its opening rows go up on this slot's stream
        HPO_HIP(hipMemcpyAsync(st.block, st.h_stage, st.upload_bytes, hipMemcpyHostToDevice, sl.stream));
        r->h_stage = st.h_stage;
        r->h_cls = st.h_cls;
        st.h_stage = nullptr;
      }
      r->offset = off;
      r->t_launch = now;
      r->batch_cells = total;
      r->batch_requests = static_cast<int32_t>(batch.size());
      HPO_HIP(hipMemcpyAsync(sl.d_y + off * ld, st.d_y, st.n * ld * sizeof(double), hipMemcpyDeviceToDevice,
                             sl.stream));
      HPO_HIP(hipMemcpyAsync(sl.d_P + off, st.d_P, st.n * sizeof(double), hipMemcpyDeviceToDevice, sl.stream));
      for (int64_t i = 0; i < st.n; ++i) sl.h_dt[off + i] = r->dt;
      off += st.n;
    }
    HPO_HIP(hipMemcpyAsync(sl.d_dt, sl.h_dt, total * sizeof(double), hipMemcpyHostToDevice, sl.stream));
    // consecutive launches of at most launch_cells packed rows each (one when 0), each with its
    // own work-queue order
    const int64_t chunk = opts.launch_cells > 0 ? opts.launch_cells : total;
    for (int64_t c0 = 0; c0 < total; c0 += chunk) {
      const int64_t cnt = std::min(chunk, total - c0);
      const int32_t* order = nullptr;
      if (opts.opts.sort_cells == 1 && cnt > 1) {
        HPO_HIP(hpo::launch_t0_order(sl.d_y + c0 * ld, cnt, ld, sl.d_owork, sl.d_order + c0, sl.stream));
        order = sl.d_order + c0;
      }
      launch_integrate(m, cnt, sl.d_y + c0 * ld, sl.d_y + c0 * ld, ld, sl.d_P + c0, sl.d_dt + c0, sl.d_status + c0,
                       nullptr, order, nullptr, opts.opts, sl.stream, nullptr, &sl.res);
    }
    for (SrvReq* r : batch) {
      SrvStream& st = *r->st;
      HPO_HIP(hipMemcpyAsync(st.d_y, sl.d_y + r->offset * ld, st.n * ld * sizeof(double), hipMemcpyDeviceToDevice,
                             sl.stream));
      HPO_HIP(hipMemcpyAsync(st.d_status, sl.d_status + r->offset, st.n * sizeof(int32_t),
                             hipMemcpyDeviceToDevice, sl.stream));
      if (r->y_out)
        HPO_HIP(hipMemcpyAsync(sl.h_y + r->offset * ld, sl.d_y + r->offset * ld, st.n * ld * sizeof(double),
                               hipMemcpyDeviceToHost, sl.stream));
    }
    HPO_HIP(hipMemcpyAsync(sl.h_status, sl.d_status, total * sizeof(int32_t), hipMemcpyDeviceToHost, sl.stream));
    HPO_HIP(hipEventRecord(sl.done, sl.stream));
  }

  // Copies the finished batch of slot k to the callers' buffers. Called WITHOUT the lock (the
  // requests are not yet marked done, so their callers are still waiting); finish() then
  // publishes the results under the lock.
  void complete(int k) {
    Slot& sl = slots[k];
    const auto now = Clock::now();
    for (SrvReq* r : sl.batch) {
      SrvStream& st = *r->st;
      const int64_t ld = st.ld;
      if (r->y_out)
        for (int64_t i = 0; i < st.n; ++i)
          std::memcpy(r->y_out + i * r->ld_out, sl.h_y + (r->offset + i) * ld, ld * sizeof(double));
      int32_t failed = 0;
      for (int64_t i = 0; i < st.n; ++i) failed += (sl.h_status[r->offset + i] != HPO_CELL_SUCCESS);
      if (r->status_out) std::memcpy(r->status_out, sl.h_status + r->offset, st.n * sizeof(int32_t));
      r->failed = failed;
      r->t_done = now;
      if (r->h_stage) hpool.put(r->h_stage, r->h_cls);
      r->h_stage = nullptr;
    }
  }

  // Marks slot k's requests done and frees the slot (lock held).
  void finish(int k) {
    Slot& sl = slots[k];
    for (SrvReq* r : sl.batch) {
      r->done = true;
      r->st->in_flight = false;
    }
    sl.batch.clear();
    sl.busy = false;
    cv_done.notify_all();
  }

  // Next batch in arrival order: the oldest pending request fixes the mechanism; later
  // requests of that mechanism join while the batch has room (lock held).
  std::vector<SrvReq*> take_batch(int64_t cap) {
    std::vector<SrvReq*> b;
    int mech = -1;
    int64_t cells = 0;
    for (auto it = pending.begin(); it != pending.end();) {
      SrvReq* r = *it;
      if ((mech < 0 || r->st->mech == mech) && cells + r->st->n <= cap) {
        mech = r->st->mech;
        cells += r->st->n;
        b.push_back(r);
        it = pending.erase(it);
      } else {
        ++it;
      }
    }
    return b;
  }

  void run() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    try {
      for (;;) {
        bool any_busy = false;
        for (int k = 0; k < static_cast<int>(slots.size()); ++k) {
          if (!slots[k].busy) continue;
          lk.unlock();
          hipError_t q = hipEventQuery(slots[k].done);
          if (q == hipSuccess) complete(k);
          lk.lock();
          if (q == hipSuccess)
            finish(k);
          else if (q == hipErrorNotReady)
            any_busy = true;
          else
            HPO_HIP(q);
        }
        if (stop && !any_busy) break;
        bool launched = false;
        if (!stop)
          for (int k = 0; k < static_cast<int>(slots.size()) && !pending.empty(); ++k) {
            if (slots[k].busy) continue;
            std::vector<SrvReq*> b = take_batch(slots[k].cap);
            if (b.empty()) break;
            slots[k].batch = std::move(b);
            slots[k].busy = true;
            lk.unlock();
            launch(k, slots[k].batch);
            lk.lock();
            launched = true;
            any_busy = true;
          }
        if (launched) continue;
        if (any_busy) {
          // poll the running batches without holding the lock
          lk.unlock();
          std::this_thread::sleep_for(std::chrono::microseconds(50));
          lk.lock();
        } else {
          cv_work.wait(lk, [&] { return stop || !pending.empty(); });
        }
      }
    } catch (const std::exception& e) {
      if (!lk.owns_lock()) lk.lock();
      thread_error = e.what();
      stop = true;
      for (auto& kv : reqs) kv.second->done = true;
      cv_done.notify_all();
    }
  }
};

extern "C" {

void hpo_server_opts_default(hpo_server_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->n_slots = 8;
  o->max_batch_cells = 262144;
  hpo_opts_default(&o->opts);
  o->opts.sort_cells = 1;   // each launch pulls its hottest cells first (an order built on the device)
  o->opts.lu_precision = 1; // FP32 Newton matrix with FP64 refinement (and the FP64 fallback)
}

int hpo_server_create(hpo_mech* const* mechs, int32_t n_mechs, const hpo_server_opts* o, hpo_server** out) {
  if (!mechs || n_mechs <= 0 || !out) return fail(HPO_E_INVALID, "bad argument");
  *out = nullptr;
  auto s = std::make_unique<hpo_server>();
  try {
    hpo_server_opts_default(&s->opts);
    if (o) s->opts = *o;
    if (s->opts.n_slots <= 0) s->opts.n_slots = 8;
    if (s->opts.n_slots > 16) throw std::invalid_argument("n_slots must be at most 16");
    if (s->opts.max_batch_cells <= 0) s->opts.max_batch_cells = 262144;
    if (s->opts.launch_cells < 0) throw std::invalid_argument("launch_cells must be 0 or positive");
    if (s->opts.cu_share_pct == 0) s->opts.cu_share_pct = 100;
    if (s->opts.cu_share_pct < 1 || s->opts.cu_share_pct > 100)
      throw std::invalid_argument("cu_share_pct must be in 0..100");
    check_opts(s->opts.opts);
    if (s->opts.opts.sort_cells == 2) throw std::invalid_argument("sort_cells = 2 needs a batch");
    HPO_HIP(hipGetDevice(&s->device));
    int64_t scratch = 0, ldmax = 0;
    for (int i = 0; i < n_mechs; ++i) {
      if (!mechs[i]) throw std::invalid_argument("null mechanism");
      hpo_mech* m = mechs[i];
      (void)use_block_layout(m, s->opts.opts);   // throws when no kernel fits the state size
      m->ensure_device();
      s->mechs.push_back(m);
      scratch = std::max<int64_t>(scratch, max_scratch_doubles(m));
      ldmax = std::max<int64_t>(ldmax, m->host.K());
    }
    size_t fr = 0, to = 0;
    HPO_HIP(hipMemGetInfo(&fr, &to));
    HPO_HIP(hipStreamCreateWithFlags(&s->io, hipStreamNonBlocking));
    s->hpool.host = true;
    s->slots.resize(s->opts.n_slots);
    for (Slot& sl : s->slots) {
      sl.cap = s->opts.max_batch_cells;
      sl.ldmax = ldmax;
      HPO_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
      HPO_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
      sl.res.d_queue = dalloc<uint32_t>(4);
      sl.res.scratch_doubles = scratch;
      sl.res.share_pct = s->opts.cu_share_pct;
      sl.res.d_scratch = dalloc<double>(static_cast<size_t>(scratch));
      sl.d_y = dalloc<double>(static_cast<size_t>(sl.cap * ldmax));
      if (s->opts.opts.lu_precision == 1 && s->opts.opts.fp64_fallback) {
        // launches run in place: the FP64 fallback restarts failed cells from a copy of the rows
        sl.res.backup_doubles = sl.cap * ldmax;
        sl.res.d_backup = dalloc<double>(static_cast<size_t>(sl.res.backup_doubles));
      }
      sl.d_P = dalloc<double>(sl.cap);
      sl.d_dt = dalloc<double>(sl.cap);
      sl.d_status = dalloc<int32_t>(sl.cap);
      sl.d_order = dalloc<int32_t>(sl.cap);
      sl.d_owork = dalloc<uint32_t>(hpo::kOrderBuckets);
      HPO_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.h_y), sl.cap * ldmax * sizeof(double), hipHostMallocDefault));
      HPO_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.h_dt), sl.cap * sizeof(double), hipHostMallocDefault));
      HPO_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.h_status), sl.cap * sizeof(int32_t), hipHostMallocDefault));
    }
    s->budget = s->opts.hbm_budget_bytes > 0 ? s->opts.hbm_budget_bytes : static_cast<int64_t>(0.9 * fr);
    s->th = std::thread([p = s.get()] { p->run(); });
    *out = s.release();
    return HPO_OK;
  } catch (...) {
    s->release();
    return map_exceptions();
  }
}

void hpo_server_destroy(hpo_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->th.joinable()) s->th.join();
  (void)hipSetDevice(s->device);
  s->release();
  delete s;
}

int hpo_server_open_stream(hpo_server* s, int32_t mech, int64_t n, const double* y0, int64_t ld, const double* P,
                           double P_scalar, int64_t* id) {
  if (!s || !y0 || !id || n <= 0) return fail(HPO_E_INVALID, "bad argument");
  if (mech < 0 || mech >= static_cast<int32_t>(s->mechs.size())) return fail(HPO_E_INVALID, "no such mechanism");
  const int64_t K = s->mechs[mech]->host.K();
  if (ld < K) return fail(HPO_E_INVALID, "ld smaller than the state length");
  if (n > s->opts.max_batch_cells) return fail(HPO_E_INVALID, "stream larger than a launch (max_batch_cells)");
  // y [n][K], then P [n], then status [n], each 256-byte aligned inside one pool block
  const int64_t y_bytes = (n * K * 8 + 255) & ~int64_t(255);
  const int64_t p_bytes = (n * 8 + 255) & ~int64_t(255);
  const int64_t cls = BlockPool::class_of(y_bytes + p_bytes + n * 4);
  auto st = std::make_shared<SrvStream>();
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (!s->thread_error.empty()) return fail(HPO_E_DEVICE, s->thread_error);
    if (s->used + cls > s->budget) return fail(HPO_E_BUSY, "HBM budget of the server exhausted");
    s->used += cls;
    st->id = s->next_stream++;
  }
  try {
    HPO_HIP(hipSetDevice(s->device));
    st->mech = mech;
    st->n = n;
    st->ld = K;
    st->bytes = cls;
    st->block = s->pool.get(cls);
    st->pool = &s->pool;
    uint8_t* base = static_cast<uint8_t*>(st->block);
    st->d_y = reinterpret_cast<double*>(base);
    st->d_P = reinterpret_cast<double*>(base + y_bytes);
    st->d_status = reinterpret_cast<int32_t*>(base + y_bytes + p_bytes);
    // staged in pinned host memory: no device work here (an upload on a stream of its own
    // waited behind the slots' launches on the shared hardware queues, for seconds under load)
    st->upload_bytes = y_bytes + n * 8;
    st->h_cls = BlockPool::class_of(st->upload_bytes);
    st->h_stage = s->hpool.get(st->h_cls);
    st->hpool = &s->hpool;
    uint8_t* hb = static_cast<uint8_t*>(st->h_stage);
    double* hy = reinterpret_cast<double*>(hb);
    double* hP = reinterpret_cast<double*>(hb + y_bytes);
    for (int64_t i = 0; i < n; ++i) std::memcpy(hy + i * K, y0 + i * ld, K * sizeof(double));
    for (int64_t i = 0; i < n; ++i) hP[i] = P ? P[i] : P_scalar;
  } catch (...) {
    int rc = map_exceptions();
    st.reset();   // returns the block to the pool
    std::lock_guard<std::mutex> g(s->mu);
    s->used -= cls;
    return rc;
  }
  std::lock_guard<std::mutex> g(s->mu);
  *id = st->id;
  s->streams[st->id] = std::move(st);
  return HPO_OK;
}

int hpo_server_close_stream(hpo_server* s, int64_t id) {
  if (!s) return fail(HPO_E_INVALID, "null server");
  std::shared_ptr<SrvStream> st;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->streams.find(id);
    if (it == s->streams.end()) return fail(HPO_E_INVALID, "no such stream");
    if (it->second->in_flight) return fail(HPO_E_BUSY, "a substep of this stream is in flight");
    st = std::move(it->second);
    s->streams.erase(it);
    s->used -= st->bytes;
  }
  // the block returns to the pool when the last reference (this one, or a request not yet
  // waited for) goes
  return HPO_OK;
}

int hpo_server_submit(hpo_server* s, int64_t sid, double dt, double* y_out, int64_t ld_out, int32_t* status_out,
                      int64_t* rid) {
  if (!s || !rid) return fail(HPO_E_INVALID, "bad argument");
  if (!(dt > 0.0)) return fail(HPO_E_INVALID, "dt must be positive");
  std::lock_guard<std::mutex> g(s->mu);
  if (!s->thread_error.empty()) return fail(HPO_E_DEVICE, s->thread_error);
  auto it = s->streams.find(sid);
  if (it == s->streams.end()) return fail(HPO_E_INVALID, "no such stream");
  SrvStream& st = *it->second;
  if (y_out && ld_out < st.ld) return fail(HPO_E_INVALID, "ld_out smaller than the state length");
  if (st.in_flight) return fail(HPO_E_BUSY, "the previous substep of this stream has not completed");
  auto r = std::make_unique<SrvReq>();
  r->id = s->next_req++;
  r->st = it->second;
  r->ncells = st.n;
  r->dt = dt;
  r->y_out = y_out;
  r->ld_out = ld_out;
  r->status_out = status_out;
  r->t_submit = Clock::now();
  st.in_flight = true;
  *rid = r->id;
  s->pending.push_back(r.get());
  s->reqs[r->id] = std::move(r);
  s->cv_work.notify_one();
  return HPO_OK;
}

int hpo_server_wait(hpo_server* s, int64_t rid, double timeout_s, hpo_request_info* info) {
  if (!s) return fail(HPO_E_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(rid);
  if (it == s->reqs.end()) return fail(HPO_E_INVALID, "no such request (or already waited for)");
  SrvReq* r = it->second.get();
  auto ready = [&] { return r->done; };
  if (timeout_s < 0)
    s->cv_done.wait(lk, ready);
  else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
    return fail(HPO_E_TIMEOUT, "request not completed");
  if (!s->thread_error.empty()) return fail(HPO_E_DEVICE, s->thread_error);
  if (info) {
    info->latency_ms = ms_between(r->t_submit, r->t_done);
    info->queue_ms = ms_between(r->t_submit, r->t_launch);
    info->run_ms = ms_between(r->t_launch, r->t_done);
    info->ncells = r->ncells;
    info->batch_cells = r->batch_cells;
    info->batch_requests = r->batch_requests;
    info->failed_cells = r->failed;
  }
  s->reqs.erase(it);
  return HPO_OK;
}

int hpo_server_stream_get(hpo_server* s, int64_t sid, double* y, int64_t ld) {
  if (!s || !y) return fail(HPO_E_INVALID, "bad argument");
  std::shared_ptr<SrvStream> st;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->streams.find(sid);
    if (it == s->streams.end()) return fail(HPO_E_INVALID, "no such stream");
    st = it->second;
    if (st->in_flight) return fail(HPO_E_BUSY, "a substep of this stream is in flight");
    // held in flight while the copy runs, so no substep is submitted over it
    st->in_flight = true;
  }
  int rc = HPO_OK;
  if (ld < st->ld) {
    rc = fail(HPO_E_INVALID, "ld smaller than the state length");
  } else {
    try {
      if (st->h_stage) {
        // not launched yet: the opening rows are still in the staging buffer
        const double* hy = static_cast<const double*>(st->h_stage);
        for (int64_t i = 0; i < st->n; ++i) std::memcpy(y + i * ld, hy + i * st->ld, st->ld * sizeof(double));
      } else {
        HPO_HIP(hipSetDevice(s->device));
        std::lock_guard<std::mutex> g(s->io_mu);
        HPO_HIP(hipMemcpy2DAsync(y, ld * sizeof(double), st->d_y, st->ld * sizeof(double), st->ld * sizeof(double),
                                 st->n, hipMemcpyDeviceToHost, s->io));
        HPO_HIP(hipStreamSynchronize(s->io));
      }
    } catch (...) {
      rc = map_exceptions();
    }
  }
  std::lock_guard<std::mutex> g(s->mu);
  st->in_flight = false;
  return rc;
}

int hpo_server_hbm_used(const hpo_server* s, int64_t* used, int64_t* budget) {
  if (!s) return fail(HPO_E_INVALID, "null server");
  auto* m = const_cast<hpo_server*>(s);
  std::lock_guard<std::mutex> g(m->mu);
  if (used) *used = m->used;
  if (budget) *budget = m->budget;
  return HPO_OK;
}

}  // extern "C"
