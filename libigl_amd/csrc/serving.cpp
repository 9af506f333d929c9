// serving.cpp -- continuous batching of winding-number queries on resident meshes.
//
// One host thread per slot, each with its own HIP stream, two pinned chunk
// buffers and a device buffer that grows in size classes.  Requests queue
// FIFO; an idle slot takes the head request when the request's device
// footprint fits in the budget left after the resident meshes and the buffers
// other slots hold (a slot that cannot fit the head releases its own idle
// buffer so another can).  Device buffers are (re)allocated outside the
// server lock, so a growing slot never stalls submissions or other slots.
// A request is processed in chunks, pipelined so that host-side narrowing and
// copy-out overlap device work.  Requests on different slots overlap on the
// device; each is bit-identical to a standalone gwn_soup_query because every
// query's result depends only on the query, the tree and the precision mode.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "gwn_internal.h"

namespace gwn {
namespace {

// A request runs in chunks of at most kChunk queries.  Two chunk buffers per
// slot (pinned host and device) let the host narrow chunk c+1 and copy out
// chunk c-1 while the device works on chunk c; per-query results do not
// depend on which other queries share a launch, so chunking leaves every
// result bit-identical to one gwn_soup_query over the whole batch.  Chunks
// are large: a launch over more queries packs them closer along the Morton
// curve (more coherent wavefronts) and pays the per-chunk sort and count once
// (measured on 64 meshes of 1k-500k triangles at 2000 requests/s: p99 4.4 ms
// against 6.8 ms with chunks of 2^18; at scale 5, 2.2e8 against 1.7e8
// queries/s sustained).
constexpr int64_t kChunk = 1 << 20;

size_t a256(size_t b) { return (b + 255) / 256 * 256; }

// Device bytes of one chunk buffer for n queries: queries, results, workspace.
size_t region_bytes(int64_t n) { return a256(12 * (size_t)n) + a256(8 * (size_t)n) + soup_query_work_bytes(n); }

// Device bytes a slot needs for a request of nq queries: one buffer up to a
// chunk, two (double-buffered) above.
size_t need_bytes(int64_t nq) { return nq <= kChunk ? region_bytes(nq) : 2 * region_bytes(kChunk); }

// Slots grow their device buffer in power-of-two classes (at least 1 MiB, at
// most the two-chunk maximum) so a slot reallocates a handful of times at most.
size_t class_bytes(int64_t nq) {
  const size_t need = need_bytes(nq), top = 2 * region_bytes(kChunk);
  size_t c = (size_t)1 << 20;
  while (c < need) c <<= 1;
  return std::max(need, std::min(c, top));
}

struct Request {
  int64_t id = 0;
  int32_t mesh = 0;
  // resolved under the server lock at submission: the worker never reads the
  // mesh list, which gwn_server_add_mesh may grow concurrently
  const gwn_soup* soup = nullptr;
  const double* Q = nullptr;
  int64_t nq = 0;
  float beta = 2.0f;
  int prec = GWN_PREC_FP32;
  double* W = nullptr;
  double t_submit = 0, t_start = 0, t_done = 0, t_uploaded = 0, t_device = 0;
  int status = GWN_OK;
  bool done = false;
};

// Persistent host threads for the narrowing and copy-out loops of large
// chunks, shared by all slots of all servers.  A loop is cut into parts that
// the caller and idle pool threads claim one at a time, so the caller never
// waits for a part nobody has started: when every pool thread is busy with
// other slots' loops it runs its own parts itself.  (With parts handed to the
// pool first-come first-served, a big request's narrowing waited behind other
// requests' parts: 4.6 ms at p99 against 0.26 ms at p50 on the 64-mesh trace
// at accuracy_scale 4.9.)
class HostPool {
 public:
  static HostPool& get() {
    static HostPool pool;
    return pool;
  }
  template <typename Fn>
  void parallel(int64_t n, Fn fn) {
    constexpr int64_t kMin = 1 << 16;
    constexpr int kParts = 8;
    if (n < kMin) {
      fn(0, n);
      return;
    }
    auto job = std::make_shared<Job>();
    job->n = n;
    job->chunk = (n + kParts - 1) / kParts;
    job->parts = (int)((n + job->chunk - 1) / job->chunk);
    job->fn = fn;
    {
      std::lock_guard<std::mutex> l(mu_);
      for (int t = 1; t < job->parts; ++t) tasks_.push_back(job);
    }
    cv_.notify_all();
    job->work();
    std::unique_lock<std::mutex> g(job->m);
    job->cv.wait(g, [&] { return job->done == job->parts; });
  }

 private:
  struct Job {
    std::atomic<int> next{0};
    int parts = 0, done = 0;  // done: under m
    int64_t n = 0, chunk = 0;
    std::function<void(int64_t, int64_t)> fn;
    std::mutex m;
    std::condition_variable cv;
    // claims and runs parts until none is left
    void work() {
      for (int p; (p = next.fetch_add(1)) < parts;) {
        const int64_t b = p * chunk, e = std::min(n, b + chunk);
        fn(b, e);
        std::lock_guard<std::mutex> g(m);
        if (++done == parts) cv.notify_all();
      }
    }
  };
  HostPool() {
    for (int i = 0; i < 8; ++i) th_.emplace_back([this] { run(); });
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> l(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  void run() {
    std::unique_lock<std::mutex> l(mu_);
    for (;;) {
      cv_.wait(l, [&] { return stop_ || !tasks_.empty(); });
      if (tasks_.empty()) return;
      auto job = std::move(tasks_.front());
      tasks_.pop_front();
      l.unlock();
      job->work();  // nothing left to claim: returns at once
      l.lock();
    }
  }
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::shared_ptr<Job>> tasks_;
  std::vector<std::thread> th_;
  bool stop_ = false;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  DevBuf<uint8_t> dev;  // one or two chunk regions: dQ | dW | workspace
  size_t dev_cap = 0;   // bytes reserved against the budget (== dev.n once allocated)
  float* hQ[2] = {nullptr, nullptr};
  double* hW[2] = {nullptr, nullptr};
  std::thread th;
};

}  // namespace
}  // namespace gwn

struct gwn_server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<gwn::Request*> pending;
  std::vector<std::unique_ptr<gwn::Request>> reqs;
  std::vector<gwn_soup*> meshes;
  int64_t resident = 0, held = 0, budget = 0;
  std::vector<std::unique_ptr<gwn::Slot>> slots;
  bool stop = false;
  int64_t outstanding = 0;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  void worker(gwn::Slot* sl);
  void run(gwn::Slot* sl, gwn::Request* r);
  ~gwn_server();
};

// Runs one admitted request on a slot whose device buffer holds
// need_bytes(r->nq); called without the server lock.
void gwn_server::run(gwn::Slot* sl, gwn::Request* r) {
  using namespace gwn;
  const gwn_soup* m = r->soup;
  const int64_t nchunk = (r->nq + kChunk - 1) / kChunk;
  const int64_t cn = std::min<int64_t>(r->nq, kChunk);
  const size_t reg = region_bytes(cn);
  auto copy_out = [&](int64_t c) {
    const int b = (int)(c & 1);
    GWN_HIP(hipEventSynchronize(sl->ev[b]));
    if (c == nchunk - 1) r->t_device = now();
    const int64_t off = c * kChunk, n = std::min(kChunk, r->nq - off);
    double* Wd = r->W + off;
    const double* hW = sl->hW[b];
    HostPool::get().parallel(n, [Wd, hW](int64_t lo, int64_t hi) {
      std::memcpy(Wd + lo, hW + lo, 8 * (size_t)(hi - lo));
    });
  };
  for (int64_t c = 0; c < nchunk; ++c) {
    const int b = (int)(c & 1);
    if (c >= 2) copy_out(c - 2);  // frees host buffer b
    const int64_t off = c * kChunk, n = std::min(kChunk, r->nq - off);
    // fast_winding_number.cpp:431-434: queries narrowed to float
    float* hQ = sl->hQ[b];
    const double* Qs = r->Q + 3 * off;
    HostPool::get().parallel(3 * n, [hQ, Qs](int64_t lo, int64_t hi) {
      for (int64_t i = lo; i < hi; ++i) hQ[i] = (float)Qs[i];
    });
    uint8_t* base = sl->dev.p + (size_t)b * reg;
    float* dQ = reinterpret_cast<float*>(base);
    double* dW = reinterpret_cast<double*>(base + a256(12 * (size_t)cn));
    uint8_t* ws = base + a256(12 * (size_t)cn) + a256(8 * (size_t)cn);
    GWN_HIP(hipMemcpyAsync(dQ, hQ, 12 * (size_t)n, hipMemcpyHostToDevice, sl->stream));
    if (c == nchunk - 1) r->t_uploaded = now();
    // The walk is chosen on the whole request as a standalone query of the
    // same batch chooses it, so the result is the standalone one.  Smaller
    // requests take the walk with one wavefront per query: sequential walks
    // cost a small request its longest path whatever its size, and with
    // per-lane walks requests of a few thousand queries held a slot for
    // 1.5-2 ms at scale 4.9 while the big requests queued behind them.
    // Larger requests take the wave-coherent walk, here in its hand-over form
    // on trees of 2^17+ slots: a request is done when its slowest wavefront
    // is, and on a 434k-triangle UV sphere (fans of thin triangles at the
    // poles) that cut a 1M-query batch at scale 4.9 from 8.0 to 5.2 ms, with
    // bitwise the same result.
    const int walk = soup_walk_for(m, r->nq, true);
    soup_query_device_ws(m, dQ, n, r->beta, r->prec, dW, sl->stream, ws, soup_query_work_bytes(cn), walk);
    GWN_HIP(hipMemcpyAsync(sl->hW[b], dW, 8 * (size_t)n, hipMemcpyDeviceToHost, sl->stream));
    GWN_HIP(hipEventRecord(sl->ev[b], sl->stream));
  }
  for (int64_t c = std::max<int64_t>(0, nchunk - 2); c < nchunk; ++c) copy_out(c);
}

void gwn_server::worker(gwn::Slot* sl) {
  using namespace gwn;
  (void)hipSetDevice(0);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    Request* r = nullptr;
    size_t grow_to = 0;  // nonzero: reallocate the device buffer (outside the lock)
    while (!stop) {
      if (!pending.empty()) {
        Request* head = pending.front();
        const size_t need = need_bytes(head->nq);
        if ((int64_t)need > budget - resident) {  // can never fit: fail it
          pending.pop_front();
          head->status = GWN_ERR_NOMEM;
          head->t_start = head->t_done = head->t_uploaded = head->t_device = now();
          head->done = true;
          --outstanding;
          cv_done.notify_all();
          continue;
        }
        if (sl->dev_cap >= need) {
          pending.pop_front();
          r = head;
          break;
        }
        const size_t cls = class_bytes(head->nq);
        const size_t want = resident + held - (int64_t)sl->dev_cap + (int64_t)cls <= budget ? cls : need;
        if (resident + held - (int64_t)sl->dev_cap + (int64_t)want <= budget) {
          pending.pop_front();
          r = head;
          held += (int64_t)want - (int64_t)sl->dev_cap;
          sl->dev_cap = want;
          grow_to = want;
          break;
        }
        if (sl->dev_cap > 0) {  // make room for another slot (rare: memory pressure)
          held -= (int64_t)sl->dev_cap;
          sl->dev.release();
          sl->dev_cap = 0;
          cv_work.notify_all();
        }
      }
      cv_work.wait(lk);
    }
    if (!r) return;
    r->t_start = now();
    lk.unlock();
    int status = GWN_OK;
    std::string err;
    try {
      if (grow_to) {
        sl->dev.release();
        sl->dev.alloc(grow_to);
      }
      run(sl, r);
    } catch (const Error& e) {
      status = e.code;
      err = e.msg;
    } catch (const std::exception& e) {
      status = GWN_ERR_INVALID;
      err = e.what();
    }
    if (status != GWN_OK) (void)hipStreamSynchronize(sl->stream);  // no copy may still read the buffers
    lk.lock();
    if (grow_to && sl->dev.p == nullptr) {  // allocation failed: return the reservation
      held -= (int64_t)sl->dev_cap;
      sl->dev_cap = 0;
    }
    if (status != GWN_OK) set_error(err);
    r->status = status;
    r->t_done = now();
    r->done = true;
    --outstanding;
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

gwn_server::~gwn_server() {
  {
    std::lock_guard<std::mutex> l(mu);
    stop = true;
  }
  cv_work.notify_all();
  for (auto& s : slots) {
    if (s->th.joinable()) s->th.join();
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (int b = 0; b < 2; ++b) {
      if (s->hQ[b]) (void)hipHostFree(s->hQ[b]);
      if (s->hW[b]) (void)hipHostFree(s->hW[b]);
      if (s->ev[b]) (void)hipEventDestroy(s->ev[b]);
    }
    s->dev.release();
    if (s->stream) (void)hipStreamDestroy(s->stream);
  }
  for (auto* m : meshes) gwn_soup_free(m);
}

using namespace gwn;

extern "C" {

int gwn_server_create(int nslots, int64_t budget, gwn_server** out) {
  try {
    if (!out || nslots < 1 || nslots > 64 || budget <= 0) {
      set_error("bad server arguments");
      return GWN_ERR_INVALID;
    }
    *out = nullptr;
    ensure_device();
    auto* s = new gwn_server();
    s->budget = budget;
    for (int i = 0; i < nslots; ++i) {
      auto sl = std::make_unique<Slot>();
      bool ok = hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) == hipSuccess;
      for (int b = 0; b < 2 && ok; ++b) {
        ok = hipEventCreateWithFlags(&sl->ev[b], hipEventDisableTiming) == hipSuccess &&
             hipHostMalloc(reinterpret_cast<void**>(&sl->hQ[b]), 12 * (size_t)kChunk) == hipSuccess &&
             hipHostMalloc(reinterpret_cast<void**>(&sl->hW[b]), 8 * (size_t)kChunk) == hipSuccess;
      }
      s->slots.push_back(std::move(sl));
      if (!ok) {
        (void)hipGetLastError();
        delete s;
        set_error("stream, event or pinned buffer creation failed");
        return GWN_ERR_HIP;
      }
    }
    for (auto& sl : s->slots) sl->th = std::thread(&gwn_server::worker, s, sl.get());
    *out = s;
    return GWN_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  }
}

void gwn_server_destroy(gwn_server* s) { delete s; }

int gwn_server_add_mesh(gwn_server* s, const double* V, int64_t nv, const int32_t* F, int64_t nf, int order,
                        int tree_kind, int32_t* mesh_id) {
  if (!s || !mesh_id) {
    set_error("null argument");
    return GWN_ERR_INVALID;
  }
  gwn_soup* m = nullptr;
  int rc = gwn_soup_build(V, nv, F, nf, order, tree_kind, &m);  // FP64 moments included
  if (rc != GWN_OK) return rc;
  gwn_soup_info info;
  gwn_soup_get_info(m, &info);
  std::lock_guard<std::mutex> l(s->mu);
  if (s->resident + info.device_bytes + s->held > s->budget) {
    gwn_soup_free(m);
    set_error("mesh does not fit in the device budget");
    return GWN_ERR_NOMEM;
  }
  s->resident += info.device_bytes;
  s->meshes.push_back(m);
  *mesh_id = (int32_t)(s->meshes.size() - 1);
  return GWN_OK;
}

int gwn_server_add_mesh_file(gwn_server* s, const char* path, int order, int tree_kind, int32_t* mesh_id) {
  std::vector<double> V;
  std::vector<int32_t> F;
  try {
    read_triangle_mesh(path, V, F);
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  }
  return gwn_server_add_mesh(s, V.data(), (int64_t)(V.size() / 3), F.data(), (int64_t)(F.size() / 3), order,
                             tree_kind, mesh_id);
}

int gwn_server_submit(gwn_server* s, int32_t mesh_id, const double* Q, int64_t nq, float accuracy_scale,
                      int precision, double* W, int64_t* request_id) {
  if (!s || !request_id || nq < 0 || (nq > 0 && (!Q || !W)) || precision < GWN_PREC_FP32 ||
      precision > GWN_PREC_FP64) {
    set_error("bad submit arguments");
    return GWN_ERR_INVALID;
  }
  std::lock_guard<std::mutex> l(s->mu);
  if (mesh_id < 0 || mesh_id >= (int32_t)s->meshes.size()) {
    set_error("unknown mesh id");
    return GWN_ERR_INVALID;
  }
  auto r = std::make_unique<Request>();
  r->id = (int64_t)s->reqs.size();
  r->mesh = mesh_id;
  r->soup = s->meshes[mesh_id];
  r->Q = Q;
  r->nq = nq;
  r->beta = accuracy_scale;
  r->prec = precision;
  r->W = W;
  r->t_submit = s->now();
  *request_id = r->id;
  Request* rp = r.get();
  s->reqs.push_back(std::move(r));
  if (nq == 0) {
    rp->t_start = rp->t_done = rp->t_uploaded = rp->t_device = rp->t_submit;
    rp->done = true;
    s->cv_done.notify_all();
    return GWN_OK;
  }
  ++s->outstanding;
  s->pending.push_back(rp);
  s->cv_work.notify_all();
  return GWN_OK;
}

int gwn_server_wait(gwn_server* s, int64_t id) {
  if (!s) return GWN_ERR_INVALID;
  std::unique_lock<std::mutex> l(s->mu);
  if (id < 0 || id >= (int64_t)s->reqs.size()) {
    set_error("unknown request id");
    return GWN_ERR_INVALID;
  }
  Request* r = s->reqs[id].get();
  s->cv_done.wait(l, [&] { return r->done; });
  return r->status;
}

int gwn_server_drain(gwn_server* s) {
  if (!s) return GWN_ERR_INVALID;
  std::unique_lock<std::mutex> l(s->mu);
  s->cv_done.wait(l, [&] { return s->outstanding == 0; });
  return GWN_OK;
}

int gwn_server_request_info(const gwn_server* cs, int64_t id, gwn_request_info* info) {
  gwn_server* s = const_cast<gwn_server*>(cs);
  if (!s || !info) return GWN_ERR_INVALID;
  std::lock_guard<std::mutex> l(s->mu);
  if (id < 0 || id >= (int64_t)s->reqs.size()) {
    set_error("unknown request id");
    return GWN_ERR_INVALID;
  }
  const Request* r = s->reqs[id].get();
  info->t_submit = r->t_submit;
  info->t_start = r->t_start;
  info->t_done = r->t_done;
  info->t_uploaded = r->t_uploaded;
  info->t_device = r->t_device;
  info->num_queries = r->nq;
  info->mesh = r->mesh;
  info->status = r->done ? r->status : -1;
  return GWN_OK;
}

int gwn_server_memory(const gwn_server* cs, int64_t* resident, int64_t* in_flight, int64_t* budget) {
  gwn_server* s = const_cast<gwn_server*>(cs);
  if (!s) return GWN_ERR_INVALID;
  std::lock_guard<std::mutex> l(s->mu);
  if (resident) *resident = s->resident;
  if (in_flight) *in_flight = s->held;
  if (budget) *budget = s->budget;
  return GWN_OK;
}

double gwn_server_now(const gwn_server* s) { return s ? s->now() : 0.0; }

int64_t gwn_server_request_bytes(int64_t num_queries) {
  return num_queries <= 0 ? 0 : (int64_t)need_bytes(num_queries);
}

int gwn_server_reserve(gwn_server* s, int64_t num_queries) {
  if (!s || num_queries <= 0) {
    set_error("bad reserve arguments");
    return GWN_ERR_INVALID;
  }
  std::lock_guard<std::mutex> l(s->mu);
  if (s->outstanding != 0) {
    set_error("gwn_server_reserve needs an idle server");
    return GWN_ERR_INVALID;
  }
  // No request is in flight and the lock is held, so no worker touches its
  // slot's buffer while it is replaced here.
  const size_t want = class_bytes(num_queries);
  try {
    for (auto& sl : s->slots) {
      if (sl->dev_cap >= want) continue;
      if (s->resident + s->held - (int64_t)sl->dev_cap + (int64_t)want > s->budget) {
        set_error("slot buffers do not fit in the device budget");
        return GWN_ERR_NOMEM;
      }
      s->held -= (int64_t)sl->dev_cap;
      sl->dev_cap = 0;
      sl->dev.alloc(want);
      sl->dev_cap = want;
      s->held += (int64_t)want;
    }
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  }
  return GWN_OK;
}

}  // extern "C"
