// serving.cpp -- continuous batching of winding-number queries on resident meshes.
//
// One host thread per slot, each with its own HIP stream, pinned staging
// buffers and grow-only device buffers.  Requests queue FIFO; an idle slot
// takes the head request when the request's device footprint fits in the
// budget left after the resident meshes and the buffers other slots hold (a
// slot that cannot fit the head releases its own idle buffers so another can).
// Requests on different slots overlap on the device; each is bit-identical to
// a standalone gwn_soup_query because it runs the same kernels on the same
// inputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "gwn_internal.h"

namespace gwn {
namespace {

struct Request {
  int64_t id = 0;
  int32_t mesh = 0;
  const double* Q = nullptr;
  int64_t nq = 0;
  float beta = 2.0f;
  int prec = GWN_PREC_FP32;
  double* W = nullptr;
  double t_submit = 0, t_start = 0, t_done = 0;
  int status = GWN_OK;
  bool done = false;
};

size_t footprint(int64_t nq) { return (size_t)nq * (12 + 8) + soup_query_work_bytes(nq) + 3 * 256; }

// Host-side copies of large batches (narrowing into pinned memory, results
// back to the caller) split over a few threads.
template <typename Fn>
void host_parallel(int64_t n, Fn fn) {
  constexpr int64_t kMin = 1 << 17;
  constexpr int kThreads = 4;
  if (n < kMin) {
    fn(0, n);
    return;
  }
  std::thread th[kThreads - 1];
  const int64_t chunk = (n + kThreads - 1) / kThreads;
  for (int t = 1; t < kThreads; ++t) {
    const int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b < e) th[t - 1] = std::thread(fn, b, e);
  }
  fn(0, std::min(n, chunk));
  for (auto& x : th)
    if (x.joinable()) x.join();
}

struct Slot {
  hipStream_t stream = nullptr;
  DevBuf<uint8_t> dev;  // dQ | dW | workspace
  size_t dev_cap = 0;
  float* hQ = nullptr;
  double* hW = nullptr;
  size_t host_cap = 0;  // queries
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
  ~gwn_server();
};

void gwn_server::worker(gwn::Slot* sl) {
  using namespace gwn;
  (void)hipSetDevice(0);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    Request* r = nullptr;
    while (!stop) {
      if (!pending.empty()) {
        Request* head = pending.front();
        const size_t need = footprint(head->nq);
        if ((int64_t)need > budget - resident) {  // can never fit: fail it
          pending.pop_front();
          head->status = GWN_ERR_NOMEM;
          head->t_start = head->t_done = now();
          head->done = true;
          --outstanding;
          cv_done.notify_all();
          continue;
        }
        if (sl->dev_cap >= need || resident + held - (int64_t)sl->dev_cap + (int64_t)need <= budget) {
          pending.pop_front();
          r = head;
          break;
        }
        if (sl->dev_cap > 0) {  // make room for another slot
          held -= (int64_t)sl->dev_cap;
          sl->dev.release();
          sl->dev_cap = 0;
          cv_work.notify_all();
        }
      }
      cv_work.wait(lk);
    }
    if (!r) return;
    const size_t need = footprint(r->nq);
    r->t_start = now();
    int status = GWN_OK;
    std::string err;
    try {
      if (sl->dev_cap < need) {
        held -= (int64_t)sl->dev_cap;
        sl->dev.release();
        sl->dev_cap = 0;
        sl->dev.alloc(need);
        sl->dev_cap = need;
        held += (int64_t)need;
      }
      gwn_soup* m = meshes[r->mesh];
      lk.unlock();
      if ((size_t)r->nq > sl->host_cap) {
        if (sl->hQ) (void)hipHostFree(sl->hQ);
        if (sl->hW) (void)hipHostFree(sl->hW);
        sl->hQ = nullptr;
        sl->hW = nullptr;
        sl->host_cap = 0;
        GWN_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl->hQ), sizeof(float) * 3 * (size_t)r->nq));
        GWN_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl->hW), sizeof(double) * (size_t)r->nq));
        sl->host_cap = (size_t)r->nq;
      }
      // fast_winding_number.cpp:431-434: queries narrowed to float
      float* hQ = sl->hQ;
      const double* Qs = r->Q;
      host_parallel(3 * r->nq, [hQ, Qs](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) hQ[i] = (float)Qs[i];
      });
      uint8_t* base = sl->dev.p;
      auto a256 = [](size_t b) { return (b + 255) / 256 * 256; };
      float* dQ = reinterpret_cast<float*>(base);
      double* dW = reinterpret_cast<double*>(base + a256(12 * (size_t)r->nq));
      uint8_t* ws = base + a256(12 * (size_t)r->nq) + a256(8 * (size_t)r->nq);
      const size_t wsb = need - (size_t)(ws - base);
      GWN_HIP(hipMemcpyAsync(dQ, sl->hQ, 12 * (size_t)r->nq, hipMemcpyHostToDevice, sl->stream));
      soup_query_device_ws(m, dQ, r->nq, r->beta, r->prec, dW, sl->stream, ws, wsb);
      GWN_HIP(hipMemcpyAsync(sl->hW, dW, 8 * (size_t)r->nq, hipMemcpyDeviceToHost, sl->stream));
      GWN_HIP(hipStreamSynchronize(sl->stream));
      double* Wd = r->W;
      const double* hW = sl->hW;
      host_parallel(r->nq, [Wd, hW](int64_t b, int64_t e) { std::memcpy(Wd + b, hW + b, 8 * (size_t)(e - b)); });
    } catch (const Error& e) {
      status = e.code;
      err = e.msg;
    } catch (const std::exception& e) {
      status = GWN_ERR_INVALID;
      err = e.what();
    }
    if (!lk.owns_lock()) lk.lock();
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
    if (s->hQ) (void)hipHostFree(s->hQ);
    if (s->hW) (void)hipHostFree(s->hW);
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
      if (hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        delete s;
        set_error("hipStreamCreate failed");
        return GWN_ERR_HIP;
      }
      s->slots.push_back(std::move(sl));
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
  int rc = gwn_soup_build(V, nv, F, nf, order, tree_kind, &m);
  if (rc != GWN_OK) return rc;
  try {
    soup_ensure_fp64(m, nullptr);
  } catch (const Error& e) {
    gwn_soup_free(m);
    set_error(e.msg);
    return e.code;
  }
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
    rp->t_start = rp->t_done = rp->t_submit;
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

}  // extern "C"
