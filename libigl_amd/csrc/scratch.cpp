// scratch.cpp -- per-call device scratch, reused in stream order.
//
// Every query entry point needs temporary device memory (sort keys, the query
// permutation, sort temp, staged inputs and outputs).  Allocating it per call
// costs a hipMalloc/hipFree pair (hipFree waits for the device); keeping one
// buffer per handle makes concurrent calls on one handle overwrite each
// other.  Instead, idle buffers sit in a process-wide pool.  A call borrows
// the smallest idle buffer that is large enough and, if that buffer was last
// used on some stream, makes its own stream wait for an event recorded there
// after that use; giving the buffer back records such an event on the
// borrower's stream.  So calls on different threads or streams never share
// live scratch, nothing waits on the host, and after warm-up no call
// allocates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "gwn_internal.h"

namespace gwn {

struct ScratchEntry {
  void* p = nullptr;
  size_t bytes = 0;
  int device = 0;
  hipEvent_t ev = nullptr;
  bool recorded = false;
};

namespace {

// At most this many idle buffers are kept; beyond it the smallest is freed.
constexpr size_t kMaxIdle = 32;

struct Pool {
  std::mutex mu;
  std::vector<ScratchEntry*> idle;
};

// never destroyed: buffers stay valid until the process ends, so no HIP call
// runs from a static destructor after the runtime has shut down
Pool& pool() {
  static Pool* p = new Pool();
  return *p;
}

// buffers come in power-of-two classes from 64 KiB, so a handful of sizes
// serve every batch size
size_t size_class(size_t b) {
  size_t c = (size_t)1 << 16;
  while (c < b) c <<= 1;
  return c;
}

void destroy(ScratchEntry* e) {
  if (e->recorded) (void)hipEventSynchronize(e->ev);
  if (e->p) (void)hipFree(e->p);
  if (e->ev) (void)hipEventDestroy(e->ev);
  (void)hipGetLastError();
  delete e;
}

// frees every idle buffer of the device (used when an allocation fails)
void trim(int device) {
  std::vector<ScratchEntry*> drop;
  {
    std::lock_guard<std::mutex> l(pool().mu);
    auto& v = pool().idle;
    for (size_t i = 0; i < v.size();) {
      if (v[i]->device == device) {
        drop.push_back(v[i]);
        v[i] = v.back();
        v.pop_back();
      } else {
        ++i;
      }
    }
  }
  for (auto* e : drop) destroy(e);
}

ScratchEntry* fresh(size_t bytes, int device) {
  auto* e = new ScratchEntry();
  e->bytes = size_class(bytes);
  e->device = device;
  hipError_t r = hipMalloc(&e->p, e->bytes);
  if (r != hipSuccess) {
    (void)hipGetLastError();
    e->p = nullptr;
    trim(device);
    r = hipMalloc(&e->p, e->bytes);
  }
  if (r != hipSuccess) {
    (void)hipGetLastError();
    const size_t want = e->bytes;
    delete e;
    throw Error{GWN_ERR_NOMEM, "hipMalloc of " + std::to_string(want) + " bytes of query scratch failed"};
  }
  if (hipEventCreateWithFlags(&e->ev, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    e->ev = nullptr;
    destroy(e);
    throw Error{GWN_ERR_HIP, "hipEventCreate for query scratch failed"};
  }
  return e;
}

}  // namespace

Scratch::Scratch(size_t bytes, hipStream_t st) : st_(st) {
  if (bytes == 0) bytes = 1;
  int dev = 0;
  GWN_HIP(hipGetDevice(&dev));
  ScratchEntry* e = nullptr;
  {
    std::lock_guard<std::mutex> l(pool().mu);
    auto& v = pool().idle;
    size_t best = v.size();
    for (size_t i = 0; i < v.size(); ++i)
      if (v[i]->device == dev && v[i]->bytes >= bytes && (best == v.size() || v[i]->bytes < v[best]->bytes)) best = i;
    if (best != v.size()) {
      e = v[best];
      v[best] = v.back();
      v.pop_back();
    }
  }
  if (e && e->recorded) {
    // the previous borrower's work on its stream comes first
    if (hipStreamWaitEvent(st, e->ev, 0) != hipSuccess) {
      (void)hipGetLastError();
      destroy(e);
      throw Error{GWN_ERR_HIP, "hipStreamWaitEvent on query scratch failed"};
    }
  }
  if (!e) e = fresh(bytes, dev);
  e_ = e;
  p = static_cast<uint8_t*>(e->p);
  capacity = e->bytes;
}

Scratch::~Scratch() {
  if (!e_) return;
  if (hipEventRecord(e_->ev, st_) == hipSuccess) {
    e_->recorded = true;
  } else {
    // cannot order a later borrower after this call's work: wait for it here
    (void)hipGetLastError();
    (void)hipStreamSynchronize(st_);
    (void)hipGetLastError();
    e_->recorded = false;
  }
  ScratchEntry* drop = nullptr;
  {
    std::lock_guard<std::mutex> l(pool().mu);
    auto& v = pool().idle;
    v.push_back(e_);
    if (v.size() > kMaxIdle) {
      auto it = std::min_element(v.begin(), v.end(),
                                 [](const ScratchEntry* a, const ScratchEntry* b) { return a->bytes < b->bytes; });
      drop = *it;
      *it = v.back();
      v.pop_back();
    }
  }
  if (drop) destroy(drop);
  e_ = nullptr;
}

uint8_t* Scratch::take(size_t bytes) {
  const size_t off = (off_ + 255) / 256 * 256;
  GWN_CHECK(off + bytes <= capacity, GWN_ERR_INVALID, "query scratch layout exceeds its buffer");
  off_ = off + bytes;
  return p + off;
}

}  // namespace gwn
