// api.cpp -- extern "C" entry points of libgwn.so (declared in include/gwn.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gwn_internal.h"
#include "wntree.h"

namespace gwn {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

void ensure_device() {
  static std::once_flag once;
  static int ok = 0;
  std::call_once(once, [] {
    int n = 0;
    if (hipGetDeviceCount(&n) == hipSuccess && n > 0) ok = 1;
    (void)hipGetLastError();
  });
  if (!ok) throw Error{GWN_ERR_NODEVICE, "no HIP device visible"};
}

// defined in the .hip translation units
void exact2d_device(const double* V, const int32_t* E, int64_t ne, const double* O, int64_t no, double* W,
                    hipStream_t st);
void exact_device(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* O, int64_t no, double* W,
                  hipStream_t st);
void wntree_upload(const WNTreeHost& T, DevBuf<uint8_t>& rec, DevBuf<double>& tri, hipStream_t st);
void soup_setup(gwn_soup* s, hipStream_t st);
void soup_finish_build(gwn_soup* s, hipStream_t st);
void soup_query_device(const gwn_soup* s, const float* Q, int64_t nq, float accuracy_scale, int precision, double* W,
                       hipStream_t st);
void cloud_build_device(gwn_cloud* c, hipStream_t st, bool moments);
void cloud_build_from_octree_device(gwn_cloud* c, const int32_t* CH, int64_t ncells, const int64_t* cell_start,
                                    const int32_t* point_list, const float* CM, const float* R, const float* EC,
                                    hipStream_t st);
void cloud_query_device(const gwn_cloud* c, const float* Q, int64_t nq, double beta, float* WN, hipStream_t st);

template <typename Fn>
static int guard(Fn&& fn) {
  try {
    fn();
    return GWN_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return GWN_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_error(e.what());
    return GWN_ERR_INVALID;
  }
}

template <typename T>
static void upload(DevBuf<T>& d, const T* h, size_t n, hipStream_t st) {
  d.alloc(n);
  if (n) GWN_HIP(hipMemcpyAsync(d.p, h, sizeof(T) * n, hipMemcpyHostToDevice, st));
}

// Host-array entry points stage their inputs and outputs in pooled scratch
// (no allocation per call once warm) on the calling thread's stream.
template <typename T>
static T* stage_in(Scratch& ws, const T* h, size_t n, hipStream_t st) {
  T* d = ws.take_n<T>(n);
  if (n) GWN_HIP(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, st));
  return d;
}
template <typename T>
static void stage_out(T* h, const T* d, size_t n, hipStream_t st) {
  if (n) GWN_HIP(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, st));
}

static void check_faces(const int32_t* F, int64_t nf, int64_t nv) {
  for (int64_t i = 0; i < 3 * nf; ++i)
    GWN_CHECK(F[i] >= 0 && F[i] < nv, GWN_ERR_INVALID, "face index out of range at F[" + std::to_string(i) + "]");
}

static void soup_build_common(gwn_soup* s, std::vector<float>& U, const double* V64, const int32_t* F, int64_t nv,
                              int64_t nf, int order, int tree_kind) {
  GWN_CHECK(order >= 0 && order <= 2, GWN_ERR_INVALID, "order must be 0, 1 or 2");
  GWN_CHECK(tree_kind == GWN_TREE_SAH4 || tree_kind == GWN_TREE_LBVH, GWN_ERR_INVALID, "unknown tree kind");
  GWN_CHECK(nv >= 0 && nf >= 0 && nf < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID, "bad sizes");
  if (nf > 0) GWN_CHECK(F != nullptr && nv > 0, GWN_ERR_INVALID, "null faces/vertices");
  check_faces(F, nf, nv);
  ensure_device();
  hipStream_t st = nullptr;
  s->nv = nv;
  s->nf = nf;
  s->order = order;
  s->tree_kind = tree_kind;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) {
      const float v = U[3 * i + a];
      if (std::isfinite(v)) {
        lo[a] = std::min(lo[a], v);
        hi[a] = std::max(hi[a], v);
      }
    }
  for (int a = 0; a < 3; ++a) {
    s->bbox[a] = std::isfinite(lo[a]) ? lo[a] : 0.0f;
    s->bbox[3 + a] = std::isfinite(hi[a]) ? hi[a] : 0.0f;
  }
  upload(s->U, U.data(), U.size(), st);
  if (V64) {
    upload(s->V64, V64, 3 * (size_t)nv, st);
  } else {
    std::vector<double> w(U.begin(), U.end());
    upload(s->V64, w.data(), w.size(), st);
  }
  upload(s->F, F, 3 * (size_t)nf, st);
  soup_setup(s, st);
  if (nf == 0) {
    GWN_HIP(hipStreamSynchronize(st));
    return;
  }
  if (tree_kind == GWN_TREE_SAH4) {
    std::vector<float> boxes(6 * (size_t)nf);
    GWN_HIP(hipMemcpyAsync(boxes.data(), s->boxes.p, sizeof(float) * boxes.size(), hipMemcpyDeviceToHost, st));
    GWN_HIP(hipStreamSynchronize(st));
    std::vector<uint32_t> nodes;
    sah4_build(boxes.data(), (uint32_t)nf, nodes);
    s->nnodes = (int64_t)(nodes.size() / 4);
    if (s->nnodes) upload(s->nodes, nodes.data(), nodes.size(), st);
  } else {
    s->nnodes = lbvh_build(s->boxes.p, nf, s->bbox, s->nodes, st);
  }
  soup_finish_build(s, st);
}

// An order-0 soup over FP64 vertices (for the antipodal scene's rays).
void soup_build_into(gwn_soup* s, const double* V, int64_t nv, const int32_t* F, int64_t nf, int tree_kind) {
  GWN_CHECK(nv == 0 || V != nullptr, GWN_ERR_INVALID, "null vertices");
  std::vector<float> U(3 * (size_t)nv);
  for (size_t i = 0; i < U.size(); ++i) U[i] = (float)V[i];
  soup_build_common(s, U, V, F, nv, nf, 0, tree_kind);
}

void antipodal_build(gwn_antipodal* a, const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* x0,
                     hipStream_t st);
void antipodal_query_device(const gwn_antipodal* a, const double* O, int64_t no, double* W, hipStream_t st);

// igl::WindingNumberAABB<double,int> hier(V,F); hier.grow()
// (winding_number.cpp:43-47), built on the host (wntree.cpp) and uploaded
void exact_tree_build(gwn_exact_tree* t, const double* V, int64_t nv, const int32_t* F, int64_t nf) {
  t->nv = nv;
  t->nf = nf;
  WNTreeHost T;
  wntree_build_host(V, nv, F, nf, T);
  t->nrec = (int64_t)T.rec.size();
  t->ntri = (int64_t)(T.tri.size() / 9);
  if (t->nrec) {
    for (int d = 0; d < 3; ++d) {
      const double lo = T.rec[0].lo[d], hi = T.rec[0].hi[d];
      t->bbox[d] = std::isfinite(lo) ? (float)lo : 0.0f;
      t->bbox[3 + d] = std::isfinite(hi) ? (float)hi : 0.0f;
    }
  }
  t->nodes.resize(5 * (size_t)t->nrec);
  for (int64_t i = 0; i < t->nrec; ++i) {
    t->nodes[5 * i] = T.nfaces[i];
    t->nodes[5 * i + 1] = T.ncap[i];
    t->nodes[5 * i + 2] = T.apex[i];
    t->nodes[5 * i + 3] = T.left[i];
    t->nodes[5 * i + 4] = T.right[i];
  }
  wntree_upload(T, t->rec, t->tri, nullptr);
}

}  // namespace gwn

using namespace gwn;

extern "C" {

const char* gwn_last_error(void) { return g_last_error.c_str(); }

const char* gwn_version(void) { return "gwn-mi355x 0.1 (gfx950)"; }

int gwn_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int gwn_exact_winding_number_device(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* O,
                                    int64_t no, double* W, void* stream) {
  return guard([&] {
    GWN_CHECK(nv >= 0 && nf >= 0 && no >= 0, GWN_ERR_INVALID, "negative size");
    GWN_CHECK(no == 0 || (O && W), GWN_ERR_INVALID, "null query/output");
    GWN_CHECK(nf == 0 || (V && F), GWN_ERR_INVALID, "null mesh");
    ensure_device();
    exact_device(V, nv, F, nf, O, no, W, (hipStream_t)stream);
  });
}

int gwn_exact_winding_number(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* O, int64_t no,
                             double* W) {
  return guard([&] {
    GWN_CHECK(nv >= 0 && nf >= 0 && no >= 0, GWN_ERR_INVALID, "negative size");
    GWN_CHECK(no == 0 || (O && W), GWN_ERR_INVALID, "null query/output");
    if (nf > 0) {
      GWN_CHECK(V && F, GWN_ERR_INVALID, "null mesh");
      check_faces(F, nf, nv);
    }
    ensure_device();
    // winding_number.cpp:39-53: the hierarchy is built for the call (drawing
    // rand() as the reference's does), then every origin is walked through it
    gwn_exact_tree t;
    exact_tree_build(&t, V, nv, F, nf);
    if (no == 0) return;
    hipStream_t st = nullptr;
    Scratch io(scratch_bytes({24 * (size_t)no, 8 * (size_t)no}), st);
    const double* dO = stage_in(io, O, 3 * (size_t)no, st);
    double* dW = io.take_n<double>((size_t)no);
    wntree_query_device(t.rec.p, t.nrec, t.tri.p, t.bbox, dO, no, dW, st);
    stage_out(W, dW, (size_t)no, st);
    GWN_HIP(hipStreamSynchronize(st));
  });
}

int gwn_exact_winding_number_bruteforce(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* O,
                                        int64_t no, double* W) {
  return guard([&] {
    GWN_CHECK(nv >= 0 && nf >= 0 && no >= 0, GWN_ERR_INVALID, "negative size");
    GWN_CHECK(no == 0 || (O && W), GWN_ERR_INVALID, "null query/output");
    if (no == 0) return;
    if (nf > 0) {
      GWN_CHECK(V && F, GWN_ERR_INVALID, "null mesh");
      check_faces(F, nf, nv);
    }
    ensure_device();
    hipStream_t st = nullptr;
    const size_t v = 3 * (size_t)nv, f = 3 * (size_t)nf, o = 3 * (size_t)no;
    Scratch io(scratch_bytes({8 * v, 4 * f, 8 * o, 8 * (size_t)no}), st);
    const double* dV = stage_in(io, V, v, st);
    const int32_t* dF = stage_in(io, F, f, st);
    const double* dO = stage_in(io, O, o, st);
    double* dW = io.take_n<double>((size_t)no);
    exact_device(dV, nv, dF, nf, dO, no, dW, st);
    stage_out(W, dW, (size_t)no, st);
    GWN_HIP(hipStreamSynchronize(st));
  });
}

int gwn_exact_tree_build(const double* V, int64_t nv, const int32_t* F, int64_t nf, gwn_exact_tree** out) {
  return guard([&] {
    GWN_CHECK(out != nullptr, GWN_ERR_INVALID, "null output handle");
    *out = nullptr;
    GWN_CHECK(nv >= 0 && nf >= 0, GWN_ERR_INVALID, "negative size");
    if (nf > 0) {
      GWN_CHECK(V && F, GWN_ERR_INVALID, "null mesh");
      check_faces(F, nf, nv);
    }
    ensure_device();
    gwn_exact_tree* t = new gwn_exact_tree();
    try {
      exact_tree_build(t, V, nv, F, nf);
    } catch (...) {
      delete t;
      throw;
    }
    *out = t;
  });
}

void gwn_exact_tree_free(gwn_exact_tree* t) { delete t; }

int gwn_exact_tree_host_nodes(const double* V, int64_t nv, const int32_t* F, int64_t nf, int64_t* nodes,
                              int64_t capacity, int64_t* num_nodes) {
  return guard([&] {
    GWN_CHECK(num_nodes != nullptr, GWN_ERR_INVALID, "null node count");
    GWN_CHECK(nv >= 0 && nf >= 0, GWN_ERR_INVALID, "negative size");
    if (nf > 0) {
      GWN_CHECK(V && F, GWN_ERR_INVALID, "null mesh");
      check_faces(F, nf, nv);
    }
    WNTreeHost T;
    wntree_build_host(V, nv, F, nf, T);
    *num_nodes = (int64_t)T.rec.size();
    GWN_CHECK(nodes == nullptr || capacity >= *num_nodes, GWN_ERR_INVALID, "node capacity too small");
    if (nodes)
      for (int64_t i = 0; i < *num_nodes; ++i) {
        nodes[5 * i] = T.nfaces[i];
        nodes[5 * i + 1] = T.ncap[i];
        nodes[5 * i + 2] = T.apex[i];
        nodes[5 * i + 3] = T.left[i];
        nodes[5 * i + 4] = T.right[i];
      }
  });
}

int gwn_exact_tree_info(const gwn_exact_tree* t, int64_t* num_nodes, int64_t* num_triangles) {
  return guard([&] {
    GWN_CHECK(t != nullptr, GWN_ERR_INVALID, "null handle");
    if (num_nodes) *num_nodes = t->nrec;
    if (num_triangles) *num_triangles = t->ntri;
  });
}

int gwn_exact_tree_nodes(const gwn_exact_tree* t, int64_t* out) {
  return guard([&] {
    GWN_CHECK(t && out, GWN_ERR_INVALID, "null argument");
    if (!t->nodes.empty()) std::memcpy(out, t->nodes.data(), sizeof(int64_t) * t->nodes.size());
  });
}

int gwn_exact_tree_query_device(const gwn_exact_tree* t, const double* O, int64_t no, double* W, void* stream) {
  return guard([&] {
    GWN_CHECK(t != nullptr, GWN_ERR_INVALID, "null handle");
    GWN_CHECK(no >= 0 && no < (int64_t)0xFFFFFFFF, GWN_ERR_INVALID, "bad query count");
    GWN_CHECK(no == 0 || (O && W), GWN_ERR_INVALID, "null query/output");
    wntree_query_device(t->rec.p, t->nrec, t->tri.p, t->bbox, O, no, W, (hipStream_t)stream);
  });
}

int gwn_exact_tree_query(const gwn_exact_tree* t, const double* O, int64_t no, double* W) {
  return guard([&] {
    GWN_CHECK(t != nullptr, GWN_ERR_INVALID, "null handle");
    GWN_CHECK(no >= 0 && no < (int64_t)0xFFFFFFFF, GWN_ERR_INVALID, "bad query count");
    GWN_CHECK(no == 0 || (O && W), GWN_ERR_INVALID, "null query/output");
    if (no == 0) return;
    hipStream_t st = nullptr;
    Scratch io(scratch_bytes({24 * (size_t)no, 8 * (size_t)no}), st);
    const double* dO = stage_in(io, O, 3 * (size_t)no, st);
    double* dW = io.take_n<double>((size_t)no);
    wntree_query_device(t->rec.p, t->nrec, t->tri.p, t->bbox, dO, no, dW, st);
    stage_out(W, dW, (size_t)no, st);
    GWN_HIP(hipStreamSynchronize(st));
  });
}

int gwn_soup_build(const double* V, int64_t nv, const int32_t* F, int64_t nf, int order, int tree_kind,
                   gwn_soup** out) {
  return guard([&] {
    GWN_CHECK(out != nullptr, GWN_ERR_INVALID, "null output handle");
    *out = nullptr;
    GWN_CHECK(nv == 0 || V != nullptr, GWN_ERR_INVALID, "null vertices");
    // fast_winding_number.cpp:392-399: the BVH keeps a float copy of V
    std::vector<float> U(3 * (size_t)nv);
    for (size_t i = 0; i < U.size(); ++i) U[i] = (float)V[i];
    gwn_soup* s = new gwn_soup();
    try {
      soup_build_common(s, U, V, F, nv, nf, order, tree_kind);
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

int gwn_soup_build_f32(const float* V, int64_t nv, const int32_t* F, int64_t nf, int order, int tree_kind,
                       gwn_soup** out) {
  return guard([&] {
    GWN_CHECK(out != nullptr, GWN_ERR_INVALID, "null output handle");
    *out = nullptr;
    GWN_CHECK(nv == 0 || V != nullptr, GWN_ERR_INVALID, "null vertices");
    std::vector<float> U(V, V + 3 * (size_t)nv);
    gwn_soup* s = new gwn_soup();
    try {
      soup_build_common(s, U, nullptr, F, nv, nf, order, tree_kind);
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

void gwn_soup_free(gwn_soup* s) { delete s; }

int gwn_soup_get_info(const gwn_soup* s, gwn_soup_info* info) {
  return guard([&] {
    GWN_CHECK(s && info, GWN_ERR_INVALID, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->num_vertices = s->nv;
    info->num_triangles = s->nf;
    info->num_nodes = s->nnodes;
    info->num_slots = s->nslots;
    info->order = s->order;
    info->tree_kind = s->tree_kind;
    info->depth = s->depth;
    info->device_bytes = (int64_t)(s->U.bytes() + s->F.bytes() + s->tri.bytes() + s->boxes.bytes() +
                                   s->nodes.bytes() + s->boxdata.bytes() + s->slots.bytes() + s->slots64.bytes() +
                                   s->V64.bytes() + s->dbox.bytes());
    std::memcpy(info->bbox, s->bbox, sizeof(s->bbox));
  });
}

int gwn_soup_set_mixed_params(gwn_soup* s, const gwn_mixed_params* p) {
  return guard([&] {
    GWN_CHECK(s && p, GWN_ERR_INVALID, "null argument");
    GWN_CHECK(p->near_dist >= 0 && p->ambiguous_band >= 0, GWN_ERR_INVALID, "negative parameter");
    s->mixed = *p;
  });
}

int gwn_soup_query_device(const gwn_soup* s, const float* Q, int64_t nq, float accuracy_scale, int precision,
                          double* W, void* stream) {
  return guard([&] {
    GWN_CHECK(s != nullptr, GWN_ERR_INVALID, "null handle");
    GWN_CHECK(nq >= 0 && nq < (int64_t)0xFFFFFFFF, GWN_ERR_INVALID, "bad query count");
    GWN_CHECK(nq == 0 || (Q && W), GWN_ERR_INVALID, "null query/output");
    GWN_CHECK(precision >= GWN_PREC_FP32 && precision <= GWN_PREC_FP64, GWN_ERR_INVALID, "unknown precision");
    soup_query_device(s, Q, nq, accuracy_scale, precision, W, (hipStream_t)stream);
  });
}

int gwn_soup_query_device_noalloc(const gwn_soup* s, const float* Q, int64_t nq, float accuracy_scale, int precision,
                                  double* W, void* stream) {
  return guard([&] {
    GWN_CHECK(s != nullptr, GWN_ERR_INVALID, "null handle");
    GWN_CHECK(nq >= 0, GWN_ERR_INVALID, "bad query count");
    GWN_CHECK(nq == 0 || (Q && W), GWN_ERR_INVALID, "null query/output");
    GWN_CHECK(precision >= GWN_PREC_FP32 && precision <= GWN_PREC_FP64, GWN_ERR_INVALID, "unknown precision");
    soup_query_device_noalloc(s, Q, nq, accuracy_scale, precision, W, (hipStream_t)stream);
  });
}

int64_t gwn_soup_noalloc_max_queries(void) { return soup_noalloc_max_queries(); }

int gwn_soup_query(const gwn_soup* s, const double* Q, int64_t nq, float accuracy_scale, int precision, double* W) {
  return guard([&] {
    GWN_CHECK(s != nullptr, GWN_ERR_INVALID, "null handle");
    GWN_CHECK(nq >= 0 && nq < (int64_t)0xFFFFFFFF, GWN_ERR_INVALID, "bad query count");
    GWN_CHECK(nq == 0 || (Q && W), GWN_ERR_INVALID, "null query/output");
    GWN_CHECK(precision >= GWN_PREC_FP32 && precision <= GWN_PREC_FP64, GWN_ERR_INVALID, "unknown precision");
    if (nq == 0) return;
    hipStream_t st = nullptr;
    const size_t n = (size_t)nq;
    Scratch io(scratch_bytes({24 * n, 12 * n, 8 * n}), st);
    const double* dQ64 = stage_in(io, Q, 3 * n, st);
    float* dQ = io.take_n<float>(3 * n);
    double* dW = io.take_n<double>(n);
    // fast_winding_number.cpp:431-434: each query coordinate narrowed to float
    to_float_device(dQ64, 3 * nq, dQ, st);
    soup_query_device(s, dQ, nq, accuracy_scale, precision, dW, st);
    stage_out(W, dW, n, st);
    GWN_HIP(hipStreamSynchronize(st));
  });
}

int gwn_sah4_tree(const float* V, int64_t nv, const int32_t* F, int64_t nf, uint32_t* nodes, int64_t capacity,
                  int64_t* num_nodes) {
  return guard([&] {
    GWN_CHECK(num_nodes != nullptr, GWN_ERR_INVALID, "null node count");
    GWN_CHECK(nv >= 0 && nf >= 0 && nf < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID, "bad sizes");
    if (nf > 0) {
      GWN_CHECK(V && F, GWN_ERR_INVALID, "null mesh");
      check_faces(F, nf, nv);
    }
    std::vector<float> boxes(6 * (size_t)nf);
    auto hmin = [](float a, float b) { return (a > b) ? b : a; };
    auto hmax = [](float a, float b) { return (a < b) ? b : a; };
    for (int64_t i = 0; i < nf; ++i)
      for (int a = 0; a < 3; ++a) {
        float lo = V[3 * (int64_t)F[3 * i] + a], hi = lo;
        for (int k = 1; k < 3; ++k) {
          const float v = V[3 * (int64_t)F[3 * i + k] + a];
          lo = hmin(lo, v);
          hi = hmax(hi, v);
        }
        boxes[6 * i + 2 * a] = lo;
        boxes[6 * i + 2 * a + 1] = hi;
      }
    std::vector<uint32_t> out;
    sah4_build(boxes.data(), (uint32_t)nf, out);
    *num_nodes = (int64_t)(out.size() / 4);
    GWN_CHECK(nodes == nullptr || capacity >= *num_nodes, GWN_ERR_INVALID, "node capacity too small");
    if (nodes && !out.empty()) std::memcpy(nodes, out.data(), sizeof(uint32_t) * out.size());
  });
}

int gwn_soup_copy_nodes(const gwn_soup* s, uint32_t* out) {
  return guard([&] {
    GWN_CHECK(s && out, GWN_ERR_INVALID, "null argument");
    if (s->nnodes)
      GWN_HIP(hipMemcpy(out, s->nodes.p, sizeof(uint32_t) * 4 * (size_t)s->nnodes, hipMemcpyDeviceToHost));
  });
}

int gwn_soup_copy_boxdata(const gwn_soup* s, float* out) {
  return guard([&] {
    GWN_CHECK(s && out, GWN_ERR_INVALID, "null argument");
    if (s->nnodes)
      GWN_HIP(hipMemcpy(out, s->boxdata.p, sizeof(float) * 92 * (size_t)s->nnodes, hipMemcpyDeviceToHost));
  });
}

// ---------------- 2D curve (UT_SubtendedAngle) ----------------
static void check_segments(const int32_t* E, int64_t ne, int64_t np) {
  for (int64_t i = 0; i < 2 * ne; ++i)
    GWN_CHECK(E[i] >= 0 && E[i] < np, GWN_ERR_INVALID, "segment index out of range at E[" + std::to_string(i) + "]");
}

// segment boxes of UT_SubtendedAngle::init (7365-7390), 6 floats each (z = 0)
static std::vector<float> segment_boxes(const float* P, const int32_t* E, int64_t ne) {
  std::vector<float> boxes(6 * (size_t)ne, 0.0f);
  auto hmin = [](float a, float b) { return (a > b) ? b : a; };
  auto hmax = [](float a, float b) { return (a < b) ? b : a; };
  for (int64_t i = 0; i < ne; ++i)
    for (int a = 0; a < 2; ++a) {
      const float u = P[2 * (int64_t)E[2 * i] + a], v = P[2 * (int64_t)E[2 * i + 1] + a];
      boxes[6 * i + 2 * a] = hmin(u, v);
      boxes[6 * i + 2 * a + 1] = hmax(u, v);
    }
  return boxes;
}

int gwn_curve_build(const float* P, int64_t np, const int32_t* E, int64_t ne, int order, int tree_kind,
                    gwn_curve** out) {
  return guard([&] {
    GWN_CHECK(out != nullptr, GWN_ERR_INVALID, "null output handle");
    *out = nullptr;
    GWN_CHECK(order >= 0 && order <= 2, GWN_ERR_INVALID, "order must be 0, 1 or 2");
    GWN_CHECK(tree_kind == GWN_TREE_SAH4 || tree_kind == GWN_TREE_LBVH, GWN_ERR_INVALID, "unknown tree kind");
    GWN_CHECK(np >= 0 && ne >= 0 && ne < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID, "bad sizes");
    GWN_CHECK(np == 0 || P != nullptr, GWN_ERR_INVALID, "null points");
    if (ne > 0) {
      GWN_CHECK(E != nullptr && np > 0, GWN_ERR_INVALID, "null segments/points");
      check_segments(E, ne, np);
    }
    ensure_device();
    hipStream_t st = nullptr;
    gwn_curve* c = new gwn_curve();
    try {
      c->np = np;
      c->ne = ne;
      c->order = order;
      c->tree_kind = tree_kind;
      float lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
      for (int64_t i = 0; i < np; ++i)
        for (int a = 0; a < 2; ++a) {
          const float v = P[2 * i + a];
          if (std::isfinite(v)) {
            lo[a] = std::min(lo[a], v);
            hi[a] = std::max(hi[a], v);
          }
        }
      for (int a = 0; a < 2; ++a) {
        c->bbox[a] = std::isfinite(lo[a]) ? lo[a] : 0.0f;
        c->bbox[2 + a] = std::isfinite(hi[a]) ? hi[a] : 0.0f;
      }
      upload(c->P, P, 2 * (size_t)np, st);
      upload(c->E, E, 2 * (size_t)ne, st);
      curve_build_device(c, st);
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

void gwn_curve_free(gwn_curve* c) { delete c; }

int gwn_curve_get_info(const gwn_curve* c, gwn_curve_info* info) {
  return guard([&] {
    GWN_CHECK(c && info, GWN_ERR_INVALID, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->num_points = c->np;
    info->num_segments = c->ne;
    info->num_nodes = c->nnodes;
    info->num_slots = c->nslots;
    info->order = c->order;
    info->tree_kind = c->tree_kind;
    info->depth = c->depth;
    info->device_bytes = (int64_t)(c->P.bytes() + c->E.bytes() + c->seg.bytes() + c->boxes.bytes() +
                                   c->nodes.bytes() + c->boxdata.bytes() + c->slots.bytes());
    std::memcpy(info->bbox, c->bbox, sizeof(c->bbox));
  });
}

int gwn_curve_query_device(const gwn_curve* c, const float* Q, int64_t nq, float accuracy_scale, float* angle,
                           void* stream) {
  return guard([&] {
    GWN_CHECK(c != nullptr, GWN_ERR_INVALID, "null curve");
    GWN_CHECK(nq >= 0 && nq < (int64_t)0xFFFFFFFF, GWN_ERR_INVALID, "bad query count");
    GWN_CHECK(nq == 0 || (Q && angle), GWN_ERR_INVALID, "null query/output");
    curve_query_device(c, Q, nq, accuracy_scale, angle, (hipStream_t)stream);
  });
}

int gwn_curve_query(const gwn_curve* c, const float* Q, int64_t nq, float accuracy_scale, float* angle) {
  return guard([&] {
    GWN_CHECK(c != nullptr, GWN_ERR_INVALID, "null curve");
    GWN_CHECK(nq >= 0 && nq < (int64_t)0xFFFFFFFF, GWN_ERR_INVALID, "bad query count");
    GWN_CHECK(nq == 0 || (Q && angle), GWN_ERR_INVALID, "null query/output");
    if (nq == 0) return;
    hipStream_t st = nullptr;
    Scratch io(scratch_bytes({8 * (size_t)nq, 4 * (size_t)nq}), st);
    const float* dQ = stage_in(io, Q, 2 * (size_t)nq, st);
    float* dA = io.take_n<float>((size_t)nq);
    curve_query_device(c, dQ, nq, accuracy_scale, dA, st);
    stage_out(angle, dA, (size_t)nq, st);
    GWN_HIP(hipStreamSynchronize(st));
  });
}

int gwn_curve_copy_nodes(const gwn_curve* c, uint32_t* out) {
  return guard([&] {
    GWN_CHECK(c && out, GWN_ERR_INVALID, "null argument");
    if (c->nnodes)
      GWN_HIP(hipMemcpy(out, c->nodes.p, sizeof(uint32_t) * 4 * (size_t)c->nnodes, hipMemcpyDeviceToHost));
  });
}

int gwn_curve_copy_boxdata(const gwn_curve* c, float* out) {
  return guard([&] {
    GWN_CHECK(c && out, GWN_ERR_INVALID, "null argument");
    if (c->nnodes)
      GWN_HIP(hipMemcpy(out, c->boxdata.p, sizeof(float) * 48 * (size_t)c->nnodes, hipMemcpyDeviceToHost));
  });
}

int gwn_sah4_tree_2d(const float* P, int64_t np, const int32_t* E, int64_t ne, uint32_t* nodes, int64_t capacity,
                     int64_t* num_nodes) {
  return guard([&] {
    GWN_CHECK(num_nodes != nullptr, GWN_ERR_INVALID, "null node count");
    GWN_CHECK(np >= 0 && ne >= 0 && ne < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID, "bad sizes");
    if (ne > 0) {
      GWN_CHECK(P && E, GWN_ERR_INVALID, "null curve");
      check_segments(E, ne, np);
    }
    const std::vector<float> boxes = segment_boxes(P, E, ne);
    std::vector<uint32_t> out;
    sah4_build(boxes.data(), (uint32_t)ne, out, 2);
    *num_nodes = (int64_t)(out.size() / 4);
    GWN_CHECK(nodes == nullptr || capacity >= *num_nodes, GWN_ERR_INVALID, "node capacity too small");
    if (nodes && !out.empty()) std::memcpy(nodes, out.data(), sizeof(uint32_t) * out.size());
  });
}

// ---------------- 2D exact winding number ----------------
int gwn_exact_winding_number_2d(const double* V, int64_t nv, const int32_t* E, int64_t ne, const double* O,
                                int64_t no, double* W) {
  return guard([&] {
    GWN_CHECK(nv >= 0 && ne >= 0 && no >= 0, GWN_ERR_INVALID, "negative size");
    GWN_CHECK(no == 0 || (O && W), GWN_ERR_INVALID, "null query/output");
    if (no == 0) return;
    if (ne > 0) {
      GWN_CHECK(V && E, GWN_ERR_INVALID, "null curve");
      for (int64_t i = 0; i < 2 * ne; ++i)
        GWN_CHECK(E[i] >= 0 && E[i] < nv, GWN_ERR_INVALID, "edge index out of range");
    }
    ensure_device();
    hipStream_t st = nullptr;
    const size_t v = 2 * (size_t)nv, e = 2 * (size_t)ne, o = 2 * (size_t)no;
    Scratch io(scratch_bytes({8 * v, 4 * e, 8 * o, 8 * (size_t)no}), st);
    const double* dV = stage_in(io, V, v, st);
    const int32_t* dE = stage_in(io, E, e, st);
    const double* dO = stage_in(io, O, o, st);
    double* dW = io.take_n<double>((size_t)no);
    exact2d_device(dV, dE, ne, dO, no, dW, st);
    stage_out(W, dW, (size_t)no, st);
    GWN_HIP(hipStreamSynchronize(st));
  });
}

// ---------------- mesh files ----------------
int gwn_read_triangle_mesh(const char* path, double** V, int64_t* nv, int32_t** F, int64_t* nf) {
  return guard([&] {
    GWN_CHECK(V && nv && F && nf, GWN_ERR_INVALID, "null output");
    *V = nullptr;
    *F = nullptr;
    *nv = *nf = 0;
    std::vector<double> v;
    std::vector<int32_t> f;
    read_triangle_mesh(path, v, f);
    double* vo = (double*)std::malloc(sizeof(double) * std::max<size_t>(1, v.size()));
    int32_t* fo = (int32_t*)std::malloc(sizeof(int32_t) * std::max<size_t>(1, f.size()));
    if (!vo || !fo) {
      std::free(vo);
      std::free(fo);
      throw Error{GWN_ERR_NOMEM, "out of host memory"};
    }
    if (!v.empty()) std::memcpy(vo, v.data(), sizeof(double) * v.size());
    if (!f.empty()) std::memcpy(fo, f.data(), sizeof(int32_t) * f.size());
    *V = vo;
    *F = fo;
    *nv = (int64_t)(v.size() / 3);
    *nf = (int64_t)(f.size() / 3);
  });
}

void gwn_free(void* p) { std::free(p); }

// ---------------- signed distance ----------------
static void sd_check(gwn_soup* s, int64_t np, int sign_type, const double* P, double* S, int32_t* I, double* C) {
  GWN_CHECK(s != nullptr, GWN_ERR_INVALID, "null handle");
  GWN_CHECK(np >= 0 && np < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID, "bad point count");
  GWN_CHECK(np == 0 || (P && S && I && C), GWN_ERR_INVALID, "null input/output");
  GWN_CHECK(sign_type == GWN_SIGN_WINDING_NUMBER || sign_type == GWN_SIGN_DEFAULT || sign_type == GWN_SIGN_UNSIGNED ||
                sign_type == GWN_SIGN_FAST_WINDING_NUMBER || sign_type == GWN_SIGN_PSEUDONORMAL,
            GWN_ERR_INVALID, "unknown sign type");
}

int gwn_signed_distance_with_normals_device(gwn_soup* s, const double* P, int64_t np, int sign_type,
                                            double lower_bound, double upper_bound, double* S, int32_t* I, double* C,
                                            double* N, void* stream) {
  return guard([&] {
    sd_check(s, np, sign_type, P, S, I, C);
    signed_distance_device(s, P, np, sign_type, lower_bound, upper_bound, S, I, C, N, (hipStream_t)stream);
  });
}

int gwn_signed_distance_device(gwn_soup* s, const double* P, int64_t np, int sign_type, double lower_bound,
                               double upper_bound, double* S, int32_t* I, double* C, void* stream) {
  return gwn_signed_distance_with_normals_device(s, P, np, sign_type, lower_bound, upper_bound, S, I, C, nullptr,
                                                 stream);
}

int gwn_signed_distance_with_normals(gwn_soup* s, const double* P, int64_t np, int sign_type, double lower_bound,
                                     double upper_bound, double* S, int32_t* I, double* C, double* N) {
  return guard([&] {
    sd_check(s, np, sign_type, P, S, I, C);
    if (np == 0) return;
    hipStream_t st = nullptr;
    const size_t n = (size_t)np;
    const bool want_n = N != nullptr && sign_type == GWN_SIGN_PSEUDONORMAL;
    Scratch io(scratch_bytes({24 * n, 8 * n, 4 * n, 24 * n, want_n ? 24 * n : 8}), st);
    const double* dP = stage_in(io, P, 3 * n, st);
    double* dS = io.take_n<double>(n);
    int32_t* dI = io.take_n<int32_t>(n);
    double* dC = io.take_n<double>(3 * n);
    double* dN = want_n ? io.take_n<double>(3 * n) : nullptr;
    signed_distance_device(s, dP, np, sign_type, lower_bound, upper_bound, dS, dI, dC, dN, st);
    stage_out(S, dS, n, st);
    stage_out(I, dI, n, st);
    stage_out(C, dC, 3 * n, st);
    if (want_n) stage_out(N, dN, 3 * n, st);
    GWN_HIP(hipStreamSynchronize(st));
  });
}

int gwn_signed_distance(gwn_soup* s, const double* P, int64_t np, int sign_type, double lower_bound,
                        double upper_bound, double* S, int32_t* I, double* C) {
  return gwn_signed_distance_with_normals(s, P, np, sign_type, lower_bound, upper_bound, S, I, C, nullptr);
}

// ---------------- antipodal method ----------------
int gwn_antipodal_build(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* x0,
                        gwn_antipodal** out) {
  return guard([&] {
    GWN_CHECK(out != nullptr, GWN_ERR_INVALID, "null output handle");
    *out = nullptr;
    GWN_CHECK(nv >= 0 && nf >= 0 && nf < (int64_t)0x7FFFFFFF && nv < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID,
              "bad sizes");
    ensure_device();
    gwn_antipodal* a = new gwn_antipodal();
    try {
      antipodal_build(a, V, nv, F, nf, x0, nullptr);
    } catch (...) {
      delete a;
      throw;
    }
    *out = a;
  });
}

void gwn_antipodal_free(gwn_antipodal* a) { delete a; }

int gwn_antipodal_info(const gwn_antipodal* a, int64_t* num_boundary_segments, int64_t* num_faces, double* x0) {
  return guard([&] {
    GWN_CHECK(a != nullptr, GWN_ERR_INVALID, "null scene");
    if (num_boundary_segments) *num_boundary_segments = a->nseg;
    if (num_faces) *num_faces = a->soup.nf;
    if (x0)
      for (int k = 0; k < 3; ++k) x0[k] = a->x0[k];
  });
}

int gwn_antipodal_query(const gwn_antipodal* a, const double* O, int64_t no, double* W) {
  return guard([&] {
    GWN_CHECK(a != nullptr && no >= 0, GWN_ERR_INVALID, "bad arguments");
    if (no == 0) return;
    GWN_CHECK(O && W, GWN_ERR_INVALID, "null input");
    hipStream_t st = nullptr;
    Scratch io(scratch_bytes({24 * (size_t)no, 8 * (size_t)no}), st);
    const double* dO = stage_in(io, O, 3 * (size_t)no, st);
    double* dW = io.take_n<double>((size_t)no);
    antipodal_query_device(a, dO, no, dW, st);
    stage_out(W, dW, (size_t)no, st);
    GWN_HIP(hipStreamSynchronize(st));
  });
}

int gwn_antipodal_query_device(const gwn_antipodal* a, const double* O, int64_t no, double* W, void* stream) {
  return guard([&] {
    GWN_CHECK(a != nullptr && no >= 0, GWN_ERR_INVALID, "bad arguments");
    if (no == 0) return;
    GWN_CHECK(O && W, GWN_ERR_INVALID, "null input");
    antipodal_query_device(a, O, no, W, static_cast<hipStream_t>(stream));
  });
}

int gwn_winding_number_antipodal(const double* V, int64_t nv, const int32_t* F, int64_t nf, const double* O,
                                 int64_t no, double* W) {
  gwn_antipodal* a = nullptr;
  int rc = gwn_antipodal_build(V, nv, F, nf, nullptr, &a);
  if (rc != GWN_OK) return rc;
  rc = gwn_antipodal_query(a, O, no, W);
  gwn_antipodal_free(a);
  return rc;
}

// ---------------- k nearest neighbours, point areas ----------------
int gwn_knn_columns(int64_t nv, int64_t k) { return (k < 1 || nv < 0) ? 0 : knn_columns(nv, k); }

int gwn_knn(const double* P, int64_t np, const double* V, int64_t nv, int64_t k, const int32_t* cand, int64_t ncand,
            int32_t* I) {
  return guard([&] {
    GWN_CHECK(np >= 0 && nv >= 0 && np < (int64_t)0x7FFFFFFF && nv < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID,
              "bad point count");
    GWN_CHECK(k >= 1 && k < (1 << 20), GWN_ERR_INVALID, "k must be in [1, 2^20)");
    GWN_CHECK(!cand || (ncand >= 0 && ncand <= nv), GWN_ERR_INVALID, "bad candidate count");
    if (np == 0 || nv == 0) return;
    GWN_CHECK(P && V && I, GWN_ERR_INVALID, "null input");
    if (cand)
      for (int64_t j = 0; j < ncand; ++j)
        GWN_CHECK(cand[j] >= 0 && cand[j] < nv, GWN_ERR_INVALID, "candidate index out of range");
    ensure_device();
    hipStream_t st = nullptr;
    const int64_t kk = knn_columns(nv, k);
    DevBuf<double> dP, dV;
    DevBuf<int32_t> dc, dI((size_t)np * kk);
    upload(dP, P, 3 * (size_t)np, st);
    upload(dV, V, 3 * (size_t)nv, st);
    if (cand) upload(dc, cand, (size_t)ncand, st);
    knn_device(dP.p, np, dV.p, nv, k, cand ? dc.p : nullptr, ncand, dI.p, st);
    GWN_HIP(hipMemcpyAsync(I, dI.p, sizeof(int32_t) * (size_t)np * kk, hipMemcpyDeviceToHost, st));
    GWN_HIP(hipStreamSynchronize(st));
  });
}

int gwn_knn_device(const double* P, int64_t np, const double* V, int64_t nv, int64_t k, const int32_t* cand,
                   int64_t ncand, int32_t* I, void* stream) {
  return guard([&] {
    GWN_CHECK(np >= 0 && nv >= 0 && np < (int64_t)0x7FFFFFFF && nv < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID,
              "bad point count");
    GWN_CHECK(k >= 1 && k < (1 << 20), GWN_ERR_INVALID, "k must be in [1, 2^20)");
    GWN_CHECK(!cand || (ncand >= 0 && ncand <= nv), GWN_ERR_INVALID, "bad candidate count");
    if (np == 0 || nv == 0) return;
    GWN_CHECK(P && V && I, GWN_ERR_INVALID, "null input");
    ensure_device();
    knn_device(P, np, V, nv, k, cand, ncand, I, static_cast<hipStream_t>(stream));
  });
}

int gwn_point_areas(const double* P, int64_t n, const int32_t* I, int64_t k, const double* N, double* A,
                    double* T) {
  return guard([&] {
    GWN_CHECK(n >= 0 && n < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID, "bad point count");
    GWN_CHECK(k >= 1 && k < (1 << 20), GWN_ERR_INVALID, "I must have 1 to 2^20 columns");
    if (n == 0) return;
    GWN_CHECK(P && I && A, GWN_ERR_INVALID, "null input");
    for (int64_t j = 0; j < n * k; ++j)
      GWN_CHECK(I[j] >= 0 && I[j] < n, GWN_ERR_INVALID, "neighbour index out of range at I[" + std::to_string(j) + "]");
    ensure_device();
    hipStream_t st = nullptr;
    DevBuf<double> dP, dN, dA(n), dT;
    DevBuf<int32_t> dI;
    upload(dP, P, 3 * (size_t)n, st);
    upload(dI, I, (size_t)n * k, st);
    if (N) upload(dN, N, 3 * (size_t)n, st);
    if (T) dT.alloc(3 * (size_t)n);
    point_areas_device(dP.p, n, dI.p, k, N ? dN.p : nullptr, dA.p, T ? dT.p : nullptr, st);
    GWN_HIP(hipMemcpyAsync(A, dA.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    if (T) GWN_HIP(hipMemcpyAsync(T, dT.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st));
    GWN_HIP(hipStreamSynchronize(st));
  });
}

int gwn_point_areas_device(const double* P, int64_t n, const int32_t* I, int64_t k, const double* N, double* A,
                           double* T, void* stream) {
  return guard([&] {
    GWN_CHECK(n >= 0 && n < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID, "bad point count");
    GWN_CHECK(k >= 1 && k < (1 << 20), GWN_ERR_INVALID, "I must have 1 to 2^20 columns");
    if (n == 0) return;
    GWN_CHECK(P && I && A, GWN_ERR_INVALID, "null input");
    ensure_device();
    point_areas_device(P, n, I, k, N, A, T, static_cast<hipStream_t>(stream));
  });
}

// ---------------- point cloud ----------------
int gwn_cloud_build_estimated_areas(const double* P, const double* N, int64_t n, int order, double* A,
                                    gwn_cloud** out) {
  return guard([&] {
    GWN_CHECK(out != nullptr, GWN_ERR_INVALID, "null output handle");
    *out = nullptr;
    GWN_CHECK(order >= 0 && order <= 2, GWN_ERR_INVALID, "order must be 0, 1 or 2");
    GWN_CHECK(n >= 0 && n < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID, "bad point count");
    GWN_CHECK(n == 0 || (P && N), GWN_ERR_INVALID, "null input");
    ensure_device();
    hipStream_t st = nullptr;
    gwn_cloud* c = new gwn_cloud();
    try {
      c->n = n;
      c->order = order;
      c->nterms = order == 0 ? 3 : (order == 1 ? 12 : 39);
      c->P.alloc(3 * (size_t)(n > 0 ? n : 1));
      c->N.alloc(3 * (size_t)(n > 0 ? n : 1));
      c->A.alloc((size_t)(n > 0 ? n : 1));
      if (n > 0) {
        // copyleft/cgal/fast_winding_number.cpp:32-35: octree, knn with k = 21, point areas
        constexpr int64_t kAreaNeighbours = 21;
        DevBuf<double> dP, dN, dA(n);
        upload(dP, P, 3 * (size_t)n, st);
        upload(dN, N, 3 * (size_t)n, st);
        DevBuf<int32_t> cand;
        const int64_t nc = distinct_points_device(dP.p, n, cand, st);
        const int64_t kk = knn_columns(n, kAreaNeighbours);
        DevBuf<int32_t> dI((size_t)n * kk);
        knn_device(dP.p, n, dP.p, n, kAreaNeighbours, cand.p, nc, dI.p, st);
        point_areas_device(dP.p, n, dI.p, kk, dN.p, dA.p, nullptr, st);
        to_float_device(dP.p, 3 * n, c->P.p, st);
        to_float_device(dN.p, 3 * n, c->N.p, st);
        to_float_device(dA.p, n, c->A.p, st);
        if (A) GWN_HIP(hipMemcpyAsync(A, dA.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        GWN_HIP(hipStreamSynchronize(st));
      }
      cloud_build_device(c, st, true);
      GWN_HIP(hipStreamSynchronize(st));
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

int gwn_cloud_build(const float* P, const float* N, const float* A, int64_t n, int order, gwn_cloud** out) {
  return guard([&] {
    GWN_CHECK(out != nullptr, GWN_ERR_INVALID, "null output handle");
    *out = nullptr;
    GWN_CHECK(order >= 0 && order <= 2, GWN_ERR_INVALID, "order must be 0, 1 or 2");
    GWN_CHECK(n >= 0 && n < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID, "bad point count");
    const bool octree_only = !N && !A;
    GWN_CHECK(n == 0 || (P && (octree_only || (N && A))), GWN_ERR_INVALID, "null input");
    ensure_device();
    hipStream_t st = nullptr;
    gwn_cloud* c = new gwn_cloud();
    try {
      c->n = n;
      c->order = order;
      c->nterms = order == 0 ? 3 : (order == 1 ? 12 : 39);
      c->octree_only = octree_only;
      upload(c->P, P, 3 * (size_t)n, st);
      if (octree_only) {
        c->N.alloc(3 * (size_t)(n > 0 ? n : 1));
        c->A.alloc((size_t)(n > 0 ? n : 1));
        GWN_HIP(hipMemsetAsync(c->N.p, 0, c->N.bytes(), st));
        GWN_HIP(hipMemsetAsync(c->A.p, 0, c->A.bytes(), st));
      } else {
        upload(c->N, N, 3 * (size_t)n, st);
        upload(c->A, A, (size_t)n, st);
      }
      cloud_build_device(c, st, !octree_only);
      GWN_HIP(hipStreamSynchronize(st));
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

int gwn_cloud_build_from_octree(const float* P, const float* N, const float* A, int64_t n, const int32_t* CH,
                                int64_t ncells, const int64_t* cell_start, const int32_t* point_list, int order,
                                const float* CM, const float* R, const float* EC, gwn_cloud** out) {
  return guard([&] {
    GWN_CHECK(out != nullptr, GWN_ERR_INVALID, "null output handle");
    *out = nullptr;
    GWN_CHECK(order >= 0 && order <= 2, GWN_ERR_INVALID, "order must be 0, 1 or 2");
    GWN_CHECK(n >= 0 && n < (int64_t)0x7FFFFFFF, GWN_ERR_INVALID, "bad point count");
    GWN_CHECK(n == 0 || (P && N && A), GWN_ERR_INVALID, "null input");
    ensure_device();
    hipStream_t st = nullptr;
    gwn_cloud* c = new gwn_cloud();
    try {
      c->n = n;
      c->order = order;
      c->nterms = order == 0 ? 3 : (order == 1 ? 12 : 39);
      upload(c->P, P, 3 * (size_t)n, st);
      upload(c->N, N, 3 * (size_t)n, st);
      upload(c->A, A, (size_t)n, st);
      cloud_build_from_octree_device(c, CH, ncells, cell_start, point_list, CM, R, EC, st);
      GWN_HIP(hipStreamSynchronize(st));
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

void gwn_cloud_free(gwn_cloud* c) { delete c; }

int gwn_cloud_num_cells(const gwn_cloud* c, int64_t* ncells) {
  return guard([&] {
    GWN_CHECK(c && ncells, GWN_ERR_INVALID, "null argument");
    *ncells = c->ncells;
  });
}

int gwn_cloud_copy_octree(const gwn_cloud* c, int32_t* CH, float* CN, float* W, int64_t* cell_start,
                          int32_t* cell_count, int32_t* point_list) {
  return guard([&] {
    GWN_CHECK(c != nullptr, GWN_ERR_INVALID, "null handle");
    GWN_CHECK(!c->user_octree, GWN_ERR_INVALID, "cloud was built over a caller's octree");
    const size_t nc = (size_t)c->ncells;
    if (CH && nc) GWN_HIP(hipMemcpy(CH, c->CH.p, sizeof(int32_t) * 8 * nc, hipMemcpyDeviceToHost));
    if (CN && nc) GWN_HIP(hipMemcpy(CN, c->CN.p, sizeof(float) * 3 * nc, hipMemcpyDeviceToHost));
    if (W && nc) GWN_HIP(hipMemcpy(W, c->Wd.p, sizeof(float) * nc, hipMemcpyDeviceToHost));
    if (cell_start && nc) GWN_HIP(hipMemcpy(cell_start, c->cstart.p, sizeof(int64_t) * nc, hipMemcpyDeviceToHost));
    if (cell_count && nc) GWN_HIP(hipMemcpy(cell_count, c->ccount.p, sizeof(int32_t) * nc, hipMemcpyDeviceToHost));
    if (point_list && c->n) GWN_HIP(hipMemcpy(point_list, c->perm.p, sizeof(int32_t) * (size_t)c->n, hipMemcpyDeviceToHost));
  });
}

int gwn_cloud_copy_expansion(const gwn_cloud* c, float* CM, float* R, float* EC) {
  return guard([&] {
    GWN_CHECK(c != nullptr, GWN_ERR_INVALID, "null handle");
    GWN_CHECK(!c->octree_only, GWN_ERR_INVALID, "cloud was built without normals and areas");
    const size_t nc = (size_t)c->ncells;
    if (CM && nc) GWN_HIP(hipMemcpy(CM, c->CM.p, sizeof(float) * 3 * nc, hipMemcpyDeviceToHost));
    if (R && nc) GWN_HIP(hipMemcpy(R, c->R.p, sizeof(float) * nc, hipMemcpyDeviceToHost));
    if (EC && nc) GWN_HIP(hipMemcpy(EC, c->EC.p, sizeof(float) * (size_t)c->nterms * nc, hipMemcpyDeviceToHost));
  });
}

int gwn_cloud_query_device(const gwn_cloud* c, const float* Q, int64_t nq, double beta, float* WN, void* stream) {
  return guard([&] {
    GWN_CHECK(c != nullptr, GWN_ERR_INVALID, "null handle");
    GWN_CHECK(nq >= 0 && nq < (int64_t)0xFFFFFFFF, GWN_ERR_INVALID, "bad query count");
    GWN_CHECK(nq == 0 || (Q && WN), GWN_ERR_INVALID, "null query/output");
    GWN_CHECK(!c->octree_only, GWN_ERR_INVALID, "cloud was built without normals and areas");
    cloud_query_device(c, Q, nq, beta, WN, (hipStream_t)stream);
  });
}

int gwn_cloud_query(const gwn_cloud* c, const float* Q, int64_t nq, double beta, float* WN) {
  return guard([&] {
    GWN_CHECK(c != nullptr, GWN_ERR_INVALID, "null handle");
    GWN_CHECK(nq >= 0 && nq < (int64_t)0xFFFFFFFF, GWN_ERR_INVALID, "bad query count");
    GWN_CHECK(nq == 0 || (Q && WN), GWN_ERR_INVALID, "null query/output");
    GWN_CHECK(!c->octree_only, GWN_ERR_INVALID, "cloud was built without normals and areas");
    if (nq == 0) return;
    hipStream_t st = nullptr;
    Scratch io(scratch_bytes({12 * (size_t)nq, 4 * (size_t)nq}), st);
    const float* dQ = stage_in(io, Q, 3 * (size_t)nq, st);
    float* dW = io.take_n<float>((size_t)nq);
    cloud_query_device(c, dQ, nq, beta, dW, st);
    stage_out(WN, dW, (size_t)nq, st);
    GWN_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"
