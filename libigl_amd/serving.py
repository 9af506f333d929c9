"""Python handle on the native continuous-batching server (gwn_server_*)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as nat
from ._native import check, lib


class Server:
    """Resident meshes + continuous batching over `slots` HIP streams within a
    device-memory budget.  submit() returns at once; the result array is
    filled when wait()/drain() says the request is done.  Give the process at
    least `slots` hardware queues (GPU_MAX_HW_QUEUES, set before it first uses
    HIP; 4 otherwise): slots that share a queue run one after the other."""

    def __init__(self, slots=4, device_budget_bytes=32 << 30):
        h = nat.vp()
        check(lib().gwn_server_create(slots, int(device_budget_bytes), C.byref(h)))
        self._h = h.value
        self._keep = {}  # request id -> (Q, W) kept alive until done

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().gwn_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_mesh(self, V, F, order=2, tree=nat.TREE_SAH4):
        V = np.ascontiguousarray(V, dtype=np.float64)
        F = np.ascontiguousarray(F, dtype=np.int32)
        mid = C.c_int32()
        check(lib().gwn_server_add_mesh(self._h, V, V.shape[0], F, F.shape[0], order, tree, C.byref(mid)))
        return mid.value

    def reserve(self, num_queries):
        """Sizes every slot's device buffer for requests of num_queries up
        front (call while no request is in flight)."""
        check(lib().gwn_server_reserve(self._h, int(num_queries)))

    @staticmethod
    def request_bytes(num_queries):
        """Device bytes a request of num_queries holds while it runs."""
        return int(lib().gwn_server_request_bytes(int(num_queries)))

    def add_mesh_file(self, path, order=2, tree=nat.TREE_SAH4):
        """Reads an .off / .obj file and keeps its tree resident."""
        mid = C.c_int32()
        check(lib().gwn_server_add_mesh_file(self._h, str(path).encode(), order, tree, C.byref(mid)))
        return mid.value

    def submit(self, mesh, Q, accuracy_scale=2.0, precision=nat.PREC_FP32, out=None):
        """Queues a request; `out` (float64, nq) receives the result if given
        (a caller that reuses its result buffers avoids first-touch page
        faults on fresh memory)."""
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(-1, 3)
        if out is None:
            W = np.empty(Q.shape[0], np.float64)
        else:
            W = out
            if W.dtype != np.float64 or not W.flags.c_contiguous or W.shape != (Q.shape[0],):
                raise ValueError("out must be a contiguous float64 array of nq elements")
        rid = C.c_int64()
        check(lib().gwn_server_submit(self._h, mesh, Q.ctypes.data, Q.shape[0], float(accuracy_scale), precision,
                                      W.ctypes.data, C.byref(rid)))
        self._keep[rid.value] = (Q, W)
        return rid.value

    def wait(self, rid):
        rc = lib().gwn_server_wait(self._h, rid)
        Q, W = self._keep.pop(rid)
        check(rc)
        return W

    def result(self, rid):
        return self._keep[rid][1]

    def drain(self):
        check(lib().gwn_server_drain(self._h))

    def info(self, rid):
        i = nat.RequestInfo()
        check(lib().gwn_server_request_info(self._h, rid, C.byref(i)))
        return i

    def memory(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().gwn_server_memory(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def now(self):
        return lib().gwn_server_now(self._h)


def poisson_trace(server, mesh_ids, batches, lam, duration, seed=0, accuracy_scale=2.0,
                  precision=nat.PREC_FP32, nq_lo=1_000, nq_hi=1_000_000, max_requests=None, out_ring=64):
    """Replays a Poisson arrival trace against a server.

    Requests arrive at rate `lam` per second for `duration` seconds; each
    names a random mesh and carries a batch of size log-uniform in
    [nq_lo, nq_hi], a window of that mesh's pre-generated queries (batches[m],
    uniform in its bbox).  Each request is submitted at its arrival time on
    the server clock; latency is measured from that scheduled arrival (not
    from the submit call) to the result in host memory.  Results land in a
    ring of `out_ring` reused output buffers, as a client that recycles its
    response buffers would do.
    Returns a dict of throughput and latency percentiles."""
    import time

    rng = np.random.default_rng(seed)
    arrivals = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / lam)
        if t >= duration or (max_requests and len(arrivals) >= max_requests):
            break
        arrivals.append(t)
    n = len(arrivals)
    meshes = rng.integers(0, len(mesh_ids), size=n)
    sizes = np.exp(rng.uniform(np.log(nq_lo), np.log(nq_hi), size=n)).astype(np.int64)
    starts = [int(rng.integers(0, batches[m].shape[0] - s + 1)) for m, s in zip(meshes, sizes)]
    ring = [np.ones(int(nq_hi), np.float64) for _ in range(max(1, min(out_ring, n)))]
    ring_rid = [None] * len(ring)
    rids = []
    t0 = server.now() + 0.01
    late = 0.0
    for k in range(n):
        slot = k % len(ring)
        if ring_rid[slot] is not None:
            server.wait(ring_rid[slot])  # buffer reuse: its previous request must be done
        target = t0 + arrivals[k]
        while True:
            d = target - server.now()
            if d <= 0:
                late = max(late, -d)
                break
            time.sleep(min(d, 0.002))
        m = meshes[k]
        Q = batches[m][starts[k]: starts[k] + sizes[k]]
        rid = server.submit(mesh_ids[m], Q, accuracy_scale, precision, out=ring[slot][: sizes[k]])
        ring_rid[slot] = rid
        rids.append(rid)
    server.drain()
    infos = [server.info(r) for r in rids]
    for r in rids:
        if r in server._keep:
            server.wait(r)
    arr = t0 + np.asarray(arrivals)
    lat = np.array([i.t_done for i in infos]) - arr
    wait = np.array([i.t_start for i in infos]) - arr
    serv = np.array([i.t_done - i.t_start for i in infos])
    q = np.array([i.num_queries for i in infos], dtype=np.float64)
    span = max(i.t_done for i in infos) - float(arr[0])
    ok = all(i.status == 0 for i in infos)
    big = q >= np.quantile(q, 0.9)
    # stages of the largest decile: narrowing and upload, device work and the
    # copy back, copy into the caller's W
    up = np.array([i.t_uploaded - i.t_start for i in infos])[big]
    dev = np.array([i.t_device - i.t_uploaded for i in infos])[big]
    out = np.array([i.t_done - i.t_device for i in infos])[big]
    stages = {k: (round(1e3 * float(np.quantile(v, 0.5)), 3), round(1e3 * float(np.quantile(v, 0.99)), 3))
              for k, v in (("upload", up), ("device", dev), ("copy_out", out))}
    return {"requests": n, "queries": int(q.sum()), "span_s": span, "queries_per_s": q.sum() / span,
            "offered_queries_per_s": q.sum() / duration, "p50_ms": 1e3 * float(np.quantile(lat, 0.5)),
            "p99_ms": 1e3 * float(np.quantile(lat, 0.99)), "max_ms": 1e3 * float(lat.max()),
            "queue_p99_ms": 1e3 * float(np.quantile(wait, 0.99)),
            "service_p99_ms": 1e3 * float(np.quantile(serv, 0.99)),
            "largest_decile_service_p50_ms": 1e3 * float(np.quantile(serv[big], 0.5)),
            "submit_lag_ms": 1e3 * late, "all_ok": ok, "lambda": lam,
            "largest_decile_stages_p50_p99_ms": stages}
