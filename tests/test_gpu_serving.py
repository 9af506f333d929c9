"""Serving loop: continuous batching over HIP streams with resident meshes and
memory-based admission.  Every result in the concurrent mix must equal the
same request run alone (stated tolerance 1e-4; observed: bit-identical)."""
import numpy as np
import pytest

from libigl_amd import datagen as D
from libigl_amd import gwn
from libigl_amd.serving import Server, poisson_trace

pytestmark = pytest.mark.gpu


def _meshes():
    out = []
    for i, nt in enumerate((1000, 5000, 20000, 60000)):
        out.append(D.uv_sphere(nt, nbumps=6, seed=100 + i))
    return out


def test_mix_equals_alone():
    meshes = _meshes()
    srv = Server(slots=4, device_budget_bytes=8 << 30)
    ids = [srv.add_mesh(V, F) for V, F in meshes]
    alone = [gwn.FastWindingNumberBVH().build(V, F) for V, F in meshes]
    rng = np.random.default_rng(0)
    reqs = []
    for k in range(48):
        m = int(rng.integers(0, len(meshes)))
        V, _ = meshes[m]
        nq = int(np.exp(rng.uniform(np.log(100), np.log(200_000))))
        Q = D.uniform_queries(V.min(0), V.max(0), nq, rng)
        prec = [gwn.PREC_FP32, gwn.PREC_MIXED][k % 2]
        beta = [2.0, 4.0][(k // 2) % 2]
        reqs.append((srv.submit(ids[m], Q, beta, prec), m, Q, beta, prec))
    srv.drain()
    for rid, m, Q, beta, prec in reqs:
        W = srv.wait(rid)
        Wa = alone[m].query(Q, beta, prec)
        assert np.array_equal(W, Wa), (rid, np.abs(W - Wa).max())
    for rid, *_ in reqs:
        i = srv.info(rid)
        assert i.status == 0 and i.t_submit <= i.t_start <= i.t_done
    srv.close()


def test_admission_by_memory():
    V, F = D.uv_sphere(4000, 4, 7)
    probe = gwn.FastWindingNumberBVH().build(V, F)
    tmp = Server(slots=1, device_budget_bytes=1 << 30)
    tmp.add_mesh(V, F)
    mesh_bytes = tmp.memory()[0]  # resident bytes include the FP64 moments
    tmp.close()
    small = Server.request_bytes(100_000)
    # a request above one chunk runs double-buffered: its footprint stays bounded
    assert Server.request_bytes(2_000_000) == Server.request_bytes(10_000_000) > Server.request_bytes(262_144)
    # budget: the mesh plus room for one 100k-query request at a time
    budget = mesh_bytes + small + small // 2
    srv = Server(slots=4, device_budget_bytes=budget)
    mid = srv.add_mesh(V, F)
    rng = np.random.default_rng(1)
    Qs = [D.uniform_queries(V.min(0), V.max(0), 100_000, rng) for _ in range(6)]
    rids = [srv.submit(mid, Q) for Q in Qs]
    srv.drain()
    for rid, Q in zip(rids, Qs):
        assert np.array_equal(srv.wait(rid), probe.query(Q))
    resident, held, budget_r = srv.memory()
    assert resident == mesh_bytes and budget_r == budget and resident + held <= budget
    # a request whose footprint can never fit is failed, not hung
    big = D.uniform_queries(V.min(0), V.max(0), 2_000_000, rng)
    assert Server.request_bytes(big.shape[0]) > budget - mesh_bytes
    rid = srv.submit(mid, big)
    srv.drain()
    assert srv.info(rid).status == gwn.GWNError(3, "").code
    with pytest.raises(gwn.GWNError):
        srv.wait(rid)
    # reserving slot buffers beyond the budget is refused
    with pytest.raises(gwn.GWNError):
        srv.reserve(1_000_000)
    # a mesh that does not fit is refused
    V2, F2 = D.uv_sphere(200_000, 4, 8)
    with pytest.raises(gwn.GWNError):
        srv.add_mesh(V2, F2)
    srv.close()


def test_chunked_requests_equal_alone():
    """Requests above one chunk (1,048,576 queries) run as a pipeline of chunks
    over two buffers; results equal one standalone query of the whole batch,
    also after reserve() pre-sized the slots."""
    V, F = D.uv_sphere(30_000, 6, 17)
    alone = gwn.FastWindingNumberBVH().build(V, F)
    srv = Server(slots=2, device_budget_bytes=4 << 30)
    mid = srv.add_mesh(V, F)
    srv.reserve(1_000_000)
    _, held, _ = srv.memory()
    assert held >= 2 * Server.request_bytes(1_000_000)  # each slot sized (to its class) for such a request
    rng = np.random.default_rng(4)
    reqs = []
    for nq, prec in ((1_048_576, gwn.PREC_FP32), (1_048_577, gwn.PREC_MIXED), (600_001, gwn.PREC_FP32),
                     (2_500_001, gwn.PREC_MIXED), (5, gwn.PREC_FP64), (70_000, gwn.PREC_FP32)):
        Q = D.uniform_queries(V.min(0) - 0.1, V.max(0) + 0.1, nq, rng)
        reqs.append((srv.submit(mid, Q, 2.0, prec), Q, prec))
    srv.drain()
    for rid, Q, prec in reqs:
        assert np.array_equal(srv.wait(rid), alone.query(Q, 2.0, prec)), (Q.shape, prec)
    srv.close()


def test_poisson_trace_small():
    meshes = _meshes()
    srv = Server(slots=4, device_budget_bytes=8 << 30)
    ids = [srv.add_mesh(V, F) for V, F in meshes]
    rng = np.random.default_rng(5)
    batches = [D.uniform_queries(V.min(0), V.max(0), 200_000, rng) for V, _ in meshes]
    r = poisson_trace(srv, ids, batches, lam=500.0, duration=0.5, seed=3, nq_hi=200_000)
    print(r)
    assert r["all_ok"] and r["requests"] > 50 and r["p99_ms"] >= r["p50_ms"] > 0
    srv.close()


def test_add_mesh_file(tmp_path):
    """A mesh read from a file (.obj with quads split into fans) serves the same
    results as the same arrays added directly."""
    V, F = D.uv_sphere(3000, 4, 21)
    p = tmp_path / "m.obj"
    with open(p, "w") as f:
        for v in V:
            f.write("v %.17g %.17g %.17g\n" % tuple(v))
        for t in F:
            f.write("f %d %d %d\n" % tuple(t + 1))
    srv = Server(slots=2, device_budget_bytes=1 << 30)
    a = srv.add_mesh_file(str(p))
    b = srv.add_mesh(V, F)
    Q = D.uniform_queries(V.min(0), V.max(0), 5000, np.random.default_rng(2))
    ra, rb = srv.submit(a, Q), srv.submit(b, Q)
    srv.drain()
    assert np.array_equal(srv.wait(ra), srv.wait(rb))
    srv.close()


def test_add_mesh_while_requests_in_flight():
    """gwn_server_add_mesh may grow the mesh list while slots serve requests
    on earlier meshes (each request holds its own mesh, resolved at submit)."""
    import threading
    meshes = _meshes()
    srv = Server(slots=4, device_budget_bytes=8 << 30)
    first = srv.add_mesh(*meshes[0])
    alone = gwn.FastWindingNumberBVH().build(*meshes[0])
    rng = np.random.default_rng(12)
    Qs = [D.uniform_queries(meshes[0][0].min(0), meshes[0][0].max(0), 150_000, rng) for _ in range(24)]
    added = []

    def add_more():
        for _ in range(6):
            for V, F in meshes[1:]:
                added.append(srv.add_mesh(V, F))

    th = threading.Thread(target=add_more)
    th.start()
    rids = [srv.submit(first, Q, 2.0) for Q in Qs]
    th.join()
    srv.drain()
    for rid, Q in zip(rids, Qs):
        assert np.array_equal(srv.wait(rid), alone.query(Q, 2.0))
    assert len(added) == 18 and len(set(added)) == 18
    Q = D.uniform_queries(meshes[3][0].min(0), meshes[3][0].max(0), 10_000, rng)
    r = srv.submit(added[-1], Q)
    assert np.array_equal(srv.wait(r), gwn.FastWindingNumberBVH().build(*meshes[3]).query(Q, 2.0))
    srv.close()

