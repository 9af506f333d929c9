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
    # budget fits the mesh and one small request at a time
    probe = gwn.FastWindingNumberBVH().build(V, F)
    mesh_bytes = probe.info().device_bytes
    srv = Server(slots=4, device_budget_bytes=mesh_bytes * 2 + (16 << 20))
    mid = srv.add_mesh(V, F)
    rng = np.random.default_rng(1)
    Qs = [D.uniform_queries(V.min(0), V.max(0), 100_000, rng) for _ in range(6)]
    rids = [srv.submit(mid, Q) for Q in Qs]
    srv.drain()
    for rid, Q in zip(rids, Qs):
        assert np.array_equal(srv.wait(rid), probe.query(Q))
    resident, held, budget = srv.memory()
    assert resident + held <= budget
    # a request that can never fit is failed, not hung
    big = D.uniform_queries(V.min(0), V.max(0), 2_000_000, rng)
    rid = srv.submit(mid, big)
    srv.drain()
    assert srv.info(rid).status == gwn.GWNError(3, "").code
    with pytest.raises(gwn.GWNError):
        srv.wait(rid)
    # a mesh that does not fit is refused
    V2, F2 = D.uv_sphere(200_000, 4, 8)
    with pytest.raises(gwn.GWNError):
        srv.add_mesh(V2, F2)
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
