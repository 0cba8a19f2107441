"""Continuous-batching query server (pglServer*).

* Every request's results are bit-identical whether it runs alone or mixed
  with concurrently submitted requests, and equal to the synchronous batch
  API (pglFieldSampleBatch) on the same queries (tolerance: exact, the query
  kernel is per-query deterministic).
* Memory admission: in-flight bytes never exceed the budget; a request larger
  than the budget is rejected.
* The server serves its snapshot while the field is updated, and the updated
  field after a refresh.
* Requests of 1 to 65 queries (less than one wave up to just over it) equal
  the batch API; an empty request is rejected.
"""
import threading

import numpy as np
import pytest

from openpgl_amd.synth import generate, queries

pytestmark = pytest.mark.gpu


def _field(dev, frame=0):
    import openpgl_amd as pgl
    f = pgl.Field(dev, max_samples_per_leaf=4000)
    s = generate(60000, seed=0, frame=frame)
    v = generate(30000, seed=1, flags=2, frame=frame)
    f.update(pgl.SampleStorage(np.concatenate([s, v])))
    return f


def _requests(count, seed, max_n=20000):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = int(rng.integers(1, max_n))
        pos, nrm, u2, is_vol, g = queries(n, seed=seed * 1000 + i)
        u1 = rng.random(n).astype(np.float32) if i % 2 else None
        reqs.append((pos, nrm, u2, u1, g, is_vol.astype(np.uint8)))
    return reqs


def _run_alone(server, reqs):
    outs = []
    for r in reqs:
        rid, out = server.submit(*r[:3], u1=r[3], g=r[4], is_volume=r[5])
        assert server.wait(rid, 60000.0)
        outs.append(out)
        server.release(rid)
    return outs


def _run_mixed(server, reqs):
    outs = [None] * len(reqs)
    ids = [None] * len(reqs)

    def worker(idx):
        for i in idx:
            r = reqs[i]
            ids[i], outs[i] = server.submit(*r[:3], u1=r[3], g=r[4], is_volume=r[5])

    threads = [threading.Thread(target=worker, args=(list(range(t, len(reqs), 3)),)) for t in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for rid in ids:
        assert server.wait(rid, 60000.0)
        server.release(rid)
    return outs


def _same(a, b):
    return all(np.asarray(x).tobytes() == np.asarray(y).tobytes() for x, y in zip(a, b))


def test_server_same_alone_and_in_mix():
    import openpgl_amd as pgl
    dev = pgl.Device()
    f = _field(dev)
    reqs = _requests(24, seed=3)
    server = pgl.Server(f, slots=3)
    alone = _run_alone(server, reqs)
    mixed = _run_mixed(server, reqs)
    st = server.stats()
    print("stats", st, flush=True)
    assert st["batches"] < 2 * len(reqs)  # the mixed run batched requests together
    for i, (a, m) in enumerate(zip(alone, mixed)):
        assert _same(a, m), i
    # the synchronous batch API computes the same queries identically
    for i in (0, 5, 11):
        r = reqs[i]
        d, pdf, ids, integ = f.sample_batch(r[0], r[1], r[2], u1=r[3], g=r[4], is_volume=r[5])
        assert _same(alone[i], (d, pdf, ids, integ)), i
    valid = np.concatenate([o[1] for o in alone]) > 0
    assert valid.mean() > 0.9
    server.close()


def test_server_memory_admission():
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import pinned_copy
    dev = pgl.Device()
    f = _field(dev)
    per_query = 68
    budget = 25000 * per_query
    server = pgl.Server(f, memory_budget=budget, slots=4)
    reqs = _requests(30, seed=5, max_n=20000)
    reqs = [tuple(None if a is None else pinned_copy(a) for a in r) for r in reqs]
    outs = _run_mixed(server, reqs)
    st = server.stats()
    print("stats", st, flush=True)
    assert st["requests"] == 30
    assert st["peak_inflight_bytes"] <= budget
    with pytest.raises(RuntimeError):
        big = _requests(1, seed=9, max_n=30000)[0]
        n = 26000
        server.submit(np.resize(big[0], (n, 3)).astype(np.float32), np.resize(big[1], (n, 3)).astype(np.float32),
                      np.resize(big[2], (n, 2)).astype(np.float32))
    ref = _run_alone(pgl.Server(f), reqs[:4])
    for i in range(4):
        assert _same(outs[i], ref[i])
    server.close()


def test_server_snapshot_and_refresh():
    import openpgl_amd as pgl
    dev = pgl.Device()
    f = _field(dev)
    reqs = _requests(4, seed=7)
    before = [f.sample_batch(r[0], r[1], r[2], u1=r[3], g=r[4], is_volume=r[5]) for r in reqs]
    server = pgl.Server(f)
    s = generate(60000, seed=0, frame=1)
    v = generate(30000, seed=1, flags=2, frame=1)
    f.update(pgl.SampleStorage(np.concatenate([s, v])))
    after = [f.sample_batch(r[0], r[1], r[2], u1=r[3], g=r[4], is_volume=r[5]) for r in reqs]
    assert not all(_same(a, b) for a, b in zip(before, after))
    snap = _run_alone(server, reqs)
    for a, b in zip(snap, before):
        assert _same(a, b)
    server.refresh()
    fresh = _run_alone(server, reqs)
    for a, b in zip(fresh, after):
        assert _same(a, b)
    server.close()


def test_server_small_and_empty_requests():
    import openpgl_amd as pgl
    dev = pgl.Device()
    f = _field(dev)
    server = pgl.Server(f, slots=2)
    reqs = []
    for i, n in enumerate([1, 2, 3, 63, 64, 65]):
        pos, nrm, u2, is_vol, g = queries(n, seed=200 + i)
        u1 = np.random.default_rng(i).random(n).astype(np.float32) if i % 2 else None
        reqs.append((pos, nrm, u2, u1, g, is_vol.astype(np.uint8)))
    outs = _run_alone(server, reqs) + _run_mixed(server, reqs)
    for i, r in enumerate(reqs):
        ref = f.sample_batch(r[0], r[1], r[2], u1=r[3], g=r[4], is_volume=r[5])
        assert _same(outs[i], ref) and _same(outs[len(reqs) + i], ref), i
    with pytest.raises(RuntimeError):
        e3 = np.zeros((0, 3), np.float32)
        server.submit(e3, e3, np.zeros((0, 2), np.float32))
    server.close()
