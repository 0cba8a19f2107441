"""Parity of the HIP path with the oracle at each benchmark workload's own
settings (default field arguments: maxSamplesPerLeaf 32000, K up to 32,
split-and-merge, deterministic; SURVEY workload generator):

  c1  200k surface samples, one Update;
  c2  8 successive incremental Updates of 1M samples, blob centres drifting
      0.01 per frame;
  c3  4096 pre-partitioned regions, sample counts log-U[64, 65536]
      (~36M samples), initial K in {8, 16, 32}, one batched launch;
  c4  query serving on the c1 field (200k surface + 200k volume samples):
      request sizes log-U[1k, 256k], 30% volume queries with an HG product;
  c5  16M surface + 4M volume samples (64 blobs): error of weights, mean
      directions, mean cosines and pdf against the FP64 oracle.

Bars: kd tree, leaf ids, region sample statistics and KNN lists bit-exact
after every Update; mixtures per region (parity_checks: median of the
per-region median relative pdf error <= 1e-4, >= 99% of the regions with
the 99th percentile < 1e-2, >= 99% with the oracle's component count);
query pdf of the GPU's sampled direction under the oracle's distribution
within 1e-3 relative for every valid query; region ids exact.
"""
import json
import os

import numpy as np
import pytest

from field_compare import OracleField, precision_report, unit_dirs
from oracle_ctypes import OracleRegion, default_config, load_hg_table_into_oracle
from parity_checks import assert_mixture_bars, assert_spatial_bit_exact, mixture_errors
from openpgl_amd.synth import batched_regions, generate, generate_large, queries

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _records(field, kind=0):
    n = field.num_regions(kind) if kind else field.num_regions()
    return [field.region(i, kind) if kind else field.region(i) for i in range(n)]


def _query_pdf_parity(gpu, orcs, n, seed, knn_u):
    """Batched GPU queries (30% volume) against the oracle: validity and ids
    exact, pdf of the GPU's sampled direction within 1e-3 for every valid query."""
    pos, nrm, u2, is_vol, g = queries(n, seed=seed)
    u1 = np.random.default_rng(seed).random(n).astype(np.float32) if knn_u else np.full(n, -1.0, np.float32)
    d, pdf, ids, _ = gpu.sample_batch(pos, nrm, u2, u1=u1, g=g, is_volume=is_vol)
    worst = 0.0
    for volume, o in enumerate(orcs):
        m = is_vol == bool(volume)
        _, opdf, oids, _ = o.query_batch(volume, pos[m], nrm[m], u2[m], u1=u1[m], g=g[m], threads=THREADS)
        valid = opdf > 0
        assert np.array_equal(pdf[m] > 0, valid)
        assert np.array_equal(ids[m][valid], oids[valid])
        idx = np.nonzero(m)[0][valid]
        rel = np.array([abs(pdf[i] - o.query_pdf(volume, pos[i], nrm[i], g[i], u1[i], d[i])[1]) /
                        o.query_pdf(volume, pos[i], nrm[i], g[i], u1[i], d[i])[1] for i in idx])
        print("volume" if volume else "surface", "valid queries", len(idx), "median rel", np.median(rel),
              "max rel", rel.max(), flush=True)
        assert len(idx) > 0.5 * m.sum()
        assert np.all(rel <= 1e-3)
        worst = max(worst, float(rel.max()))
    return worst


def test_c1_single_update():
    import openpgl_amd as pgl
    load_hg_table_into_oracle()
    dev = pgl.Device()
    gpu = pgl.Field(dev, max_samples_per_leaf=32000, deterministic=True)
    orc = OracleField(default_config(32000))
    s = generate(200000, seed=0)
    print("c1: 200000 samples", flush=True)
    gpu.update(pgl.SampleStorage(s))
    orc.update(s, threads=THREADS)
    R = assert_spatial_bit_exact(gpu, orc, generate(20000, seed=9, frame=1)["pos"])
    print("regions", R, flush=True)
    assert R >= 4
    assert_mixture_bars(mixture_errors(_records(gpu), _records(orc), unit_dirs(512)), "c1")
    # surface queries only (c1 trains the surface field)
    pos, nrm, u2, _, _ = queries(20000, seed=5)
    u1 = np.random.default_rng(6).random(len(pos)).astype(np.float32)
    d, pdf, ids, _ = gpu.sample_batch(pos, nrm, u2, u1=u1)
    _, opdf, oids, _ = orc.query_batch(0, pos, nrm, u2, u1=u1, threads=THREADS)
    valid = opdf > 0
    assert np.array_equal(pdf > 0, valid) and np.array_equal(ids[valid], oids[valid])
    rel = np.array([abs(pdf[i] / orc.query_pdf(0, pos[i], nrm[i], 0.0, u1[i], d[i])[1] - 1.0)
                    for i in np.nonzero(valid)[0]])
    print("c1 queries valid", int(valid.sum()), "median rel", np.median(rel), "max rel", rel.max(), flush=True)
    assert valid.mean() > 0.9 and np.all(rel <= 1e-3)


def test_c2_eight_incremental_frames():
    import openpgl_amd as pgl
    dev = pgl.Device()
    gpu = pgl.Field(dev, max_samples_per_leaf=32000, deterministic=True)
    orc = OracleField(default_config(32000))
    dirs = unit_dirs(512)
    summary = []
    for f in range(8):
        s = generate(1_000_000, seed=0, frame=f)
        print("c2: frame", f, flush=True)
        gpu.update(pgl.SampleStorage(s))
        orc.update(s, threads=THREADS)
        R = assert_spatial_bit_exact(gpu, orc, generate(20000, seed=13, frame=f + 1)["pos"])
        st = assert_mixture_bars(mixture_errors(_records(gpu), _records(orc), dirs), "c2 frame %d regions %d" % (f, R))
        summary.append(st)
    assert summary[-1]["regions"] > 50


def _oracle_region_records(recs, counts, stats, init_k, regions):
    import concurrent.futures
    offs = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])

    def one(r):
        cfg = default_config(32000)
        cfg.initK = int(init_k[r])
        o = OracleRegion(cfg)
        o.step(recs[offs[r]:offs[r + 1]], stats[r, 0:3], stats[r, 3:6], float(stats[r, 6]))
        return o.record()

    with concurrent.futures.ThreadPoolExecutor(max_workers=THREADS) as ex:
        return list(ex.map(one, regions))


_BATCH_KEYS = ["K", "weights", "kappas", "meanDirections", "normalizations", "meanCosines", "distances",
               "sumOfWeightedDirections", "sumOfWeightedStats", "sumOfDistanceWeights", "chiSquareMCEstimates",
               "splitCovariances", "emIterations", "partialEmIterations", "numSplits", "numMerges", "valid"]


def test_c3_batched_4096_regions():
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray
    dev = pgl.Device()
    recs, counts, stats, init_k = batched_regions(4096, seed=0)
    print("c3: regions", len(counts), "samples", len(recs), flush=True)
    d = DeviceArray.from_numpy(recs)
    batch = pgl.RegionBatch(dev)
    batch.step_device(d.value, counts, stats, init_k)
    gpu = [batch.region(r) for r in range(len(counts))]
    # every region: valid, weights sum to 1, components within [1, 32]
    for r, a in enumerate(gpu):
        assert a["valid"], r
        K = int(a["K"])
        assert 1 <= K <= 32
        assert abs(float(a["weights"][:K].astype(np.float64).sum()) - 1.0) < 1e-5, r
    # every 16th region against the oracle
    sel = list(range(0, len(counts), 16))
    orc = _oracle_region_records(recs, counts, stats, init_k, sel)
    assert_mixture_bars(mixture_errors([gpu[r] for r in sel], orc, unit_dirs(512)), "c3 (every 16th region)")
    # another schedule (forced large teams) gives the same state for all regions
    b2 = pgl.RegionBatch(dev)
    b2.set_schedule(per_region_launches=0, balance=8.0, max_team=64, min_chunks_per_rank=1)
    b2.step_device(d.value, counts, stats, init_k)
    for r in range(len(counts)):
        a, b = gpu[r], b2.region(r)
        for k in _BATCH_KEYS:
            assert np.asarray(a[k]).tobytes() == np.asarray(b[k]).tobytes(), (r, k)


def _c1_field_with_volume(dev, knn=True):
    import openpgl_amd as pgl
    gpu = pgl.Field(dev, max_samples_per_leaf=32000, deterministic=True, knn=knn)
    cfg = default_config(32000)
    cfg.knnLookup = 1 if knn else 0
    s = generate(200000, seed=0)
    v = generate(200000, seed=1, flags=2)
    gpu.update(pgl.SampleStorage(np.concatenate([s, v])))
    orc_s, orc_v = OracleField(cfg), OracleField(cfg)
    orc_s.update(s, threads=THREADS)
    orc_v.update(v, threads=THREADS)
    return gpu, orc_s, orc_v


def test_c4_server_on_c1_field():
    import threading
    import openpgl_amd as pgl
    load_hg_table_into_oracle()
    dev = pgl.Device()
    gpu, orc_s, orc_v = _c1_field_with_volume(dev)
    # queries with u1 = -1 (the workload's two randoms) and with a KNN random
    _query_pdf_parity(gpu, (orc_s, orc_v), 8000, seed=21, knn_u=False)
    _query_pdf_parity(gpu, (orc_s, orc_v), 8000, seed=22, knn_u=True)
    rng = np.random.default_rng(4)
    sizes = np.floor(np.exp(rng.uniform(np.log(1000), np.log(256000), 12))).astype(int)
    reqs = []
    for i, n in enumerate(sizes):
        pos, nrm, u2, is_vol, g = queries(int(n), seed=300 + i)
        u1 = rng.random(int(n)).astype(np.float32) if i % 3 == 0 else None
        reqs.append((pos, nrm, u2, u1, g, is_vol.astype(np.uint8)))
    print("c4: request sizes", sizes.tolist(), flush=True)
    server = pgl.Server(gpu, slots=4)
    alone = []
    for r in reqs:
        rid, out = server.submit(*r[:3], u1=r[3], g=r[4], is_volume=r[5])
        assert server.wait(rid, 60000.0)
        server.release(rid)
        alone.append(out)
    mixed, ids = [None] * len(reqs), [None] * len(reqs)

    def worker(idx):
        for i in idx:
            r = reqs[i]
            ids[i], mixed[i] = server.submit(*r[:3], u1=r[3], g=r[4], is_volume=r[5])

    ths = [threading.Thread(target=worker, args=(list(range(t, len(reqs), 4)),)) for t in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    for rid in ids:
        assert server.wait(rid, 60000.0)
        server.release(rid)
    print("c4 server stats", server.stats(), flush=True)
    for i, (a, m) in enumerate(zip(alone, mixed)):
        assert all(np.asarray(x).tobytes() == np.asarray(y).tobytes() for x, y in zip(a, m)), i
    # the served results equal the batch API's, which the oracle pins above
    for i in (0, 5):
        r = reqs[i]
        ref = gpu.sample_batch(r[0], r[1], r[2], u1=r[3], g=r[4], is_volume=r[5])
        assert all(np.asarray(x).tobytes() == np.asarray(y).tobytes() for x, y in zip(alone[i], ref)), i
    server.close()


@pytest.mark.timeout(600)
def test_c5_large_mixed_error_against_fp64():
    import openpgl_amd as pgl
    dev = pgl.Device()
    s = generate_large(16_000_000, seed=0, num_blobs=64)
    v = generate_large(4_000_000, seed=1, num_blobs=64, flags=2)
    print("c5: 16M surface + 4M volume records", flush=True)
    gpu = pgl.Field(dev, max_samples_per_leaf=32000, deterministic=True)
    both = np.concatenate([s, v])
    gpu.update(pgl.SampleStorage(both))
    del both
    cfg = default_config(32000)
    dirs = unit_dirs(512)
    report = {}
    for kind, data in ((0, s), (1, v)):
        o32, o64 = OracleField(cfg, 0), OracleField(cfg, 1)
        o32.update(data, threads=THREADS)
        o64.update(data, threads=THREADS)
        print("c5 kind", kind, "oracles trained", flush=True)
        R = gpu.num_regions(kind)
        assert R == o64.num_regions() == o32.num_regions()
        g = _records(gpu, kind)
        r32 = _records(o32)
        r64 = _records(o64)
        name = "volume" if kind else "surface"
        report[name + " GPU vs FP64"] = precision_report(g, r64, dirs)
        report[name + " FP32 oracle vs FP64"] = precision_report(r32, r64, dirs)
        assert_mixture_bars(mixture_errors(g, r32, dirs), "c5 %s GPU vs FP32 oracle" % name)
    print("c5 precision:", json.dumps(report), flush=True)
    for name in ("surface", "volume"):
        gpu_rep, o32_rep = report[name + " GPU vs FP64"], report[name + " FP32 oracle vs FP64"]
        assert gpu_rep["pdf_rel_median"] <= 1e-4
        assert gpu_rep["pdf_rel_median"] <= 2 * o32_rep["pdf_rel_median"] + 1e-6
