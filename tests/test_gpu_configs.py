"""Parity of the HIP path with the oracle at each benchmark workload's own
settings (default field arguments: maxSamplesPerLeaf 32000, K up to 32,
split-and-merge, deterministic; the blob generator of openpgl_amd/synth.py):

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
after every Update; mixtures three-way against the FP32 and FP64 oracles
(tests/parity_checks.py); queries: the query path alone (the oracle's
trained field served by the GPU through a field file) gives the oracle's
region ids exactly and the pdf of every valid query within 1e-3 relative;
end to end (GPU-trained field against oracle-trained fields) three-way as
the mixtures (tests/parity_checks.py assert_query_pdfs).
"""
import json
import os

import numpy as np
import pytest

from field_compare import OracleField, canonical_ids, load_oracle_field_on_gpu, precision_report, unit_dirs
from oracle_ctypes import OracleRegion, default_config, load_hg_table_into_oracle
from parity_checks import assert_mixture_parity, assert_query_pdfs, assert_spatial_bit_exact, query_pdf_errors
from openpgl_amd.synth import batched_regions, generate, generate_large, queries

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _records(field, kind=0):
    n = field.num_regions(kind) if kind else field.num_regions()
    return [field.region(i, kind) if kind else field.region(i) for i in range(n)]


def _query_rel(gpu, orcs, n, seed, knn_u, canonical=False, orcs64=None):
    """Batched GPU queries (30% volume) against the oracle fields (surface,
    volume; None = untrained): validity and region ids exact; returns the
    relative pdf errors of the GPU's sampled directions (query_pdf_errors)"""
    pos, nrm, u2, is_vol, g = queries(n, seed=seed)
    u1 = np.random.default_rng(seed).random(n).astype(np.float32) if knn_u else np.full(n, -1.0, np.float32)
    d, pdf, ids, _ = gpu.sample_batch(pos, nrm, u2, u1=u1, g=g, is_volume=is_vol)
    out = [[], [], []]
    for volume, o in enumerate(orcs):
        if o is None:
            continue
        m = is_vol == bool(volume)
        _, opdf, oids, _ = o.query_batch(volume, pos[m], nrm[m], u2[m], u1=u1[m], g=g[m], threads=THREADS)
        valid = opdf > 0
        assert np.array_equal(pdf[m] > 0, valid)
        expect = canonical_ids(o)[oids[valid]] if canonical else oids[valid]
        assert np.array_equal(ids[m][valid], expect)
        assert valid.sum() > 0.5 * m.sum()
        o64 = orcs64[volume] if orcs64 is not None else None
        for k, r in enumerate(query_pdf_errors(d, pdf, np.nonzero(m)[0][valid], volume, pos, nrm, g, u1, o, o64)):
            if r is not None:
                out[k].append(r)
    return tuple(np.concatenate(x) if x else None for x in out)


def test_c1_single_update(tmp_path):
    import openpgl_amd as pgl
    load_hg_table_into_oracle()
    dev = pgl.Device()
    gpu = pgl.Field(dev, max_samples_per_leaf=32000, deterministic=True)
    cfg = default_config(32000)
    o32, o64 = OracleField(cfg, 0), OracleField(cfg, 1)
    s = generate(200000, seed=0)
    print("c1: 200000 samples", flush=True)
    gpu.update(pgl.SampleStorage(s))
    o32.update(s, threads=THREADS)
    o64.update(s, threads=THREADS)
    R = assert_spatial_bit_exact(gpu, o32, generate(20000, seed=9, frame=1)["pos"])
    assert R >= 4
    assert_mixture_parity(_records(gpu), _records(o32), _records(o64), unit_dirs(512), "c1 regions %d" % R)
    # end to end (surface only: c1 trains the surface field)
    r32, r64, r3264 = _query_rel(gpu, (o32, None), 20000, seed=5, knn_u=True, orcs64=(o64, None))
    assert_query_pdfs(r32, "c1 queries end to end", r64, r3264)
    # the query path alone: the oracle's field served by the GPU
    served = load_oracle_field_on_gpu(dev, str(tmp_path / "c1.pgl"), cfg, 1, 32000, surface=o32)
    rel = _query_rel(served, (o32, None), 20000, seed=5, knn_u=True, canonical=True)[0]
    print("c1 queries on the oracle's field: median rel", np.median(rel), "max", rel.max(), flush=True)
    assert np.all(rel <= 1e-3)


def test_c2_eight_incremental_frames():
    import openpgl_amd as pgl
    dev = pgl.Device()
    gpu = pgl.Field(dev, max_samples_per_leaf=32000, deterministic=True)
    cfg = default_config(32000)
    o32, o64 = OracleField(cfg, 0), OracleField(cfg, 1)
    dirs = unit_dirs(512)
    for f in range(8):
        s = generate(1_000_000, seed=0, frame=f)
        print("c2: frame", f, flush=True)
        gpu.update(pgl.SampleStorage(s))
        o32.update(s, threads=THREADS)
        o64.update(s, threads=THREADS)
        R = assert_spatial_bit_exact(gpu, o32, generate(20000, seed=13, frame=f + 1)["pos"])
        assert_mixture_parity(_records(gpu), _records(o32), _records(o64), dirs, "c2 frame %d regions %d" % (f, R))
    assert R > 200


def _oracle_region_records(recs, counts, stats, init_k, regions, precision=0):
    import concurrent.futures
    offs = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])

    def one(r):
        cfg = default_config(32000)
        cfg.initK = int(init_k[r])
        o = OracleRegion(cfg, precision)
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
    o32 = _oracle_region_records(recs, counts, stats, init_k, sel, 0)
    o64 = _oracle_region_records(recs, counts, stats, init_k, sel, 1)
    assert_mixture_parity([gpu[r] for r in sel], o32, o64, unit_dirs(512), "c3 (every 16th region)")
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
    orcs = []
    for prec in (0, 1):
        orc_s, orc_v = OracleField(cfg, prec), OracleField(cfg, prec)
        orc_s.update(s, threads=THREADS)
        orc_v.update(v, threads=THREADS)
        orcs.append((orc_s, orc_v))
    return gpu, orcs[0], orcs[1]


def test_c4_server_on_c1_field(tmp_path):
    import threading
    import openpgl_amd as pgl
    load_hg_table_into_oracle()
    dev = pgl.Device()
    gpu, (orc_s, orc_v), o64s = _c1_field_with_volume(dev)
    # the query path alone: the oracle's c1 surface and volume fields served by the GPU,
    # queries with u1 = -1 (the workload's two randoms) and with a KNN random
    served = load_oracle_field_on_gpu(dev, str(tmp_path / "c1sv.pgl"), default_config(32000), 1, 32000,
                                      surface=orc_s, volume=orc_v)
    for knn_u, seed in ((False, 21), (True, 22)):
        rel = _query_rel(served, (orc_s, orc_v), 8000, seed=seed, knn_u=knn_u, canonical=True)[0]
        print("c4 queries on the oracle's field, knn", knn_u, "median rel", np.median(rel), "max", rel.max(),
              flush=True)
        assert np.all(rel <= 1e-3)
    # end to end on the GPU-trained field
    for knn_u, seed in ((False, 23), (True, 24)):
        r32, r64, r3264 = _query_rel(gpu, (orc_s, orc_v), 8000, seed=seed, knn_u=knn_u, orcs64=o64s)
        assert_query_pdfs(r32, "c4 end to end knn %s" % knn_u, r64, r3264)
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
        assert_mixture_parity(g, r32, r64, dirs, "c5 %s regions %d" % (name, R))
    print("c5 precision:", json.dumps(report), flush=True)
