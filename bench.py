"""Benchmark of the MI355X guiding field.

Workload (headline line): 8 successive frames of incremental Field::Update,
1M surface samples per frame from the synthetic generator (16 Gaussian blobs,
sigma 0.05, centres drifting 0.01 per frame, 4-lobe vMF directions), default
Open PGL settings (maxSamplesPerLeaf 32000, max depth 32, initK 16, maxK 32,
split-and-merge, deterministic).  Frames 0..W-1 are warm-up (frame 0 builds
the field), frames W..W+K-1 are timed.  All frames' records are resident in
HBM before timing (pglFieldUpdateSurfaceDevice).

value = timed samples / timed seconds.  The roofline object reports the
region kernel (the dominant kernel: weighted EM + split/merge for all regions
in one launch) against the FP32 vector peak; its time comes from HIP events
recorded inside the library around the kernel launch on the kernel's stream.
traffic (HBM bytes per launch) is read from profiles/k_region_update_traffic.txt,
which holds the rocprofv3 FETCH_SIZE / WRITE_SIZE measurements of the same
workload and the copy-kernel calibration (--calibrate-copy).  cpu_baseline
times the oracle (a CPU restatement of the reference, std::thread over
regions) on a bounded sample of the same workload.  A closed-loop serving run
(batches of log-uniform size in [1k, 256k] queries) reports guided samples/s
and per-batch p50/p99 latency.
"""
import argparse
import json
import os
import sys
import time

# Serving next to training needs a hardware queue per concurrently busy stream
# (a field uses 3 streams, a query server slots + 1); HIP's default is 4 per
# process, and streams sharing a queue serialise (a query batch would wait
# behind a region kernel).  A smaller value in the environment (such as the
# default 4 exported explicitly) is raised.  Must be set before the HIP
# runtime starts.
try:
    _hwq = int(os.environ.get("GPU_MAX_HW_QUEUES", "0"))
except ValueError:
    _hwq = 0
if _hwq < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector (packed) peak
PEAK_HBM_GBS = 8000.0
N_PER_FRAME = 1_000_000
# algorithmic FP32 flops per (sample, component) evaluation of each pass
FLOPS_ESTEP = 20.0
FLOPS_SPLIT = 47.0
FLOPS_DIST = 20.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_frames(n, frames, seed=0):
    from openpgl_amd.synth import generate
    return [generate(n, seed=seed, frame=f) for f in range(frames)]


def run_gpu(frames_host, warmup, steps, schedule=None):
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray, synchronize
    dev = pgl.Device()
    field = pgl.Field(dev, max_samples_per_leaf=32000, deterministic=True)
    if schedule:
        field.set_region_schedule(**schedule)
    dframes = [DeviceArray.from_numpy(f) for f in frames_host]
    synchronize()
    for f in range(warmup):
        field.update_surface_device(dframes[f].value, len(frames_host[f]))
        log("warmup frame", f, "regions", field.num_regions(), field.last_update_info())
    synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        f = warmup + s
        field.update_surface_device(dframes[f].value, len(frames_host[f]))
    synchronize()
    t1 = time.perf_counter()
    # updates are queued; their figures are collected afterwards
    infos = field.update_history(steps)
    for s, i in enumerate(infos):
        log("timed frame", warmup + s, i)
    return field, dev, (t1 - t0), infos


def render_loop(frames_host, warmup, steps, queries_per_frame=262144):
    """The drop-in use in a guided render loop, synchronous: per frame the
    records arrive in a host SampleStorage (pglFieldUpdate copies them over
    PCIe), the update completes, then one batch of guided queries
    (pglFieldSampleBatch from host arrays, results back in host memory) runs
    on the new field before the next frame.  Nothing overlaps across frames."""
    import openpgl_amd as pgl
    from openpgl_amd.synth import queries
    dev = pgl.Device()
    field = pgl.Field(dev, max_samples_per_leaf=32000, deterministic=True)
    stores = [pgl.SampleStorage(f) for f in frames_host]
    pos, nrm, u2, is_vol, g = queries(queries_per_frame, seed=17)
    is_vol[:] = False  # the loop trains the surface field only
    upd, qry = [], []
    for f in range(warmup + steps):
        t0 = time.perf_counter()
        field.update(stores[f])
        field.num_regions()  # waits for the update
        t1 = time.perf_counter()
        field.sample_batch(pos, nrm, u2, g=g, is_volume=is_vol)
        t2 = time.perf_counter()
        if f >= warmup:
            upd.append(t1 - t0)
            qry.append(t2 - t1)
    n = sum(len(frames_host[warmup + s]) for s in range(steps))
    frame_s = sum(upd) + sum(qry)
    return {"mode": "synchronous: host SampleStorage -> pglFieldUpdate (PCIe copy included) -> wait -> "
                    "pglFieldSampleBatch of %d host queries -> next frame" % queries_per_frame,
            "update_samples_per_s": n / sum(upd), "frame_samples_per_s": n / frame_s,
            "update_ms_avg": 1e3 * float(np.mean(upd)), "query_batch_ms_avg": 1e3 * float(np.mean(qry)),
            "frames_timed": steps}


def serving(field, n_batches=40, seed=0):
    """Closed loop: each batch is a log-uniform [1k, 256k] set of queries (30%
    volume is not trained here -> surface only), inputs resident on device."""
    from openpgl_amd.hip_runtime import DeviceArray, hip, synchronize
    from openpgl_amd.synth import queries
    import ctypes
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(1000), np.log(256000), n_batches)).astype(int)
    nmax = int(sizes.max())
    pos, nrm, u2, _, _ = queries(nmax, seed=seed)
    u1 = rng.random(nmax).astype(np.float32)
    dpos, dnrm, du2, du1 = (DeviceArray.from_numpy(a) for a in (pos, nrm, u2, u1))
    ddir = DeviceArray(nmax * 12)
    dpdf = DeviceArray(nmax * 4)
    lib = field.lib
    synchronize()
    # warm-up
    lib.pglFieldSampleBatchDevice(field.h, 1000, dpos.ptr, dnrm.ptr, None, du2.ptr, du1.ptr, None, ddir.ptr,
                                  dpdf.ptr, None, None, None)
    synchronize()
    lat = []
    total = 0
    t0 = time.perf_counter()
    for n in sizes:
        a = time.perf_counter()
        lib.pglFieldSampleBatchDevice(field.h, int(n), dpos.ptr, dnrm.ptr, None, du2.ptr, du1.ptr, None, ddir.ptr,
                                      dpdf.ptr, None, None, None)
        synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
        total += int(n)
    t1 = time.perf_counter()
    pdf = dpdf.to_numpy(np.float32, nmax)
    return {"guided_samples_per_s": total / (t1 - t0), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "batches": int(n_batches), "queries": int(total),
            "valid_fraction": float(np.mean(pdf[:sizes[-1]] > 0)), "mode": "closed-loop, device-resident inputs"}


def serving_field():
    """Field of the query-serving workload: one Update with 200k surface records
    (the single-update generator configuration) and 200k volume records."""
    import openpgl_amd as pgl
    from openpgl_amd.synth import generate
    dev = pgl.Device()
    f = pgl.Field(dev, max_samples_per_leaf=32000, deterministic=True)
    recs = np.concatenate([generate(200000, seed=0), generate(200000, seed=1, flags=2)])
    f.update(pgl.SampleStorage(recs))
    return dev, f


def serving_poisson(field, offered_qps, duration_s=1.5, seed=0, pool=1 << 22, ring=1 << 23, slots=4):
    """Open-loop serving: requests of floor(exp(U[ln 1k, ln 256k])) queries arrive
    as a Poisson process at offered_qps / E[size] requests/s; each query is a
    position near a blob, a normal uniform on the sphere, 2 uniforms, no KNN
    random (u1 = -1); 30% are volume queries with g uniform in [-0.9, 0.9].
    Inputs are slices of a pinned pool, outputs go to a pinned ring.
    Latency = scheduled Poisson arrival -> results in host memory, per
    request (any lag of the submitting thread behind the schedule counts)."""
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import pinned_copy, pinned_empty
    from openpgl_amd.synth import queries
    rng = np.random.default_rng(seed)
    pos, nrm, u2, is_vol, g = queries(pool, seed=seed + 11)
    pos, nrm, u2, g = (pinned_copy(a) for a in (pos, nrm, u2, g))
    vol = pinned_copy(is_vol.astype(np.uint8))
    odir = pinned_empty((ring, 3), np.float32)
    opdf = pinned_empty(ring, np.float32)
    oid = pinned_empty(ring, np.uint32)
    oint = pinned_empty(ring, np.float32)
    mean_size = (256000 - 1000) / np.log(256.0)
    rate = offered_qps / mean_size
    count = max(8, int(duration_s * rate))
    sizes = np.floor(np.exp(rng.uniform(np.log(1000), np.log(256000), count))).astype(np.int64)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    starts = rng.integers(0, pool - sizes)
    server = pgl.Server(field, memory_budget=1 << 30, slots=slots)
    # warm-up request (arena allocation)
    rid, _ = server.submit(pos[:256000], nrm[:256000], u2[:256000], g=g[:256000], is_volume=vol[:256000],
                           out=(odir[:256000], opdf[:256000], oid[:256000], oint[:256000]))
    server.wait(rid)
    server.release(rid)
    ids, live, lag = [], [], []
    rpos = 0
    t0 = time.perf_counter()
    for i in range(count):
        n = int(sizes[i])
        if rpos + n > ring:
            rpos = 0
        # the ring slice must not be in use by an unfinished request (slices are
        # taken in order, so the oldest live request is the next one to overlap)
        while live and live[0][0] < rpos + n and live[0][1] > rpos:
            server.wait(live[0][2])
            live.pop(0)
        target = t0 + arrivals[i]
        while True:
            now = time.perf_counter()
            if now >= target:
                break
            if target - now > 2e-4:
                time.sleep((target - now) - 1e-4)
        a = int(starts[i])
        sl = slice(a, a + n)
        o = slice(rpos, rpos + n)
        lag.append(time.perf_counter() - target)
        rid, _ = server.submit(pos[sl], nrm[sl], u2[sl], g=g[sl], is_volume=vol[sl],
                               out=(odir[o], opdf[o], oid[o], oint[o]))
        ids.append(rid)
        live.append((rpos, rpos + n, rid))
        rpos += n
    server.drain()
    t = np.array([server.timing(r) for r in ids])
    lat_submit = t[:, 2] - t[:, 0]
    # from the scheduled (Poisson) arrival: host lag before the submit included
    lat = lat_submit + 1e3 * np.asarray(lag)
    wall = (t[:, 2].max() - t[:, 0].min()) * 1e-3
    st = server.stats()
    valid = float(np.mean(opdf[:min(ring, rpos)] > 0))
    server.close()
    return {"offered_queries_per_s": offered_qps, "achieved_guided_samples_per_s": float(sizes.sum() / wall),
            "requests": int(count), "queries": int(sizes.sum()), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
            "latency_from": "scheduled arrival",
            "p99_ms_from_submit": float(np.percentile(lat_submit, 99)),
            "mean_batch_queries": st["queries"] / max(1.0, st["batches"]), "batches": int(st["batches"]),
            "peak_pending_requests": int(st["peak_pending"]), "valid_fraction": valid,
            "mode": "open loop, Poisson arrivals, continuous batching, host pinned in/out, %d slots" % slots}


def _device_mem_used():
    import ctypes
    from openpgl_amd.hip_runtime import hip
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total))
    return total.value - free.value


def mixed_workload(frames=9, rounds=1, surf=16_000_000, vol=4_000_000, offered_qps=1.0e8, server_budget=1 << 30):
    """Large mixed workload: 16M surface + 4M volume records per frame (64 blobs),
    frame 0 builds the field; frames 1.. (`rounds` times over, so that the
    serving window holds several hundred requests) are timed (a) alone and
    (b) while the query server runs an open Poisson load of offered_qps on the
    previous update's field (refreshed after every update).  The region kernel
    keeps one workgroup per CU in (b) so that query waves fit beside it."""
    import threading
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray, pinned_copy, pinned_empty, synchronize
    from openpgl_amd.synth import generate_large, queries
    t0 = time.perf_counter()
    fs = [generate_large(surf, seed=0, frame=f, num_blobs=64) for f in range(frames)]
    fv = [generate_large(vol, seed=1, frame=f, num_blobs=64, flags=2) for f in range(frames)]
    log("mixed workload: %d frames of %d + %d records generated in %.1f s" % (frames, surf, vol,
                                                                                time.perf_counter() - t0))
    dev = pgl.Device()
    mem0 = _device_mem_used()
    ds = [DeviceArray.from_numpy(a) for a in fs]
    dv = [DeviceArray.from_numpy(a) for a in fv]
    del fs, fv
    synchronize()
    mem_inputs = _device_mem_used() - mem0
    out = {"workload": "%d surface + %d volume records/frame, 64 blobs, %d frames (frame 0 builds; frames 1-%d "
                       "timed%s)" % (surf, vol, frames, frames - 1, "" if rounds == 1 else " %d times over" % rounds)}
    # (a) alone
    fa = pgl.Field(dev, max_samples_per_leaf=32000, deterministic=True)
    fa.update_surface_device(ds[0].value, surf)
    fa.update_volume_device(dv[0].value, vol)
    synchronize()
    t = time.perf_counter()
    for _ in range(rounds):
        for f in range(1, frames):
            fa.update_surface_device(ds[f].value, surf)
            fa.update_volume_device(dv[f].value, vol)
    synchronize()
    alone_s = time.perf_counter() - t
    mem_field = _device_mem_used() - mem0 - mem_inputs
    regions = (fa.num_regions(0), fa.num_regions(1))
    del fa
    # (b) with serving
    fb = pgl.Field(dev, max_samples_per_leaf=32000, deterministic=True)
    fb.set_region_schedule(grid=256)
    fb.update_surface_device(ds[0].value, surf)
    fb.update_volume_device(dv[0].value, vol)
    synchronize()
    server = pgl.Server(fb, memory_budget=server_budget, slots=4)
    pool = 1 << 22
    pos, nrm, u2, is_vol, g = queries(pool, seed=5, num_blobs=64)
    pos, nrm, u2, g = (pinned_copy(a) for a in (pos, nrm, u2, g))
    volf = pinned_copy(is_vol.astype(np.uint8))
    ring = 1 << 23
    odir, opdf = pinned_empty((ring, 3), np.float32), pinned_empty(ring, np.float32)
    oid, oint = pinned_empty(ring, np.uint32), pinned_empty(ring, np.float32)
    # warm-up request (arena allocation, first launches), as in the serving-alone run
    rid, _ = server.submit(pos[:256000], nrm[:256000], u2[:256000], g=g[:256000], is_volume=volf[:256000],
                           out=(odir[:256000], opdf[:256000], oid[:256000], oint[:256000]))
    server.wait(rid)
    server.release(rid)
    rng = np.random.default_rng(3)
    mean_size = (256000 - 1000) / np.log(256.0)
    rate = offered_qps / mean_size
    stop = threading.Event()
    ids = []

    def load():
        rpos = 0
        live = []
        tn = time.perf_counter()
        while not stop.is_set():
            tn += rng.exponential(1.0 / rate)
            n = int(np.floor(np.exp(rng.uniform(np.log(1000), np.log(256000)))))
            if rpos + n > ring:
                rpos = 0
            while live and live[0][0] < rpos + n and live[0][1] > rpos:
                server.wait(live[0][2])
                live.pop(0)
            while time.perf_counter() < tn and not stop.is_set():
                time.sleep(1e-4)
            a = int(rng.integers(0, pool - n))
            lagMs = 1e3 * (time.perf_counter() - tn)
            rid, _ = server.submit(pos[a:a + n], nrm[a:a + n], u2[a:a + n], g=g[a:a + n], is_volume=volf[a:a + n],
                                   out=(odir[rpos:rpos + n], opdf[rpos:rpos + n], oid[rpos:rpos + n],
                                        oint[rpos:rpos + n]))
            ids.append((rid, n, lagMs))
            live.append((rpos, rpos + n, rid))
            rpos += n

    th = threading.Thread(target=load)
    th.start()
    time.sleep(0.05)
    t = time.perf_counter()
    for _ in range(rounds):
        for f in range(1, frames):
            fb.update_surface_device(ds[f].value, surf)
            fb.update_volume_device(dv[f].value, vol)
            server.refresh()
    synchronize()
    mixed_s = time.perf_counter() - t
    stop.set()
    th.join()
    server.drain()
    mem_peak = _device_mem_used() - mem0
    tm = np.array([server.timing(r) for r, _, _ in ids])
    lat = tm[:, 2] - tm[:, 0] + np.array([lg for _, _, lg in ids])  # from the scheduled arrival
    nq = sum(n for _, n, _ in ids)
    span = (tm[:, 2].max() - tm[:, 0].min()) * 1e-3
    # the load the Poisson generator actually produced (queries over the submission window)
    offered_span = (tm[:, 0].max() - tm[:, 0].min()) * 1e-3
    server.close()
    samples = rounds * (frames - 1) * (surf + vol)
    out.update({"regions_surface_volume": regions,
                "update_alone_samples_per_s": samples / alone_s,
                "update_while_serving_samples_per_s": samples / mixed_s,
                "serving_during_updates": {"offered_queries_per_s": offered_qps, "requests": len(ids),
                                           "generated_queries_per_s": nq / offered_span,
                                           "achieved_guided_samples_per_s": nq / span,
                                           "p50_ms": float(np.percentile(lat, 50)),
                                           "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
                                           "latency_from": "scheduled arrival"},
                "hbm_bytes": {"input_frames": int(mem_inputs), "field": int(mem_field),
                              "server_budget": int(server_budget), "peak_total": int(mem_peak)}})
    return out


def host_cpu():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def cpu_threads():
    """Worker threads of the CPU baselines: every core of this GPU's host
    share (16 per GPU on the measurement hosts, whatever nproc reports for the
    whole machine), or nproc where that is smaller."""
    return min(os.cpu_count() or 1, 16)


def cpu_baseline(frames_host, threads, timed_frames=3):
    """Oracle on a bounded sample: frame 0 (build, untimed), then the incremental
    updates of frames 1..timed_frames timed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from field_compare import OracleField
    from oracle_ctypes import default_config
    o = OracleField(default_config(32000))
    o.update(frames_host[0], threads)
    t0 = time.perf_counter()
    n = 0
    for f in range(1, timed_frames + 1):
        o.update(frames_host[f], threads)
        n += len(frames_host[f])
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "samples/s", "cores": threads, "kind": "port",
            "sample": "oracle (CPU restatement of Open PGL 0.6.0, std::thread over regions), incremental "
                      "Updates of frames 1-%d (%d samples) after the frame-0 build; %.2f s" % (timed_frames, n, dt),
            "host": host_cpu()}


def batched_em(num_regions=4096, cpu=True, threads=16):
    """Batched uneven EM: num_regions pre-partitioned regions, samples per region
    log-uniform in [64, 65536], initial K in {8, 16, 32}, one fit step in one
    persistent launch.  The autotuner re-runs the launch under every candidate
    schedule (results are schedule-independent); the plain per-region path
    (one launch of one workgroup per region) is timed the same way."""
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray, synchronize
    from openpgl_amd.synth import batched_regions
    t0 = time.perf_counter()
    recs, counts, stats7, init_k = batched_regions(num_regions, seed=0)
    log("batched EM: %d regions, %d samples generated in %.1f s" % (num_regions, len(recs), time.perf_counter() - t0))
    dev = pgl.Device()
    d = DeviceArray.from_numpy(recs)
    synchronize()
    batch = pgl.RegionBatch(dev, max_samples_per_leaf=32000, deterministic=True)
    batch.step_device(d.value, counts, stats7, init_k)
    first = batch.last_info()
    best, table = batch.autotune(None, reps=2)
    _, per_region = batch.autotune([dict(per_region_launches=1)], reps=1)
    _, per_region_teams = batch.autotune([dict(per_region_launches=2)], reps=2)
    # leave the tuned schedule selected for reporting
    batch.set_schedule(**table[best][0])
    best_ms = table[best][1]
    default_ms = [ms for sc, ms in table if sc == {"per_region_launches": 0, "grid": 0, "balance": 1.0,
                                                    "min_chunks_per_rank": 1, "max_team": 64, "cooperative": 0}]
    n = int(len(recs))
    flops = FLOPS_ESTEP * first["estep_evals"] + FLOPS_SPLIT * first["split_evals"] + FLOPS_DIST * first["dist_evals"]
    out = {"workload": "%d regions, samples log-U[64, 65536] (%d total), initial K in {8,16,32}, fit step"
                       % (num_regions, n),
           "samples": n, "batched_ms": best_ms, "batched_samples_per_s": n / (best_ms * 1e-3),
           "default_schedule_ms": default_ms[0] if default_ms else None,
           "per_region_path_ms": per_region_teams[0][1],
           "per_region_path": "one launch per region, a team of one workgroup per 1024-sample chunk (up to 64), "
                              "each region alone on the device",
           "speedup_vs_per_region": per_region_teams[0][1] / best_ms,
           "one_workgroup_per_region_launch_ms": per_region[0][1],
           "tuned_schedule": table[best][0],
           "autotune_top5": sorted([(ms, sc) for sc, ms in table], key=lambda x: x[0])[:5],
           "achieved_tflops": flops / (best_ms * 1e-3) / 1e12,
           "em_iterations": first["em_iterations"]}
    if cpu:
        out["cpu"] = batched_cpu_baseline(recs, counts, stats7, init_k, threads)
        out["speedup_vs_cpu"] = out["batched_samples_per_s"] / out["cpu"]["value"]
        ref = batched_cpu_baseline(recs, counts, stats7, init_k, threads, reference=True)
        if ref is not None:
            out["cpu_reference"] = ref
            out["speedup_vs_cpu_reference"] = out["batched_samples_per_s"] / ref["value"]
    return out


def batched_cpu_baseline(recs, counts, stats7, init_k, threads, sample_regions=256, reference=False):
    """Region fits on the CPU (threads via ctypes, GIL released) on every
    (num_regions / sample_regions)-th region: the oracle restatement, or with
    reference=True the reference's own directional code (its AVX2 CPU_8
    build in oracle/_ref: ParallaxAwareVonMisesFisherWeightedEMFactory with
    adaptive split-and-merge, one region per call; None where not built)."""
    import concurrent.futures
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ctypes import OracleRegion, RefRegion, default_config, reference_available
    if reference and not reference_available():
        return None
    offs = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    step = max(1, len(counts) // sample_regions)
    regions = list(range(0, len(counts), step))
    jobs = []
    for r in regions:
        cfg = default_config(32000)
        cfg.initK = int(init_k[r])
        jobs.append((RefRegion(cfg) if reference else OracleRegion(cfg), recs[offs[r]:offs[r + 1]], stats7[r]))
    n = int(sum(len(j[1]) for j in jobs))
    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(lambda j: j[0].step(j[1], j[2][0:3], j[2][3:6], float(j[2][6])), jobs))
    dt = time.perf_counter() - t0
    who = ("the reference's own directional code (CPU_8 AVX2 build)" if reference
           else "oracle (CPU restatement)")
    return {"value": n / dt, "unit": "samples/s", "cores": threads, "kind": "reference" if reference else "port",
            "sample": "%s region fits of %d of the regions (every %d-th, %d samples), %d threads, %.2f s"
                      % (who, len(regions), step, n, threads, dt),
            "host": host_cpu()}


def path_segments(num_paths=1_000_000, reps=5, cpu=True):
    """Batched path segments -> training samples (pglPathSegmentBatchPrepareSamplesDevice)
    on 1M recorded paths of up to 12 vertices, inputs resident in HBM; the CPU
    line is the oracle's per-path restatement on one core over 50k paths."""
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray
    from openpgl_amd.synth import random_paths
    seg, off = random_paths(num_paths, seed=21, max_len=12)
    log("path segments:", num_paths, "paths,", len(seg), "vertices")
    dseg, doff = DeviceArray.from_numpy(seg), DeviceArray.from_numpy(off)
    cap = len(seg)
    ds, dv = DeviceArray(40 * cap), DeviceArray(40 * cap)
    batch = pgl.PathSegmentBatch()
    r = batch.prepare_device(dseg.value, len(seg), doff.value, num_paths, ds.value, cap, dv.value, cap)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r = batch.prepare_device(dseg.value, len(seg), doff.value, num_paths, ds.value, cap, dv.value, cap)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    ns, nv = r
    # algorithmic bytes: every 136 B vertex read once, every 40 B record written once
    nbytes = 136 * len(seg) + 40 * (ns + nv)
    out = {"paths": num_paths, "vertices": int(len(seg)), "samples": int(ns + nv), "ms": dt * 1e3,
           "paths_per_s": num_paths / dt, "samples_per_s": (ns + nv) / dt,
           "hbm_gbs_algorithmic": nbytes / dt / 1e9, "hbm_frac": nbytes / dt / 1e9 / PEAK_HBM_GBS}
    if cpu:
        import ctypes
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_ctypes import oracle
        lib = oracle()
        fn = lib.orc_path_prepare_batch
        fn.restype = ctypes.c_size_t
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                       ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t]
        sample_paths = 200_000
        buf = np.zeros(int(off[sample_paths]) * 40 + 40, np.uint8)
        t0 = time.perf_counter()
        k = fn(seg.ctypes.data, off.ctypes.data, sample_paths, 1e6, 0, 0, 1, buf.ctypes.data, len(buf) // 40)
        ct = time.perf_counter() - t0
        out["cpu"] = {"value": sample_paths / ct, "unit": "paths/s", "cores": 1, "kind": "port",
                      "sample": "oracle prepareSamples over the first 200k of the same paths, one thread (%d samples)"
                                % k}
        out["speedup_vs_cpu"] = out["paths_per_s"] / out["cpu"]["value"]
    return out


def read_traffic():
    """HBM bytes per region-kernel launch from the committed rocprofv3 summary."""
    path = os.path.join(ROOT, "profiles", "k_region_update_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and not line.lstrip().startswith("#"):
            k, v = line.split(":", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except (ValueError, IndexError):
                pass
    if "traffic_bytes_per_launch" not in vals:
        return None
    return vals["traffic_bytes_per_launch"]


def calibrate_copy(nbytes=1 << 30, reps=5):
    """Plain device-to-device copy of known size (the FETCH_SIZE / WRITE_SIZE
    calibration kernel for rocprofv3 --pmc)."""
    from openpgl_amd.hip_runtime import DeviceArray, hip, synchronize
    a = DeviceArray(nbytes)
    b = DeviceArray(nbytes)
    for _ in range(reps):
        hip().hipMemcpy(b.ptr, a.ptr, nbytes, 3)
    synchronize()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--samples", type=int, default=N_PER_FRAME)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched uneven EM workload")
    ap.add_argument("--no-mixed", action="store_true", help="skip the large mixed update + serving workload")
    ap.add_argument("--no-paths", action="store_true", help="skip the path-segment workload")
    ap.add_argument("--no-render-loop", action="store_true", help="skip the synchronous render-loop line")
    ap.add_argument("--region-grid", type=int, default=None, help="region kernel workgroups (default: all co-resident)")
    ap.add_argument("--region-balance", type=float, default=None, help="region team balance (default 1)")
    ap.add_argument("--region-cooperative", type=int, default=None, help="1: cooperative region launch (default 0: ordinary launch)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (rocprofv3 --pmc calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    nframes = args.warmup + args.steps
    log("generating", nframes, "frames of", args.samples, "samples")
    frames = make_frames(args.samples, nframes)
    sched = {}
    if args.region_grid is not None:
        sched["grid"] = args.region_grid
    if args.region_balance is not None:
        sched["balance"] = args.region_balance
    if args.region_cooperative is not None:
        sched["cooperative"] = args.region_cooperative
    field, dev, dt, infos = run_gpu(frames, args.warmup, args.steps, sched or None)
    timed = sum(len(frames[args.warmup + s]) for s in range(args.steps))
    value = timed / dt
    # roofline of the region kernel over the timed frames
    kern_ms = sum(i["region_kernel_ms"] for i in infos)
    flops = sum(FLOPS_ESTEP * i["estep_evals"] + FLOPS_SPLIT * i["split_evals"] + FLOPS_DIST * i["dist_evals"]
                for i in infos)
    achieved = flops / (kern_ms * 1e-3) / 1e12
    per_launch = {
        "avg_region_kernel_ms": kern_ms / len(infos),
        "algorithmic_flops_per_launch": flops / len(infos),
        "spatial_ms_avg": float(np.mean([i["spatial_ms"] for i in infos])),
        "sort_ms_avg": float(np.mean([i["sort_ms"] for i in infos])),
        "gather_ms_avg": float(np.mean([i["gather_ms"] for i in infos])),
        "knn_ms_avg": float(np.mean([i["knn_ms"] for i in infos])),
        "wall_ms_avg": float(np.mean([i["wall_ms"] for i in infos])),
        "regions_last": int(infos[-1]["regions"]),
        "directional_ms_avg": float(np.mean([i["directional_ms"] for i in infos])),
    }
    out = {
        "metric": "Field::Update samples/s; guided samples/s + p50/p99 at fixed PDF error",
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (16 drifting Gaussian blobs, 4-lobe vMF directions, seed 0)",
        "config": {"workload": "8 frames x 1M surface samples, incremental Field::Update, blobs drifting 0.01/frame, "
                               "split-and-merge on, maxK 32, kd max depth 32, maxSamplesPerLeaf 32000",
                   "samples_per_frame": args.samples, "frames_timed": args.steps, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": read_traffic(), "kernel": "k_region_update",
                     **per_launch},
    }
    if not args.no_render_loop:
        out["render_loop"] = render_loop(frames, args.warmup, args.steps)
    if not args.no_serving:
        out["serving"] = serving(field)
        sdev, sfield = serving_field()
        out["serving_poisson"] = [serving_poisson(sfield, q) for q in (1.0e8, 2.0e8)]
    if not args.no_batched:
        out["batched_em"] = batched_em(cpu=not args.no_cpu, threads=cpu_threads())
    if not args.no_mixed:
        out["mixed_update_and_serving"] = mixed_workload()
    if not args.no_paths:
        out["path_segments"] = path_segments(cpu=not args.no_cpu)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(frames, cpu_threads(), timed_frames=min(3, nframes - 1))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
