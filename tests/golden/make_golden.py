"""Writes the golden fixtures in tests/golden/ from the reference's own code
compiled in oracle/_ref (`make -C oracle ref`).  Run in a container that has
the reference tree; the fixtures are data (inputs and the reference's outputs):
  hg_table.npy       128x18 HG phase-function representation table
  int_stats.npz      IntegerSampleStatistics of point sets (raw int64 + float)
  stats_ops.npz      SampleStatistics::merge / split results
  rcp.npz            embree::rcp (rcpss + Newton) of this host
  region_kat.npz     per-region fit/update trajectories (samples + outputs)
  query_kat.npz      init + cosine/HG product + sample + pdf on a trained region
  region_kat_np.npz  the same trajectories and queries for the plain VMM
                     field type (no parallax compensation)
Usage: make_golden.py [fixture-name ...]   (default: all)
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle_ctypes import RefRegion, default_config, ptr, reference, region_sample_stats  # noqa: E402
from openpgl_amd.synth import region_samples  # noqa: E402


def hg_table(ref):
    t = np.zeros(128 * 18, np.float32)
    ref.ref_hg_table(ptr(t))
    return t.reshape(128, 18)


def plain_vmm_kat():
    """Trajectories and queries of the VMM field type (cfg.parallax = 0)."""
    cfg = default_config(32000)
    cfg.minSamplesForSplitting = 1000
    cfg.minSamplesForPartialRefitting = 1000
    cfg.minSamplesForMerging = 2000
    cfg.parallax = 0
    dirs = np.random.default_rng(98).normal(size=(256, 3))
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    kat = {"dirs": dirs}
    for reg in range(2):
        rr = RefRegion(cfg)
        for step in range(3):
            s = region_samples(2500, seed=300 + 100 * reg + step)
            m, v, n = region_sample_stats(s)
            split = 1 if step == 2 else 0
            rr.step(s, m, v, n, split)
            rec = rr.record()
            key = "r%d_s%d_" % (reg, step)
            kat[key + "samples"] = s
            kat[key + "ss"] = np.concatenate([m, v, [n]]).astype(np.float32)
            kat[key + "split"] = np.int32(split)
            for f in ["K", "weights", "kappas", "meanDirections", "distances", "pivot", "numSplits", "numMerges",
                      "emIterations", "valid"]:
                kat[key + f] = np.asarray(rec[f])
            kat[key + "pdf"] = np.array([rr.pdf(d) for d in dirs], np.float32)
        qrng = np.random.default_rng(7 + reg)
        qp = (m + qrng.normal(scale=0.5, size=(100, 3))).astype(np.float32)
        qn = qrng.normal(size=(100, 3))
        qn = (qn / np.linalg.norm(qn, axis=1, keepdims=True)).astype(np.float32)
        qu = qrng.random((100, 2)).astype(np.float32)
        qg = qrng.uniform(-0.9, 0.9, 100).astype(np.float32)
        kat["q%d_pos" % reg], kat["q%d_nrm" % reg], kat["q%d_u2" % reg], kat["q%d_g" % reg] = qp, qn, qu, qg
        for vol in (0, 1):
            pa = np.zeros(100, np.float32); d = np.zeros((100, 3), np.float32); p = np.zeros(100, np.float32)
            for i in range(100):
                _, pa[i], d[i], p[i] = rr.query(vol, qp[i], qn[i], qg[i], qu[i], dirs[i])
            kat["q%d_pdf_at_dir_%d" % (reg, vol)] = pa
            kat["q%d_dir_%d" % (reg, vol)] = d
            kat["q%d_pdf_%d" % (reg, vol)] = p
    np.savez_compressed(os.path.join(HERE, "region_kat_np.npz"), **kat)


def main():
    which = sys.argv[1:]
    if which == ["region_kat_np"]:
        plain_vmm_kat()
        return
    ref = reference()
    np.save(os.path.join(HERE, "hg_table.npy"), hg_table(ref))

    rng = np.random.default_rng(1234)
    sets, bounds, raw, flt = [], [], [], []
    for i in range(24):
        n = int(rng.integers(1, 400))
        lo = rng.uniform(-2, 0.5, 3).astype(np.float32)
        hi = (lo + rng.uniform(0.01, 3, 3)).astype(np.float32)
        p = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
        if i % 5 == 0:
            p[:3] = lo  # exact corner points
        r = np.zeros(13, np.int64)
        f = np.zeros(13, np.float32)
        ref.ref_integer_stats(ptr(p), n, ptr(lo), ptr(hi), ptr(f), ptr(r))
        sets.append(np.pad(p, ((0, 400 - n), (0, 0))))
        bounds.append(np.concatenate([lo, hi, [n]]).astype(np.float32))
        raw.append(r)
        flt.append(f)
    np.savez(os.path.join(HERE, "int_stats.npz"), points=np.array(sets), bounds=np.array(bounds),
             raw=np.array(raw), stats=np.array(flt))

    a_all, b_all, dims, poss, m_all, l_all, r_all = [], [], [], [], [], [], []
    for i in range(32):
        def rand_stats():
            mean = rng.uniform(0, 1, 3)
            var = rng.uniform(0.0, 0.2, 3) * rng.uniform(1, 500)
            n = rng.uniform(1, 5000)
            lo = mean - rng.uniform(0.01, 0.5, 3)
            hi = mean + rng.uniform(0.01, 0.5, 3)
            return np.concatenate([mean, var, [n], lo, hi]).astype(np.float32)
        a, b = rand_stats(), rand_stats()
        dim = int(rng.integers(0, 3))
        pos = np.float32(a[dim] + rng.uniform(-0.05, 0.05))
        m = np.zeros(13, np.float32); l = np.zeros(13, np.float32); r = np.zeros(13, np.float32)
        ref.ref_stats_merge_split(ptr(a), ptr(b), dim, float(pos), ptr(m), ptr(l), ptr(r))
        a_all.append(a); b_all.append(b); dims.append(dim); poss.append(pos); m_all.append(m); l_all.append(l); r_all.append(r)
    np.savez(os.path.join(HERE, "stats_ops.npz"), a=np.array(a_all), b=np.array(b_all), dim=np.array(dims),
             pos=np.array(poss, np.float32), merged=np.array(m_all), left=np.array(l_all), right=np.array(r_all))

    xs = np.concatenate([rng.uniform(1e-3, 1e3, 4000), rng.uniform(0.01, 3, 1000)]).astype(np.float32)
    ys = np.array([ref.ref_rcp(float(x)) for x in xs], np.float32)
    np.savez(os.path.join(HERE, "rcp.npz"), x=xs, y=ys)

    # region trajectories: fit + 2 updates (second with a spatial-split decay)
    cfg = default_config(32000)
    cfg.minSamplesForSplitting = 1000
    cfg.minSamplesForPartialRefitting = 1000
    cfg.minSamplesForMerging = 2000
    dirs = np.random.default_rng(99).normal(size=(256, 3))
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    kat = {"dirs": dirs}
    steps = []
    for reg in range(2):
        rr = RefRegion(cfg)
        for step in range(3):
            s = region_samples(2500, seed=100 * reg + step)
            m, v, n = region_sample_stats(s)
            split = 1 if step == 2 else 0
            rr.step(s, m, v, n, split)
            rec = rr.record()
            key = "r%d_s%d_" % (reg, step)
            kat[key + "samples"] = s
            kat[key + "ss"] = np.concatenate([m, v, [n]]).astype(np.float32)
            kat[key + "split"] = np.int32(split)
            for f in ["K", "weights", "kappas", "meanDirections", "meanCosines", "distances", "pivot",
                      "sumOfWeightedStats", "sumWeights", "numSamples", "overallNumSamples", "chiSquareMCEstimates",
                      "splitNumSamples", "numSplits", "numMerges", "emIterations", "partialEmIterations", "valid"]:
                kat[key + f] = np.asarray(rec[f])
            kat[key + "pdf"] = np.array([rr.pdf(d) for d in dirs], np.float32)
            steps.append(key)
        if reg == 0:
            qrng = np.random.default_rng(5)
            qp = (m + qrng.normal(scale=0.05, size=(200, 3))).astype(np.float32)
            qn = qrng.normal(size=(200, 3)); qn = (qn / np.linalg.norm(qn, axis=1, keepdims=True)).astype(np.float32)
            qu = qrng.random((200, 2)).astype(np.float32)
            qg = qrng.uniform(-0.9, 0.9, 200).astype(np.float32)
            out = {"pos": qp, "nrm": qn, "u2": qu, "g": qg, "pdf_dir": dirs[:200]}
            for vol in (0, 1):
                integ = np.zeros(200, np.float32); pa = np.zeros(200, np.float32)
                d = np.zeros((200, 3), np.float32); p = np.zeros(200, np.float32)
                for i in range(200):
                    integ[i], pa[i], d[i], p[i] = rr.query(vol, qp[i], qn[i], qg[i], qu[i], dirs[i])
                out["integral_%d" % vol] = integ; out["pdf_at_dir_%d" % vol] = pa
                out["dir_%d" % vol] = d; out["pdf_%d" % vol] = p
            for k, v in list(rr.record().items()):
                if k in ("K", "weights", "kappas", "meanDirections", "normalizations", "eMinus2Kappa", "meanCosines",
                         "distances", "pivot"):
                    out["vmm_" + k] = np.asarray(v)
            np.savez_compressed(os.path.join(HERE, "query_kat.npz"), **out)
    kat["steps"] = np.array(steps)
    np.savez_compressed(os.path.join(HERE, "region_kat.npz"), **kat)
    plain_vmm_kat()
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
