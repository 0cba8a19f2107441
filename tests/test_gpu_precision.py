"""Error of the GPU's FP32 training (FP64 only in mean-cosine -> kappa and the
log-likelihood test, Kahan-compensated cross-chunk sums) against the FP64
oracle, next to the error of the reference's own FP32 arithmetic (FP32
oracle) against the same FP64 oracle.  Printed figures are the ones the
design notes quote.

Bars: the GPU's median pdf relative error vs FP64 <= 1e-3 (the pdf
tolerance), and not worse than 10x the FP32 oracle's median error plus 1e-5.
"""
import json

import numpy as np
import pytest

from field_compare import OracleField, precision_report, unit_dirs
from oracle_ctypes import OracleRegion, default_config
from openpgl_amd.synth import batched_regions, generate

pytestmark = pytest.mark.gpu


def test_field_error_against_fp64_oracle():
    import openpgl_amd as pgl
    dev = pgl.Device()
    gpu = pgl.Field(dev, max_samples_per_leaf=4000, deterministic=True)
    cfg = default_config(4000)
    o32, o64 = OracleField(cfg, 0), OracleField(cfg, 1)
    for f in range(2):
        s = generate(100000, seed=0, frame=f)
        gpu.update(pgl.SampleStorage(s))
        o32.update(s)
        o64.update(s)
    R = gpu.num_regions()
    assert R == o64.num_regions()
    dirs = unit_dirs(512)
    g = [gpu.region(i) for i in range(R)]
    r32 = [o32.region(i) for i in range(R)]
    r64 = [o64.region(i) for i in range(R)]
    rep_gpu = precision_report(g, r64, dirs)
    rep_o32 = precision_report(r32, r64, dirs)
    print("field GPU vs FP64:", json.dumps(rep_gpu))
    print("field FP32 oracle vs FP64:", json.dumps(rep_o32))
    assert rep_gpu["pdf_rel_median"] <= 1e-3
    assert rep_gpu["pdf_rel_median"] <= 10 * rep_o32["pdf_rel_median"] + 1e-5


def test_batched_regions_error_against_fp64_oracle():
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray
    dev = pgl.Device()
    recs, counts, stats, init_k = batched_regions(48, seed=11, nmin=64, nmax=16000)
    batch = pgl.RegionBatch(dev)
    d = DeviceArray.from_numpy(recs)
    batch.step_device(d.value, counts, stats, init_k)
    offs = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    r32, r64 = [], []
    for r in range(len(counts)):
        cfg = default_config(32000)
        cfg.initK = int(init_k[r])
        for prec, out in ((0, r32), (1, r64)):
            o = OracleRegion(cfg, prec)
            o.step(recs[offs[r]:offs[r + 1]], stats[r, 0:3], stats[r, 3:6], float(stats[r, 6]))
            out.append(o.record())
    g = [batch.region(r) for r in range(len(counts))]
    dirs = unit_dirs(512)
    rep_gpu = precision_report(g, r64, dirs)
    rep_o32 = precision_report(r32, r64, dirs)
    print("batched GPU vs FP64:", json.dumps(rep_gpu))
    print("batched FP32 oracle vs FP64:", json.dumps(rep_o32))
    assert rep_gpu["pdf_rel_median"] <= 1e-3
    assert rep_gpu["pdf_rel_median"] <= 10 * rep_o32["pdf_rel_median"] + 1e-5
