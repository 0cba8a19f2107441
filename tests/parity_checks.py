"""Checks shared by the GPU parity tests: bit-exact spatial structure and the
per-region mixture comparison against the oracle.  Test infrastructure only.

Mixture bars (used by every field and batch parity test):
  * median over regions of the per-region median relative pdf error
    (512 directions) <= MIX_MEDIAN;
  * at least MIX_FRAC_REGIONS of the regions with their 99th percentile
    below MIX_P99;
  * at least K_EQUAL_FRAC of the regions end with the same number of
    components as the oracle.  A region whose EM stopping iteration, split
    or merge sits exactly at its threshold can take the other discrete
    decision under a different (compensated, parallel) summation order.
"""
import numpy as np

from field_compare import vmm_pdf

MIX_MEDIAN = 1e-4
MIX_P99 = 1e-2
MIX_FRAC_REGIONS = 0.99
K_EQUAL_FRAC = 0.99


def assert_spatial_bit_exact(gpu, orc, fresh_positions=None, knn=True):
    """kd tree (canonical order), region sample statistics and bounds,
    samples per region of the last update, leaf ids of fresh positions and
    KNN neighbour lists identical to the oracle's."""
    gp, gd, gc = gpu.nodes()
    op, od, oc = orc.nodes()
    assert len(gp) == len(op), (len(gp), len(op))
    assert np.array_equal(gd, od)
    assert np.array_equal(gc, oc)
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32))
    R = gpu.num_regions()
    assert R == orc.num_regions()
    for i in range(R):
        a, b = gpu.region(i), orc.region(i)
        for key in ["ssMean", "ssVariance", "ssLower", "ssUpper", "regionLower", "regionUpper"]:
            assert np.array_equal(a[key].view(np.uint32), b[key].view(np.uint32)), (i, key, a[key], b[key])
        assert np.float32(a["ssNumSamples"]) == np.float32(b["ssNumSamples"]), i
        assert a["numSamplesThisUpdate"] == b["numSamplesThisUpdate"], i
    if fresh_positions is not None:
        assert np.array_equal(gpu.leaf_ids(fresh_positions), orc.leaf_ids(fresh_positions))
    if knn:
        assert np.array_equal(gpu.knn(), orc.knn())
    return R


def mixture_errors(gpu_records, orc_records, dirs):
    """per valid region: (median rel pdf error, 99th percentile, K gpu, K oracle)"""
    errs = []
    for i, (a, b) in enumerate(zip(gpu_records, orc_records)):
        assert a["valid"] == b["valid"], i
        if not b["valid"]:
            continue
        pa, pb = vmm_pdf(a, dirs), vmm_pdf(b, dirs)
        rel = np.abs(pa - pb) / pb
        errs.append((np.median(rel), np.percentile(rel, 99), a["K"], b["K"]))
    return np.array(errs)


def mixture_summary(e):
    return {"regions": int(len(e)), "median_of_medians": float(np.median(e[:, 0])),
            "max_median": float(e[:, 0].max()), "frac_p99_below": float(np.mean(e[:, 1] < MIX_P99)),
            "worst_p99": float(e[:, 1].max()), "K_equal": float(np.mean(e[:, 2] == e[:, 3]))}


def assert_mixture_bars(e, label=""):
    s = mixture_summary(e)
    print(label, s, flush=True)
    assert s["median_of_medians"] <= MIX_MEDIAN, (label, s)
    assert s["frac_p99_below"] >= MIX_FRAC_REGIONS, (label, s)
    assert s["K_equal"] >= K_EQUAL_FRAC, (label, s)
    return s
