"""Field-file layout (pglFieldStoreToFile / pglDeviceNewFieldFromFile) on the CPU.

The Python restatement of the format (tests/field_file.py) is pinned against
the reference's own serialisation code compiled in oracle/_ref: a storage
entry (Region + Range) trained and written by the reference for SIMD widths
4 and 8 must parse into exactly the reference's state and re-serialise to the
same bytes, and bytes written by the restatement must deserialise, through
the reference's code, into the state they were written from.  The factory
configuration block is pinned the same way.  Field-level members around the
storage entries (Field.h:240-267) are restated member by member from the
source; their order is exercised end to end by the GPU tests.
"""
import ctypes

import numpy as np
import pytest

import field_file as ff
from oracle_ctypes import (RefFileRegion, default_config, ref_config_bytes, reference_available)
from openpgl_amd.synth import generate

needs_ref = pytest.mark.skipif(not reference_available(), reason="oracle/_ref not built (reference tree absent)")

RECORD_KEYS = ["K", "weights", "kappas", "meanDirections", "normalizations", "eMinus2Kappa", "meanCosines",
               "distances", "pivot", "sumOfWeightedDirections", "sumOfWeightedStats", "sumOfDistanceWeights",
               "sumWeights", "numSamples", "overallNumSamples", "suffK", "normalized", "chiSquareMCEstimates",
               "splitMeans", "splitCovariances", "splitNumSamples", "splitSumWeights", "splitSumAssignedSamples",
               "splitK", "numSamplesAfterLastSplit", "numSamplesAfterLastMerge", "ssMean", "ssVariance",
               "ssNumSamples", "ssLower", "ssUpper", "regionLower", "regionUpper", "splitFlag", "valid"]


def _same(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype.kind == "f" or b.dtype.kind == "f":
        return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))
    return np.array_equal(a.astype(np.int64), b.astype(np.int64))


def _stats13(s):
    p = np.stack([s["pos"]["x"], s["pos"]["y"], s["pos"]["z"]], 1) if s["pos"].dtype.names else s["pos"]
    p = np.asarray(p, np.float64)
    m = p.mean(0)
    v = ((p - m) ** 2).sum(0)
    return np.concatenate([m, v, [len(p)], p.min(0), p.max(0)]).astype(np.float32)


def _trained_ref_region(W, steps=2, seed=3):
    cfg = default_config(4000)
    r = RefFileRegion(cfg, width=W)
    for f in range(steps):
        s = generate(3000, seed=seed, frame=f, num_blobs=2)
        ss = _stats13(s)
        r.step(s, ss, np.array([-1, -1, -1, 1, 1, 1], np.float32) * 2.0, split_flag=int(f == 1))
    return r


@needs_ref
@pytest.mark.parametrize("W", [4, 8])
def test_region_entry_matches_reference_serializer(W):
    r = _trained_ref_region(W)
    data = r.serialize()
    assert len(data) == ff.REGION_BYTES
    inp = ff._In(data)
    parsed = ff.read_region(inp, W)
    assert inp.o == len(data)
    ref = r.record()
    assert ref["K"] >= 1 and ref["valid"] == 1
    for k in RECORD_KEYS:
        assert _same(parsed[k], ref[k]), (W, k, parsed[k], ref[k])
    assert parsed["rangeBegin"] == 0 and parsed["rangeEnd"] == 0
    out = ff._Out()
    ff.write_region(out, parsed, W)
    assert out.data() == data


@needs_ref
@pytest.mark.parametrize("W", [4, 8])
def test_region_entry_written_by_restatement_reads_back_in_reference(W):
    r = _trained_ref_region(W, seed=5)
    parsed = ff.read_region(ff._In(r.serialize()), W)
    # distinct values in every slot, so that a transposed block or lane shows
    rng = np.random.default_rng(W)
    for k in ["weights", "kappas", "meanDirections", "normalizations", "eMinus2Kappa", "meanCosines", "distances",
              "sumOfWeightedDirections", "sumOfWeightedStats", "sumOfDistanceWeights", "chiSquareMCEstimates",
              "splitCovariances", "splitNumSamples", "splitSumWeights", "splitSumAssignedSamples"]:
        parsed[k] = rng.random(parsed[k].shape).astype(np.float32)
    parsed["splitMeans"] = rng.random((32, 2)).astype(np.float32)
    parsed["K"] = 27
    parsed["numSamplesAfterLastSplit"] = (1 << 33) + 5
    out = ff._Out()
    ff.write_region(out, parsed, W)
    back = RefFileRegion(width=W, data=out.data())
    rec = back.record()
    for k in RECORD_KEYS:
        assert _same(rec[k], parsed[k]), (W, k)


@needs_ref
def test_widths_differ_only_in_block_interleave():
    r = _trained_ref_region(8, seed=9)
    parsed = ff.read_region(ff._In(r.serialize()), 8)
    o4, o8 = ff._Out(), ff._Out()
    ff.write_region(o4, parsed, 4)
    ff.write_region(o8, parsed, 8)
    assert len(o4.data()) == len(o8.data()) == ff.REGION_BYTES
    assert o4.data() != o8.data()
    back = ff.read_region(ff._In(o4.data()), 4)
    for k in RECORD_KEYS:
        assert _same(back[k], parsed[k]), k


@needs_ref
def test_factory_configuration_block_matches_reference():
    cfg = default_config(12000)
    cfg.initK = 11
    cfg.maxEMIterations = 77
    cfg.minSamplesForSplitting = 123
    cfg.minSamplesForMerging = 456
    cfg.minSamplesForPartialRefitting = 789
    cfg.partialReFit = 1
    data = ref_config_bytes(cfg)
    inp = ff._In(data)
    c = ff.read_config(inp)
    assert inp.o == len(data)
    assert c["initK"] == 11 and c["maxEMIterrations"] == 77 and c["maxK"] == cfg.maxK
    assert c["minSamplesForSplitting"] == 123 and c["minSamplesForMerging"] == 456
    assert c["minSamplesForPartialRefitting"] == 789 and c["partialReFit"] == 1 and c["maxSplitItr"] == 1
    assert np.float32(c["maxKappa"]) == np.float32(cfg.maxKappa)
    assert np.float32(c["splittingThreshold"]) == np.float32(cfg.splittingThreshold)
    out = ff._Out()
    ff.write_config(out, c)
    assert out.data() == data


def test_restated_file_round_trip():
    cfg = {k: (1 if f in "BQi" else 0.5) for k, f in ff.CONFIG_FIELDS}
    surf = ff.empty_field(True, cfg, 100, 32000, 32)
    vol = ff.empty_field(False, cfg, 100, 32000, 32, knn=False)
    surf.update({"iteration": 2, "initialized": 1, "treeInit": 1, "treeBounds": (-1, -2, -3, 1, 2, 3),
                 "nodes": np.array([(0.5, (0 << 30) | 1), (0.0, (3 << 30) | 0), (0.0, (3 << 30) | 1)],
                                   dtype=[("pivot", "<f4"), ("bits", "<u4")]), "knnBuilt": 1})
    rng = np.random.default_rng(0)
    regs = []
    for i in range(2):
        r = {"valid": 1, "K": 3, "suffK": 3, "splitK": 3, "normalized": 1, "splitFlag": 0,
             "numSamplesAfterLastSplit": 10, "numSamplesAfterLastMerge": 20, "sumWeights": 1.0, "numSamples": 2.0,
             "overallNumSamples": 3.0, "ssNumSamples": 4.0}
        for name, _ in ff.VMM_ARRAYS + ff.SUFF_ARRAYS + ff.SPLIT_ARRAYS:
            r[name] = rng.random(ff.ARRAY_SHAPES.get(name, (32,))).astype(np.float32)
        for name in ["pivot", "regionLower", "regionUpper", "ssMean", "ssVariance", "ssLower", "ssUpper"]:
            r[name] = rng.random(3).astype(np.float32)
        regs.append(r)
    surf["regions"] = regs
    surf["knnPoints"] = np.array([[*regs[0]["ssMean"], 0], [*regs[1]["ssMean"], 0]], np.float32)
    d = {"iteration": 2, "surface": surf, "volume": vol}
    for W in (4, 8, 16):
        data = ff.write_file(d, W)
        back = ff.read_file(data, W)
        assert back["header"] == ff.HEADER
        assert ff.write_file(back, W) == data
        assert len(back["surface"]["regions"]) == 2 and back["volume"]["regions"] == []
        for k in RECORD_KEYS[:-2]:
            if k in regs[1]:
                assert _same(back["surface"]["regions"][1][k], regs[1][k]), k


def test_file_entry_points_reject_null_arguments():
    import openpgl_amd as pgl
    lib = pgl.load_library()
    assert not lib.pglDeviceNewFieldFromFile(None, b"/nonexistent")
    assert not lib.pglFieldStoreToFile(None, b"/tmp/x")
