"""GPU parity of Field::Update and the query path against the oracle.

Tolerances:
  * kd tree, leaf ids, region sample statistics, KNN lists: bit-exact.
  * mixtures: three-way against the FP32 and FP64 oracles
    (tests/parity_checks.py).
  * queries end to end (GPU-trained field against the oracle-trained field):
    region ids exact, pdfs three-way as the mixtures (the query path alone
    is pinned to 1e-3 for every query in tests/test_gpu_configs.py,
    tests/test_gpu_field_file.py and tests/test_gpu_scalar_api.py).
"""
import numpy as np
import pytest

from field_compare import OracleField, unit_dirs, vmm_pdf
from oracle_ctypes import default_config, load_hg_table_into_oracle
from parity_checks import assert_mixture_parity, assert_query_pdfs, query_pdf_errors
from openpgl_amd.synth import generate, queries

pytestmark = pytest.mark.gpu


def _fields(max_samples, n, frames=1, seed=0, knn=True):
    import openpgl_amd as pgl
    dev = pgl.Device()
    gpu = pgl.Field(dev, max_samples_per_leaf=max_samples, deterministic=True, knn=knn)
    cfg = default_config(max_samples)
    cfg.knnLookup = 1 if knn else 0
    orc = OracleField(cfg)
    for f in range(frames):
        s = generate(n, seed=seed, frame=f)
        print("frame", f, "samples", n, flush=True)
        gpu.update(pgl.SampleStorage(s))
        orc.update(s)
    return dev, gpu, orc


def _tree_equal(gpu, orc):
    gp, gd, gc = gpu.nodes()
    op, od, oc = orc.nodes()
    assert len(gp) == len(op), (len(gp), len(op))
    assert np.array_equal(gd, od)
    assert np.array_equal(gc, oc)
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32))


@pytest.mark.parametrize("frames", [1, 3])
def test_kdtree_bit_exact(frames):
    dev, gpu, orc = _fields(2000, 40000, frames=frames)
    _tree_equal(gpu, orc)
    R = gpu.num_regions()
    assert R == orc.num_regions() and R > 4
    for i in range(R):
        a, b = gpu.region(i), orc.region(i)
        for key in ["ssMean", "ssVariance", "ssLower", "ssUpper", "regionLower", "regionUpper"]:
            assert np.array_equal(a[key].view(np.uint32), b[key].view(np.uint32)), (i, key, a[key], b[key])
        assert np.float32(a["ssNumSamples"]) == np.float32(b["ssNumSamples"])
        assert a["numSamplesThisUpdate"] == b["numSamplesThisUpdate"], (i, a["numSamplesThisUpdate"], b["numSamplesThisUpdate"])
    s = generate(5000, seed=9, frame=frames)
    assert np.array_equal(gpu.leaf_ids(s["pos"]), orc.leaf_ids(s["pos"]))
    assert np.array_equal(gpu.knn(), orc.knn())


def test_kdtree_bit_exact_large_tree():
    """A tree of more than 2048 nodes: the descending statistics pass walks
    it from global memory instead of LDS; the second frame descends it."""
    dev, gpu, orc = _fields(64, 80000, frames=2, seed=3, knn=False)
    gp, _, _ = gpu.nodes()
    print("nodes", len(gp), flush=True)
    assert len(gp) > 2048
    _tree_equal(gpu, orc)
    s = generate(5000, seed=11, frame=2)
    assert np.array_equal(gpu.leaf_ids(s["pos"]), orc.leaf_ids(s["pos"]))


def test_kdtree_bit_exact_beyond_lds_tables():
    """More than 4096 nodes: the push-down rounds read the node -> frontier
    table from global memory instead of LDS (and the descending pass walks the
    tree from global memory, the statistics run in several slot tiles); region
    statistics and leaf ids bit-exact over two frames."""
    dev, gpu, orc = _fields(64, 250000, frames=2, seed=5, knn=False)
    gp, _, _ = gpu.nodes()
    print("nodes", len(gp), flush=True)
    assert len(gp) > 4096
    _tree_equal(gpu, orc)
    R = gpu.num_regions()
    assert R == orc.num_regions()
    for i in range(0, R, 7):
        a, b = gpu.region(i), orc.region(i)
        for key in ["ssMean", "ssVariance", "ssLower", "ssUpper", "regionLower", "regionUpper"]:
            assert np.array_equal(a[key].view(np.uint32), b[key].view(np.uint32)), (i, key, a[key], b[key])
        assert a["numSamplesThisUpdate"] == b["numSamplesThisUpdate"]
    s = generate(5000, seed=13, frame=2)
    assert np.array_equal(gpu.leaf_ids(s["pos"]), orc.leaf_ids(s["pos"]))


@pytest.mark.parametrize("frames", [1, 3])
def test_mixture_parity(frames):
    import openpgl_amd as pgl
    dev = pgl.Device()
    gpu = pgl.Field(dev, max_samples_per_leaf=2000, deterministic=True)
    cfg = default_config(2000)
    o32, o64 = OracleField(cfg, 0), OracleField(cfg, 1)
    for f in range(frames):
        s = generate(40000, seed=0, frame=f)
        gpu.update(pgl.SampleStorage(s))
        o32.update(s)
        o64.update(s)
    R = gpu.num_regions()
    rec = lambda fld: [fld.region(i) for i in range(R)]
    assert_mixture_parity(rec(gpu), rec(o32), rec(o64), unit_dirs(512), "frames %d regions %d" % (frames, R))


def test_query_parity_surface_and_volume():
    import openpgl_amd as pgl
    load_hg_table_into_oracle()
    dev = pgl.Device()
    gpu = pgl.Field(dev, max_samples_per_leaf=2000)
    cfg = default_config(2000)
    s = generate(40000, seed=0)
    v = generate(20000, seed=1, flags=2)
    gpu.update(pgl.SampleStorage(np.concatenate([s, v])))
    orcs = {}
    for prec in (0, 1):
        for volume, data in ((0, s), (1, v)):
            o = OracleField(cfg, prec)
            o.update(data)
            orcs[prec, volume] = o
    pos, nrm, u2, is_vol, g = queries(20000, seed=3)
    u1 = np.random.default_rng(5).random(len(pos)).astype(np.float32)
    d, pdf, ids, integ = gpu.sample_batch(pos, nrm, u2, u1=u1, g=g, is_volume=is_vol)
    for volume in (0, 1):
        o = orcs[0, volume]
        m = is_vol == bool(volume)
        od, opdf, oids, ointeg = o.query_batch(volume, pos[m], nrm[m], u2[m], u1=u1[m], g=g[m])
        valid = opdf > 0
        assert np.array_equal(pdf[m] > 0, valid)
        assert np.array_equal(ids[m][valid], oids[valid])
        # pdf of the GPU's sampled direction under the oracles' distributions, every valid query
        r32, r64, r3264 = query_pdf_errors(d, pdf, np.nonzero(m)[0][valid], volume, pos, nrm, g, u1, o,
                                           orcs[1, volume])
        assert_query_pdfs(r32, "volume" if volume else "surface", r64, r3264)


_DIRECTIONAL_KEYS = ["K", "weights", "kappas", "meanDirections", "normalizations", "meanCosines", "distances",
                     "pivot", "sumOfWeightedDirections", "sumOfWeightedStats", "sumOfDistanceWeights",
                     "chiSquareMCEstimates", "splitCovariances", "splitNumSamples", "emIterations",
                     "partialEmIterations", "numSplits", "numMerges", "valid"]


def _directional_state(field):
    out = []
    for i in range(field.num_regions()):
        r = field.region(i)
        out.append({k: np.asarray(r[k]).copy() for k in _DIRECTIONAL_KEYS})
    return out


def test_schedule_independence():
    """The directional update is bit-identical for the per-region path, the
    default team schedule and forced large teams (chunking and fold order do
    not depend on the team size)."""
    import openpgl_amd as pgl
    dev = pgl.Device()
    frames = [generate(120000, seed=4, frame=f) for f in range(3)]
    schedules = [dict(per_region_launches=1),
                 dict(per_region_launches=2),
                 dict(per_region_launches=0),
                 dict(per_region_launches=0, balance=64.0, min_chunks_per_rank=1, max_team=64),
                 dict(per_region_launches=0, grid=7, balance=2.0)]
    states, teams = [], []
    for sc in schedules:
        f = pgl.Field(dev, max_samples_per_leaf=8000, deterministic=True)
        f.set_region_schedule(**sc)
        print("schedule", sc, flush=True)
        for fr in frames:
            f.update(pgl.SampleStorage(fr))
            teams.append(f.last_update_info()["max_team"])
        states.append(_directional_state(f))
    print("largest teams per update", teams)
    assert max(teams) > 1
    for s in states[1:]:
        assert len(s) == len(states[0])
        for i, (a, b) in enumerate(zip(states[0], s)):
            for k in _DIRECTIONAL_KEYS:
                assert np.asarray(a[k]).tobytes() == np.asarray(b[k]).tobytes(), (i, k, a[k], b[k])


@pytest.mark.parametrize("device_type", ["cpu4", "cpu16"])
def test_device_width_arithmetic(device_type):
    """A field created on a CPU_4 (SSE4) or CPU_16 (AVX-512) device follows
    that device's arithmetic where it decides discrete results: the kd-tree
    reciprocals (embree::rcp per ISA) and the KNN squared distances (unfused
    under SSE4).  Tree, region statistics and KNN-picked region ids equal
    the oracle configured for the same width."""
    import openpgl_amd as pgl
    width = 4 if device_type == "cpu4" else 16
    if width == 16:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
        if "avx512vl" not in flags:
            pytest.skip("host CPU without AVX-512VL")
    dev = pgl.Device(pgl.PGL_DEVICE_TYPE_CPU_4 if width == 4 else pgl.PGL_DEVICE_TYPE_CPU_16)
    gpu = pgl.Field(dev, max_samples_per_leaf=2000, deterministic=True)
    cfg = default_config(2000)
    cfg.vectorWidth = width
    orc, o64 = OracleField(cfg, 0), OracleField(cfg, 1)
    for f in range(2):
        s = generate(40000, seed=17, frame=f)
        gpu.update(pgl.SampleStorage(s))
        orc.update(s)
        o64.update(s)
    _tree_equal(gpu, orc)
    R = gpu.num_regions()
    for i in range(R):
        a, b = gpu.region(i), orc.region(i)
        for key in ["ssMean", "ssVariance", "ssLower", "ssUpper"]:
            assert np.array_equal(a[key].view(np.uint32), b[key].view(np.uint32)), (i, key)
    assert np.array_equal(gpu.knn(), orc.knn())
    # the merge loop stops at K <= width (VMMChiSquareComponentMerger.h:72)
    rec = lambda fld: [fld.region(i) for i in range(R)]
    assert_mixture_parity(rec(gpu), rec(orc), rec(o64), unit_dirs(512), "%s regions %d" % (device_type, R))
    assert min(r["K"] for r in rec(gpu) if r["valid"]) <= width or width == 16
    pos, nrm, u2, _, _ = queries(20000, seed=8)
    u1 = np.random.default_rng(9).random(len(pos)).astype(np.float32)
    d, pdf, ids, _ = gpu.sample_batch(pos, nrm, u2, u1=u1)
    _, opdf, oids, _ = orc.query_batch(0, pos, nrm, u2, u1=u1)
    valid = opdf > 0
    assert np.array_equal(pdf > 0, valid)
    assert np.array_equal(ids[valid], oids[valid])
    r32, r64, r3264 = query_pdf_errors(d, pdf, np.nonzero(valid)[0], 0, pos, nrm, np.zeros(len(pos)), u1, orc, o64)
    assert_query_pdfs(r32, device_type + " queries end to end", r64, r3264)
