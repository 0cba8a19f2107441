"""GPU parity with the oracle on degenerate inputs of Field::Update.

  * coincident positions: most samples at one point, so every split of the
    leaf holding them leaves them together and the kd tree grows to its
    maximum depth (one push-down round per level);
  * tiny updates: 1 to 7 samples per update (less than one sample step of a
    chunk) before and after a larger one; a field whose first update holds a
    single sample keeps every region invalid in the oracle, and so must the
    GPU;
  * identical directions: every sample of a blob shares one direction, so
    the fitted mean cosines reach their clamp;
  * zero weights: 30 % of the samples (the fit holds), or all of them (the
    oracle's mixtures become NaN and every region invalid: so must the GPU's).

Bars as in test_gpu_field.py: kd tree, region sample statistics and leaf ids
bit-exact; mixtures three-way against the FP32 and FP64 oracles
(tests/parity_checks.py).
"""
import numpy as np
import pytest

from field_compare import OracleField, unit_dirs
from oracle_ctypes import default_config
from parity_checks import assert_mixture_parity
from openpgl_amd.synth import generate

pytestmark = pytest.mark.gpu


def _train(frames, max_samples):
    import openpgl_amd as pgl
    dev = pgl.Device()
    gpu = pgl.Field(dev, max_samples_per_leaf=max_samples, deterministic=True)
    cfg = default_config(max_samples)
    o32, o64 = OracleField(cfg, 0), OracleField(cfg, 1)
    for i, s in enumerate(frames):
        print("frame", i, "samples", len(s), flush=True)
        gpu.update(pgl.SampleStorage(s))
        o32.update(s)
        o64.update(s)
    return dev, gpu, o32, o64


def _check(gpu, o32, o64, probe, label):
    gp, gd, gc = gpu.nodes()
    op, od, oc = o32.nodes()
    assert len(gp) == len(op), (label, len(gp), len(op))
    assert np.array_equal(gd, od) and np.array_equal(gc, oc), label
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)), label
    R = gpu.num_regions()
    assert R == o32.num_regions(), label
    for i in range(R):
        a, b = gpu.region(i), o32.region(i)
        for key in ["ssMean", "ssVariance", "ssLower", "ssUpper", "regionLower", "regionUpper"]:
            assert np.array_equal(a[key].view(np.uint32), b[key].view(np.uint32)), (label, i, key)
        assert a["numSamplesThisUpdate"] == b["numSamplesThisUpdate"], (label, i)
    assert np.array_equal(gpu.leaf_ids(probe), o32.leaf_ids(probe)), label
    rec = lambda fld: [fld.region(i) for i in range(R)]
    g, a, b = rec(gpu), rec(o32), rec(o64)
    valid = [r["valid"] for r in a]
    assert [r["valid"] for r in g] == valid, label
    if any(valid):
        assert_mixture_parity(g, a, b, unit_dirs(512), "%s regions %d" % (label, R))
    return R, sum(valid)


def test_coincident_positions_reach_max_depth():
    frames = []
    for f in range(2):
        s = generate(6000, seed=21, frame=f)
        s["pos"][:3500] = np.float32([0.37, 0.52, 0.61])
        frames.append(s)
    dev, gpu, o32, o64 = _train(frames, 500)
    piv, dims, code = gpu.nodes()
    p = np.float32([0.37, 0.52, 0.61])
    node, depth = 0, 1
    while dims[node] < 3:  # inner node: children code, code + 1
        node = code[node] + (0 if p[dims[node]] < piv[node] else 1)
        depth += 1
    print("nodes", len(dims), "depth of the point's leaf", depth, flush=True)
    assert depth == default_config(500).maxDepth
    probe = np.concatenate([generate(2000, seed=22, frame=2)["pos"], np.float32([[0.37, 0.52, 0.61]] * 4)])
    R, V = _check(gpu, o32, o64, probe, "coincident")
    assert R > 30 and V > 0


@pytest.mark.parametrize("seq,max_samples,valid", [([1, 2, 7, 2000, 1], 100, False),
                                                   ([3, 1, 7, 20000, 2], 1000, True)])
def test_tiny_updates(seq, max_samples, valid):
    frames = [generate(n, seed=30 + n, frame=i) for i, n in enumerate(seq)]
    dev, gpu, o32, o64 = _train(frames, max_samples)
    R, V = _check(gpu, o32, o64, generate(1000, seed=35, frame=len(seq))["pos"], "tiny %s" % seq)
    assert R > 1 and (V > 0) == valid


def test_identical_directions():
    frames = []
    for f in range(2):
        s = generate(30000, seed=40, frame=f)
        s["dir"][: len(s) // 2] = np.float32([0.0, 0.6, 0.8])
        frames.append(s)
    dev, gpu, o32, o64 = _train(frames, 2000)
    R, V = _check(gpu, o32, o64, generate(2000, seed=41, frame=2)["pos"], "identical directions")
    assert V > 0


@pytest.mark.parametrize("frac,valid", [(0.3, True), (1.0, False)])
def test_zero_weights(frac, valid):
    frames = []
    for f in range(2):
        s = generate(30000, seed=50, frame=f)
        s["weight"][: int(len(s) * frac)] = 0.0
        frames.append(s)
    dev, gpu, o32, o64 = _train(frames, 2000)
    R, V = _check(gpu, o32, o64, generate(2000, seed=51, frame=2)["pos"], "zero weights %.1f" % frac)
    assert (V > 0) == valid
