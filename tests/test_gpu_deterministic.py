"""Deterministic mode: Field::fitRegions / updateRegions sort every region's
samples with SampleDataLess before fitting (field/Field.h:409-412, 469-472;
order: weight, pdf, distance, position xyz, direction xyz), so the trained
field does not depend on the order in which the samples were added.

Checks (GPU through the C ABI):
  * a field trained on randomly permuted copies of the same frames is
    bit-identical (pglFieldCompare, canonical order) to one trained on the
    original order, for continuous weights and for inputs with many equal
    weights, pdfs, distances and whole duplicated records (the refinement
    passes of the device sort);
  * with ties the field still matches the oracle (which sorts with
    std::sort(SampleDataLess)): kd tree and region sample statistics
    bit-exact, mixture pdfs within the field tolerance;
  * the batched region API orders each region the same way.
"""
import numpy as np
import pytest

from field_compare import OracleField, unit_dirs, vmm_pdf
from oracle_ctypes import OracleRegion, default_config
from openpgl_amd.synth import batched_regions, generate

pytestmark = pytest.mark.gpu


def _tied_frame(n, seed, frame):
    """generate() with weights, pdfs and distances drawn from a few values
    and 10% of the records duplicated (whole records, at other positions)."""
    s = generate(n, seed=seed, frame=frame)
    rng = np.random.default_rng([seed, frame, 99])
    s["weight"] = rng.choice(np.array([0.5, 1.0, 2.0], np.float32), size=n)
    s["pdf"] = rng.choice(np.array([0.25, 0.0795774715], np.float32), size=n)
    s["distance"] = rng.choice(np.array([0.5, 1.0, 4.0, np.inf], np.float32), size=n)
    dup = rng.choice(n, size=n // 10, replace=False)
    src = rng.choice(n, size=n // 10, replace=False)
    s[dup] = s[src]
    # equal positions differing only in direction (last fields of the comparator)
    same = rng.choice(n, size=n // 20, replace=False)
    s["pos"][same] = s["pos"][same[0]]
    # -0 and +0 compare equal in SampleDataLess
    z = rng.choice(n, size=64, replace=False)
    s["pos"][z, 0] = np.where(np.arange(64) % 2 == 0, np.float32(0.0), np.float32(-0.0))
    return s


def _train(dev, frames, perm_seed=None, max_samples=8000):
    import openpgl_amd as pgl
    f = pgl.Field(dev, max_samples_per_leaf=max_samples, deterministic=True)
    rng = np.random.default_rng(perm_seed) if perm_seed is not None else None
    for s in frames:
        if rng is not None:
            s = s[rng.permutation(len(s))]
        f.update(pgl.SampleStorage(s))
    return f


@pytest.mark.parametrize("tied", [False, True])
def test_permuted_input_gives_identical_field(tied):
    import openpgl_amd as pgl
    dev = pgl.Device()
    make = _tied_frame if tied else (lambda n, seed, frame: generate(n, seed=seed, frame=frame))
    frames = [make(150000, 21, f) for f in range(3)]
    print("tied" if tied else "continuous", "frames", len(frames), flush=True)
    a = _train(dev, frames)
    b = _train(dev, frames, perm_seed=5)
    c = _train(dev, frames, perm_seed=6)
    assert a.num_regions() > 8
    assert a.compare(b)
    assert a.compare(c)


def test_tied_input_matches_oracle():
    import openpgl_amd as pgl
    dev = pgl.Device()
    frames = [_tied_frame(120000, 23, f) for f in range(2)]
    gpu = _train(dev, frames, max_samples=8000)
    orc = OracleField(default_config(8000))
    for s in frames:
        orc.update(s)
    gp, gd, gc = gpu.nodes()
    op, od, oc = orc.nodes()
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)) and np.array_equal(gd, od) and np.array_equal(gc, oc)
    dirs = unit_dirs(512)
    med = []
    for i in range(gpu.num_regions()):
        a, b = gpu.region(i), orc.region(i)
        for key in ["ssMean", "ssVariance", "ssLower", "ssUpper"]:
            assert np.array_equal(a[key].view(np.uint32), b[key].view(np.uint32)), (i, key)
        assert a["valid"] == b["valid"]
        if b["valid"]:
            rel = np.abs(vmm_pdf(a, dirs) - vmm_pdf(b, dirs)) / vmm_pdf(b, dirs)
            med.append(np.median(rel))
    print("regions", len(med), "median-of-median rel", np.median(med), "max", np.max(med), flush=True)
    assert np.median(med) <= 2e-4


def test_region_batch_order_independent():
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray
    dev = pgl.Device()
    recs, counts, stats, init_k = batched_regions(32, seed=7, nmin=64, nmax=12000)
    # ties inside regions as well
    rng = np.random.default_rng(3)
    recs["weight"][rng.random(len(recs)) < 0.3] = np.float32(1.0)
    offs = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    perm = recs.copy()
    for r in range(len(counts)):
        seg = perm[offs[r]:offs[r + 1]]
        perm[offs[r]:offs[r + 1]] = seg[rng.permutation(len(seg))]
    keys = ["K", "weights", "kappas", "meanDirections", "meanCosines", "distances", "sumOfWeightedDirections",
            "sumOfWeightedStats", "chiSquareMCEstimates", "splitCovariances", "emIterations", "numSplits",
            "numMerges"]
    states = []
    for data in (recs, perm):
        b = pgl.RegionBatch(dev)
        d = DeviceArray.from_numpy(data)
        b.step_device(d.value, counts, stats, init_k)
        states.append([{k: np.asarray(b.region(r)[k]).tobytes() for k in keys} for r in range(len(counts))])
    assert states[0] == states[1]
    # and the sorted order is the oracle's
    dirs = unit_dirs(256)
    med = []
    for r in range(0, len(counts), 4):
        cfg = default_config(32000)
        cfg.initK = int(init_k[r])
        o = OracleRegion(cfg)
        o.step(perm[offs[r]:offs[r + 1]], stats[r, 0:3], stats[r, 3:6], float(stats[r, 6]))
        a, bb = b.region(r), o.record()
        if bb["valid"]:
            med.append(np.median(np.abs(vmm_pdf(a, dirs) - vmm_pdf(bb, dirs)) / vmm_pdf(bb, dirs)))
    print("oracle regions", len(med), "median-of-median rel", np.median(med), flush=True)
    assert np.median(med) <= 2e-4
