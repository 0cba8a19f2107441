"""Batched directional training of pre-partitioned regions (pglRegionBatch*)
against the oracle, region by region, and schedule independence of the
batched launch, the per-region path and the autotuner.

Tolerance (as for the field's mixtures): three-way against the FP32 and
FP64 oracles (tests/parity_checks.py).
"""
import numpy as np
import pytest

from field_compare import unit_dirs, vmm_pdf
from oracle_ctypes import OracleRegion, default_config
from parity_checks import assert_mixture_parity
from openpgl_amd.synth import batched_regions

pytestmark = pytest.mark.gpu


def _oracle_regions(recs, counts, stats7, init_k, steps, precision=0):
    out = []
    offs = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    for r in range(len(counts)):
        cfg = default_config(32000)
        cfg.initK = int(init_k[r])
        o = OracleRegion(cfg, precision)
        out.append(o)
    for st in steps:
        recs_s, stats_s = st
        for r, o in enumerate(out):
            seg = recs_s[offs[r]:offs[r + 1]]
            o.step(seg, stats_s[r, 0:3], stats_s[r, 3:6], float(stats_s[r, 6]))
    return out


def test_region_batch_matches_oracle():
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray
    dev = pgl.Device()
    recs0, counts, stats0, init_k = batched_regions(40, seed=3, nmin=64, nmax=12000)
    recs1, _, stats1, _ = batched_regions(40, seed=3, nmin=64, nmax=12000, step=1)
    batch = pgl.RegionBatch(dev, max_samples_per_leaf=32000)
    dirs = unit_dirs(512)
    steps = []
    for recs, stats in [(recs0, stats0), (recs1, stats1)]:
        print("step", len(steps), "regions", len(counts), "samples", len(recs), flush=True)
        d = DeviceArray.from_numpy(recs)
        batch.step_device(d.value, counts, stats, init_k)
        steps.append((recs, stats))
        o32 = _oracle_regions(recs0, counts, stats0, init_k, steps, 0)
        o64 = _oracle_regions(recs0, counts, stats0, init_k, steps, 1)
        assert_mixture_parity([batch.region(r) for r in range(len(counts))], [o.record() for o in o32],
                              [o.record() for o in o64], dirs, "step %d" % (len(steps) - 1))


_KEYS = ["K", "weights", "kappas", "meanDirections", "normalizations", "meanCosines", "distances",
         "sumOfWeightedDirections", "sumOfWeightedStats", "sumOfDistanceWeights", "chiSquareMCEstimates",
         "splitCovariances", "emIterations", "partialEmIterations", "numSplits", "numMerges", "valid"]


def _state(batch):
    return [{k: np.asarray(batch.region(r)[k]).tobytes() for k in _KEYS} for r in range(batch.num_regions())]


def test_region_batch_schedules_and_autotune_agree():
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray
    dev = pgl.Device()
    recs, counts, stats, init_k = batched_regions(96, seed=5, nmin=64, nmax=20000)
    d = DeviceArray.from_numpy(recs)
    ref = None
    for sc in [dict(per_region_launches=1), dict(per_region_launches=2), dict(per_region_launches=0),
               dict(per_region_launches=0, balance=8.0, max_team=64)]:
        b = pgl.RegionBatch(dev)
        b.set_schedule(**sc)
        print("schedule", sc, flush=True)
        b.step_device(d.value, counts, stats, init_k)
        st = _state(b)
        if ref is None:
            ref = st
        else:
            assert st == ref
    b = pgl.RegionBatch(dev)
    b.step_device(d.value, counts, stats, init_k)
    cands = [dict(per_region_launches=0, balance=0.5), dict(per_region_launches=0, balance=4.0, max_team=32),
             dict(per_region_launches=0, grid=1, balance=1.0)]
    best, table = b.autotune(cands, reps=1)
    print("autotune", best, table, flush=True)
    assert 0 <= best < len(cands)
    assert all(ms > 0 for _, ms in table)
    assert b.schedule()["balance"] == table[best][0]["balance"]
    assert _state(b) == ref
