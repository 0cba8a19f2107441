"""Batched path segments -> training samples on the GPU
(pglPathSegmentBatchPrepareSamplesDevice, pglFieldUpdateFromPathSegmentsDevice)
against the oracle's PathSegmentDataStorage::prepareSamples applied path by
path: the records must be identical bit for bit (same float operations, no
contraction), in per-path propagation order, surface and volume apart; a
field trained from path segments on the device must equal, state for state,
the field trained from the same samples through a host sample storage.
"""
import numpy as np
import pytest

from oracle_ctypes import oracle_path_samples
from openpgl_amd.synth import PATH_SEGMENT_DTYPE, SAMPLE_DTYPE, random_paths

pytestmark = pytest.mark.gpu


def _device_samples(batch, seg, off, **opt):
    from openpgl_amd.hip_runtime import DeviceArray
    dseg = DeviceArray.from_numpy(seg if len(seg) else np.zeros(1, PATH_SEGMENT_DTYPE))
    doff = DeviceArray.from_numpy(off)
    cap = max(1, len(seg))
    ds, dv = DeviceArray(cap * 40), DeviceArray(cap * 40)
    r = batch.prepare_device(dseg.value, len(seg), doff.value, len(off) - 1, ds.value, cap, dv.value, cap, **opt)
    assert r is not None
    ns, nv = r
    s = ds.to_numpy(SAMPLE_DTYPE, (ns,)) if ns else np.zeros(0, SAMPLE_DTYPE)
    v = dv.to_numpy(SAMPLE_DTYPE, (nv,)) if nv else np.zeros(0, SAMPLE_DTYPE)
    return s, v


@pytest.mark.parametrize("nee,guide,rr_direct,max_distance", [(False, False, True, 1e6), (True, False, True, 1e6),
                                                              (False, True, False, 2.0), (True, True, True, 0.5)])
def test_batch_equals_oracle(nee, guide, rr_direct, max_distance):
    import openpgl_amd as pgl
    seg, off = random_paths(40000, seed=7, max_len=16)
    print("paths", len(off) - 1, "vertices", len(seg), "options", nee, guide, rr_direct, max_distance, flush=True)
    batch = pgl.PathSegmentBatch()
    opt = dict(max_distance=max_distance, use_nee_mi_weights=nee, guide_direct_light=guide,
               rr_affects_direct=rr_direct)
    s, v = _device_samples(batch, seg, off, **opt)
    ws, wv = oracle_path_samples(seg, off, **opt)
    assert len(s) == len(ws) and len(v) == len(wv)
    assert len(s) > 1000 and len(v) > 100
    assert s.tobytes() == ws.tobytes()
    assert v.tobytes() == wv.tobytes()


def test_batch_long_and_empty_paths():
    """Paths reaching far beyond a 256-vertex tile (read from global memory
    past the staged vertices) and tiles holding more paths than the offsets
    kept in LDS (runs of empty paths) give the oracle's records too."""
    import openpgl_amd as pgl
    batch = pgl.PathSegmentBatch()
    seg, off = random_paths(300, seed=11, max_len=700)
    print("long paths", len(off) - 1, "vertices", len(seg), "longest", int(np.diff(off).max()), flush=True)
    s, v = _device_samples(batch, seg, off)
    ws, wv = oracle_path_samples(seg, off)
    assert len(s) == len(ws) and len(v) == len(wv) and len(s) > 1000
    assert s.tobytes() == ws.tobytes() and v.tobytes() == wv.tobytes()
    seg, off = random_paths(3000, seed=12, max_len=10)
    off2 = np.concatenate([np.repeat(off[:-1], 25), off[-1:]]).astype(np.uint32)  # 24 empty paths before each
    print("with empty paths", len(off2) - 1, "vertices", len(seg), flush=True)
    s, v = _device_samples(batch, seg, off2)
    ws, wv = oracle_path_samples(seg, off2)
    assert len(s) == len(ws) and len(v) == len(wv) and len(s) > 1000
    assert s.tobytes() == ws.tobytes() and v.tobytes() == wv.tobytes()


def test_batch_edge_cases():
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray
    batch = pgl.PathSegmentBatch()
    # no paths; only empty and one-vertex paths
    s, v = _device_samples(batch, np.zeros(0, PATH_SEGMENT_DTYPE), np.zeros(1, np.uint32))
    assert len(s) == 0 and len(v) == 0
    seg, _ = random_paths(10, seed=1)
    one = seg[:3]
    off = np.array([0, 0, 1, 1, 2, 3, 3], np.uint32)
    s, v = _device_samples(batch, one, off)
    assert len(s) == 0 and len(v) == 0
    # too small an output: refused, nothing written
    seg, off = random_paths(2000, seed=2)
    ws, wv = oracle_path_samples(seg, off)
    dseg, doff = DeviceArray.from_numpy(seg), DeviceArray.from_numpy(off)
    ds, dv = DeviceArray(40 * len(ws)), DeviceArray(40 * len(wv))
    assert batch.prepare_device(dseg.value, len(seg), doff.value, len(off) - 1, ds.value, len(ws) - 1, dv.value,
                                len(wv)) is None
    # a range beyond the vertices (or a decreasing one): refused, nothing read
    assert batch.prepare_device(dseg.value, len(seg) - 1, doff.value, len(off) - 1, ds.value, len(ws), dv.value,
                                len(wv)) is None
    r = batch.prepare_device(dseg.value, len(seg), doff.value, len(off) - 1, ds.value, len(ws), dv.value, len(wv))
    assert r == (len(ws), len(wv))
    assert ds.to_numpy(SAMPLE_DTYPE, (len(ws),)).tobytes() == ws.tobytes()


def test_field_update_from_path_segments():
    """Three frames of path segments on the device train the field exactly as
    the same samples through a host sample storage do."""
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray
    dev = pgl.Device()
    a = pgl.Field(dev, max_samples_per_leaf=2000)
    b = pgl.Field(dev, max_samples_per_leaf=2000)
    for frame in range(3):
        seg, off = random_paths(60000, seed=100 + frame, max_len=10)
        print("frame", frame, "paths", len(off) - 1, "vertices", len(seg), flush=True)
        dseg, doff = DeviceArray.from_numpy(seg), DeviceArray.from_numpy(off)
        a.update_from_path_segments_device(dseg.value, len(seg), doff.value, len(off) - 1)
        dseg.free()  # the path arrays are consumed when the call returns
        doff.free()
        ws, wv = oracle_path_samples(seg, off)
        st = pgl.SampleStorage(np.concatenate([ws, wv]))
        b.update(st)
        assert a.iteration() == b.iteration() == frame + 1
    assert a.num_regions(0) > 4 and a.num_regions(1) > 1
    assert a.compare(b)
    for kind in (0, 1):
        for i in range(0, a.num_regions(kind), 7):
            ra, rb = a.region(i, kind), b.region(i, kind)
            for key in ("K", "weights", "kappas", "meanDirections", "ssMean", "numSamplesThisUpdate"):
                assert np.array_equal(np.asarray(ra[key]), np.asarray(rb[key])), (kind, i, key)
