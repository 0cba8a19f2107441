"""The allocation-free path batch (pglPathSegmentsPrepareSamplesDeviceNoAlloc)
against the allocating one (pglPathSegmentBatchPrepareSamplesDevice): the
same records bit for bit, on two calls in a row sharing the static scratch,
the second one smaller (a stale count or record from the first call would
show), and a call above the maximum refused before any work is queued.
"""
import numpy as np
import pytest

from oracle_ctypes import oracle_path_samples
from openpgl_amd.synth import PATH_SEGMENT_DTYPE, SAMPLE_DTYPE, random_paths

pytestmark = pytest.mark.gpu


def _both(seg, off, **opt):
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray
    dseg = DeviceArray.from_numpy(seg)
    doff = DeviceArray.from_numpy(off)
    cap = max(1, len(seg))
    out = []
    for which in ("allocating", "no-alloc"):
        ds, dv = DeviceArray(cap * 40), DeviceArray(cap * 40)
        args = (dseg.value, len(seg), doff.value, len(off) - 1, ds.value, cap, dv.value, cap)
        print(which, "paths", len(off) - 1, "vertices", len(seg), opt, flush=True)
        if which == "allocating":
            r = pgl.PathSegmentBatch().prepare_device(*args, **opt)
        else:
            r = pgl.prepare_path_samples_device_noalloc(*args, **opt)
        assert r is not None, which
        ns, nv = r
        s = ds.to_numpy(SAMPLE_DTYPE, (ns,)) if ns else np.zeros(0, SAMPLE_DTYPE)
        v = dv.to_numpy(SAMPLE_DTYPE, (nv,)) if nv else np.zeros(0, SAMPLE_DTYPE)
        out.append((s, v))
    return out


def test_no_alloc_matches_allocating_entry_point():
    import openpgl_amd as pgl
    # first call near the maximum, second call much smaller
    seg, off = random_paths(20000, seed=31, max_len=24)
    assert 200000 < len(seg) <= pgl.NO_ALLOC_MAX_VERTICES
    (a_s, a_v), (b_s, b_v) = _both(seg, off)
    assert len(a_s) > 100000 and len(a_v) > 1000
    assert a_s.tobytes() == b_s.tobytes() and a_v.tobytes() == b_v.tobytes()
    small, soff = random_paths(700, seed=32, max_len=12)
    opt = dict(max_distance=2.0, use_nee_mi_weights=True, guide_direct_light=True, rr_affects_direct=False)
    (a_s, a_v), (b_s, b_v) = _both(small, soff, **opt)
    assert 0 < len(a_s) < 20000
    assert a_s.tobytes() == b_s.tobytes() and a_v.tobytes() == b_v.tobytes()
    # and both are the oracle's records
    ws, wv = oracle_path_samples(small, soff, **opt)
    assert b_s.tobytes() == ws.tobytes() and b_v.tobytes() == wv.tobytes()


def test_no_alloc_refuses_more_than_the_maximum():
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray
    n = pgl.NO_ALLOC_MAX_VERTICES + 1
    # buffers sized for the request, so nothing would be out of bounds anyway
    dseg = DeviceArray.from_numpy(np.zeros(n, PATH_SEGMENT_DTYPE))
    doff = DeviceArray.from_numpy(np.array([0, n], np.uint32))
    ds, dv = DeviceArray(n * 40), DeviceArray(n * 40)
    print("over the maximum", n, flush=True)
    assert pgl.prepare_path_samples_device_noalloc(dseg.value, n, doff.value, 1, ds.value, n, dv.value, n) is None
    # the entry point still works afterwards
    seg, off = random_paths(300, seed=33, max_len=10)
    (a_s, a_v), (b_s, b_v) = _both(seg, off)
    assert a_s.tobytes() == b_s.tobytes() and a_v.tobytes() == b_v.tobytes()
