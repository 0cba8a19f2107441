"""Field statistics and the failed-update state.

FieldStatistics (field/FieldStatistics.h:16-97, Field.h:604-638): region
counts and sizes, the last update's times (the copy of host records
included), KDTreeStatistics (KDTree.h:651-674: a leaf counts depth 1) and
DirectionalDistributionStatistics (min / max / mean / root mean square of
the component counts); checked against figures recomputed from the exported
tree and regions, and against the oracle's region and node counts.

Failed update: a directional update whose kernel reports a failure (a team
barrier that timed out) is injected from the host with
pglFieldDebugFailNextUpdate -- every kernel still runs to completion -- and
the field must refuse every later update, query, export and file write
until pglFieldReset, after which it trains as a fresh field does.
"""
import os
import tempfile

import numpy as np
import pytest

from field_compare import OracleField
from oracle_ctypes import default_config
from openpgl_amd.synth import generate, queries

pytestmark = pytest.mark.gpu


def _depth(nodes_dims, nodes_child, i=0):
    # canonical export: dims 3 = leaf; children of an inner node at child, child + 1
    if nodes_dims[i] == 3:
        return 1
    c = int(nodes_child[i])
    return 1 + max(_depth(nodes_dims, nodes_child, c), _depth(nodes_dims, nodes_child, c + 1))


def test_field_statistics():
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray
    dev = pgl.Device()
    f = pgl.Field(dev, max_samples_per_leaf=2000, deterministic=True)
    orc = OracleField(default_config(2000))
    for fr in range(2):
        s = generate(40000, seed=2, frame=fr)
        f.update(pgl.SampleStorage(s))
        orc.update(s)
    st, text, header, csv = f.statistics(0)
    print(text, flush=True)
    R = f.num_regions()
    assert st["numCacheRegions"] == R == orc.num_regions()
    assert st["numCacheRegionsReserved"] >= R
    assert st["sizePerCacheRegions"] == 4096
    piv, dims, child = f.nodes()
    assert st["numberOfNodes"] == len(piv) == len(orc.nodes()[0])
    assert st["maxDepth"] == _depth(dims, child)
    K = np.array([f.region(i)["K"] for i in range(R)])
    # float accumulation in region order, as Field::getStatistics
    avg, sec = np.float32(0), np.float32(0)
    for k in K:
        avg = np.float32(avg + np.float32(k))
        sec = np.float32(sec + np.float32(k * k))
    assert st["minNumberOfComponents"] == K.min() and st["maxNumberOfComponents"] == K.max()
    assert np.float32(st["averageNumberOfComponents"]) == np.float32(avg / np.float32(R))
    assert np.float32(st["secondMomentNumberOfComponents"]) == np.float32(np.sqrt(np.float32(sec / np.float32(R))))
    assert st["timeLastUpdate"] > 0 and st["timeLastUpdateCopySamples"] > 0
    assert st["timeLastUpdateSpatialStructureUpdate"] > 0 and st["timeLastUpdateDirectionalDistriubtionUpdate"] > 0
    assert "numCacheRegions = %d" % R in text and "maxDepth" in text
    assert len(header.split(" , ")) == len(csv.split(" , ")) == 24
    # records already in device memory: no copy step
    d = DeviceArray.from_numpy(generate(40000, seed=2, frame=2))
    f.update_surface_device(d.value, 40000)
    st2 = f.statistics(0)[0]
    assert st2["timeLastUpdateCopySamples"] == 0 and st2["timeLastUpdate"] > 0
    # the empty volume field
    v = f.statistics(1)[0]
    assert v["numCacheRegions"] == 0 and v["numberOfNodes"] == 0


def test_failed_update_is_refused_until_reset():
    import openpgl_amd as pgl
    dev = pgl.Device()
    frames = [generate(40000, seed=6, frame=fr) for fr in range(3)]
    f = pgl.Field(dev, max_samples_per_leaf=2000, deterministic=True)
    f.update(pgl.SampleStorage(frames[0]))
    assert f.validate()
    print("injecting a failure into the second update", flush=True)
    f.debug_fail_next_update(0)
    f.update(pgl.SampleStorage(frames[1]))  # queued; completes in the next call
    assert not f.validate()                  # reports the failure
    it = f.iteration()
    assert not f.validate()                  # and keeps refusing
    assert f.statistics(0) is None
    assert f.num_regions() == 0
    with tempfile.TemporaryDirectory() as tmp:
        assert not f.store(os.path.join(tmp, "failed.field"))
    pos, nrm, u2, _, _ = queries(1000, seed=1)
    d, pdf, ids, _ = f.sample_batch(pos, nrm, u2)
    assert not np.any(pdf > 0)
    f.update(pgl.SampleStorage(frames[2]))
    assert f.iteration() == it and not f.validate()
    print("reset", flush=True)
    f.reset()
    f.update(pgl.SampleStorage(frames[0]))
    assert f.validate()
    g = pgl.Field(dev, max_samples_per_leaf=2000, deterministic=True)
    g.update(pgl.SampleStorage(frames[0]))
    assert f.compare(g)
