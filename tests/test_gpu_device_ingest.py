"""Device ingest with routing (pglFieldUpdateDevice): mixed surface and
volume records in device memory are split by their EInsideVolume flag on the
device, each part in input order, as SampleDataStorage::addSamples routes
them (data/SampleDataStorage.h:72-90), then both fields are updated
(Field::updateField).

Checks (GPU through the C ABI):
  * the field equals (pglFieldCompare, and the directional state bit for
    bit) one trained from the same records through a host sample storage
    and pglFieldUpdate, over several frames;
  * against the oracle trained on each part: kd trees, region statistics
    and KNN lists bit-exact, mixtures within the three-way bars;
  * an input with no volume records leaves the volume field empty, and an
    empty input changes no field.
"""
import numpy as np
import pytest

from field_compare import OracleField, unit_dirs
from oracle_ctypes import default_config
from parity_checks import assert_mixture_parity, assert_spatial_bit_exact
from openpgl_amd.synth import generate

pytestmark = pytest.mark.gpu


def _mixed(n_surf, n_vol, seed, frame):
    s = generate(n_surf, seed=seed, frame=frame)
    v = generate(n_vol, seed=seed + 1, frame=frame, flags=2)
    rng = np.random.default_rng([seed, frame])
    # splatted records of both kinds (flag bit 0 does not route)
    s["flags"][rng.random(len(s)) < 0.1] |= 1
    v["flags"][rng.random(len(v)) < 0.1] |= 1
    m = np.concatenate([s, v])
    return m[rng.permutation(len(m))]


def test_device_routing_equals_host_storage_and_oracle():
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray
    dev = pgl.Device()
    a = pgl.Field(dev, max_samples_per_leaf=4000, deterministic=True)
    b = pgl.Field(dev, max_samples_per_leaf=4000, deterministic=True)
    cfg = default_config(4000)
    o32 = {k: OracleField(cfg, 0) for k in (0, 1)}
    o64 = {k: OracleField(cfg, 1) for k in (0, 1)}
    for frame in range(3):
        m = _mixed(90000, 40000, 5, frame)
        print("frame", frame, "records", len(m), "volume", int(np.sum((m["flags"] & 2) != 0)), flush=True)
        d = DeviceArray.from_numpy(m)
        a.update_device(d.value, len(m))
        d.free()  # consumed when the call returns
        b.update(pgl.SampleStorage(m))
        vol = (m["flags"] & 2) != 0
        for k, part in ((0, m[~vol]), (1, m[vol])):
            o32[k].update(part)
            o64[k].update(part)
        assert a.iteration() == b.iteration() == frame + 1
    assert a.compare(b)
    dirs = unit_dirs(512)
    for kind in (0, 1):
        R = a.num_regions(kind)
        assert R == o32[kind].num_regions() and R > 2, (kind, R)
        ga = [a.region(i, kind) for i in range(R)]
        gb = [b.region(i, kind) for i in range(R)]
        for ra, rb in zip(ga, gb):
            for key in ("K", "weights", "kappas", "meanDirections", "ssMean", "ssVariance", "numSamplesThisUpdate"):
                assert np.asarray(ra[key]).tobytes() == np.asarray(rb[key]).tobytes(), (kind, key)
        orc = [o32[kind].region(i) for i in range(R)]
        for ra, ro in zip(ga, orc):
            for key in ("ssMean", "ssVariance", "ssLower", "ssUpper", "regionLower", "regionUpper"):
                assert np.array_equal(np.asarray(ra[key]).view(np.uint32), np.asarray(ro[key]).view(np.uint32)), \
                    (kind, key)
            assert ra["numSamplesThisUpdate"] == ro["numSamplesThisUpdate"]
        assert_mixture_parity(ga, orc, [o64[kind].region(i) for i in range(R)], dirs,
                              "routed %s regions %d" % ("volume" if kind else "surface", R))
    # kd trees and KNN lists of both fields against the oracle
    assert_spatial_bit_exact(_KindView(a, 0), o32[0])
    assert_spatial_bit_exact(_KindView(a, 1), o32[1])


class _KindView:
    """one kind of a field with the single-kind accessors the shared checks use"""

    def __init__(self, field, kind):
        self.f, self.k = field, kind

    def nodes(self):
        return self.f.nodes(self.k)

    def num_regions(self):
        return self.f.num_regions(self.k)

    def region(self, i):
        return self.f.region(i, self.k)

    def knn(self):
        return self.f.knn(self.k)

    def leaf_ids(self, pos):
        return self.f.leaf_ids(pos, self.k)


def test_device_routing_edge_cases():
    import openpgl_amd as pgl
    from openpgl_amd.hip_runtime import DeviceArray
    dev = pgl.Device()
    f = pgl.Field(dev, max_samples_per_leaf=4000, deterministic=True)
    s = generate(30000, seed=9)
    d = DeviceArray.from_numpy(s)
    print("surface only", len(s), flush=True)
    f.update_device(d.value, len(s))
    assert f.num_regions(0) > 2 and f.num_regions(1) == 0
    before = [f.region(i, 0)["weights"].tobytes() for i in range(f.num_regions(0))]
    print("empty update", flush=True)
    f.update_device(d.value, 0)
    assert f.iteration() == 2
    assert [f.region(i, 0)["weights"].tobytes() for i in range(f.num_regions(0))] == before
    v = generate(20000, seed=10, flags=2)
    dv = DeviceArray.from_numpy(v)
    print("volume only", len(v), flush=True)
    f.update_device(dv.value, len(v))
    assert f.num_regions(1) > 1
    assert [f.region(i, 0)["weights"].tobytes() for i in range(f.num_regions(0))] == before
