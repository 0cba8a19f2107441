"""Path segments -> training samples (PathSegmentDataStorage, data/
PathSegmentDataStorage.h:160-393) on the CPU: the oracle's restatement and the
library's host path-segment storage (pglPathSegmentStorage*).

The reference's PathSegmentDataStorage.h includes SampleDataStorage.h, which
needs oneTBB, so it is not built here; the oracle is pinned by hand-derived
known-answer paths (each case below states the value the reference's code
gives and why), by an FP64 evaluation of the contribution algebra in product
form (a different formulation from the running recurrence), and the vertex
layout by the reference's own header compiled in oracle/_ref.  The library's
host storage must equal the oracle bit for bit.
"""
import numpy as np
import pytest

from oracle_ctypes import (oracle_path_samples, oracle_pixel_estimate, ref_path_segment_layout,
                           reference_available)
from openpgl_amd.synth import PATH_SEGMENT_DTYPE, SAMPLE_DTYPE, random_paths


def _vertex(pos, **kw):
    v = np.zeros(1, PATH_SEGMENT_DTYPE)[0]
    v["position"] = pos
    v["directionIn"] = (0, 1, 0)
    v["directionOut"] = (1, 0, 0)
    v["normal"] = (0, 0, 1)
    v["pdfDirectionIn"] = 1.0
    v["scatteringWeight"] = (1, 1, 1)
    v["transmittanceWeight"] = (1, 1, 1)
    v["miWeight"] = 1.0
    v["russianRouletteProbability"] = 1.0
    v["eta"] = 1.0
    v["roughness"] = 1.0
    for k, x in kw.items():
        v[k] = x
    return v


def _path(*verts):
    return np.array(list(verts), PATH_SEGMENT_DTYPE)


def _samples(path, **opt):
    s, v = oracle_path_samples(path, [0, len(path)], **opt)
    return np.concatenate([s, v]) if len(v) else s


# ------------------------------------------------------------- known answers
def test_single_bounce_sample():
    # v0 diffuse, 5 units from v1 which scatters (0.3, 0.6, 0.9): one sample at
    # v0 with weight mean(0.3, 0.6, 0.9) / pdf = 0.6 / 0.5
    p = _path(_vertex((0, 0, 0), pdfDirectionIn=0.5, directionIn=(0.6, 0.8, 0)),
              _vertex((3, 4, 0), scatteredContribution=(0.3, 0.6, 0.9)))
    s = _samples(p)
    assert len(s) == 1
    assert np.allclose(s["pos"][0], 0) and np.allclose(s["dir"][0], (0.6, 0.8, 0))
    assert s["distance"][0] == np.float32(5.0) and s["pdf"][0] == np.float32(0.5)
    assert abs(s["weight"][0] - 1.2) < 1e-6 and s["flags"][0] == 0


@pytest.mark.parametrize("nee,guide,expected", [(False, False, None), (False, True, 1.0), (True, False, None),
                                                (True, True, 0.25)])
def test_direct_light_at_the_next_vertex(nee, guide, expected):
    # emission (1,1,1) with MIS weight 0.25 seen from the next vertex counts
    # only with guideDirectLight, MIS-weighted only with useNEEMiWeights
    p = _path(_vertex((0, 0, 0)), _vertex((0, 0, 1), directContribution=(1, 1, 1), miWeight=0.25))
    s = _samples(p, use_nee_mi_weights=nee, guide_direct_light=guide)
    if expected is None:
        assert len(s) == 0
    else:
        assert len(s) == 1 and abs(s["weight"][0] - expected) < 1e-7


def test_direct_light_further_along_is_mis_weighted():
    # v2's emission (1,1,1), MIS 0.5, reaches v0 through v1's weight 2: 2 * 0.5 = 1;
    # v1 sees it as "next vertex" emission and, by default, gets no sample
    p = _path(_vertex((0, 0, 0)), _vertex((0, 0, 1), scatteringWeight=(2, 2, 2)),
              _vertex((0, 0, 2), directContribution=(1, 1, 1), miWeight=0.5))
    s = _samples(p)
    assert len(s) == 1 and np.allclose(s["pos"][0], 0) and abs(s["weight"][0] - 1.0) < 1e-7


def test_throughput_clamped_at_ten():
    # transmittance (100, 1, 1) is clamped to (10, 1, 1) before it scales the
    # scattered (1,1,1): weight (10 + 1 + 1) / 3 = 4
    p = _path(_vertex((0, 0, 0), transmittanceWeight=(100, 1, 1)),
              _vertex((0, 0, 1), scatteredContribution=(1, 1, 1)))
    s = _samples(p)
    assert len(s) == 1 and abs(s["weight"][0] - 4.0) < 1e-6


def test_russian_roulette():
    mid = dict(russianRouletteProbability=0.5)
    far = dict(scatteredContribution=(1, 1, 1))
    p = _path(_vertex((0, 0, 0)), _vertex((0, 0, 1), **mid), _vertex((0, 0, 2), **far))
    s = _samples(p)
    # v1 (one step): 1; v0 through v1's survival 1 / 0.5: 2 (v1 comes first: vertices go backwards)
    assert len(s) == 2 and abs(s["weight"][0] - 1.0) < 1e-7 and abs(s["weight"][1] - 2.0) < 1e-7
    # probability 0 terminates the throughput: no sample at v0
    p0 = _path(_vertex((0, 0, 0)), _vertex((0, 0, 1), russianRouletteProbability=0.0), _vertex((0, 0, 2), **far))
    s0 = _samples(p0)
    assert len(s0) == 1 and np.allclose(s0["pos"][0], (0, 0, 1))
    # emission beyond a roulette vertex: rrAffectsDirectContribution=False scales it by that probability
    pe = _path(_vertex((0, 0, 0)), _vertex((0, 0, 1), **mid), _vertex((0, 0, 2), directContribution=(1, 1, 1)))
    a = _samples(pe)
    b = _samples(pe, rr_affects_direct=False)
    assert abs(a["weight"][0] - 2.0) < 1e-7 and abs(b["weight"][0] - 1.0) < 1e-7


def test_specular_chain_carries_distance():
    far = dict(scatteredContribution=(1, 1, 1))
    v0 = _vertex((0, 0, 0))
    v1 = _vertex((0, 0, 2), isDelta=True, normal=(0, 0, 1), directionOut=(0, 0, -1), directionIn=(0, 0, 1))
    v2 = _vertex((0, 0, 5))
    v3 = _vertex((0, 0, 6), **far)
    s = _samples(_path(v0, v1, v2, v3))
    # v2: 1 to v3; v1 is delta (no sample) and passes on 3; v0: 2 + 3 = 5
    assert len(s) == 2 and s["distance"][0] == 1.0 and s["distance"][1] == 5.0
    # refraction (eta 1.5) scales the carried length by |cos_out / (cos_in eta)| = 1 / 1.5: 2 + 2 = 4
    v1["eta"] = 1.5
    s = _samples(_path(v0, v1, v2, v3))
    assert abs(s["distance"][1] - 4.0) < 1e-6
    # a glossy vertex (roughness 0.2) yields a sample and also carries its length
    v1["isDelta"] = False
    v1["eta"] = 1.0
    v1["roughness"] = 0.2
    s = _samples(_path(v0, v1, v2, v3))
    assert len(s) == 3 and list(s["distance"]) == [1.0, 3.0, 5.0]
    # roughness <= 0.01 gives no sample (but carries)
    v1["roughness"] = 0.005
    s = _samples(_path(v0, v1, v2, v3))
    assert len(s) == 2 and list(s["distance"]) == [1.0, 5.0]
    # the carried length is capped at twice the maximum distance
    s = _samples(_path(v0, v1, v2, v3), max_distance=2.0)
    assert list(s["distance"]) == [1.0, 4.0]


def test_pdf_floor_and_volume_flag():
    p = _path(_vertex((0, 0, 0), pdfDirectionIn=0.01, volumeScatter=True),
              _vertex((0, 0, 1), scatteredContribution=(1, 1, 1)))
    sur, vol = oracle_path_samples(p, [0, 2])
    assert len(sur) == 0 and len(vol) == 1
    assert vol["pdf"][0] == np.float32(0.1) and abs(vol["weight"][0] - 10.0) < 1e-5 and vol["flags"][0] == 2


def test_short_paths_give_nothing():
    assert len(_samples(_path())) == 0
    assert len(_samples(_path(_vertex((0, 0, 0), scatteredContribution=(1, 1, 1))))) == 0


def test_pixel_estimate_known_values():
    one = _path(_vertex((0, 0, 0), directContribution=(1, 2, 3), scatteredContribution=(0.5, 0, 0)))
    assert np.allclose(oracle_pixel_estimate(one), (1.5, 2, 3))
    two = _path(_vertex((0, 0, 0), directContribution=(1, 0, 0), scatteredContribution=(0, 1, 0),
                        scatteringWeight=(0.5, 0.5, 0.5)),
                _vertex((0, 0, 1), scatteredContribution=(2, 2, 2), directContribution=(1, 1, 1), miWeight=0.5))
    # own (1,1,0) + 0.5 * (2 + 0.5)
    assert np.allclose(oracle_pixel_estimate(two), (2.25, 2.25, 1.25))


# ------------------------------------------------------------- FP64 product form
def _fp64_contribution(seg, i, nee, guide, rr_direct):
    """sum over j > i of min(T_ij, 10) * (scattered_j + direct terms), with
    T_ij = prod_{m=i}^{j-1} transmittance_m * prod_{m=i+1}^{j-1} scattering_m / rr_m."""
    n = len(seg)
    total = np.zeros(3)
    for j in range(i + 1, n):
        t = np.ones(3)
        for m in range(i, j):
            t = t * seg["transmittanceWeight"][m].astype(np.float64)
        dead = False
        for m in range(i + 1, j):
            rr = float(seg["russianRouletteProbability"][m])
            if rr <= 0:
                dead = True
            t = t * seg["scatteringWeight"][m].astype(np.float64) / (rr if rr > 0 else 1.0)
        if dead:
            t = np.zeros(3)
        ct = np.minimum(t, 10.0)
        total += ct * seg["scatteredContribution"][j]
        direct = seg["directContribution"][j].astype(np.float64)
        if not rr_direct:
            direct = direct * (float(seg["russianRouletteProbability"][j - 1]) if j - 1 > i else 1.0)
        if j == i + 1 and not nee:
            if guide:
                total += ct * direct
        elif j > i + 1 or guide:
            total += ct * float(seg["miWeight"][j]) * direct
    return total


@pytest.mark.parametrize("nee,guide,rr_direct", [(False, False, True), (True, True, False), (False, True, True)])
def test_contribution_algebra_against_fp64(nee, guide, rr_direct):
    seg, off = random_paths(400, seed=3)
    worst = 0.0
    for p in range(len(off) - 1):
        path = seg[off[p]:off[p + 1]]
        got = oracle_path_samples(path, [0, len(path)], use_nee_mi_weights=nee, guide_direct_light=guide,
                                  rr_affects_direct=rr_direct)
        want = ([], [])
        for i in range(len(path) - 2, -1, -1):
            v = path[i]
            if v["isDelta"] or not v["roughness"] > 0.01:
                continue
            c = _fp64_contribution(path, i, nee, guide, rr_direct)
            if not (c > 0).any():
                continue
            pdf = max(0.1, float(v["pdfDirectionIn"]))
            want[1 if v["volumeScatter"] else 0].append((v["position"], c.mean() / pdf))
        for kind in (0, 1):
            assert len(want[kind]) == len(got[kind]), (p, kind)
            for k, (pos, w) in enumerate(want[kind]):
                assert np.array_equal(got[kind]["pos"][k], pos), (p, kind, k)
                worst = max(worst, abs(float(got[kind]["weight"][k]) - w) / w)
    assert worst < 1e-5, worst


# ------------------------------------------------------------- layout
def test_vertex_layout():
    import ctypes
    import openpgl_amd as pgl
    names = PATH_SEGMENT_DTYPE.names
    assert PATH_SEGMENT_DTYPE.itemsize == 136 and ctypes.sizeof(pgl.PGLPathSegmentData) == 136
    for name in names:
        assert getattr(pgl.PGLPathSegmentData, name).offset == PATH_SEGMENT_DTYPE.fields[name][1], name
    if reference_available():
        ref = ref_path_segment_layout()
        assert ref[0] == 136
        assert list(ref[1:]) == [PATH_SEGMENT_DTYPE.fields[n][1] for n in names]


# ------------------------------------------------------------- library host storage
@pytest.mark.parametrize("nee,guide,rr_direct", [(False, False, True), (True, False, True), (False, True, False),
                                                (True, True, True)])
def test_host_storage_equals_oracle(nee, guide, rr_direct):
    import openpgl_amd as pgl
    seg, off = random_paths(300, seed=11)
    st = pgl.PathSegmentStorage(capacity=64, max_distance=3.0)
    for p in range(len(off) - 1):
        path = seg[off[p]:off[p + 1]]
        st.clear()
        st.add_segments(path)
        assert st.num_segments() == len(path)
        n = st.prepare_samples(nee, guide, rr_direct)
        got = st.samples()
        assert n == len(got)
        s, v = oracle_path_samples(path, [0, len(path)], max_distance=3.0, use_nee_mi_weights=nee,
                                   guide_direct_light=guide, rr_affects_direct=rr_direct)
        want = np.concatenate([s, v])
        order = np.argsort(np.concatenate([np.nonzero((got["flags"] & 2) == 0)[0], np.nonzero(got["flags"] & 2)[0]]))
        assert got.tobytes() == want[order].tobytes() if len(got) else len(want) == 0, p
        assert np.array_equal(st.pixel_estimate(rr_direct).view(np.uint32),
                              oracle_pixel_estimate(path, rr_direct).view(np.uint32)), p


def test_host_storage_propagate_capacity_and_defaults():
    import ctypes
    import openpgl_amd as pgl
    seg, off = random_paths(50, seed=5)
    storage = pgl.SampleStorage()
    st = pgl.PathSegmentStorage(capacity=32)
    want_s, want_v = [], []
    for p in range(len(off) - 1):
        path = seg[off[p]:off[p + 1]]
        st.add_segments(path)
        assert st.validate_segments()
        st.propagate(storage)  # prepares, appends to the sample storage, clears the path
        assert st.num_segments() == 0 and st.num_samples() == 0
        s, v = oracle_path_samples(path, [0, len(path)])
        want_s.append(s)
        want_v.append(v)
    assert storage.size_surface() == sum(len(x) for x in want_s)
    assert storage.size_volume() == sum(len(x) for x in want_v)
    # capacity: vertices beyond it are dropped; a fresh vertex has the reference's defaults
    small = pgl.PathSegmentStorage(capacity=2)
    a = small.next_segment()
    b = small.next_segment()
    assert a and b and small.next_segment() == 0 and small.num_segments() == 2
    v = pgl.PGLPathSegmentData.from_address(a)
    assert (v.directionIn.y, v.directionOut.x, v.normal.z, v.pdfDirectionIn, v.miWeight, v.eta, v.roughness,
            v.russianRouletteProbability, v.scatteringWeight.x, v.transmittanceWeight.z) == (1, 1, 1, 1, 1, 1, 1, 1, 1, 1)
    assert not v.isDelta and not v.volumeScatter and v.directContribution.x == 0
    bad = _path(_vertex((0, 0, 0), roughness=1.5))
    small.clear()
    small.add_segments(bad)
    assert not small.validate_segments()
