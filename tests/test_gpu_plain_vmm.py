"""GPU parity of the plain VMM field type (PGL_DIRECTIONAL_DISTRIBUTION_VMM,
device/Device.h:139-184) against the oracle run with parallax compensation
off (itself pinned against the reference's VMM<8, 32, false> build):
no reprojection of the training samples, no shift of the statistics and
mixture before an update (the pivot stays where the first fit put it), no
shift at a query.

Tolerances as for the parallax-aware field: kd tree and sample statistics
bit-exact, pivots bit-exact, mixtures three-way against the FP32 and FP64
oracles (tests/parity_checks.py), query pdfs end to end likewise.
"""
import numpy as np
import pytest

import field_file as ff
from field_compare import OracleField, unit_dirs, vmm_pdf
from oracle_ctypes import default_config, load_hg_table_into_oracle
from parity_checks import assert_mixture_parity, assert_query_pdfs, query_pdf_errors
from openpgl_amd.synth import generate, queries

pytestmark = pytest.mark.gpu


def _fields(frames, n=40000, max_samples=2000):
    import openpgl_amd as pgl
    dev = pgl.Device()
    gpu = pgl.Field(dev, max_samples_per_leaf=max_samples, deterministic=True,
                    directional_type=pgl.PGL_DIRECTIONAL_DISTRIBUTION_VMM)
    cfg = default_config(max_samples)
    cfg.parallax = 0
    orc = OracleField(cfg)
    o64 = OracleField(cfg, 1)
    for f in range(frames):
        s = generate(n, seed=4, frame=f)
        print("plain VMM frame", f, "samples", n, flush=True)
        gpu.update(pgl.SampleStorage(s))
        orc.update(s)
        o64.update(s)
    return dev, gpu, orc, o64


@pytest.mark.parametrize("frames", [1, 3])
def test_plain_vmm_mixture_parity(frames):
    dev, gpu, orc, o64 = _fields(frames)
    gp, gd, gc = gpu.nodes()
    op, od, oc = orc.nodes()
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)) and np.array_equal(gc, oc)
    dirs = unit_dirs(512)
    errs = []
    for i in range(gpu.num_regions()):
        a, b = gpu.region(i), orc.region(i)
        assert np.array_equal(a["ssMean"].view(np.uint32), b["ssMean"].view(np.uint32)), i
        assert a["valid"] == b["valid"], i
        if not b["valid"]:
            continue
        # the pivot is the mean at the region's first fit, never shifted
        assert np.array_equal(np.asarray(a["pivot"]).view(np.uint32), np.asarray(b["pivot"]).view(np.uint32)), i
    R = gpu.num_regions()
    assert_mixture_parity([gpu.region(i) for i in range(R)], [orc.region(i) for i in range(R)],
                          [o64.region(i) for i in range(R)], dirs, "plain VMM frames %d regions %d" % (frames, R))
    if frames > 1:
        moved = [i for i in range(gpu.num_regions())
                 if not np.array_equal(gpu.region(i)["pivot"], gpu.region(i)["ssMean"])]
        assert moved, "after updates some pivots differ from the current sample means"


def test_plain_vmm_query_parity():
    import openpgl_amd as pgl
    load_hg_table_into_oracle()
    dev = pgl.Device()
    gpu = pgl.Field(dev, max_samples_per_leaf=2000, directional_type=pgl.PGL_DIRECTIONAL_DISTRIBUTION_VMM)
    cfg = default_config(2000)
    cfg.parallax = 0
    orc, orcv = OracleField(cfg), OracleField(cfg)
    o64, o64v = OracleField(cfg, 1), OracleField(cfg, 1)
    for f in range(2):
        s = generate(40000, seed=0, frame=f)
        v = generate(20000, seed=1, frame=f, flags=2)
        gpu.update(pgl.SampleStorage(np.concatenate([s, v])))
        orc.update(s)
        orcv.update(v)
        o64.update(s)
        o64v.update(v)
    pos, nrm, u2, is_vol, g = queries(20000, seed=3)
    u1 = np.random.default_rng(5).random(len(pos)).astype(np.float32)
    d, pdf, ids, integ = gpu.sample_batch(pos, nrm, u2, u1=u1, g=g, is_volume=is_vol)
    for volume, o, oo in [(0, orc, o64), (1, orcv, o64v)]:
        m = is_vol == bool(volume)
        od, opdf, oids, ointeg = o.query_batch(volume, pos[m], nrm[m], u2[m], u1=u1[m], g=g[m])
        valid = opdf > 0
        assert np.array_equal(pdf[m] > 0, valid)
        assert np.array_equal(ids[m][valid], oids[valid])
        r32, r64, r3264 = query_pdf_errors(d, pdf, np.nonzero(m)[0][valid], volume, pos, nrm, g, u1, o, oo)
        assert_query_pdfs(r32, "plain VMM " + ("volume" if volume else "surface"), r64, r3264)


def test_plain_vmm_field_file(tmp_path):
    import openpgl_amd as pgl
    dev, gpu, orc, _ = _fields(2, n=30000)
    path = str(tmp_path / "plain.pgl")
    assert gpu.store(path)
    parsed = ff.read_file(open(path, "rb").read(), 8)
    assert parsed["directionalType"] == 2  # PGL_DIRECTIONAL_DISTRIBUTION_VMM
    back = pgl.Field.from_file(dev, path)
    assert back.compare(gpu)
    pos, nrm, u2, is_vol, g = queries(5000, seed=8)
    a = gpu.sample_batch(pos, nrm, u2, g=g, is_volume=is_vol)
    b = back.sample_batch(pos, nrm, u2, g=g, is_volume=is_vol)
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))
