"""CPU tests: the oracle is pinned against golden vectors written by the
reference's own code (tests/golden, from oracle/_ref), against closed forms,
brute force and invariants; when the reference build is present (this
container) it is also compared live.  No GPU is used here."""
import ctypes
import os
import platform

import numpy as np
import pytest

from field_compare import OracleField, unit_dirs, vmm_pdf
from oracle_ctypes import (OracleRegion, RefRegion, default_config, load_hg_table_into_oracle, oracle, oracle_full,
                           ptr, reference, reference_available, region_query, region_sample_stats)
from openpgl_amd.synth import SAMPLE_DTYPE, generate, region_samples

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def _cpu_vendor():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("vendor_id"):
                    return line.split(":")[1].strip()
    except OSError:
        pass
    return platform.processor()


# ---------------------------------------------------------------- statistics
@pytest.mark.skipif(_cpu_vendor() != "GenuineIntel", reason="rcpss tables differ between CPU vendors")
def test_rcp_matches_reference_vectors():
    g = _load("rcp.npz")
    lib = oracle()
    got = np.array([lib.orc_rcp(float(x)) for x in g["x"]], np.float32)
    assert np.array_equal(got.view(np.uint32), g["y"].view(np.uint32))


def test_rcp_is_newton_refined_reciprocal():
    # rcpss has 12 bits; one Newton step gives ~22 bits: |rcp(x) * x - 1| < 2^-21
    lib = oracle()
    xs = np.random.default_rng(0).uniform(1e-3, 1e3, 2000).astype(np.float32)
    r = np.array([lib.orc_rcp(float(x)) for x in xs], np.float64)
    assert np.max(np.abs(r * xs - 1.0)) < 2.0 ** -21


@pytest.mark.skipif(_cpu_vendor() != "GenuineIntel", reason="rcpss tables differ between CPU vendors")
def test_integer_statistics_match_reference_vectors():
    g = _load("int_stats.npz")
    lib = oracle()
    for pts, b, raw, st in zip(g["points"], g["bounds"], g["raw"], g["stats"]):
        n = int(b[6])
        p = np.ascontiguousarray(pts[:n], np.float32)
        lo, hi = np.ascontiguousarray(b[:3]), np.ascontiguousarray(b[3:6])
        out = np.zeros(13, np.float32)
        r = np.zeros(13, np.int64)
        lib.orc_integer_stats(ptr(p), n, ptr(lo), ptr(hi), ptr(out), ptr(r))
        assert np.array_equal(r, raw)
        assert np.array_equal(out.view(np.uint32), st.view(np.uint32))


def test_integer_statistics_brute_force():
    """Integer statistics recomputed in numpy from the definition: t = (p - c) * rcp(h),
    sums of trunc(4096 t) and trunc(4096 t^2) in int64, min/max of trunc(4096 t)."""
    lib = oracle()
    rng = np.random.default_rng(3)
    for _ in range(10):
        n = int(rng.integers(1, 300))
        lo = rng.uniform(-1, 0, 3).astype(np.float32)
        hi = (lo + rng.uniform(0.1, 2, 3)).astype(np.float32)
        p = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
        half = ((hi - lo) * np.float32(0.5)).astype(np.float32)
        inv = np.array([lib.orc_rcp(float(h)) for h in half], np.float32)
        c = (lo + half).astype(np.float32)
        t = ((p - c).astype(np.float32) * inv).astype(np.float32)
        m = np.trunc((t * np.float32(4096.0)).astype(np.float32)).astype(np.int64)
        v = np.trunc(((t * t).astype(np.float32) * np.float32(4096.0)).astype(np.float32)).astype(np.int64)
        out = np.zeros(13, np.float32)
        r = np.zeros(13, np.int64)
        lib.orc_integer_stats(ptr(p), n, ptr(lo), ptr(hi), ptr(out), ptr(r))
        assert np.array_equal(r[0:3], m.sum(0)) and np.array_equal(r[3:6], v.sum(0))
        assert r[6] == n
        assert np.array_equal(r[7:10], m.min(0)) and np.array_equal(r[10:13], m.max(0))
        # mean is the bin-quantised mean of the points
        assert np.allclose(out[0:3], p.mean(0), atol=float((hi - lo).max()) / 4096 * 2)


def test_stats_merge_split_match_reference_vectors():
    g = _load("stats_ops.npz")
    lib = oracle_full()
    for a, b, d, pos, m, l, r in zip(g["a"], g["b"], g["dim"], g["pos"], g["merged"], g["left"], g["right"]):
        om, ol, orr = np.zeros(13, np.float32), np.zeros(13, np.float32), np.zeros(13, np.float32)
        lib.orc_stats_merge_split(ptr(np.ascontiguousarray(a)), ptr(np.ascontiguousarray(b)), int(d), float(pos),
                                  ptr(om), ptr(ol), ptr(orr))
        for x, y in [(om, m), (ol, l), (orr, r)]:
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


# ----------------------------------------------------------------- kd tree
def _oracle_field(max_samples, samples, frames=1):
    f = OracleField(default_config(max_samples))
    for s in samples[:frames]:
        f.update(s)
    return f


def test_kdtree_invariants():
    """Every sample sits in exactly one leaf; lookup (p >= pivot -> right) agrees
    with the partition (p < pivot -> left) used for the statistics; leaves hold
    at most maxSamples unless at maximum depth; each region's sample count is
    the number of samples whose lookup lands in it."""
    s = generate(30000, seed=11)
    f = _oracle_field(1500, [s])
    piv, dims, cod = f.nodes()
    R = f.num_regions()
    ids = f.leaf_ids(s["pos"])
    counts = np.bincount(ids, minlength=R)
    recs = [f.region(i) for i in range(R)]
    assert counts.sum() == len(s)
    for i, r in enumerate(recs):
        assert r["numSamplesThisUpdate"] == counts[i]
        assert r["numSamplesThisUpdate"] <= 1500
        # region bounds contain its samples (positions strictly inside the tree bounds)
        p = s["pos"][ids == i]
        if len(p):
            assert np.all(p >= r["regionLower"] - 1e-6) and np.all(p <= r["regionUpper"] + 1e-6)
    # inner nodes: dim in 0..2, children inside the array
    inner = dims < 3
    assert np.all(cod[inner] + 1 < len(piv))


def test_kdtree_split_rule_on_tiny_input():
    """Brute force on a tiny set: with maxSamples = 3 the root must split at the
    mean of the axis of largest variance of the integer statistics."""
    pts = np.array([[0.1, 0.5, 0.5], [0.2, 0.5, 0.52], [0.9, 0.55, 0.5], [0.8, 0.45, 0.48], [0.5, 0.5, 0.5]],
                   np.float32)
    s = np.zeros(len(pts), SAMPLE_DTYPE)
    s["pos"] = pts
    s["dir"] = [0, 0, 1]
    s["weight"] = 1
    s["pdf"] = 1 / (4 * np.pi)
    s["distance"] = 1
    f = OracleField(default_config(3))
    lo = np.array([0, 0, 0], np.float32)
    hi = np.array([1, 1, 1], np.float32)
    f.set_scene_bounds(lo, hi)
    f.update(s)
    piv, dims, cod = f.nodes()
    assert dims[0] == 0  # x has by far the largest variance
    out = np.zeros(13, np.float32)
    oracle().orc_integer_stats(ptr(pts), len(pts), ptr(lo), ptr(hi), ptr(out), None)
    assert piv[0] == out[0]


def test_scene_bounds_skip_last_sample_and_scale():
    """estimateSceneBounds: AABB of samples[0..n-2] scaled 3x about its centre."""
    s = generate(1000, seed=2)
    s["pos"][-1] = [50.0, 50.0, 50.0]  # must be ignored
    f = _oracle_field(32000, [s])
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    oracle().orc_field_get_scene_bounds(f.h, ptr(lo), ptr(hi))
    p = s["pos"][:-1]
    a, b = p.min(0), p.max(0)
    c = ((a + b) * np.float32(0.5)).astype(np.float32)
    assert np.array_equal(lo, (c + np.float32(3) * (a - c)).astype(np.float32))
    assert np.array_equal(hi, (c + np.float32(3) * (b - c)).astype(np.float32))


def test_knn_brute_force():
    s = generate(30000, seed=4)
    f = _oracle_field(1000, [s])
    R = f.num_regions()
    means = np.array([f.region(i)["ssMean"] for i in range(R)], np.float32)
    nb = f.knn()
    for i in range(R):
        d = ((means - means[i]) ** 2).sum(1)
        order = np.lexsort((np.arange(R), d))[:8]
        assert set(nb[i]) == set(order), i


# ------------------------------------------------------------- directional
def test_region_trajectories_match_reference_vectors():
    g = _load("region_kat.npz")
    cfg = default_config(32000)
    cfg.minSamplesForSplitting = 1000
    cfg.minSamplesForPartialRefitting = 1000
    cfg.minSamplesForMerging = 2000
    dirs = g["dirs"]
    worst = 0.0
    for reg in range(2):
        o = OracleRegion(cfg)
        for step in range(3):
            k = "r%d_s%d_" % (reg, step)
            samples = np.ascontiguousarray(g[k + "samples"])
            ss = g[k + "ss"]
            o.step(samples, ss[0:3], ss[3:6], ss[6], int(g[k + "split"]))
            rec = o.record()
            assert rec["K"] == g[k + "K"], k
            if step > 0:  # the reference's FittingStatistics counters are only set by update()
                assert rec["emIterations"] == g[k + "emIterations"], k
                assert rec["numSplits"] == g[k + "numSplits"] and rec["numMerges"] == g[k + "numMerges"], k
            pdf = np.array([o.pdf(d) for d in dirs])
            rel = np.abs(pdf - g[k + "pdf"]) / g[k + "pdf"]
            worst = max(worst, float(np.median(rel)))
            assert np.median(rel) < 1e-4 and np.max(rel) < 5e-3, (k, np.median(rel), np.max(rel))
    print("worst median pdf rel error vs reference", worst)


@pytest.mark.skipif(not reference_available(), reason="reference build absent")
def test_region_live_against_reference_build():
    cfg = default_config(32000)
    for seed in range(4):
        o, r = OracleRegion(cfg), RefRegion(cfg)
        for step in range(3):
            s = region_samples(4000, seed=1000 + 10 * seed + step)
            m, v, n = region_sample_stats(s)
            o.step(s, m, v, n, 0)
            r.step(s, m, v, n, 0)
            a, b = o.record(), r.record()
            assert a["K"] == b["K"]
            dirs = unit_dirs(256, seed)
            pa = np.array([o.pdf(d) for d in dirs])
            pb = np.array([r.pdf(d) for d in dirs])
            assert np.median(np.abs(pa - pb) / pb) < 1e-4


def test_fp64_twin_agrees_with_fp32():
    cfg = default_config(32000)
    o32, o64 = OracleRegion(cfg, 0), OracleRegion(cfg, 1)
    s = region_samples(6000, seed=77)
    m, v, n = region_sample_stats(s)
    o32.step(s, m, v, n)
    o64.step(s, m, v, n)
    dirs = unit_dirs(512)
    p32 = np.array([o32.pdf(d) for d in dirs])
    p64 = np.array([o64.pdf(d) for d in dirs])
    assert o32.record()["K"] == o64.record()["K"]
    assert np.median(np.abs(p32 - p64) / p64) < 1e-4


def test_k1_closed_form_mle():
    """K = 1 without split/merge: mean direction = normalised sum of w d, mean
    cosine = MAP blend of |sum w d| / sum w with prior (0, 0.2)."""
    cfg = default_config(32000)
    cfg.initK = 1
    cfg.useSplitAndMerge = 0
    o = OracleRegion(cfg, 1)
    s = region_samples(3000, seed=5, lobes=1)
    s["distance"] = np.inf  # reprojection keeps directions
    m, v, n = region_sample_stats(s)
    o.step(s, m, v, n)
    rec = o.record()
    d = s["dir"].astype(np.float64)
    w = s["weight"].astype(np.float64)
    sd = (w[:, None] * d).sum(0)
    mu = sd / np.linalg.norm(sd)
    rbar = np.linalg.norm(sd) / w.sum()
    N = len(s)
    wk = (0.01 + N) / (0.01 * 1 + N)  # MAP weight with prior 0.01: (a + N) / (a K + N)
    partial = wk * N
    mc = (0.0 * 0.2 + rbar * partial) / (0.2 + partial)
    assert rec["K"] == 1
    assert np.allclose(rec["meanDirections"][0], mu, atol=1e-6)
    assert abs(rec["meanCosines"][0] - mc) < 1e-6
    kappa = (3 * mc - mc ** 3) / (1 - mc * mc)
    assert abs(rec["kappas"][0] - kappa) / kappa < 1e-5


def test_vmf_normalisation_and_pdf_integral():
    """A trained mixture's pdf integrates to 1 over the sphere (Fibonacci quadrature)."""
    cfg = default_config(32000)
    o = OracleRegion(cfg)
    s = region_samples(6000, seed=8)
    m, v, n = region_sample_stats(s)
    o.step(s, m, v, n)
    rec = o.record()
    N = 200000
    i = np.arange(N) + 0.5
    z = 1 - 2 * i / N
    phi = np.pi * (1 + 5 ** 0.5) * i
    r = np.sqrt(1 - z * z)
    dirs = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    integral = vmm_pdf(rec, dirs).mean() * 4 * np.pi
    assert abs(integral - 1) < 2e-3


def test_kappa_mean_cosine_inverse_pair():
    for k in [0.01, 0.5, 2.18853, 10, 100, 1000, 31999]:
        r = 1 / np.tanh(k) - 1 / k
        k2 = (3 * r - r ** 3) / (1 - r * r)
        # Banerjee et al.'s approximation inverts coth(k) - 1/k within a few percent
        assert abs(k2 - k) / k < 0.1 or k < 1


# ------------------------------------------------------------------- query
def test_hg_table_matches_reference_data():
    t = load_hg_table_into_oracle()
    assert t.shape == (128, 18)
    assert np.all(t[:, 0] == 3)
    # the product's data file carries exactly the same values
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "openpgl_amd", "csrc",
                       "hg_table.inc")
    vals = []
    for line in open(inc):
        line = line.strip()
        if not line or line.startswith("//"):
            continue
        vals += [float.fromhex(v.strip().rstrip("f")) if "0x" in v else float(v.strip().rstrip("f"))
                 for v in line.rstrip(",").split(",")]
    assert np.array_equal(np.array(vals, np.float32), t.reshape(-1))
    if reference_available():
        live = np.zeros(128 * 18, np.float32)
        reference().ref_hg_table(ptr(live))
        assert np.array_equal(live, t.reshape(-1))


def test_query_kat_matches_reference_vectors():
    load_hg_table_into_oracle()
    g = _load("region_kat.npz")
    q = _load("query_kat.npz")
    cfg = default_config(32000)
    cfg.minSamplesForSplitting = 1000
    cfg.minSamplesForPartialRefitting = 1000
    cfg.minSamplesForMerging = 2000
    o = OracleRegion(cfg)
    for step in range(3):
        k = "r0_s%d_" % step
        ss = g[k + "ss"]
        o.step(np.ascontiguousarray(g[k + "samples"]), ss[0:3], ss[3:6], ss[6], int(g[k + "split"]))
    for vol in (0, 1):
        rel_int, rel_pdf, ang = [], [], []
        for i in range(200):
            integ, pa, d, p = region_query(o, vol, q["pos"][i], q["nrm"][i], q["g"][i], q["u2"][i], q["pdf_dir"][i])
            rel_int.append(abs(integ - q["integral_%d" % vol][i]) / q["integral_%d" % vol][i])
            rel_pdf.append(abs(pa - q["pdf_at_dir_%d" % vol][i]) / q["pdf_at_dir_%d" % vol][i])
            ang.append(float(np.dot(d, q["dir_%d" % vol][i])))
        rel_int, rel_pdf, ang = map(np.array, (rel_int, rel_pdf, ang))
        assert np.median(rel_int) < 1e-4 and np.median(rel_pdf) < 1e-4, (vol, np.median(rel_int), np.median(rel_pdf))
        assert np.mean(rel_pdf < 1e-3) > 0.97
        # Sampled directions are not compared one to one: merge costs of near-duplicate
        # components sit at the rounding floor (~1e-7), so the component order after
        # merging may differ from the reference while the mixture is the same; sample()
        # then maps u to a different, equally distributed direction.  Sampling is
        # checked in distribution by test_sampling_follows_pdf.
        print("vol", vol, "identical sample directions", np.mean(ang > 0.9999))


def _plain_cfg():
    cfg = default_config(32000)
    cfg.minSamplesForSplitting = 1000
    cfg.minSamplesForPartialRefitting = 1000
    cfg.minSamplesForMerging = 2000
    cfg.parallax = 0
    return cfg


def test_plain_vmm_trajectories_match_reference_vectors():
    """PGL_DIRECTIONAL_DISTRIBUTION_VMM: no reprojection, no shift before an
    update (the pivot stays where the fit put it) or at a query."""
    load_hg_table_into_oracle()
    g = _load("region_kat_np.npz")
    dirs = g["dirs"]
    for reg in range(2):
        o, o64 = OracleRegion(_plain_cfg()), OracleRegion(_plain_cfg(), 1)
        for step in range(3):
            k = "r%d_s%d_" % (reg, step)
            ss = g[k + "ss"]
            for x in (o, o64):
                x.step(np.ascontiguousarray(g[k + "samples"]), ss[0:3], ss[3:6], ss[6], int(g[k + "split"]))
            rec = o.record()
            assert rec["K"] == g[k + "K"], k
            if step > 0:
                assert rec["emIterations"] == g[k + "emIterations"], k
                assert rec["numSplits"] == g[k + "numSplits"] and rec["numMerges"] == g[k + "numMerges"], k
            assert np.array_equal(rec["pivot"], g[k + "pivot"]), k
            assert np.array_equal(rec["pivot"], g["r%d_s0_ss" % reg][0:3]), k  # fit-time mean, never shifted
            pdf = np.array([o.pdf(d) for d in dirs])
            rel = np.abs(pdf - g[k + "pdf"]) / g[k + "pdf"]
            # Where the reference's own FP32 result is off the FP64 twin (the
            # third step here: 1.4e-4 after a spatial-split decay), the bar is
            # the twin's distance to the reference, not a fixed 1e-4.
            pdf64 = np.array([o64.pdf(d) for d in dirs])
            ref_err = np.median(np.abs(pdf64 - g[k + "pdf"]) / g[k + "pdf"])
            bar = max(1e-4, 2.0 * ref_err)
            assert np.median(rel) < bar and np.max(rel) < 5e-3, (k, np.median(rel), np.max(rel), ref_err)
        for vol in (0, 1):
            rel = []
            for i in range(100):
                _, pa, _, _ = region_query(o, vol, g["q%d_pos" % reg][i], g["q%d_nrm" % reg][i], g["q%d_g" % reg][i],
                                           g["q%d_u2" % reg][i], dirs[i])
                ref = g["q%d_pdf_at_dir_%d" % (reg, vol)][i]
                rel.append(abs(pa - ref) / ref)
            rel = np.array(rel)
            assert np.median(rel) < bar and np.mean(rel < 1e-3) > 0.97, (reg, vol, np.median(rel), bar)


def test_plain_vmm_differs_from_parallax_aware():
    """The switch matters on these inputs: the parallax-aware fit of the same
    samples is a different mixture."""
    g = _load("region_kat_np.npz")
    cfg_p = _plain_cfg()
    cfg_p.parallax = 1
    a, b = OracleRegion(_plain_cfg()), OracleRegion(cfg_p)
    for step in range(2):
        k = "r0_s%d_" % step
        ss = g[k + "ss"]
        for o in (a, b):
            o.step(np.ascontiguousarray(g[k + "samples"]), ss[0:3], ss[3:6], ss[6], int(g[k + "split"]))
    dirs = g["dirs"]
    pa = np.array([a.pdf(d) for d in dirs])
    pb = np.array([b.pdf(d) for d in dirs])
    assert np.median(np.abs(pa - pb) / pb) > 1e-2


@pytest.mark.skipif(not reference_available(), reason="reference build absent")
def test_plain_vmm_live_against_reference_build():
    """Six regions x three steps.  A split decision near its threshold can
    flip between any two FP32 evaluation orders (libm vs the reference's
    polynomial exp/log here) and the trajectory then departs; as for the GPU,
    the bar is per case: equal K and median pdf error below
    max(1e-4, 2 x the FP64 twin's distance to the reference) in >= 15 of 18."""
    cfg = _plain_cfg()
    good, cases = 0, []
    for seed in range(6):
        o, o64, r = OracleRegion(cfg), OracleRegion(cfg, 1), RefRegion(cfg)
        for step in range(3):
            s = region_samples(4000, seed=2000 + 10 * seed + step)
            m, v, n = region_sample_stats(s)
            for x in (o, o64, r):
                x.step(s, m, v, n, int(step == 2))
            dirs = unit_dirs(256, seed)
            pa = np.array([o.pdf(d) for d in dirs])
            pb = np.array([r.pdf(d) for d in dirs])
            pc = np.array([o64.pdf(d) for d in dirs])
            bar = max(1e-4, 2.0 * np.median(np.abs(pc - pb) / pb))
            err = np.median(np.abs(pa - pb) / pb)
            ok = o.record()["K"] == r.record()["K"] and err < bar
            good += ok
            cases.append((seed, step, o.record()["K"], r.record()["K"], float(err), float(bar)))
    assert good >= 15, cases


def test_sampling_follows_pdf():
    """Chi-square test: directions drawn by init + cosine product + sample fall
    into equal-area sphere cells with the probabilities the pdf integrates to."""
    load_hg_table_into_oracle()
    cfg = default_config(32000)
    o = OracleRegion(cfg)
    s = region_samples(5000, seed=21)
    m, v, n = region_sample_stats(s)
    o.step(s, m, v, n)
    pos = m + np.float32(0.01)
    for vol, nrm, gg in [(0, np.array([0.0, 0.6, 0.8], np.float32), 0.0), (1, np.array([0.0, 0.0, 1.0], np.float32), 0.5)]:
        rng = np.random.default_rng(3 + vol)
        N = 20000
        u = rng.random((N, 2)).astype(np.float32)
        dirs = np.array([region_query(o, vol, pos, nrm, gg, u[i], nrm)[2] for i in range(N)])
        # 12 x 24 equal-area cells in (z, phi)
        nz, nphi = 12, 24
        def cell(d):
            zi = np.clip(((d[:, 2] + 1) / 2 * nz).astype(int), 0, nz - 1)
            pi = np.clip(((np.arctan2(d[:, 1], d[:, 0]) + np.pi) / (2 * np.pi) * nphi).astype(int), 0, nphi - 1)
            return zi * nphi + pi
        counts = np.bincount(cell(dirs), minlength=nz * nphi)
        # expected: pdf integrated over each cell with a fine Fibonacci grid
        M = 60000
        i = np.arange(M) + 0.5
        z = 1 - 2 * i / M
        ph = np.pi * (1 + 5 ** 0.5) * i
        r = np.sqrt(1 - z * z)
        grid = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1).astype(np.float32)
        pdfs = np.array([region_query(o, vol, pos, nrm, gg, np.array([0.5, 0.5], np.float32), d)[1] for d in grid])
        expected = np.bincount(cell(grid), weights=pdfs, minlength=nz * nphi)
        expected = expected / expected.sum() * N
        ok = expected > 5
        chi2 = (((counts - expected) ** 2) / expected)[ok].sum()
        dof = ok.sum() - 1
        # 99.9% quantile of chi2(dof) ~ dof + 3.1 sqrt(2 dof) for dof ~ 100-300
        assert chi2 < dof + 4 * np.sqrt(2 * dof), (vol, chi2, dof)


# --------------------------------------------------------------- the library
def test_library_exports_every_declared_symbol():
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    header = open(os.path.join(root, "include", "openpgl_amd", "openpgl.h")).read()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    declared = set(re.findall(r"\b(pgl[A-Z][A-Za-z0-9]*)\s*\(", header))
    import openpgl_amd
    lib = openpgl_amd.load_library()
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared) >= 70
