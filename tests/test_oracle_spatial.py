"""CPU tests pinning the oracle's kd-tree update over several frames and its
KNN region pick, from the reference's rules rather than from the oracle's
own code.  No GPU is used here.

kd update (KDTreeBuilder.h:76-118, 218-350): an independent Python
re-derivation of one Update walks the tree the oracle exported after the
previous frame and predicts the next one, using only primitives pinned
elsewhere against the reference's own code (IntegerSampleStatistics and
SampleStatistics merge/split, tests/test_oracle.py):
  * samples reaching an existing leaf carry integer statistics taken with
    the bounds of the node being partitioned (the leaf's parent; the tree
    bounds at the root), KDTreeBuilder.h:300-330;
  * a leaf splits iff depth < maxDepth and float(n_region) + count >
    maxSamples (strict, in float), :252;
  * the split plane is the mean of the axis of largest variance of the
    region's statistics merged with the new ones, tie rule
    v[v1 > v0] > v2 ? (v1 > v0) : 2, :218-230, 254-256; both halves start
    from the region's previous statistics split and decayed, :259-260;
  * a leaf that does not split merges the new statistics into its history, :286.
The prediction must equal the oracle's exported tree and region statistics
bit for bit after every frame, and each rule, mutated, must break the match.

KNN pick (KNN.h:22-28, 116-149): hand-made neighbourhoods with known
squared distances: fewer than 4 neighbours, empty slots, equal distances
ordered by slot, distances equal above the low 3 mantissa bits ordered by
slot against their true order, fused (CPU_8 / CPU_16) against unfused
(CPU_4) squared distances that order two neighbours differently, and the
in-place rescaling of the random number.
"""
import ctypes
from fractions import Fraction

import numpy as np
import pytest

from field_compare import OracleField
from oracle_ctypes import default_config, oracle, oracle_full, ptr, reference_available
from openpgl_amd.synth import SAMPLE_DTYPE, generate

F32 = np.float32


# ------------------------------------------------------------ kd update
def _istats(lib, pts, lo, hi):
    out = np.zeros(13, F32)
    p = np.ascontiguousarray(pts, F32)
    lib.orc_integer_stats(ptr(p), len(p), ptr(np.ascontiguousarray(lo, F32)), ptr(np.ascontiguousarray(hi, F32)),
                          ptr(out), None)
    return out


def _merge_split(lib, a, b, dim=0, pos=0.0):
    m, l, r = np.zeros(13, F32), np.zeros(13, F32), np.zeros(13, F32)
    lib.orc_stats_merge_split(ptr(np.ascontiguousarray(a, F32)), ptr(np.ascontiguousarray(b, F32)), int(dim),
                              float(pos), ptr(m), ptr(l), ptr(r))
    return m, l, r


def _max_dim(v):
    a = 1 if v[1] > v[0] else 0
    return a if v[a] > v[2] else 2


class Leaf:
    def __init__(self, ss, lower, upper):
        self.ss, self.lower, self.upper = ss, lower, upper


class Inner:
    def __init__(self, dim, pivot, left, right):
        self.dim, self.pivot, self.left, self.right = dim, pivot, left, right


def _export(orc):
    """the oracle's state as a nested tree (canonical export)"""
    piv, dims, cod = orc.nodes()
    recs = [orc.region(i) for i in range(orc.num_regions())]

    def build(i):
        if dims[i] == 3:
            r = recs[cod[i]]
            ss = np.concatenate([r["ssMean"], r["ssVariance"], [r["ssNumSamples"]], r["ssLower"], r["ssUpper"]])
            return Leaf(ss.astype(F32), r["regionLower"].copy(), r["regionUpper"].copy())
        return Inner(int(dims[i]), F32(piv[i]), build(cod[i]), build(cod[i] + 1))

    return build(0)


def _predict(lib, tree, tree_lo, tree_hi, pts, max_samples, max_depth, mutate=None):
    """one Update of the exported tree with the frame's positions (KDTreeBuilder.h:76-118, 233-350)"""
    ms = F32(max_samples)

    def visit(node, depth, lo, hi, p, stats):
        if len(p) == 0:
            return node
        if isinstance(node, Leaf):
            total = F32(node.ss[6]) + F32(len(p))
            over = total >= ms if mutate == "ge" else total > ms
            if depth < max_depth and over:
                base = np.zeros(13, F32) if mutate == "no_history" else node.ss
                merged, _, _ = _merge_split(lib, base, stats)
                var = (merged[3:6] / merged[6]).astype(F32)
                dim = _max_dim(var)
                pos = F32(merged[dim])
                _, left_ss, right_ss = _merge_split(lib, node.ss, node.ss, dim, pos)
                lu, rl = node.upper.copy(), node.lower.copy()
                lu[dim] = pos
                rl[dim] = pos
                node = Inner(dim, pos, Leaf(left_ss, node.lower.copy(), lu), Leaf(right_ss, rl, node.upper.copy()))
            else:
                base = np.zeros(13, F32) if mutate == "no_history" else node.ss
                node.ss, _, _ = _merge_split(lib, base, stats)
                return node
        d, pv = node.dim, node.pivot
        left = p[:, d] < pv
        llo, lhi, rlo, rhi = lo.copy(), hi.copy(), lo.copy(), hi.copy()
        lhi[d] = pv
        rlo[d] = pv
        sl = sr = None
        if isinstance(node.left, Leaf) or isinstance(node.right, Leaf):
            if mutate == "child_bounds":
                sl, sr = _istats(lib, p[left], llo, lhi), _istats(lib, p[~left], rlo, rhi)
            else:
                sl, sr = _istats(lib, p[left], lo, hi), _istats(lib, p[~left], lo, hi)
        node.left = visit(node.left, depth + 1, llo, lhi, p[left], sl)
        node.right = visit(node.right, depth + 1, rlo, rhi, p[~left], sr)
        return node

    st = _istats(lib, pts, tree_lo, tree_hi) if isinstance(tree, Leaf) else None
    return visit(tree, 1, tree_lo.copy(), tree_hi.copy(), pts, st)


def _same(a, b):
    if isinstance(a, Leaf) != isinstance(b, Leaf):
        return False
    if isinstance(a, Leaf):
        return (a.ss.view(np.uint32).tolist() == b.ss.view(np.uint32).tolist()
                and np.array_equal(a.lower.view(np.uint32), b.lower.view(np.uint32))
                and np.array_equal(a.upper.view(np.uint32), b.upper.view(np.uint32)))
    return (a.dim == b.dim and F32(a.pivot).view(np.uint32) == F32(b.pivot).view(np.uint32)
            and _same(a.left, b.left) and _same(a.right, b.right))


def _count_leaves(t):
    return 1 if isinstance(t, Leaf) else _count_leaves(t.left) + _count_leaves(t.right)


def _frames_check(frames, max_samples, max_depth=32, mutations=()):
    lib = oracle_full()
    cfg = default_config(max_samples)
    cfg.maxDepth = max_depth
    cfg.knnLookup = 0
    orc = OracleField(cfg)
    orc.update(frames[0])
    lo, hi = np.zeros(3, F32), np.zeros(3, F32)
    lib.orc_field_get_scene_bounds(orc.h, ptr(lo), ptr(hi))
    broken = {m: False for m in mutations}
    for f in range(1, len(frames)):
        before = _export(orc)
        pts = np.ascontiguousarray(frames[f]["pos"], F32)
        orc.update(frames[f])
        after = _export(orc)
        assert _same(_predict(lib, _export_copy(before), lo, hi, pts, max_samples, max_depth), after), f
        for m in mutations:
            if not _same(_predict(lib, _export_copy(before), lo, hi, pts, max_samples, max_depth, m), after):
                broken[m] = True
    return orc, after, broken


def _export_copy(t):
    if isinstance(t, Leaf):
        return Leaf(t.ss.copy(), t.lower.copy(), t.upper.copy())
    return Inner(t.dim, t.pivot, _export_copy(t.left), _export_copy(t.right))


def test_kd_update_three_frames_with_splits_of_existing_leaves():
    frames = [generate(20000, seed=31, frame=f) for f in range(4)]
    orc, tree, broken = _frames_check(frames, 1500, mutations=("child_bounds", "no_history"))
    print("leaves", _count_leaves(tree), "mutations detected", broken)
    assert _count_leaves(tree) > 20
    assert all(broken.values()), broken


def test_kd_update_depth_limit():
    # all frames concentrated in a tiny ball: leaves stop splitting at maxDepth
    frames = []
    for f in range(3):
        s = generate(8000, seed=33, frame=f)
        s["pos"] = (F32(0.5) + (s["pos"] - F32(0.5)) * F32(1e-3)).astype(F32)
        frames.append(s)
    orc, tree, _ = _frames_check(frames, 500, max_depth=5)
    assert _count_leaves(tree) <= 2 ** 4


def _one_leaf_samples(n, seed):
    s = generate(n, seed=seed)
    return s


def test_kd_split_threshold_is_strict():
    """float(n) + count > maxSamples: exactly maxSamples samples in total do
    not split a leaf, one more does (first frame and a later frame)."""
    lib = oracle_full()
    ms = 1000

    def field():
        cfg = default_config(ms)
        cfg.knnLookup = 0
        f = OracleField(cfg)
        f.set_scene_bounds(np.zeros(3, F32), np.ones(3, F32))
        return f

    # first frame: exactly maxSamples -> root stays a leaf; maxSamples + 1 -> split
    a = field()
    a.update(generate(ms, seed=1))
    assert a.num_regions() == 1
    b = field()
    b.update(generate(ms + 1, seed=1))
    assert b.num_regions() == 2
    # later frame: 600 + 400 == maxSamples -> no split; the next sample splits
    c = field()
    c.update(generate(600, seed=2))
    c.update(generate(400, seed=3))
    assert c.num_regions() == 1
    assert c.region(0)["ssNumSamples"] == 1000.0
    c.update(generate(1, seed=4))
    assert c.num_regions() == 2
    # the predictor with ">=" would have split at the boundary
    lo, hi = np.zeros(3, F32), np.ones(3, F32)
    d = field()
    d.update(generate(600, seed=2))
    before = _export(d)
    pts = np.ascontiguousarray(generate(400, seed=3)["pos"], F32)
    d.update(generate(400, seed=3))
    after = _export(d)
    assert _same(_predict(lib, _export_copy(before), lo, hi, pts, ms, 32), after)
    assert not _same(_predict(lib, _export_copy(before), lo, hi, pts, ms, 32, "ge"), after)


# ------------------------------------------------------------ KNN pick
def _pick(ids, pts, size, p, u, fused=True):
    lib = oracle()
    ids = np.ascontiguousarray(ids, np.uint32)
    pts = np.ascontiguousarray(pts, F32)
    p = np.ascontiguousarray(p, F32)
    uu = np.array([u], F32)
    r = lib.orc_knn_pick(ptr(ids), ptr(pts), int(size), ptr(p), ptr(uu), int(fused))
    return int(r), F32(uu[0])


def _neigh(points, ids):
    """8 slots: pts24 = x[8] y[8] z[8]; unused slots hold ~0 ids"""
    pts = np.zeros(24, F32)
    idv = np.full(8, 0xFFFFFFFF, np.uint32)
    for s, (pt, i) in enumerate(zip(points, ids)):
        if pt is None:
            continue
        pts[s], pts[8 + s], pts[16 + s] = pt
        idv[s] = i
    return idv, pts


def test_knn_fewer_than_four_neighbours_and_u_rescaling():
    idv, pts = _neigh([(1, 0, 0), (2, 0, 0)], [10, 20])
    p = (0, 0, 0)
    # draw(): sel = floor(u * size), u' = (u - sel / size) * size
    assert _pick(idv, pts, 2, p, 0.3) == (10, F32((F32(0.3) - F32(0) / F32(2)) * F32(2)))
    assert _pick(idv, pts, 2, p, 0.75) == (20, F32((F32(0.75) - F32(1) / F32(2)) * F32(2)))
    assert _pick(idv, pts, 1, p, 0.99)[0] == 10  # one neighbour
    # u just below 1 picks the last of min(4, size)
    assert _pick(idv, pts, 2, p, np.nextafter(F32(1), F32(0)))[0] == 20


def test_knn_only_the_four_nearest_and_empty_slots():
    # slots hold distances 3, 1, -, 5, 2, -, 4, 6 from p (two empty slots)
    d = [3, 1, None, 5, 2, None, 4, 6]
    idv, pts = _neigh([None if x is None else (x, 0, 0) for x in d], [100 + s for s in range(8)])
    p = (0, 0, 0)
    order = [101, 104, 100, 106]  # ids of the 4 nearest in distance order
    for k in range(4):
        u = (k + 0.5) / 4
        assert _pick(idv, pts, 6, p, u)[0] == order[k]


def test_knn_equal_distances_ordered_by_slot():
    # slots 2 and 5 at exactly the same distance (mirror points): slot 2 first
    pts_list = [None] * 8
    pts_list[5] = (0.0, 0.75, 0.0)
    pts_list[2] = (0.0, -0.75, 0.0)
    pts_list[0] = (3.0, 0, 0)
    pts_list[7] = (4.0, 0, 0)
    idv, pts = _neigh(pts_list, [200 + s for s in range(8)])
    assert _pick(idv, pts, 4, (0, 0, 0), 0.1)[0] == 202
    assert _pick(idv, pts, 4, (0, 0, 0), 0.3)[0] == 205


def _f32_round(q):
    """round-to-nearest-even of an exact rational to float32"""
    if q == 0:
        return F32(0)
    sign = -1 if q < 0 else 1
    q = abs(q)
    e = q.numerator.bit_length() - q.denominator.bit_length()
    if Fraction(2) ** e > q:
        e -= 1
    e = max(e, -126)
    scale = Fraction(2) ** (e - 23)
    m = q / scale
    fl = m.numerator // m.denominator
    rem = m - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return F32(sign * float(fl * scale))


def _d2(dx, dy, dz, fused):
    x, y, z = (Fraction(float(v)) for v in (dx, dy, dz))
    if fused:  # madd(x, x, madd(y, y, z * z)) with FMA
        zz = _f32_round(z * z)
        inner = _f32_round(y * y + Fraction(float(zz)))
        return _f32_round(x * x + Fraction(float(inner)))
    zz, yy, xx = _f32_round(z * z), _f32_round(y * y), _f32_round(x * x)
    inner = _f32_round(Fraction(float(yy)) + Fraction(float(zz)))
    return _f32_round(Fraction(float(xx)) + Fraction(float(inner)))


def test_knn_low_mantissa_bits_replaced_by_slot():
    """two squared distances equal above the low 3 mantissa bits: the one in
    the lower slot is picked first even when it is farther"""
    rng = np.random.default_rng(5)
    for _ in range(100000):
        a = F32(rng.uniform(0.5, 0.6))
        b = np.nextafter(a, F32(1))
        da, db = _d2(a, 0, 0, True), _d2(b, 0, 0, True)
        ka, kb = int(da.view(np.uint32)), int(db.view(np.uint32))
        if da < db and (ka >> 3) == (kb >> 3):
            break
    else:
        pytest.fail("no candidate pair found")
    pts_list = [None] * 8
    pts_list[6] = (float(a), 0, 0)  # nearer, higher slot
    pts_list[1] = (float(b), 0, 0)  # farther, lower slot
    idv, pts = _neigh(pts_list, [300 + s for s in range(8)])
    assert _pick(idv, pts, 2, (0, 0, 0), 0.1)[0] == 301
    assert _pick(idv, pts, 2, (0, 0, 0), 0.9)[0] == 306


def test_knn_fused_and_unfused_squared_distances():
    """a neighbourhood where the fused (CPU_8/CPU_16) and the unfused (CPU_4)
    squared distances order two neighbours differently; the oracle picks
    accordingly for each device type"""
    # neighbours at permuted coordinates of one offset: the same exact squared
    # distance, rounded differently by the two evaluation orders
    rng = np.random.default_rng(7)
    found = None
    for _ in range(20000):
        da_ = rng.uniform(-1, 1, 3).astype(F32)
        for perm in [(1, 0, 2), (0, 2, 1), (2, 1, 0), (1, 2, 0), (2, 0, 1)]:
            db_ = da_[list(perm)]
            fa, fb = _d2(*da_, True), _d2(*db_, True)
            ua, ub = _d2(*da_, False), _d2(*db_, False)
            kf = (int(fa.view(np.uint32)) >> 3, int(fb.view(np.uint32)) >> 3)
            ku = (int(ua.view(np.uint32)) >> 3, int(ub.view(np.uint32)) >> 3)
            if kf[0] < kf[1] and ku[0] > ku[1]:
                found = (da_, db_)
                break
        if found is not None:
            break
    assert found is not None
    da_, db_ = found
    pts_list = [None] * 8
    pts_list[0] = tuple(float(v) for v in da_)
    pts_list[1] = tuple(float(v) for v in db_)
    idv, pts = _neigh(pts_list, [400, 401])
    p = (0.0, 0.0, 0.0)
    assert _pick(idv, pts, 2, p, 0.1, fused=True)[0] == 400
    assert _pick(idv, pts, 2, p, 0.1, fused=False)[0] == 401


# ------------------------------------------------------------ per-width rcp
def _ref_isa(width):
    import os
    from oracle_ctypes import ROOT
    path = os.path.join(ROOT, "oracle", "_ref", "libpglref_w%d.so" % width)
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.ref_isa_rcp.restype = ctypes.c_float
    lib.ref_isa_rcp.argtypes = [ctypes.c_float]
    lib.ref_isa_integer_stats.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p]
    return lib


def _has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
        return "avx512f" in flags and "avx512vl" in flags
    except OSError:
        return False


@pytest.mark.parametrize("width", [4, 16])
def test_rcp_per_device_width_against_reference_build(width):
    if width == 16 and not _has_avx512():
        pytest.skip("host CPU without AVX-512VL")
    ref = _ref_isa(width)
    if ref is None:
        pytest.skip("reference tree not built here (oracle/_ref)")
    lib = oracle()
    rng = np.random.default_rng(width)
    xs = np.concatenate([np.exp(rng.uniform(-30, 30, 20000)), rng.uniform(1e-3, 2.0, 20000)]).astype(F32)
    got = np.array([lib.orc_rcp_width(float(x), width) for x in xs], F32)
    exp = np.array([ref.ref_isa_rcp(float(x)) for x in xs], F32)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    # the device types really differ from CPU_8
    w8 = np.array([lib.orc_rcp_width(float(x), 8) for x in xs], F32)
    assert np.mean(w8 != got) > 0.05
    # and the integer statistics built on them
    for k in range(5):
        n = 400
        lo = rng.uniform(-1, 0, 3).astype(F32)
        hi = (lo + rng.uniform(0.1, 3, 3)).astype(F32)
        p = rng.uniform(lo, hi, (n, 3)).astype(F32)
        a, b = np.zeros(13, np.int64), np.zeros(13, np.int64)
        lib.orc_integer_stats_width(ptr(p), n, ptr(lo), ptr(hi), width, ptr(a))
        ref.ref_isa_integer_stats(ptr(p), n, ptr(lo), ptr(hi), ptr(b))
        assert np.array_equal(a, b)


@pytest.mark.parametrize("width", [4, 8, 16])
def test_rcp_per_width_is_a_refined_reciprocal(width):
    if width == 16 and not _has_avx512():
        pytest.skip("host CPU without AVX-512VL")
    lib = oracle()
    xs = np.random.default_rng(1).uniform(1e-3, 1e3, 4000).astype(F32)
    r = np.array([lib.orc_rcp_width(float(x), width) for x in xs], np.float64)
    # rcpss (12 bits) or rcp14 (14 bits) and one Newton step
    assert np.max(np.abs(r * xs - 1.0)) < 2.0 ** -21
