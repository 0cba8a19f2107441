"""ctypes bindings of the oracle (oracle/liboracle.so) and of the reference
harness (oracle/_ref/libpglref.so).  Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libpglref.so")

F32P = ctypes.POINTER(ctypes.c_float)
U32P = ctypes.POINTER(ctypes.c_uint32)


class Config(ctypes.Structure):
    _fields_ = [(n, t) for n, t in [
        ("knnLookup", ctypes.c_uint32), ("maxSamples", ctypes.c_uint32), ("maxDepth", ctypes.c_uint32),
        ("initK", ctypes.c_uint32), ("initKappa", ctypes.c_float), ("maxK", ctypes.c_uint32),
        ("maxEMIterations", ctypes.c_uint32), ("maxKappa", ctypes.c_float),
        ("convergenceThreshold", ctypes.c_float), ("weightPrior", ctypes.c_float),
        ("meanCosinePriorStrength", ctypes.c_float), ("meanCosinePrior", ctypes.c_float),
        ("useSplitAndMerge", ctypes.c_uint32), ("splittingThreshold", ctypes.c_float),
        ("mergingThreshold", ctypes.c_float), ("partialReFit", ctypes.c_uint32),
        ("minSamplesForSplitting", ctypes.c_int32), ("minSamplesForPartialRefitting", ctypes.c_int32),
        ("minSamplesForMerging", ctypes.c_int32), ("decayOnSpatialSplit", ctypes.c_float),
        ("deterministic", ctypes.c_uint32), ("vectorWidth", ctypes.c_uint32), ("parallax", ctypes.c_uint32)]]


F32 = ctypes.c_float
A32 = F32 * 32
A32x3 = (F32 * 3) * 32
A32x2 = (F32 * 2) * 32


class RegionRecord(ctypes.Structure):
    _fields_ = [
        ("K", ctypes.c_uint32), ("weights", A32), ("kappas", A32), ("meanDirections", A32x3),
        ("normalizations", A32), ("eMinus2Kappa", A32), ("meanCosines", A32), ("distances", A32),
        ("pivot", F32 * 3),
        ("sumOfWeightedDirections", A32x3), ("sumOfWeightedStats", A32), ("sumOfDistanceWeights", A32),
        ("sumWeights", F32), ("numSamples", F32), ("overallNumSamples", F32), ("suffK", ctypes.c_uint32),
        ("normalized", ctypes.c_uint32),
        ("chiSquareMCEstimates", A32), ("splitMeans", A32x2), ("splitCovariances", A32x3),
        ("splitNumSamples", A32), ("splitSumWeights", A32), ("splitSumAssignedSamples", A32),
        ("splitK", ctypes.c_uint32),
        ("numSamplesAfterLastSplit", ctypes.c_uint64), ("numSamplesAfterLastMerge", ctypes.c_uint64),
        ("ssMean", F32 * 3), ("ssVariance", F32 * 3), ("ssNumSamples", F32), ("ssLower", F32 * 3),
        ("ssUpper", F32 * 3), ("regionLower", F32 * 3), ("regionUpper", F32 * 3),
        ("splitFlag", ctypes.c_uint32), ("valid", ctypes.c_uint32), ("numSamplesThisUpdate", ctypes.c_uint32),
        ("emIterations", ctypes.c_uint32), ("partialEmIterations", ctypes.c_uint32),
        ("numSplits", ctypes.c_uint32), ("numMerges", ctypes.c_uint32)]

    def to_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = np.array(np.ctypeslib.as_array(v)) if hasattr(v, "_length_") else v
        return out


def ptr(a, t=ctypes.c_void_p):
    return ctypes.cast(a.ctypes.data, t)


_orc = None
_ref = None


def oracle():
    global _orc
    if _orc is None:
        lib = ctypes.CDLL(ORACLE_SO)
        vp = ctypes.c_void_p
        lib.orc_field_new.restype = vp
        lib.orc_field_new.argtypes = [ctypes.POINTER(Config), ctypes.c_int]
        for n in ["orc_field_release", "orc_field_iteration", "orc_field_num_nodes", "orc_field_num_regions"]:
            getattr(lib, n).argtypes = [vp]
        lib.orc_field_iteration.restype = ctypes.c_uint32
        lib.orc_field_num_nodes.restype = ctypes.c_uint32
        lib.orc_field_num_regions.restype = ctypes.c_uint32
        lib.orc_field_update.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        lib.orc_field_set_scene_bounds.argtypes = [vp, vp, vp]
        lib.orc_field_get_scene_bounds.argtypes = [vp, vp, vp]
        lib.orc_field_export_nodes.argtypes = [vp, vp, vp, vp]
        lib.orc_field_export_region.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(RegionRecord)]
        lib.orc_field_region_canonical_map.argtypes = [vp, vp]
        lib.orc_field_sample_leaf_ids.argtypes = [vp, vp, ctypes.c_size_t, vp]
        lib.orc_field_knn.argtypes = [vp, vp]
        lib.orc_field_query.restype = ctypes.c_int
        lib.orc_field_query.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_float, vp, ctypes.c_float, vp, vp, vp, vp]
        lib.orc_field_query_pdf.restype = ctypes.c_int
        lib.orc_field_query_pdf.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_float, ctypes.c_float, vp, vp]
        lib.orc_field_scalar_sequence.restype = ctypes.c_int
        lib.orc_field_scalar_sequence.argtypes = [vp, ctypes.c_int, vp, vp, vp, ctypes.c_float, ctypes.c_int, vp, vp,
                                                  vp, vp]
        lib.orc_field_query_pdf_dirs.restype = ctypes.c_int
        lib.orc_field_query_pdf_dirs.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_float, ctypes.c_size_t, vp, vp]
        lib.orc_field_query_batch.argtypes = [vp, ctypes.c_int, ctypes.c_size_t, vp, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int]
        lib.orc_region_new.restype = vp
        lib.orc_region_new.argtypes = [ctypes.POINTER(Config), ctypes.c_int]
        lib.orc_region_release.argtypes = [vp]
        lib.orc_region_step.argtypes = [vp, vp, ctypes.c_size_t, vp, vp, ctypes.c_float, ctypes.c_int]
        lib.orc_region_export.argtypes = [vp, ctypes.POINTER(RegionRecord)]
        lib.orc_region_pdf.restype = ctypes.c_float
        lib.orc_region_pdf.argtypes = [vp, vp]
        lib.orc_integer_stats.argtypes = [vp, ctypes.c_size_t, vp, vp, vp, vp]
        lib.orc_rcp.restype = ctypes.c_float
        lib.orc_rcp.argtypes = [ctypes.c_float]
        lib.orc_rcp_width.restype = ctypes.c_float
        lib.orc_rcp_width.argtypes = [ctypes.c_float, ctypes.c_int]
        lib.orc_integer_stats_width.argtypes = [vp, ctypes.c_size_t, vp, vp, ctypes.c_int, vp]
        lib.orc_knn_pick.restype = ctypes.c_uint32
        lib.orc_knn_pick.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, ctypes.c_int]
        lib.orc_set_hg_table.argtypes = [vp]
        lib.orc_hg_rep.restype = ctypes.c_int
        lib.orc_hg_rep.argtypes = [ctypes.c_float, vp, vp, vp, vp]
        lib.orc_config_defaults.argtypes = [ctypes.POINTER(Config), ctypes.c_uint32]
        _orc = lib
    return _orc


def reference_available():
    return os.path.exists(REF_SO)


def reference():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_SO)
        vp = ctypes.c_void_p
        lib.ref_region_new.restype = vp
        lib.ref_region_new.argtypes = [ctypes.POINTER(Config)]
        lib.ref_region_release.argtypes = [vp]
        lib.ref_region_step.argtypes = [vp, vp, ctypes.c_size_t, vp, vp, ctypes.c_float, ctypes.c_int]
        lib.ref_region_export.argtypes = [vp, ctypes.POINTER(RegionRecord)]
        lib.ref_region_pdf.restype = ctypes.c_float
        lib.ref_region_pdf.argtypes = [vp, vp]
        lib.ref_region_query.restype = ctypes.c_float
        lib.ref_region_query.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_float, vp, vp, vp, vp, vp]
        lib.ref_integer_stats.argtypes = [vp, ctypes.c_size_t, vp, vp, vp, vp]
        lib.ref_stats_merge_split.argtypes = [vp, vp, ctypes.c_int, ctypes.c_float, vp, vp, vp]
        lib.ref_rcp.restype = ctypes.c_float
        lib.ref_rcp.argtypes = [ctypes.c_float]
        lib.ref_hg_table.argtypes = [vp]
        lib.ref_region_np_new.restype = vp
        lib.ref_region_np_new.argtypes = [ctypes.POINTER(Config)]
        lib.ref_region_np_release.argtypes = [vp]
        lib.ref_region_np_step.argtypes = [vp, vp, ctypes.c_size_t, vp, vp, ctypes.c_float, ctypes.c_int]
        lib.ref_region_np_export.argtypes = [vp, ctypes.POINTER(RegionRecord)]
        lib.ref_region_np_pdf.restype = ctypes.c_float
        lib.ref_region_np_pdf.argtypes = [vp, vp]
        lib.ref_region_np_query.restype = ctypes.c_float
        lib.ref_region_np_query.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_float, vp, vp, vp, vp, vp]
        lib.ref_file_region_new.restype = vp
        lib.ref_file_region_new.argtypes = [ctypes.POINTER(Config), ctypes.c_int]
        lib.ref_file_region_step.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t, vp, vp, ctypes.c_int]
        lib.ref_file_region_serialize.restype = ctypes.c_size_t
        lib.ref_file_region_serialize.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        lib.ref_file_region_deserialize.restype = vp
        lib.ref_file_region_deserialize.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int]
        lib.ref_file_region_export.argtypes = [vp, ctypes.c_int, ctypes.POINTER(RegionRecord)]
        lib.ref_file_region_pdf.restype = ctypes.c_float
        lib.ref_file_region_pdf.argtypes = [vp, ctypes.c_int, vp]
        lib.ref_file_region_release.argtypes = [vp, ctypes.c_int]
        lib.ref_file_config_serialize.restype = ctypes.c_size_t
        lib.ref_file_config_serialize.argtypes = [ctypes.POINTER(Config), ctypes.c_char_p, ctypes.c_size_t]
        _ref = lib
    return _ref


def default_config(max_samples=32000):
    c = Config()
    oracle().orc_config_defaults(ctypes.byref(c), max_samples)
    return c


class OracleRegion:
    """One region driven through fit (first step) and update (later steps)."""

    def __init__(self, cfg, precision=0):
        self.lib = oracle()
        self.h = self.lib.orc_region_new(ctypes.byref(cfg), precision)

    def step(self, samples, mean, var, n_stats, split_flag=0):
        m = np.ascontiguousarray(mean, np.float32)
        v = np.ascontiguousarray(var, np.float32)
        self.lib.orc_region_step(self.h, ptr(samples), len(samples), ptr(m), ptr(v), float(n_stats), int(split_flag))

    def record(self):
        r = RegionRecord()
        self.lib.orc_region_export(self.h, ctypes.byref(r))
        return r.to_dict()

    def pdf(self, d):
        d = np.ascontiguousarray(d, np.float32)
        return self.lib.orc_region_pdf(self.h, ptr(d))

    def __del__(self):
        try:
            self.lib.orc_region_release(self.h)
        except Exception:
            pass


class RefRegion:
    """One region of the reference's CPU_8 device, parallax-aware VMM
    (cfg.parallax = 1) or plain VMM (cfg.parallax = 0)."""

    def __init__(self, cfg):
        self.lib = reference()
        self.pre = "ref_region_" if getattr(cfg, "parallax", 1) else "ref_region_np_"
        self.h = getattr(self.lib, self.pre + "new")(ctypes.byref(cfg))

    def _fn(self, name):
        return getattr(self.lib, self.pre + name)

    def step(self, samples, mean, var, n_stats, split_flag=0):
        m = np.ascontiguousarray(mean, np.float32)
        v = np.ascontiguousarray(var, np.float32)
        self._fn("step")(self.h, ptr(samples), len(samples), ptr(m), ptr(v), float(n_stats), int(split_flag))

    def record(self):
        r = RegionRecord()
        self._fn("export")(self.h, ctypes.byref(r))
        return r.to_dict()

    def pdf(self, d):
        d = np.ascontiguousarray(d, np.float32)
        return self._fn("pdf")(self.h, ptr(d))

    def query(self, volume, pos, nrm, g, u2, pdf_dir):
        pos = np.ascontiguousarray(pos, np.float32); nrm = np.ascontiguousarray(nrm, np.float32)
        u2 = np.ascontiguousarray(u2, np.float32); pdf_dir = np.ascontiguousarray(pdf_dir, np.float32)
        pa = np.zeros(1, np.float32); d = np.zeros(3, np.float32); p = np.zeros(1, np.float32)
        integ = self._fn("query")(self.h, int(volume), ptr(pos), ptr(nrm), float(g), ptr(u2), ptr(pdf_dir),
                                  ptr(pa), ptr(d), ptr(p))
        return integ, float(pa[0]), d, float(p[0])

    def __del__(self):
        try:
            self._fn("release")(self.h)
        except Exception:
            pass

    def __del__(self):
        try:
            self.lib.ref_region_release(self.h)
        except Exception:
            pass


def region_sample_stats(samples):
    """Float sample statistics of a region's positions, as the kd build keeps them
    (mean, variance * n, n) — computed in FP64 and rounded (test input only)."""
    p = samples["pos"].astype(np.float64)
    mean = p.mean(axis=0)
    var = ((p - mean) ** 2).sum(axis=0)
    return mean.astype(np.float32), var.astype(np.float32), float(len(samples))


def _extra_sigs(lib):
    vp = ctypes.c_void_p
    lib.orc_region_query.restype = ctypes.c_float
    lib.orc_region_query.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_float, vp, vp, vp, vp, vp]
    lib.orc_stats_merge_split.argtypes = [vp, vp, ctypes.c_int, ctypes.c_float, vp, vp, vp]


def oracle_full():
    lib = oracle()
    _extra_sigs(lib)
    return lib


def load_hg_table_into_oracle():
    """The HG phase-function table comes from the golden fixture (reference data)."""
    t = np.load(os.path.join(ROOT, "tests", "golden", "hg_table.npy"))
    t = np.ascontiguousarray(t, np.float32)
    oracle().orc_set_hg_table(ptr(t))
    return t


def region_query(region, volume, pos, nrm, g, u2, pdf_dir):
    lib = oracle_full()
    pos = np.ascontiguousarray(pos, np.float32); nrm = np.ascontiguousarray(nrm, np.float32)
    u2 = np.ascontiguousarray(u2, np.float32); pdf_dir = np.ascontiguousarray(pdf_dir, np.float32)
    pa = np.zeros(1, np.float32); d = np.zeros(3, np.float32); p = np.zeros(1, np.float32)
    integ = lib.orc_region_query(region.h, int(volume), ptr(pos), ptr(nrm), float(g), ptr(u2), ptr(pdf_dir),
                                 ptr(pa), ptr(d), ptr(p))
    return integ, float(pa[0]), d, float(p[0])


class RefFileRegion:
    """A field-file storage entry (Region + Range) of a reference device of
    SIMD width W (4 or 8), trained and (de)serialised by the reference's code."""

    def __init__(self, cfg=None, width=8, data=None):
        self.lib = reference()
        self.W = int(width)
        if data is not None:
            self.h = self.lib.ref_file_region_deserialize(data, len(data), self.W)
        else:
            self.h = self.lib.ref_file_region_new(ctypes.byref(cfg), self.W)
        if not self.h:
            raise RuntimeError("reference region could not be created")

    def step(self, samples, ss13, bounds6, split_flag=0):
        ss = np.ascontiguousarray(ss13, np.float32)
        b = np.ascontiguousarray(bounds6, np.float32)
        self.lib.ref_file_region_step(self.h, self.W, ptr(samples), len(samples), ptr(ss), ptr(b), int(split_flag))

    def serialize(self):
        n = self.lib.ref_file_region_serialize(self.h, self.W, None, 0)
        buf = ctypes.create_string_buffer(n)
        self.lib.ref_file_region_serialize(self.h, self.W, buf, n)
        return buf.raw[:n]

    def record(self):
        r = RegionRecord()
        self.lib.ref_file_region_export(self.h, self.W, ctypes.byref(r))
        return r.to_dict()

    def pdf(self, d):
        d = np.ascontiguousarray(d, np.float32)
        return self.lib.ref_file_region_pdf(self.h, self.W, ptr(d))

    def __del__(self):
        try:
            self.lib.ref_file_region_release(self.h, self.W)
        except Exception:
            pass


def ref_config_bytes(cfg):
    """AdaptiveSplitAndMergeFactory::Configuration::serialize of cfg, by the reference's code."""
    lib = reference()
    n = lib.ref_file_config_serialize(ctypes.byref(cfg), None, 0)
    buf = ctypes.create_string_buffer(n)
    lib.ref_file_config_serialize(ctypes.byref(cfg), buf, n)
    return buf.raw[:n]


def oracle_path_samples(segments, offsets, max_distance=1e6, use_nee_mi_weights=False, guide_direct_light=False,
                        rr_affects_direct=True):
    """The oracle's PathSegmentDataStorage::prepareSamples applied to each path
    [offsets[p], offsets[p+1]) in order; returns (surface, volume) records in
    the order per-path propagation into one sample storage appends them."""
    from openpgl_amd.synth import SAMPLE_DTYPE
    lib = oracle()
    fn = lib.orc_path_prepare_samples
    fn.restype = ctypes.c_size_t
    fn.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                   ctypes.c_void_p, ctypes.c_size_t]
    segments = np.ascontiguousarray(segments)
    out = np.zeros(max(1, len(segments)), SAMPLE_DTYPE)
    parts = []
    base = segments.ctypes.data
    for p in range(len(offsets) - 1):
        b, e = int(offsets[p]), int(offsets[p + 1])
        n = e - b
        k = fn(base + 136 * b, n, float(max_distance), int(use_nee_mi_weights), int(guide_direct_light),
               int(rr_affects_direct), out.ctypes.data, len(out))
        if k:
            parts.append(out[:k].copy())
    allv = np.concatenate(parts) if parts else np.zeros(0, SAMPLE_DTYPE)
    vol = (allv["flags"] & 2) != 0
    return allv[~vol], allv[vol]


def oracle_pixel_estimate(segments, rr_affects_direct=True):
    lib = oracle()
    fn = lib.orc_path_pixel_estimate
    fn.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    segments = np.ascontiguousarray(segments)
    out = np.zeros(3, np.float32)
    fn(segments.ctypes.data if len(segments) else None, len(segments), int(rr_affects_direct), out.ctypes.data)
    return out


def ref_path_segment_layout():
    lib = reference()
    out = np.zeros(17, np.uint64)
    lib.ref_path_segment_layout.argtypes = [ctypes.c_void_p]
    lib.ref_path_segment_layout(out.ctypes.data)
    return out
