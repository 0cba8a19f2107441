"""Helpers comparing a GPU field (through the C ABI) with the oracle field.
Test infrastructure only."""
import ctypes

import numpy as np

from oracle_ctypes import RegionRecord, oracle, ptr


class OracleField:
    def __init__(self, cfg, precision=0):
        self.lib = oracle()
        self.h = self.lib.orc_field_new(ctypes.byref(cfg), precision)

    def update(self, samples, threads=8):
        s = np.ascontiguousarray(samples)
        self.lib.orc_field_update(self.h, ptr(s), len(s), threads)

    def set_scene_bounds(self, lo, hi):
        lo = np.asarray(lo, np.float32); hi = np.asarray(hi, np.float32)
        self.lib.orc_field_set_scene_bounds(self.h, ptr(lo), ptr(hi))

    def nodes(self):
        n = self.lib.orc_field_num_nodes(self.h)
        piv = np.zeros(n, np.float32); dims = np.zeros(n, np.uint32); cod = np.zeros(n, np.uint32)
        self.lib.orc_field_export_nodes(self.h, ptr(piv), ptr(dims), ptr(cod))
        return piv, dims, cod

    def num_regions(self):
        return self.lib.orc_field_num_regions(self.h)

    def region(self, i):
        r = RegionRecord()
        self.lib.orc_field_export_region(self.h, i, ctypes.byref(r))
        return r.to_dict()

    def leaf_ids(self, pos):
        p = np.ascontiguousarray(pos, np.float32)
        out = np.zeros(len(p), np.uint32)
        self.lib.orc_field_sample_leaf_ids(self.h, ptr(p), len(p), ptr(out))
        return out

    def knn(self):
        out = np.zeros(self.num_regions() * 8, np.uint32)
        self.lib.orc_field_knn(self.h, ptr(out))
        return out.reshape(-1, 8)

    def query_batch(self, volume, pos, nrm, u2, u1=None, g=None, threads=8):
        n = len(pos)
        pos = np.ascontiguousarray(pos, np.float32); nrm = np.ascontiguousarray(nrm, np.float32)
        u2 = np.ascontiguousarray(u2, np.float32)
        u1 = None if u1 is None else np.ascontiguousarray(u1, np.float32)
        g = None if g is None else np.ascontiguousarray(g, np.float32)
        d = np.zeros((n, 3), np.float32); pdf = np.zeros(n, np.float32); ids = np.zeros(n, np.uint32)
        integ = np.zeros(n, np.float32)
        self.lib.orc_field_query_batch(self.h, int(volume), n, ptr(pos), ptr(nrm), ptr(g) if g is not None else None,
                                       ptr(u2), ptr(u1) if u1 is not None else None, ptr(d), ptr(pdf), ptr(ids),
                                       ptr(integ), threads)
        return d, pdf, ids, integ

    def query_pdf(self, volume, pos, nrm, g, u1, d):
        out = np.zeros(1, np.float32)
        p = np.ascontiguousarray(pos, np.float32); nn = np.ascontiguousarray(nrm, np.float32)
        dd = np.ascontiguousarray(d, np.float32)
        ok = self.lib.orc_field_query_pdf(self.h, int(volume), ptr(p), ptr(nn), float(g), float(u1), ptr(dd), ptr(out))
        return ok, float(out[0])

    def scalar_sequence(self, volume, pos, u1, nrm, g, products, u2, eval_dir):
        """Init + products + SamplePDF + PDF/IncomingRadiancePDF at eval_dir:
        (ok, u1 after Init, id, dir, pdf(dir), pdf(eval_dir), pdfLi(eval_dir), integral)."""
        p = np.ascontiguousarray(pos, np.float32)
        u = np.array([u1], np.float32)
        nn = np.ascontiguousarray(nrm, np.float32)
        uu = np.ascontiguousarray(u2, np.float32)
        e = np.ascontiguousarray(eval_dir, np.float32)
        out = np.zeros(7, np.float32)
        i = np.zeros(1, np.uint32)
        ok = self.lib.orc_field_scalar_sequence(self.h, int(volume), ptr(p), ptr(u), ptr(nn), float(g), int(products),
                                                ptr(uu), ptr(e), ptr(out), ptr(i))
        return bool(ok), float(u[0]), int(i[0]), out[:3].copy(), float(out[3]), float(out[4]), float(out[5]), float(out[6])

    def pdf_at_dirs(self, volume, pos, nrm, g, dirs):
        p = np.ascontiguousarray(pos, np.float32)
        nn = np.ascontiguousarray(nrm, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        out = np.zeros(len(d), np.float32)
        ok = self.lib.orc_field_query_pdf_dirs(self.h, int(volume), ptr(p), ptr(nn), float(g), len(d), ptr(d), ptr(out))
        return bool(ok), out

    def __del__(self):
        try:
            self.lib.orc_field_release(self.h)
        except Exception:
            pass


def vmm_pdf(rec, dirs):
    """pdf of an exported region record's mixture at unit directions (FP64 numpy)."""
    K = int(rec["K"])
    mu = rec["meanDirections"][:K].astype(np.float64)
    kap = rec["kappas"][:K].astype(np.float64)
    w = rec["weights"][:K].astype(np.float64)
    nrm = rec["normalizations"][:K].astype(np.float64)
    c = dirs.astype(np.float64) @ mu.T
    return (w * nrm * np.exp(kap * np.minimum(c - 1.0, 0.0))).sum(axis=1)


def unit_dirs(n, seed=7):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def precision_report(records_test, records_ref, dirs):
    """Errors of a set of region records against reference records (same
    region order): per region the pdf relative error at dirs (median and 99th
    percentile over directions); where both have the same K and the same EM /
    split / merge history, per component (matched by a minimum-cost assignment)
    the weight absolute error, the mean direction angle (radians) and the mean
    cosine relative error.  Returns a
    dict of medians / 99th percentiles over regions and the fraction of
    regions whose discrete history agrees."""
    pdf_med, pdf_p99, w_err, ang, mc_rel = [], [], [], [], []
    agree = 0
    valid = 0
    for a, b in zip(records_test, records_ref):
        if not (a["valid"] and b["valid"]):
            continue
        valid += 1
        pa, pb = vmm_pdf(a, dirs), vmm_pdf(b, dirs)
        rel = np.abs(pa - pb) / pb
        pdf_med.append(np.median(rel))
        pdf_p99.append(np.percentile(rel, 99))
        same = (a["K"] == b["K"] and a["emIterations"] == b["emIterations"] and a["numSplits"] == b["numSplits"]
                and a["numMerges"] == b["numMerges"] and a["partialEmIterations"] == b["partialEmIterations"])
        if not same:
            continue
        agree += 1
        K = int(a["K"])
        # components may sit in other slots (merge choices at tied costs):
        # match them by a minimum-cost assignment on angle and weight
        from scipy.optimize import linear_sum_assignment
        da = a["meanDirections"][:K].astype(np.float64)
        db = b["meanDirections"][:K].astype(np.float64)
        cosm = np.clip(da @ db.T, -1.0, 1.0)
        cost = np.arccos(cosm) + 10.0 * np.abs(a["weights"][:K][:, None] - b["weights"][:K][None, :])
        ia, ib = linear_sum_assignment(cost)
        w_err.extend(np.abs(a["weights"][:K][ia].astype(np.float64) - b["weights"][:K][ib]).tolist())
        ang.extend(np.arccos(cosm[ia, ib]).tolist())
        mb = b["meanCosines"][:K][ib].astype(np.float64)
        mc_rel.extend((np.abs(a["meanCosines"][:K][ia] - mb) / np.maximum(mb, 1e-30)).tolist())
    q = lambda v, p: float(np.percentile(v, p)) if len(v) else float("nan")
    return {"regions": valid, "history_agrees": agree / max(valid, 1),
            "pdf_rel_median": q(pdf_med, 50), "pdf_rel_p99_of_regions_p99": q(pdf_p99, 99),
            "weight_abs_median": q(w_err, 50), "weight_abs_p99": q(w_err, 99),
            "mean_dir_angle_median": q(ang, 50), "mean_dir_angle_p99": q(ang, 99),
            "mean_cosine_rel_median": q(mc_rel, 50), "mean_cosine_rel_p99": q(mc_rel, 99)}


def _oracle_kind_state(ff, orc, is_surface, fcfg, iteration, max_samples, knn):
    piv, dims, cod = orc.nodes()
    nodes = np.zeros(len(piv), dtype=[("pivot", "<f4"), ("bits", "<u4")])
    nodes["pivot"] = piv
    nodes["bits"] = (dims.astype(np.uint32) << 30) | cod
    regions = []
    for i in range(orc.num_regions()):
        r = orc.region(i)
        r["splitMeans"] = np.zeros((32, 2), np.float32)
        r["rangeBegin"] = r["rangeEnd"] = 0
        regions.append(r)
    lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
    orc.lib.orc_field_get_scene_bounds(orc.h, lo.ctypes.data, hi.ctypes.data)
    st = ff.empty_field(is_surface, fcfg, 100, max_samples, 32, knn=knn)
    st.update({"iteration": iteration, "isSceneBoundsSet": 1, "sceneBounds": (*lo, *hi), "initialized": 1,
               "treeInit": 1, "treeBounds": (*lo, *hi), "nodes": nodes, "regions": regions, "knnBuilt": int(knn),
               "knnPoints": (np.array([[*r["ssMean"], 0.0] for r in regions], np.float32) if knn
                             else np.zeros((0, 4), np.float32))})
    return st


def oracle_field_file(cfg, iteration, max_samples, surface=None, volume=None, W=8):
    """Bytes of a field file (the reference's stream layout, restated in
    tests/field_file.py) holding the oracle's CPU-trained surface and/or
    volume fields, so the GPU library can serve exactly the oracle's mixtures."""
    import field_file as ff
    fcfg = {"initK": cfg.initK, "initKappa": cfg.initKappa, "maxK": cfg.maxK, "maxEMIterrations": cfg.maxEMIterations,
            "maxKappa": cfg.maxKappa, "maxMeanCosine": float(np.float32(1.0 / np.tanh(np.float32(cfg.maxKappa))
                                                                         - 1.0 / np.float32(cfg.maxKappa))),
            "convergenceThreshold": cfg.convergenceThreshold, "weightPrior": cfg.weightPrior,
            "meanCosinePriorStrength": cfg.meanCosinePriorStrength, "meanCosinePrior": cfg.meanCosinePrior,
            "splittingThreshold": cfg.splittingThreshold, "mergingThreshold": cfg.mergingThreshold,
            "partialReFit": cfg.partialReFit, "maxSplitItr": 1, "minSamplesForSplitting": cfg.minSamplesForSplitting,
            "minSamplesForMerging": cfg.minSamplesForMerging,
            "minSamplesForPartialRefitting": cfg.minSamplesForPartialRefitting}
    knn = bool(cfg.knnLookup)
    surf = (_oracle_kind_state(ff, surface, True, fcfg, iteration, max_samples, knn) if surface is not None
            else ff.empty_field(True, fcfg, 100, max_samples, 32, knn=knn))
    vol = (_oracle_kind_state(ff, volume, False, fcfg, iteration, max_samples, knn) if volume is not None
           else ff.empty_field(False, fcfg, 100, max_samples, 32, knn=knn))
    return ff.write_file({"iteration": iteration, "surface": surf, "volume": vol}, W)


def load_oracle_field_on_gpu(dev, path, cfg, iteration, max_samples, surface=None, volume=None):
    import openpgl_amd as pgl
    with open(path, "wb") as fh:
        fh.write(oracle_field_file(cfg, iteration, max_samples, surface, volume))
    return pgl.Field.from_file(dev, path)


def canonical_ids(orc):
    """storage -> canonical region index map of an oracle field (the file holds canonical order)"""
    s2c = np.zeros(orc.num_regions(), np.uint32)
    orc.lib.orc_field_region_canonical_map(orc.h, s2c.ctypes.data)
    return s2c
