"""Open PGL 0.6.0 field files restated in Python (test infrastructure).

Reader and writer of the stream layout written by the reference's
SurfaceVolumeField::storeToFile (field/SurfaceVolumeField.h:197-213) and read
by Device::newFieldFromFile (device/Device.h:217-271), member by member:

  Field::serialize                   field/Field.h:240-267
  weighted-EM Configuration          directional/vmm/ParallaxAwareVonMisesFisherWeightedEMFactory.h:1335-1350
  split-and-merge Configuration      directional/vmm/AdaptiveSplitandMergeFactory.h:246-259
  KDTreePartitionBuilder::Settings   spatial/kdtree/KDTreeBuilder.h:377-382
  KDTree / KDNode                    spatial/kdtree/KDTree.h:592-602, 136-140
  Region + Range                     spatial/Region.h:63-71, data/Range.h:29-32
  mixture                            directional/vmm/ParallaxAwareVonMisesFisherMixture.h:386-401
  sufficient statistics              directional/vmm/ParallaxAwareVonMisesFisherWeightedEMFactory.h:346-358
  split statistics                   directional/vmm/VMMChiSquareComponentSplitter.h:871-884
  training statistics                directional/vmm/AdaptiveSplitandMergeFactory.h:141-149
  SampleStatistics                   data/SampleStatistics.h:157-163
  KNearestRegionsSearchTree          spatial/KNN.h:320-331

Per-component arrays are vfloat<W> blocks: block b holds lanes [bW, bW+W) of
each array of the block in member order.  Region records use the key names of
the library's PGLRegionRecord export.
"""
import struct

import numpy as np

HEADER = b"OPENPGL_0.6.0_FIELD\x00"

VMM_ARRAYS = [("weights", None), ("kappas", None), ("meanDirections", 0), ("meanDirections", 1),
              ("meanDirections", 2), ("normalizations", None), ("eMinus2Kappa", None), ("meanCosines", None),
              ("distances", None)]
SUFF_ARRAYS = [("sumOfWeightedDirections", 0), ("sumOfWeightedDirections", 1), ("sumOfWeightedDirections", 2),
               ("sumOfWeightedStats", None), ("sumOfDistanceWeights", None)]
SPLIT_ARRAYS = [("chiSquareMCEstimates", None), ("splitMeans", 0), ("splitMeans", 1), ("splitCovariances", 0),
                ("splitCovariances", 1), ("splitCovariances", 2), ("splitNumSamples", None),
                ("splitSumWeights", None), ("splitSumAssignedSamples", None)]
ARRAY_SHAPES = {"meanDirections": (32, 3), "sumOfWeightedDirections": (32, 3), "splitMeans": (32, 2),
                "splitCovariances": (32, 3)}

# bytes of one storage entry (Region + Range); the same for every W since the
# 32 component slots are always stored in full
REGION_BYTES = (1 + 9 * 128 + 8 + 12 + 24 + 5 * 128 + 12 + 8 + 1 + 9 * 128 + 8 + 16 + 52 + 1 + 16)


class _In:
    def __init__(self, data, offset=0):
        self.b = memoryview(data)
        self.o = offset

    def take(self, n):
        if self.o + n > len(self.b):
            raise ValueError("field file truncated at byte %d" % self.o)
        v = self.b[self.o:self.o + n]
        self.o += n
        return bytes(v)

    def unpack(self, fmt):
        v = struct.unpack("<" + fmt, self.take(struct.calcsize("<" + fmt)))
        return v if len(v) > 1 else v[0]

    def floats(self, n):
        return np.frombuffer(self.take(4 * n), np.float32).copy()


class _Out:
    def __init__(self):
        self.parts = []

    def pack(self, fmt, *v):
        self.parts.append(struct.pack("<" + fmt, *v))

    def floats(self, a):
        self.parts.append(np.ascontiguousarray(a, np.float32).tobytes())

    def data(self):
        return b"".join(self.parts)


def _empty_arrays(names):
    out = {}
    for name, _ in names:
        out.setdefault(name, np.zeros(ARRAY_SHAPES.get(name, (32,)), np.float32))
    return out


def _read_blocks(inp, W, names, rec):
    arrays = _empty_arrays(names)
    for b in range(32 // W):
        for name, col in names:
            v = inp.floats(W)
            if col is None:
                arrays[name][b * W:(b + 1) * W] = v
            else:
                arrays[name][b * W:(b + 1) * W, col] = v
    rec.update(arrays)


def _write_blocks(out, W, names, rec):
    for b in range(32 // W):
        for name, col in names:
            a = np.asarray(rec[name], np.float32)
            out.floats(a[b * W:(b + 1) * W] if col is None else a[b * W:(b + 1) * W, col])


def read_region(inp, W):
    """One storage entry: Region::serialize then Range::serialize."""
    r = {}
    r["valid"] = inp.unpack("B")
    _read_blocks(inp, W, VMM_ARRAYS, r)
    r["K"] = inp.unpack("Q")
    r["pivot"] = inp.floats(3)
    r["regionLower"] = inp.floats(3)
    r["regionUpper"] = inp.floats(3)
    _read_blocks(inp, W, SUFF_ARRAYS, r)
    r["sumWeights"], r["numSamples"], r["overallNumSamples"] = inp.unpack("3f")
    r["suffK"] = inp.unpack("Q")
    r["normalized"] = inp.unpack("B")
    _read_blocks(inp, W, SPLIT_ARRAYS, r)
    r["splitK"] = inp.unpack("Q")
    r["numSamplesAfterLastSplit"], r["numSamplesAfterLastMerge"] = inp.unpack("2Q")
    r["ssMean"] = inp.floats(3)
    r["ssVariance"] = inp.floats(3)
    r["ssNumSamples"] = inp.unpack("f")
    r["ssLower"] = inp.floats(3)
    r["ssUpper"] = inp.floats(3)
    r["splitFlag"] = inp.unpack("B")
    r["rangeBegin"], r["rangeEnd"] = inp.unpack("2Q")
    return r


def write_region(out, r, W):
    out.pack("B", int(r["valid"]))
    _write_blocks(out, W, VMM_ARRAYS, r)
    out.pack("Q", int(r["K"]))
    out.floats(r["pivot"])
    out.floats(r["regionLower"])
    out.floats(r["regionUpper"])
    _write_blocks(out, W, SUFF_ARRAYS, r)
    out.pack("3f", r["sumWeights"], r["numSamples"], r["overallNumSamples"])
    out.pack("Q", int(r["suffK"]))
    out.pack("B", int(r["normalized"]))
    _write_blocks(out, W, SPLIT_ARRAYS, r)
    out.pack("Q", int(r["splitK"]))
    out.pack("2Q", int(r["numSamplesAfterLastSplit"]), int(r["numSamplesAfterLastMerge"]))
    out.floats(r["ssMean"])
    out.floats(r["ssVariance"])
    out.pack("f", r["ssNumSamples"])
    out.floats(r["ssLower"])
    out.floats(r["ssUpper"])
    out.pack("B", int(r["splitFlag"]))
    out.pack("2Q", int(r.get("rangeBegin", 0)), int(r.get("rangeEnd", 0)))


CONFIG_FIELDS = [("initK", "Q"), ("initKappa", "f"), ("maxK", "Q"), ("maxEMIterrations", "Q"), ("maxKappa", "f"),
                 ("maxMeanCosine", "f"), ("convergenceThreshold", "f"), ("weightPrior", "f"),
                 ("meanCosinePriorStrength", "f"), ("meanCosinePrior", "f"), ("splittingThreshold", "f"),
                 ("mergingThreshold", "f"), ("partialReFit", "B"), ("maxSplitItr", "i"),
                 ("minSamplesForSplitting", "i"), ("minSamplesForMerging", "i"),
                 ("minSamplesForPartialRefitting", "i")]
FIELD_HEAD = [("isSurface", "B"), ("decayOnSpatialSplit", "f"), ("iteration", "Q"), ("totalSPP", "Q"),
              ("deterministic", "B"), ("fitRegions", "B"), ("isSceneBoundsSet", "B"), ("sceneBounds", "6f"),
              ("initialized", "B"), ("timeLastUpdate", "f"), ("timeLastUpdateCopySamples", "f"),
              ("timeLastUpdateSpatialStructureUpdate", "f"), ("timeLastUpdateDirectionalDistriubtionUpdate", "f")]


def read_config(inp):
    return {k: inp.unpack(f) for k, f in CONFIG_FIELDS}


def write_config(out, cfg):
    for k, f in CONFIG_FIELDS:
        out.pack(f, cfg[k])


def read_field(inp, W):
    """Field::deserialize of one kind (surface or volume)."""
    f = {}
    for k, fmt in FIELD_HEAD:
        f[k] = inp.unpack(fmt)
    f["config"] = read_config(inp)
    f["minSamples"], f["maxSamples"], f["maxDepth"] = inp.unpack("3Q")
    f["treeInit"] = inp.unpack("B")
    f["treeBounds"] = inp.unpack("6f")
    n = inp.unpack("Q")
    raw = np.frombuffer(inp.take(8 * n), dtype=np.dtype([("pivot", "<f4"), ("bits", "<u4")])).copy()
    f["nodes"] = raw
    nr = inp.unpack("Q")
    f["regions"] = [read_region(inp, W) for _ in range(nr)]
    f["useStochasticNNLookUp"] = inp.unpack("B")
    f["knnBuilt"] = inp.unpack("B")
    if f["knnBuilt"]:
        npts = inp.unpack("I")
        f["knnPoints"] = inp.floats(4 * npts).reshape(npts, 4)
    else:
        f["knnPoints"] = np.zeros((0, 4), np.float32)
    return f


def write_field(out, f, W):
    for k, fmt in FIELD_HEAD:
        v = f[k]
        if isinstance(v, (tuple, list, np.ndarray)):
            out.pack(fmt, *[float(x) for x in v])
        else:
            out.pack(fmt, v)
    write_config(out, f["config"])
    out.pack("3Q", f["minSamples"], f["maxSamples"], f["maxDepth"])
    out.pack("B", f["treeInit"])
    out.pack("6f", *[float(x) for x in f["treeBounds"]])
    nodes = np.asarray(f["nodes"])
    out.pack("Q", len(nodes))
    out.parts.append(nodes.astype(np.dtype([("pivot", "<f4"), ("bits", "<u4")])).tobytes())
    out.pack("Q", len(f["regions"]))
    for r in f["regions"]:
        write_region(out, r, W)
    out.pack("B", f["useStochasticNNLookUp"])
    out.pack("B", f["knnBuilt"])
    if f["knnBuilt"]:
        pts = np.asarray(f["knnPoints"], np.float32)
        out.pack("I", len(pts))
        out.floats(pts)


def read_file(data, W):
    """Whole file: header, structure types, SurfaceVolumeField members, surface then volume field."""
    inp = _In(data)
    header = inp.take(len(HEADER))
    out = {"header": header}
    out["spatialType"], out["directionalType"] = inp.unpack("2i")
    out["iteration"], out["totalSPP"] = inp.unpack("2Q")
    out["surface"] = read_field(inp, W)
    out["volume"] = read_field(inp, W)
    if inp.o != len(data):
        raise ValueError("%d trailing bytes" % (len(data) - inp.o))
    return out


def write_file(d, W):
    out = _Out()
    out.parts.append(HEADER)
    out.pack("2i", d.get("spatialType", 0), d.get("directionalType", 0))
    out.pack("2Q", d["iteration"], d.get("totalSPP", 0))
    write_field(out, d["surface"], W)
    write_field(out, d["volume"], W)
    return out.data()


def node_fields(nodes):
    """(pivot, dim (3 = leaf), child-or-region) of the kd node encoding (KDTree.h:21-31)."""
    bits = np.asarray(nodes["bits"], np.uint32)
    return np.asarray(nodes["pivot"], np.float32), bits >> 30, bits & 0x3FFFFFFF


def canonical(nodes):
    """KDTree::rearrangeNodesForCompare (spatial/kdtree/KDTree.h:255-282):
    depth-first renumbering of the node pairs and leaves.  Returns the
    canonical (pivot, dim, child-or-region) arrays and, per canonical region,
    its storage index."""
    nodes = np.asarray(nodes)
    if len(nodes) == 0:
        return (np.zeros(0, np.float32), np.zeros(0, np.uint32), np.zeros(0, np.uint32)), []
    pivot, dim, idx = node_fields(nodes)
    out_p, out_d, out_i = [pivot[0]], [dim[0]], [idx[0]]
    order = []

    def visit(src, dst):
        if dim[src] != 3:
            left = int(idx[src])
            new = len(out_p)
            out_i[dst] = new
            for c in (left, left + 1):
                out_p.append(pivot[c]); out_d.append(dim[c]); out_i.append(idx[c])
            visit(left, new)
            visit(left + 1, new + 1)
        else:
            out_i[dst] = len(order)
            order.append(int(idx[src]))

    visit(0, 0)
    return (np.array(out_p, np.float32), np.array(out_d, np.uint32), np.array(out_i, np.uint32)), order


def empty_field(is_surface, cfg, min_samples, max_samples, max_depth, deterministic=True, knn=True):
    """Field state of a kind that was never updated."""
    return {"isSurface": int(is_surface), "decayOnSpatialSplit": 0.25, "iteration": 0, "totalSPP": 0,
            "deterministic": int(deterministic), "fitRegions": 1, "isSceneBoundsSet": 0,
            "sceneBounds": (0.0,) * 6, "initialized": 0, "timeLastUpdate": 0.0, "timeLastUpdateCopySamples": 0.0,
            "timeLastUpdateSpatialStructureUpdate": 0.0, "timeLastUpdateDirectionalDistriubtionUpdate": 0.0,
            "config": dict(cfg), "minSamples": min_samples, "maxSamples": max_samples, "maxDepth": max_depth,
            "treeInit": 0, "treeBounds": (0.0,) * 6, "nodes": np.zeros(0, np.dtype([("pivot", "<f4"), ("bits", "<u4")])),
            "regions": [], "useStochasticNNLookUp": int(knn), "knnBuilt": 0, "knnPoints": np.zeros((0, 4), np.float32)}
