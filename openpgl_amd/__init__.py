"""Python bindings of libopenpgl_amd.so, mirroring Open PGL's C++ wrapper
classes (openpgl/include/openpgl/cpp/*.h: Device, Field, SampleStorage,
SurfaceSamplingDistribution, VolumeSamplingDistribution) over the C ABI.

The library is the MI355X implementation; importing this module never falls
back to a CPU path: if the shared library is missing, loading raises.
"""
import ctypes
import weakref
import os

import numpy as np

from .synth import PATH_SEGMENT_DTYPE, SAMPLE_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libopenpgl_amd.so")

PGL_DEVICE_TYPE_CPU_4, PGL_DEVICE_TYPE_CPU_8, PGL_DEVICE_TYPE_CPU_16, PGL_DEVICE_TYPE_NONE, PGL_DEVICE_TYPE_GPU_MI355X = range(5)
PGL_SPATIAL_STRUCTURE_KDTREE = 0
PGL_DIRECTIONAL_DISTRIBUTION_PARALLAX_AWARE_VMM = 0
PGL_DIRECTIONAL_DISTRIBUTION_QUADTREE = 1
PGL_DIRECTIONAL_DISTRIBUTION_VMM = 2

_vp = ctypes.c_void_p
_f = ctypes.c_float
_u32 = ctypes.c_uint32


class Vec3(ctypes.Structure):
    _fields_ = [("x", _f), ("y", _f), ("z", _f)]


class Vec2(ctypes.Structure):
    _fields_ = [("x", _f), ("y", _f)]


class Box3(ctypes.Structure):
    _fields_ = [("lower", Vec3), ("upper", Vec3)]


class PGLSampleData(ctypes.Structure):
    _fields_ = [("position", Vec3), ("direction", Vec3), ("weight", _f), ("pdf", _f), ("distance", _f),
                ("flags", _u32)]


class PGLPathSegmentData(ctypes.Structure):
    """include/openpgl/data.h:50-113 (136 bytes)."""
    _fields_ = [("position", Vec3), ("directionIn", Vec3), ("directionOut", Vec3), ("normal", Vec3),
                ("volumeScatter", ctypes.c_bool), ("pdfDirectionIn", ctypes.c_float), ("isDelta", ctypes.c_bool),
                ("scatteringWeight", Vec3), ("transmittanceWeight", Vec3), ("directContribution", Vec3),
                ("miWeight", ctypes.c_float), ("scatteredContribution", Vec3),
                ("russianRouletteProbability", ctypes.c_float), ("eta", ctypes.c_float),
                ("roughness", ctypes.c_float), ("regionPtr", ctypes.c_void_p)]


class PGLDebugArguments(ctypes.Structure):
    _fields_ = [("fitRegions", ctypes.c_bool)]


class PGLFieldArguments(ctypes.Structure):
    _fields_ = [("spatialStructureType", ctypes.c_int), ("spatialSturctureArguments", _vp),
                ("directionalDistributionType", ctypes.c_int), ("directionalDistributionArguments", _vp),
                ("deterministic", ctypes.c_bool), ("debugArguments", PGLDebugArguments)]


class PGLKDTreeArguments(ctypes.Structure):
    _fields_ = [("knnLookup", ctypes.c_bool), ("minSamples", ctypes.c_size_t), ("maxSamples", ctypes.c_size_t),
                ("maxDepth", ctypes.c_size_t)]


class PGLVMMFactoryArguments(ctypes.Structure):
    _fields_ = [("initK", ctypes.c_size_t), ("initKappa", _f), ("maxK", ctypes.c_size_t),
                ("maxEMIterrations", ctypes.c_size_t), ("maxKappa", _f), ("convergenceThreshold", _f),
                ("weightPrior", _f), ("meanCosinePriorStrength", _f), ("meanCosinePrior", _f),
                ("useSplitAndMerge", ctypes.c_bool), ("splittingThreshold", _f), ("mergingThreshold", _f),
                ("partialReFit", ctypes.c_bool), ("maxSplitItr", ctypes.c_int),
                ("minSamplesForSplitting", ctypes.c_int), ("minSamplesForPartialRefitting", ctypes.c_int),
                ("minSamplesForMerging", ctypes.c_int)]


_A32 = _f * 32
_A32x3 = (_f * 3) * 32
_A32x2 = (_f * 2) * 32


class PGLRegionRecord(ctypes.Structure):
    _fields_ = [
        ("K", _u32), ("weights", _A32), ("kappas", _A32), ("meanDirections", _A32x3), ("normalizations", _A32),
        ("eMinus2Kappa", _A32), ("meanCosines", _A32), ("distances", _A32), ("pivot", _f * 3),
        ("sumOfWeightedDirections", _A32x3), ("sumOfWeightedStats", _A32), ("sumOfDistanceWeights", _A32),
        ("sumWeights", _f), ("numSamples", _f), ("overallNumSamples", _f), ("suffK", _u32), ("normalized", _u32),
        ("chiSquareMCEstimates", _A32), ("splitMeans", _A32x2), ("splitCovariances", _A32x3),
        ("splitNumSamples", _A32), ("splitSumWeights", _A32), ("splitSumAssignedSamples", _A32), ("splitK", _u32),
        ("numSamplesAfterLastSplit", ctypes.c_uint64), ("numSamplesAfterLastMerge", ctypes.c_uint64),
        ("ssMean", _f * 3), ("ssVariance", _f * 3), ("ssNumSamples", _f), ("ssLower", _f * 3), ("ssUpper", _f * 3),
        ("regionLower", _f * 3), ("regionUpper", _f * 3), ("splitFlag", _u32), ("valid", _u32),
        ("numSamplesThisUpdate", _u32), ("emIterations", _u32), ("partialEmIterations", _u32),
        ("numSplits", _u32), ("numMerges", _u32)]

    def to_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = np.array(np.ctypeslib.as_array(v)) if hasattr(v, "_length_") else v
        return out


class PGLString(ctypes.Structure):
    _fields_ = [("m_str", ctypes.c_void_p), ("m_size", ctypes.c_size_t)]


# name: (restype, argtypes)
_SIGS = {
    "pglReleaseString": (None, [PGLString]),
    "pglFieldGetSurfaceStatistics": (_vp, [_vp]),
    "pglFieldGetVolumeStatistics": (_vp, [_vp]),
    "pglReleaseFieldStatistics": (None, [_vp]),
    "pglFieldStatisticsToString": (PGLString, [_vp]),
    "pglFieldStatisticsHeaderCSVString": (PGLString, [_vp]),
    "pglFieldStatisticsToCSVString": (PGLString, [_vp]),
    "pglFieldStatisticsGetValues": (None, [_vp, _vp, ctypes.c_size_t]),
    "pglFieldDebugFailNextUpdate": (None, [_vp, ctypes.c_int]),
    "pglFieldArgumentsSetDefaults": (None, [ctypes.POINTER(PGLFieldArguments), ctypes.c_int, ctypes.c_int,
                                            ctypes.c_bool, ctypes.c_size_t]),
    "pglNewDevice": (_vp, [ctypes.c_int, ctypes.c_size_t]),
    "pglReleaseDevice": (None, [_vp]),
    "pglDeviceNewField": (_vp, [_vp, PGLFieldArguments]),
    "pglReleaseField": (None, [_vp]),
    "pglDeviceNewFieldFromFile": (_vp, [_vp, ctypes.c_char_p]),
    "pglFieldStoreToFile": (ctypes.c_bool, [_vp, ctypes.c_char_p]),
    "pglFieldGetIteration": (ctypes.c_size_t, [_vp]),
    "pglFieldSetSceneBounds": (None, [_vp, Box3]),
    "pglFieldGetSceneBounds": (Box3, [_vp]),
    "pglFieldUpdate": (None, [_vp, _vp]),
    "pglFieldUpdateSurface": (None, [_vp, _vp]),
    "pglFieldUpdateVolume": (None, [_vp, _vp]),
    "pglFieldUpdateSurfaceDevice": (None, [_vp, _vp, ctypes.c_size_t]),
    "pglFieldUpdateDevice": (None, [_vp, _vp, ctypes.c_size_t]),
    "pglFieldUpdateVolumeDevice": (None, [_vp, _vp, ctypes.c_size_t]),
    "pglFieldReset": (None, [_vp]),
    "pglFieldValidate": (ctypes.c_bool, [_vp]),
    "pglFieldCompare": (ctypes.c_bool, [_vp, _vp]),
    "pglFieldNewSurfaceSamplingDistribution": (_vp, [_vp]),
    "pglFieldInitSurfaceSamplingDistribution": (ctypes.c_bool, [_vp, _vp, Vec3, ctypes.POINTER(_f)]),
    "pglFieldNewVolumeSamplingDistribution": (_vp, [_vp]),
    "pglFieldInitVolumeSamplingDistribution": (ctypes.c_bool, [_vp, _vp, Vec3, ctypes.POINTER(_f)]),
    "pglReleaseSurfaceSamplingDistribution": (None, [_vp]),
    "pglSurfaceSamplingDistributionApplyCosineProduct": (None, [_vp, Vec3]),
    "pglSurfaceSamplingDistributionSample": (Vec3, [_vp, Vec2]),
    "pglSurfaceSamplingDistributionPDF": (_f, [_vp, Vec3]),
    "pglSurfaceSamplingDistributionSamplePDF": (_f, [_vp, Vec2, ctypes.POINTER(Vec3)]),
    "pglSurfaceSamplingDistributionIncomingRadiancePDF": (_f, [_vp, Vec3]),
    "pglSurfaceSamplingDistributionGetId": (_u32, [_vp]),
    "pglReleaseVolumeSamplingDistribution": (None, [_vp]),
    "pglVolumeSamplingDistributionApplySingleLobeHenyeyGreensteinProduct": (None, [_vp, Vec3, _f]),
    "pglVolumeSamplingDistributionSample": (Vec3, [_vp, Vec2]),
    "pglVolumeSamplingDistributionPDF": (_f, [_vp, Vec3]),
    "pglVolumeSamplingDistributionIncomingRadiancePDF": (_f, [_vp, Vec3]),
    "pglVolumeSamplingDistributionSamplePDF": (_f, [_vp, Vec2, ctypes.POINTER(Vec3)]),
    "pglVolumeSamplingDistributionGetId": (_u32, [_vp]),
    "pglNewSampleStorage": (_vp, []),
    "pglReleaseSampleStorage": (None, [_vp]),
    "pglSampleStorageAddSamples": (None, [_vp, _vp, ctypes.c_size_t]),
    "pglSampleStorageGetSizeSurface": (ctypes.c_size_t, [_vp]),
    "pglSampleStorageGetSizeVolume": (ctypes.c_size_t, [_vp]),
    "pglSampleStorageClear": (None, [_vp]),
    "pglSampleStorageStoreToFile": (ctypes.c_bool, [_vp, ctypes.c_char_p]),
    "pglNewSampleStorageFromFile": (_vp, [ctypes.c_char_p]),
    "pglSampleStorageCompare": (ctypes.c_bool, [_vp, _vp]),
    "pglSampleStorageValidate": (ctypes.c_bool, [_vp]),
    "pglNewPathSegmentStorage": (_vp, []),
    "pglReleasePathSegmentStorage": (None, [_vp]),
    "pglPathSegmentStorageReserve": (None, [_vp, ctypes.c_size_t]),
    "pglPathSegmentStorageClear": (None, [_vp]),
    "pglPathSegmentSetMaxDistance": (None, [_vp, _f]),
    "pglPathSegmentGetMaxDistance": (_f, [_vp]),
    "pglPathSegmentGetNumSegments": (ctypes.c_int, [_vp]),
    "pglPathSegmentGetNumSamples": (ctypes.c_int, [_vp]),
    "pglPathSegmentStoragePrepareSamples": (ctypes.c_size_t, [_vp, ctypes.c_bool, ctypes.c_bool, ctypes.c_bool]),
    "pglPathSegmentStorageCalculatePixelEstimate": (Vec3, [_vp, ctypes.c_bool]),
    "pglPathSegmentStorageGetSamples": (_vp, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    "pglPathSegmentStorageAddSample": (None, [_vp, PGLSampleData]),
    "pglPathSegmentStorageNextSegment": (_vp, [_vp]),
    "pglPathSegmentStorageAddSegment": (None, [_vp, PGLPathSegmentData]),
    "pglPathSegmentStorageValidateSamples": (ctypes.c_bool, [_vp]),
    "pglPathSegmentStorageValidateSegments": (ctypes.c_bool, [_vp]),
    "pglPathSegmentStoragePropagateSamples": (None, [_vp, _vp, ctypes.c_bool, ctypes.c_bool, ctypes.c_bool]),
    "pglNewPathSegmentBatch": (_vp, []),
    "pglReleasePathSegmentBatch": (None, [_vp]),
    "pglPathSegmentBatchPrepareSamplesDevice": (ctypes.c_bool, [_vp, _vp, ctypes.c_size_t, _vp, _u32, _f,
                                                                 ctypes.c_bool, ctypes.c_bool, ctypes.c_bool, _vp,
                                                                 ctypes.c_size_t, _vp, ctypes.c_size_t, _vp, _vp]),
    "pglFieldUpdateFromPathSegmentsDevice": (None, [_vp, _vp, ctypes.c_size_t, _vp, _u32, _f, ctypes.c_bool,
                                                    ctypes.c_bool, ctypes.c_bool]),
    "pglPathSegmentsPrepareSamplesDeviceNoAlloc": (ctypes.c_bool, [_vp, ctypes.c_size_t, _vp, _u32, _f, ctypes.c_bool,
                                                                   ctypes.c_bool, ctypes.c_bool, _vp, ctypes.c_size_t,
                                                                   _vp, ctypes.c_size_t, _vp, _vp, _vp]),
    "pglFieldSampleBatch": (None, [_vp, ctypes.c_size_t] + [_vp] * 10),
    "pglFieldSampleBatchDevice": (None, [_vp, ctypes.c_size_t] + [_vp] * 11),
    "pglFieldGetNumNodes": (_u32, [_vp, ctypes.c_int]),
    "pglFieldGetNumRegions": (_u32, [_vp, ctypes.c_int]),
    "pglFieldExportNodes": (None, [_vp, ctypes.c_int, _vp, _vp, _vp]),
    "pglFieldExportRegion": (None, [_vp, ctypes.c_int, _u32, ctypes.POINTER(PGLRegionRecord)]),
    "pglFieldSampleLeafIds": (None, [_vp, ctypes.c_int, _vp, ctypes.c_size_t, _vp]),
    "pglFieldGetKNN": (None, [_vp, ctypes.c_int, _vp]),
    "pglFieldGetLastUpdateInfo": (None, [_vp, ctypes.c_int, _vp, ctypes.c_size_t]),
    "pglFieldSetRegionSchedule": (None, [_vp, _vp]),
    "pglFieldGetRegionSchedule": (None, [_vp, _vp]),
    "pglFieldGetRegionProfile": (None, [_vp, ctypes.c_int, _vp]),
    "pglFieldGetUpdateHistory": (ctypes.c_size_t, [_vp, ctypes.c_int, _vp, ctypes.c_size_t]),
    "pglNewRegionBatch": (_vp, [_vp, PGLFieldArguments]),
    "pglReleaseRegionBatch": (None, [_vp]),
    "pglRegionBatchStep": (None, [_vp, _u32, _vp, _vp, _vp, _vp, _vp]),
    "pglRegionBatchGetNumRegions": (_u32, [_vp]),
    "pglRegionBatchExportRegion": (None, [_vp, _u32, ctypes.POINTER(PGLRegionRecord)]),
    "pglRegionBatchSetSchedule": (None, [_vp, _vp]),
    "pglRegionBatchGetSchedule": (None, [_vp, _vp]),
    "pglRegionScheduleCandidates": (ctypes.c_size_t, [_vp, ctypes.c_size_t]),
    "pglRegionBatchAutotune": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, ctypes.c_int, _vp]),
    "pglRegionBatchGetLastInfo": (None, [_vp, _vp, ctypes.c_size_t]),
    "pglNewServer": (_vp, [_vp, _vp]),
    "pglReleaseServer": (None, [_vp]),
    "pglServerSubmit": (ctypes.c_uint64, [_vp, ctypes.c_size_t, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pglServerPoll": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "pglServerWait": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_double]),
    "pglServerDrain": (None, [_vp]),
    "pglServerGetTiming": (ctypes.c_int, [_vp, ctypes.c_uint64, _vp]),
    "pglServerForget": (None, [_vp, ctypes.c_uint64]),
    "pglServerRefresh": (None, [_vp]),
    "pglServerGetStats": (None, [_vp, _vp, ctypes.c_size_t]),
}



class RegionSchedule(ctypes.Structure):
    """PGLRegionSchedule (include/openpgl_amd/openpgl.h)."""
    _fields_ = [("perRegionLaunches", ctypes.c_int), ("grid", ctypes.c_uint32), ("balance", ctypes.c_float),
                ("minChunksPerRank", ctypes.c_uint32), ("maxTeam", ctypes.c_uint32), ("cooperative", ctypes.c_int)]


_lib = None


def _own(obj, release):
    """Ties the library handle obj.h to obj: released when obj is collected,
    by obj.close(), or at interpreter exit.  The exit pass runs before the
    interpreter tears its modules down (whose order is arbitrary) and long
    before the HIP runtime's own teardown, in reverse order of creation, so
    a server is released before its field and a field before its device."""
    obj._fin = weakref.finalize(obj, release, obj.h)


def _close(obj):
    fin = getattr(obj, "_fin", None)
    if fin is not None:
        fin()
    obj.h = None


def load_library():
    """Load libopenpgl_amd.so from the package directory (raises if absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libopenpgl_amd.so is not built: run `python -m openpgl_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class Device:
    def __init__(self, device_type=PGL_DEVICE_TYPE_GPU_MI355X, num_threads=0):
        self.lib = load_library()
        self.h = self.lib.pglNewDevice(device_type, num_threads)
        if not self.h:
            raise RuntimeError("pglNewDevice failed (no GPU?)")
        _own(self, self.lib.pglReleaseDevice)

    def close(self):
        _close(self)


class SampleStorage:
    def __init__(self, samples=None):
        self.lib = load_library()
        self.h = self.lib.pglNewSampleStorage()
        if not self.h:
            raise RuntimeError("pglNewSampleStorage failed")
        _own(self, self.lib.pglReleaseSampleStorage)
        if samples is not None:
            self.add_samples(samples)

    def add_samples(self, samples):
        samples = np.ascontiguousarray(samples, dtype=SAMPLE_DTYPE)
        self.lib.pglSampleStorageAddSamples(self.h, _ptr(samples), len(samples))

    def size_surface(self):
        return self.lib.pglSampleStorageGetSizeSurface(self.h)

    def size_volume(self):
        return self.lib.pglSampleStorageGetSizeVolume(self.h)

    def clear(self):
        self.lib.pglSampleStorageClear(self.h)

    def close(self):
        _close(self)


def _field_arguments(lib, max_samples_per_leaf, deterministic, knn=True, overrides=None,
                     directional_type=PGL_DIRECTIONAL_DISTRIBUTION_PARALLAX_AWARE_VMM):
    args = PGLFieldArguments()
    lib.pglFieldArgumentsSetDefaults(ctypes.byref(args), PGL_SPATIAL_STRUCTURE_KDTREE, directional_type,
                                     deterministic, max_samples_per_leaf)
    kd = ctypes.cast(args.spatialSturctureArguments, ctypes.POINTER(PGLKDTreeArguments)).contents
    kd.knnLookup = knn
    if overrides:
        v = ctypes.cast(args.directionalDistributionArguments, ctypes.POINTER(PGLVMMFactoryArguments)).contents
        for k, val in overrides.items():
            if hasattr(kd, k):
                setattr(kd, k, val)
            else:
                setattr(v, k, val)
    return args


_INFO_KEYS = ["wall_ms", "spatial_ms", "directional_ms", "knn_ms", "gather_ms", "sort_ms", "kd_rounds",
              "region_launches", "estep_evals", "dist_evals", "em_iterations", "samples", "split_evals", "regions",
              "region_jobs", "queue_items", "max_team", "grid", "region_kernel_ms"]


def _schedule_dict(sc):
    return {"per_region_launches": sc.perRegionLaunches, "grid": sc.grid, "balance": sc.balance,
            "min_chunks_per_rank": sc.minChunksPerRank, "max_team": sc.maxTeam, "cooperative": sc.cooperative}


def _schedule_struct(d):
    return RegionSchedule(int(d.get("per_region_launches", 0)), int(d.get("grid", 0)), float(d.get("balance", 1.0)),
                          int(d.get("min_chunks_per_rank", 1)), int(d.get("max_team", 64)),
                          int(d.get("cooperative", 0)))


class ServerConfig(ctypes.Structure):
    """PGLServerConfig (include/openpgl_amd/openpgl.h)."""
    _fields_ = [("memoryBudgetBytes", ctypes.c_size_t), ("slots", ctypes.c_uint32),
                ("maxBatchQueries", ctypes.c_uint32)]


class Server:
    """Continuous-batching query server (pglServer*).  Arrays passed to submit
    must stay alive (and unchanged) until the request is done; pinned arrays
    (hip_runtime.pinned_empty) give asynchronous copies."""

    def __init__(self, field, memory_budget=0, slots=0, max_batch=0):
        self.lib = load_library()
        self.field = field
        cfg = ServerConfig(memory_budget, slots, max_batch)
        self.h = self.lib.pglNewServer(field.h, ctypes.byref(cfg))
        if not self.h:
            raise RuntimeError("pglNewServer failed")
        _own(self, self.lib.pglReleaseServer)
        self._live = {}

    def close(self):
        """Waits for the submitted requests, then releases the server."""
        _close(self)
        self._live.clear()

    def submit(self, pos, nrm, u2, u1=None, g=None, is_volume=None, out=None):
        """Returns (request id, (dir, pdf, ids, integral) output arrays)."""
        n = len(pos)
        if out is None:
            out = (np.zeros((n, 3), np.float32), np.zeros(n, np.float32), np.zeros(n, np.uint32),
                   np.zeros(n, np.float32))
        rid = self.lib.pglServerSubmit(self.h, n, _ptr(pos), _ptr(nrm), _ptr(g), _ptr(u2), _ptr(u1),
                                       _ptr(is_volume), _ptr(out[0]), _ptr(out[1]), _ptr(out[2]), _ptr(out[3]))
        if rid == 0:
            raise RuntimeError("pglServerSubmit rejected the request")
        self._live[rid] = (pos, nrm, u2, u1, g, is_volume, out)
        return rid, out

    def wait(self, rid, timeout_ms=-1.0):
        ok = self.lib.pglServerWait(self.h, rid, timeout_ms)
        if ok < 0:
            raise RuntimeError("pglServerWait failed")
        return ok == 1

    def poll(self, rid):
        return self.lib.pglServerPoll(self.h, rid) == 1

    def timing(self, rid):
        t = np.zeros(3, np.float64)
        if self.lib.pglServerGetTiming(self.h, rid, _ptr(t)) != 1:
            return None
        return t

    def release(self, rid):
        self.lib.pglServerForget(self.h, rid)
        self._live.pop(rid, None)

    def drain(self):
        self.lib.pglServerDrain(self.h)

    def refresh(self):
        self.lib.pglServerRefresh(self.h)

    def stats(self):
        o = np.zeros(6, np.float64)
        self.lib.pglServerGetStats(self.h, _ptr(o), 6)
        return dict(zip(["requests", "queries", "batches", "peak_inflight_bytes", "peak_pending", "max_batch"],
                        o.tolist()))


class RegionBatch:
    """Batched directional training of pre-partitioned regions (pglRegionBatch*)."""

    def __init__(self, device, max_samples_per_leaf=32000, deterministic=True, overrides=None):
        self.lib = load_library()
        self.device = device
        args = _field_arguments(self.lib, max_samples_per_leaf, deterministic, True, overrides)
        self.h = self.lib.pglNewRegionBatch(device.h, args)
        if not self.h:
            raise RuntimeError("pglNewRegionBatch failed")
        _own(self, self.lib.pglReleaseRegionBatch)

    def close(self):
        _close(self)

    def step_device(self, device_ptr, counts, stats7, init_k=None, split_flags=None):
        """device_ptr: region-contiguous PGLSampleData in device memory."""
        counts = np.ascontiguousarray(counts, np.uint32)
        stats7 = np.ascontiguousarray(stats7, np.float32).reshape(-1, 7)
        ik = None if init_k is None else np.ascontiguousarray(init_k, np.uint32)
        sf = None if split_flags is None else np.ascontiguousarray(split_flags, np.int32)
        self.lib.pglRegionBatchStep(self.h, len(counts), ctypes.c_void_p(device_ptr), _ptr(counts), _ptr(stats7),
                                    _ptr(ik), _ptr(sf))

    def num_regions(self):
        return self.lib.pglRegionBatchGetNumRegions(self.h)

    def region(self, idx):
        r = PGLRegionRecord()
        self.lib.pglRegionBatchExportRegion(self.h, idx, ctypes.byref(r))
        return r.to_dict()

    def set_schedule(self, **kw):
        sc = RegionSchedule()
        self.lib.pglRegionBatchGetSchedule(self.h, ctypes.byref(sc))
        d = _schedule_dict(sc)
        d.update(kw)
        self.lib.pglRegionBatchSetSchedule(self.h, ctypes.byref(_schedule_struct(d)))

    def schedule(self):
        sc = RegionSchedule()
        self.lib.pglRegionBatchGetSchedule(self.h, ctypes.byref(sc))
        return _schedule_dict(sc)

    def autotune(self, candidates=None, reps=2):
        """Returns (best index, [(schedule dict, kernel ms)])."""
        if candidates is None:
            n = self.lib.pglRegionScheduleCandidates(None, 0)
            arr = (RegionSchedule * n)()
            self.lib.pglRegionScheduleCandidates(arr, n)
        else:
            n = len(candidates)
            arr = (RegionSchedule * n)(*[_schedule_struct(c) for c in candidates])
        ms = np.zeros(n, np.float32)
        best = self.lib.pglRegionBatchAutotune(self.h, arr, n, reps, _ptr(ms))
        if best < 0:
            raise RuntimeError("pglRegionBatchAutotune failed")
        return best, [(_schedule_dict(arr[i]), float(ms[i])) for i in range(n)]

    def last_info(self):
        o = np.zeros(19, np.float64)
        self.lib.pglRegionBatchGetLastInfo(self.h, _ptr(o), 19)
        return dict(zip(_INFO_KEYS, o.tolist()))


class Field:
    """openpgl::cpp::Field over the MI355X backend."""

    def __init__(self, device, max_samples_per_leaf=32000, deterministic=True, knn=True, overrides=None,
                 directional_type=PGL_DIRECTIONAL_DISTRIBUTION_PARALLAX_AWARE_VMM):
        self.lib = load_library()
        self.device = device
        args = _field_arguments(self.lib, max_samples_per_leaf, deterministic, knn, overrides, directional_type)
        self.h = self.lib.pglDeviceNewField(device.h, args)
        if not self.h:
            raise RuntimeError("pglDeviceNewField failed")
        _own(self, self.lib.pglReleaseField)

    @classmethod
    def from_file(cls, device, path):
        """pglDeviceNewFieldFromFile: a field read from an Open PGL field file
        (in the layout of the device's SIMD width)."""
        self = cls.__new__(cls)
        self.lib = load_library()
        self.device = device
        self.h = self.lib.pglDeviceNewFieldFromFile(device.h, os.fsencode(path))
        if not self.h:
            raise RuntimeError("pglDeviceNewFieldFromFile failed: %s" % path)
        _own(self, self.lib.pglReleaseField)
        return self

    def close(self):
        _close(self)

    def store(self, path):
        """pglFieldStoreToFile (layout of the field's device's SIMD width)."""
        return bool(self.lib.pglFieldStoreToFile(self.h, os.fsencode(path)))

    def compare(self, other):
        return bool(self.lib.pglFieldCompare(self.h, other.h))

    def set_scene_bounds(self, lower, upper):
        b = Box3(Vec3(*lower), Vec3(*upper))
        self.lib.pglFieldSetSceneBounds(self.h, b)

    def scene_bounds(self):
        b = self.lib.pglFieldGetSceneBounds(self.h)
        return (b.lower.x, b.lower.y, b.lower.z), (b.upper.x, b.upper.y, b.upper.z)

    def update(self, storage):
        self.lib.pglFieldUpdate(self.h, storage.h)

    def update_surface_device(self, device_ptr, n):
        self.lib.pglFieldUpdateSurfaceDevice(self.h, ctypes.c_void_p(device_ptr), n)

    def update_volume_device(self, device_ptr, n):
        self.lib.pglFieldUpdateVolumeDevice(self.h, ctypes.c_void_p(device_ptr), n)

    def update_device(self, device_ptr, n):
        """pglFieldUpdate with mixed records in device memory (split by the
        EInsideVolume flag on the device)."""
        self.lib.pglFieldUpdateDevice(self.h, ctypes.c_void_p(device_ptr), n)

    def update_from_path_segments_device(self, segments_ptr, num_segments, offsets_ptr, num_paths, max_distance=1e6,
                                         use_nee_mi_weights=False, guide_direct_light=False, rr_affects_direct=True):
        self.lib.pglFieldUpdateFromPathSegmentsDevice(self.h, ctypes.c_void_p(segments_ptr), num_segments,
                                                      ctypes.c_void_p(offsets_ptr), num_paths, float(max_distance),
                                                      use_nee_mi_weights, guide_direct_light, rr_affects_direct)

    def iteration(self):
        return self.lib.pglFieldGetIteration(self.h)

    def validate(self):
        return self.lib.pglFieldValidate(self.h)

    def reset(self):
        self.lib.pglFieldReset(self.h)

    _STAT_KEYS = ["numCacheRegions", "numCacheRegionsReserved", "sizePerCacheRegions", "timeLastUpdate",
                  "timeLastUpdateCopySamples", "timeLastUpdateSpatialStructureUpdate",
                  "timeLastUpdateDirectionalDistriubtionUpdate", "maxDepth", "numberOfNodes",
                  "numberOfReservedNodes", "sizePerNode", "sizePerDistribution", "minNumberOfComponents",
                  "maxNumberOfComponents", "averageNumberOfComponents", "secondMomentNumberOfComponents"]

    def statistics(self, kind=0):
        """pglFieldGet{Surface,Volume}Statistics: (values by name, toString,
        headerCSVString, toCSVString), or None on error."""
        h = (self.lib.pglFieldGetVolumeStatistics if kind else self.lib.pglFieldGetSurfaceStatistics)(self.h)
        if not h:
            return None
        try:
            v = np.zeros(16, np.float64)
            self.lib.pglFieldStatisticsGetValues(h, v.ctypes.data, 16)
            strings = []
            for fn in (self.lib.pglFieldStatisticsToString, self.lib.pglFieldStatisticsHeaderCSVString,
                       self.lib.pglFieldStatisticsToCSVString):
                ps = fn(h)
                strings.append(ctypes.string_at(ps.m_str, ps.m_size).decode() if ps.m_str else "")
                self.lib.pglReleaseString(ps)
        finally:
            self.lib.pglReleaseFieldStatistics(h)
        return (dict(zip(self._STAT_KEYS, v.tolist())), *strings)

    def debug_fail_next_update(self, kind=0):
        """pglFieldDebugFailNextUpdate (testing)."""
        self.lib.pglFieldDebugFailNextUpdate(self.h, kind)

    def set_region_schedule(self, per_region_launches=None, grid=None, balance=None, min_chunks_per_rank=None,
                            max_team=None, cooperative=None):
        """Directional scheduling (PGLRegionSchedule); None keeps the current value."""
        sc = RegionSchedule()
        self.lib.pglFieldGetRegionSchedule(self.h, ctypes.byref(sc))
        if per_region_launches is not None:
            sc.perRegionLaunches = int(per_region_launches)
        if grid is not None:
            sc.grid = int(grid)
        if balance is not None:
            sc.balance = float(balance)
        if min_chunks_per_rank is not None:
            sc.minChunksPerRank = int(min_chunks_per_rank)
        if max_team is not None:
            sc.maxTeam = int(max_team)
        if cooperative is not None:
            sc.cooperative = int(cooperative)
        self.lib.pglFieldSetRegionSchedule(self.h, ctypes.byref(sc))

    def update_history(self, count=256, kind=0):
        """Figures of the last `count` updates (oldest first), as last_update_info dicts."""
        o = np.zeros((count, 19), np.float64)
        n = self.lib.pglFieldGetUpdateHistory(self.h, kind, _ptr(o), count)
        return [dict(zip(_INFO_KEYS, o[i].tolist())) for i in range(n)]

    def region_profile(self, kind=0):
        """(regions, 12): us of total, E-step, split, distance passes, barrier waits, merging,
        splitting logic, [7] team size, [8] E-step passes, [9] M-steps, [10] folds, [11] samples.
        Storage order."""
        out = np.zeros((self.num_regions(kind), 12), np.float32)
        self.lib.pglFieldGetRegionProfile(self.h, kind, _ptr(out))
        return out

    def region_schedule(self):
        sc = RegionSchedule()
        self.lib.pglFieldGetRegionSchedule(self.h, ctypes.byref(sc))
        return {"per_region_launches": sc.perRegionLaunches, "grid": sc.grid, "balance": sc.balance,
                "min_chunks_per_rank": sc.minChunksPerRank, "max_team": sc.maxTeam, "cooperative": sc.cooperative}

    # --- introspection
    def num_regions(self, kind=0):
        return self.lib.pglFieldGetNumRegions(self.h, kind)

    def nodes(self, kind=0):
        n = self.lib.pglFieldGetNumNodes(self.h, kind)
        piv = np.zeros(n, np.float32)
        dims = np.zeros(n, np.uint32)
        cod = np.zeros(n, np.uint32)
        self.lib.pglFieldExportNodes(self.h, kind, _ptr(piv), _ptr(dims), _ptr(cod))
        return piv, dims, cod

    def region(self, idx, kind=0):
        r = PGLRegionRecord()
        self.lib.pglFieldExportRegion(self.h, kind, idx, ctypes.byref(r))
        return r.to_dict()

    def leaf_ids(self, positions, kind=0):
        p = np.ascontiguousarray(positions, np.float32)
        out = np.zeros(len(p), np.uint32)
        self.lib.pglFieldSampleLeafIds(self.h, kind, _ptr(p), len(p), _ptr(out))
        return out

    def knn(self, kind=0):
        out = np.zeros(self.num_regions(kind) * 8, np.uint32)
        self.lib.pglFieldGetKNN(self.h, kind, _ptr(out))
        return out.reshape(-1, 8)

    def last_update_info(self, kind=0):
        o = np.zeros(19, np.float64)
        self.lib.pglFieldGetLastUpdateInfo(self.h, kind, _ptr(o), 19)
        return dict(zip(_INFO_KEYS, o.tolist()))

    # --- queries
    def sample_batch(self, pos, nrm, u2, u1=None, g=None, is_volume=None):
        n = len(pos)
        pos = np.ascontiguousarray(pos, np.float32)
        nrm = np.ascontiguousarray(nrm, np.float32)
        u2 = np.ascontiguousarray(u2, np.float32)
        u1 = None if u1 is None else np.ascontiguousarray(u1, np.float32)
        g = None if g is None else np.ascontiguousarray(g, np.float32)
        vol = None if is_volume is None else np.ascontiguousarray(is_volume, np.uint8)
        d = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        ids = np.zeros(n, np.uint32)
        integ = np.zeros(n, np.float32)
        self.lib.pglFieldSampleBatch(self.h, n, _ptr(pos), _ptr(nrm), _ptr(g), _ptr(u2), _ptr(u1), _ptr(vol),
                                     _ptr(d), _ptr(pdf), _ptr(ids), _ptr(integ))
        return d, pdf, ids, integ

    def new_surface_distribution(self):
        return SurfaceSamplingDistribution(self)

    def new_volume_distribution(self):
        return VolumeSamplingDistribution(self)


class _Dist:
    _kind = "Surface"

    def __init__(self, field):
        self.lib = field.lib
        self.field = field
        self.h = getattr(self.lib, "pglFieldNew%sSamplingDistribution" % self._kind)(field.h)
        if not self.h:
            raise RuntimeError("pglFieldNew%sSamplingDistribution failed" % self._kind)
        _own(self, getattr(self.lib, "pglRelease%sSamplingDistribution" % self._kind))

    def init(self, pos, sample1d=-1.0):
        u = ctypes.c_float(sample1d)
        ok = getattr(self.lib, "pglFieldInit%sSamplingDistribution" % self._kind)(self.field.h, self.h, Vec3(*pos),
                                                                                 ctypes.byref(u))
        return ok, u.value

    def sample(self, u):
        v = getattr(self.lib, "pgl%sSamplingDistributionSample" % self._kind)(self.h, Vec2(*u))
        return np.array([v.x, v.y, v.z], np.float32)

    def pdf(self, d):
        return getattr(self.lib, "pgl%sSamplingDistributionPDF" % self._kind)(self.h, Vec3(*[float(x) for x in d]))

    def sample_pdf(self, u):
        out = Vec3()
        p = getattr(self.lib, "pgl%sSamplingDistributionSamplePDF" % self._kind)(self.h, Vec2(*u), ctypes.byref(out))
        return np.array([out.x, out.y, out.z], np.float32), p

    def id(self):
        return getattr(self.lib, "pgl%sSamplingDistributionGetId" % self._kind)(self.h)

    def close(self):
        _close(self)


class SurfaceSamplingDistribution(_Dist):
    _kind = "Surface"

    def apply_cosine_product(self, n):
        self.lib.pglSurfaceSamplingDistributionApplyCosineProduct(self.h, Vec3(*[float(x) for x in n]))

    def incoming_radiance_pdf(self, d):
        return self.lib.pglSurfaceSamplingDistributionIncomingRadiancePDF(self.h, Vec3(*[float(x) for x in d]))


class VolumeSamplingDistribution(_Dist):
    _kind = "Volume"

    def incoming_radiance_pdf(self, d):
        return self.lib.pglVolumeSamplingDistributionIncomingRadiancePDF(self.h, Vec3(*[float(x) for x in d]))

    def apply_hg_product(self, d, g):
        self.lib.pglVolumeSamplingDistributionApplySingleLobeHenyeyGreensteinProduct(
            self.h, Vec3(*[float(x) for x in d]), float(g))


class PathSegmentStorage:
    """pglPathSegmentStorage*: one recorded path on the host (the reference's
    per-render-thread storage)."""

    def __init__(self, capacity=64, max_distance=None):
        self.lib = load_library()
        self.h = self.lib.pglNewPathSegmentStorage()
        if not self.h:
            raise RuntimeError("pglNewPathSegmentStorage failed")
        _own(self, self.lib.pglReleasePathSegmentStorage)
        self.lib.pglPathSegmentStorageReserve(self.h, capacity)
        if max_distance is not None:
            self.lib.pglPathSegmentSetMaxDistance(self.h, float(max_distance))

    def close(self):
        _close(self)

    def add_segments(self, segments):
        seg = np.ascontiguousarray(segments, PATH_SEGMENT_DTYPE)
        raw = seg.view(np.uint8).reshape(len(seg), 136)
        for i in range(len(seg)):
            self.lib.pglPathSegmentStorageAddSegment(self.h, PGLPathSegmentData.from_buffer_copy(raw[i].tobytes()))

    def next_segment(self):
        """pglPathSegmentStorageNextSegment: address of a default vertex (0 when full)."""
        return self.lib.pglPathSegmentStorageNextSegment(self.h) or 0

    def clear(self):
        self.lib.pglPathSegmentStorageClear(self.h)

    def num_segments(self):
        return self.lib.pglPathSegmentGetNumSegments(self.h)

    def num_samples(self):
        return self.lib.pglPathSegmentGetNumSamples(self.h)

    def prepare_samples(self, use_nee_mi_weights=False, guide_direct_light=False, rr_affects_direct=True):
        return self.lib.pglPathSegmentStoragePrepareSamples(self.h, use_nee_mi_weights, guide_direct_light,
                                                            rr_affects_direct)

    def samples(self):
        n = ctypes.c_size_t(0)
        p = self.lib.pglPathSegmentStorageGetSamples(self.h, ctypes.byref(n))
        if n.value == 0:
            return np.zeros(0, SAMPLE_DTYPE)
        return np.frombuffer((ctypes.c_char * (40 * n.value)).from_address(p), SAMPLE_DTYPE).copy()

    def pixel_estimate(self, rr_affects_direct=True):
        v = self.lib.pglPathSegmentStorageCalculatePixelEstimate(self.h, rr_affects_direct)
        return np.array([v.x, v.y, v.z], np.float32)

    def propagate(self, storage, guide_direct_light=False, use_nee_mi_weights=False, rr_affects_direct=True):
        self.lib.pglPathSegmentStoragePropagateSamples(self.h, storage.h, guide_direct_light, use_nee_mi_weights,
                                                       rr_affects_direct)

    def validate_segments(self):
        return self.lib.pglPathSegmentStorageValidateSegments(self.h)

    def validate_samples(self):
        return self.lib.pglPathSegmentStorageValidateSamples(self.h)


NO_ALLOC_MAX_VERTICES = 1 << 18


def prepare_path_samples_device_noalloc(segments_ptr, num_segments, offsets_ptr, num_paths, surface_ptr,
                                        surface_capacity, volume_ptr, volume_capacity, max_distance=1e6,
                                        use_nee_mi_weights=False, guide_direct_light=False, rr_affects_direct=True,
                                        stream=None):
    """EXTENSION: PathSegmentBatch.prepare_device without any allocation
    (pglPathSegmentsPrepareSamplesDeviceNoAlloc; at most
    NO_ALLOC_MAX_VERTICES vertices per call; calls overlapping on one GPU
    race).  Returns (surface count, volume count), or None on error."""
    lib = load_library()
    ns, nv = ctypes.c_size_t(0), ctypes.c_size_t(0)
    ok = lib.pglPathSegmentsPrepareSamplesDeviceNoAlloc(
        ctypes.c_void_p(segments_ptr), num_segments, ctypes.c_void_p(offsets_ptr), num_paths, float(max_distance),
        use_nee_mi_weights, guide_direct_light, rr_affects_direct, ctypes.c_void_p(surface_ptr), surface_capacity,
        ctypes.c_void_p(volume_ptr), volume_capacity, ctypes.byref(ns), ctypes.byref(nv), ctypes.c_void_p(stream))
    return (ns.value, nv.value) if ok else None


class PathSegmentBatch:
    """EXTENSION: training samples of many recorded paths, on the GPU."""

    def __init__(self):
        self.lib = load_library()
        self.h = self.lib.pglNewPathSegmentBatch()
        if not self.h:
            raise RuntimeError("pglNewPathSegmentBatch failed")
        _own(self, self.lib.pglReleasePathSegmentBatch)

    def close(self):
        _close(self)

    def prepare_device(self, segments_ptr, num_segments, offsets_ptr, num_paths, surface_ptr, surface_capacity,
                       volume_ptr, volume_capacity, max_distance=1e6, use_nee_mi_weights=False,
                       guide_direct_light=False, rr_affects_direct=True):
        """Returns (surface count, volume count), or None on error (capacity too small, bad ranges)."""
        ns, nv = ctypes.c_size_t(0), ctypes.c_size_t(0)
        ok = self.lib.pglPathSegmentBatchPrepareSamplesDevice(
            self.h, ctypes.c_void_p(segments_ptr), num_segments, ctypes.c_void_p(offsets_ptr), num_paths,
            float(max_distance),
            use_nee_mi_weights, guide_direct_light, rr_affects_direct, ctypes.c_void_p(surface_ptr),
            surface_capacity, ctypes.c_void_p(volume_ptr), volume_capacity, ctypes.byref(ns), ctypes.byref(nv))
        return (ns.value, nv.value) if ok else None
