"""Synthetic guiding-sample generator for the benchmark and test workloads.

Records follow PGLSampleData (40 bytes: pos f32x3, dir f32x3, weight, pdf,
distance f32, flags u32).  Positions: 80% from Gaussian blobs (sigma 0.05)
with centres in U[0.1, 0.9]^3, 20% uniform in the unit cube.  Each blob owns
a 4-lobe vMF mixture (equal weights, mean directions uniform on the sphere,
kappa log-uniform in [5, 500]); uniform-position records draw from the
nearest blob's lobes.  weight ~ lognormal(0, 1), pdf = 1/(4 pi),
distance ~ Exp(1).  All draws come from numpy's PCG64 with the given seed.
"""
import numpy as np

SAMPLE_DTYPE = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("weight", "<f4"),
                         ("pdf", "<f4"), ("distance", "<f4"), ("flags", "<u4")])
assert SAMPLE_DTYPE.itemsize == 40


def _unit_vectors(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def sample_vmf(rng, mu, kappa):
    """Draw one direction per row from vMF(mu[i], kappa[i]) (stable inverse CDF)."""
    n = mu.shape[0]
    u = rng.random(n)
    # cos(theta) = 1 + log(u + (1 - u) exp(-2 kappa)) / kappa
    w = 1.0 + np.log(u + (1.0 - u) * np.exp(-2.0 * kappa)) / kappa
    w = np.clip(w, -1.0, 1.0)
    phi = 2.0 * np.pi * rng.random(n)
    s = np.sqrt(np.maximum(0.0, 1.0 - w * w))
    local = np.stack([s * np.cos(phi), s * np.sin(phi), w], axis=1)
    # orthonormal frame around mu
    a = np.where(np.abs(mu[:, 0:1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(mu, a)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(mu, t)
    d = local[:, 0:1] * t + local[:, 1:2] * b + local[:, 2:3] * mu
    return d / np.linalg.norm(d, axis=1, keepdims=True)


class SceneModel:
    """Blob centres, drift directions and per-blob lobes of one synthetic scene."""

    def __init__(self, seed=0, num_blobs=16, lobes=4):
        rng = np.random.Generator(np.random.PCG64(seed))
        self.centres = rng.uniform(0.1, 0.9, size=(num_blobs, 3))
        self.drift = _unit_vectors(rng, num_blobs)
        self.lobe_mu = _unit_vectors(rng, num_blobs * lobes).reshape(num_blobs, lobes, 3)
        self.lobe_kappa = np.exp(rng.uniform(np.log(5.0), np.log(500.0), size=(num_blobs, lobes)))
        self.num_blobs = num_blobs
        self.lobes = lobes

    def centres_at(self, frame, drift_per_frame=0.01):
        return self.centres + drift_per_frame * frame * self.drift


def generate(n, seed=0, frame=0, num_blobs=16, flags=0, scene_seed=None, blob_fraction=0.8):
    """Return a structured array of n PGLSampleData records."""
    scene = SceneModel(seed if scene_seed is None else scene_seed, num_blobs)
    rng = np.random.Generator(np.random.PCG64([seed, frame, flags, 7919]))
    centres = scene.centres_at(frame)
    nb = int(round(n * blob_fraction))
    blob = rng.integers(0, num_blobs, size=n)
    pos = np.empty((n, 3))
    pos[:nb] = centres[blob[:nb]] + rng.normal(scale=0.05, size=(nb, 3))
    pos[nb:] = rng.random((n - nb, 3))
    if n - nb > 0:
        # nearest blob centre (first minimum on ties)
        q = pos[nb:]
        best = np.full(n - nb, np.inf)
        arg = np.zeros(n - nb, dtype=blob.dtype)
        for b in range(num_blobs):
            d2 = (q[:, 0] - centres[b, 0]) ** 2 + (q[:, 1] - centres[b, 1]) ** 2 + (q[:, 2] - centres[b, 2]) ** 2
            closer = d2 < best
            best[closer] = d2[closer]
            arg[closer] = b
        blob[nb:] = arg
    lobe = rng.integers(0, scene.lobes, size=n)
    mu = scene.lobe_mu[blob, lobe]
    kappa = scene.lobe_kappa[blob, lobe]
    d = sample_vmf(rng, mu, kappa)
    out = np.zeros(n, dtype=SAMPLE_DTYPE)
    out["pos"] = pos.astype(np.float32)
    out["dir"] = d.astype(np.float32)
    out["weight"] = np.exp(rng.normal(size=n)).astype(np.float32)
    out["pdf"] = np.float32(1.0 / (4.0 * np.pi))
    out["distance"] = rng.exponential(1.0, size=n).astype(np.float32)
    out["flags"] = flags
    perm = rng.permutation(n)
    return out[perm]


def generate_large(n, seed=0, frame=0, num_blobs=64, flags=0, block=1 << 20, threads=None):
    """generate() for large n: blocks of `block` records from independent random
    streams over the same scene (scene_seed = seed), produced on a thread pool
    (the result does not depend on the thread count)."""
    import concurrent.futures
    import os
    counts = [min(block, n - i) for i in range(0, n, block)]
    workers = threads or min(16, os.cpu_count() or 1)

    def part(b):
        return generate(counts[b], seed=seed * 1000003 + b + 1, frame=frame, num_blobs=num_blobs, flags=flags,
                        scene_seed=seed)

    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        parts = list(ex.map(part, range(len(counts))))
    return np.concatenate(parts)


def region_samples(n, seed=0, lobes=4):
    """Samples of one synthetic region: positions in a 0.05-sigma blob at the
    origin, directions from a 4-lobe vMF mixture (kappa log-U[5, 500])."""
    rng = np.random.Generator(np.random.PCG64([seed, 4242]))
    mu = _unit_vectors(rng, lobes)
    kappa = np.exp(rng.uniform(np.log(5.0), np.log(500.0), size=lobes))
    lobe = rng.integers(0, lobes, size=n)
    out = np.zeros(n, dtype=SAMPLE_DTYPE)
    out["pos"] = rng.normal(scale=0.05, size=(n, 3)).astype(np.float32)
    out["dir"] = sample_vmf(rng, mu[lobe], kappa[lobe]).astype(np.float32)
    out["weight"] = np.exp(rng.normal(size=n)).astype(np.float32)
    out["pdf"] = np.float32(1.0 / (4.0 * np.pi))
    out["distance"] = rng.exponential(1.0, size=n).astype(np.float32)
    return out


def queries(n, seed=0, scene_seed=0, num_blobs=16, volume_fraction=0.3):
    """Query batch: position near a blob, normal uniform on the sphere, 2
    uniforms; a volume_fraction of them carry an HG mean cosine g in [-0.9, 0.9]."""
    scene = SceneModel(scene_seed, num_blobs)
    rng = np.random.Generator(np.random.PCG64([seed, 31337]))
    blob = rng.integers(0, num_blobs, size=n)
    pos = (scene.centres[blob] + rng.normal(scale=0.05, size=(n, 3))).astype(np.float32)
    nrm = _unit_vectors(rng, n).astype(np.float32)
    u2 = rng.random((n, 2)).astype(np.float32)
    is_vol = rng.random(n) < volume_fraction
    g = rng.uniform(-0.9, 0.9, size=n).astype(np.float32)
    return pos, nrm, u2, is_vol, g


def _region_block(counts, mu, kappa, seed, step, block):
    """Records of consecutive regions (one RNG stream per block of regions)."""
    rng = np.random.Generator(np.random.PCG64([seed, 778, step, block]))
    total = int(counts.sum())
    region = np.repeat(np.arange(len(counts)), counts)
    lobe = rng.integers(0, 4, size=total)
    out = np.zeros(total, dtype=SAMPLE_DTYPE)
    out["pos"] = rng.normal(scale=0.05, size=(total, 3)).astype(np.float32)
    out["dir"] = sample_vmf(rng, mu[region, lobe], kappa[region, lobe]).astype(np.float32)
    out["weight"] = np.exp(rng.normal(size=total)).astype(np.float32)
    out["pdf"] = np.float32(1.0 / (4.0 * np.pi))
    out["distance"] = rng.exponential(1.0, size=total).astype(np.float32)
    stats = np.zeros((len(counts), 7), np.float32)
    off = 0
    for r, n in enumerate(counts):
        n = int(n)
        if n:
            p = out["pos"][off:off + n].astype(np.float64)
            m = p.mean(axis=0)
            stats[r, 0:3] = m
            stats[r, 3:6] = ((p - m) ** 2).sum(axis=0)
        stats[r, 6] = n
        off += n
    return out, stats


def batched_regions(num_regions=4096, seed=0, nmin=64, nmax=65536, ks=(8, 16, 32), step=0, threads=None):
    """Pre-partitioned regions for batched directional training.  Region r has
    floor(exp(U[ln nmin, ln nmax])) samples (fixed per region across steps),
    an initial component count drawn from ks, positions in a 0.05-sigma blob
    and directions from its own 4-lobe vMF mixture (kappa log-U[5, 500]);
    weights lognormal(0, 1), pdf 1/(4 pi), distance Exp(1).  step selects an
    independent draw of new samples from the same mixtures.  Generated in
    blocks of 64 regions (one random stream each, so the result does not
    depend on the thread count).  Returns (records region-contiguous,
    counts u32, stats7 f32 (mean[3], variance * n[3], n) of the positions,
    initK u32)."""
    import concurrent.futures
    import os
    rng = np.random.Generator(np.random.PCG64([seed, 777]))
    counts = np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=num_regions))).astype(np.uint32)
    init_k = rng.choice(np.asarray(ks, np.uint32), size=num_regions)
    mu = _unit_vectors(rng, num_regions * 4).reshape(num_regions, 4, 3)
    kappa = np.exp(rng.uniform(np.log(5.0), np.log(500.0), size=(num_regions, 4)))
    B = 64
    blocks = [(b, slice(b * B, min(num_regions, (b + 1) * B))) for b in range((num_regions + B - 1) // B)]
    workers = threads or min(16, os.cpu_count() or 1)
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        parts = list(ex.map(lambda bs: _region_block(counts[bs[1]], mu[bs[1]], kappa[bs[1]], seed, step, bs[0]),
                            blocks))
    records = np.concatenate([p[0] for p in parts])
    stats7 = np.concatenate([p[1] for p in parts])
    return records, counts, stats7, init_k


# PGLPathSegmentData (include/openpgl/data.h:50-113), 136 bytes
PATH_SEGMENT_DTYPE = np.dtype({
    "names": ["position", "directionIn", "directionOut", "normal", "volumeScatter", "pdfDirectionIn", "isDelta",
              "scatteringWeight", "transmittanceWeight", "directContribution", "miWeight", "scatteredContribution",
              "russianRouletteProbability", "eta", "roughness", "regionPtr"],
    "formats": [("<f4", 3), ("<f4", 3), ("<f4", 3), ("<f4", 3), "u1", "<f4", "u1", ("<f4", 3), ("<f4", 3),
                ("<f4", 3), "<f4", ("<f4", 3), "<f4", "<f4", "<f4", "<u8"],
    "offsets": [0, 12, 24, 36, 48, 52, 56, 60, 72, 84, 96, 100, 112, 116, 120, 128],
    "itemsize": 136})


def random_paths(num_paths, seed=0, max_len=12, scene_seed=0, num_blobs=16):
    """Recorded paths of a synthetic scene: vertex positions near the scene's
    blobs, consecutive vertices joined by their directions, a mix of diffuse,
    glossy and delta (some refractive) vertices, volume vertices, sparse
    emission and scattered contributions, Russian roulette after the third
    vertex (sometimes terminating with probability 0).  Returns the vertex
    array (PATH_SEGMENT_DTYPE) and the path offsets (num_paths + 1, uint32)."""
    rng = np.random.Generator(np.random.PCG64([seed, 4243]))
    scene = SceneModel(scene_seed, num_blobs)
    lengths = rng.integers(0, max_len + 1, size=num_paths)
    offsets = np.zeros(num_paths + 1, np.uint32)
    offsets[1:] = np.cumsum(lengths)
    n = int(offsets[-1])
    # built as 34 float32 columns (one 136-byte record per row): single-column
    # strided writes are fast, 3-wide sub-array field writes are not
    raw = np.zeros((n, 34), np.float32)
    if n == 0:
        return raw.view(PATH_SEGMENT_DTYPE).reshape(0), offsets
    col = {name: PATH_SEGMENT_DTYPE.fields[name][1] // 4 for name in PATH_SEGMENT_DTYPE.names}

    def put(name, values):
        values = np.asarray(values, np.float32)
        if values.ndim == 1:
            raw[:, col[name]] = values
        else:
            for k in range(values.shape[1]):
                raw[:, col[name] + k] = values[:, k]

    def unit(v):
        return v / np.sqrt((v * v).sum(1))[:, None]

    path_of = np.repeat(np.arange(num_paths), lengths)
    index = np.arange(n) - offsets[path_of].astype(np.int64)
    blob = rng.integers(0, num_blobs, size=n)
    pos = scene.centres[blob] + rng.normal(scale=0.08, size=(n, 3))
    put("position", pos)
    last = index == lengths[path_of] - 1
    nxt = np.where(last, np.arange(n), np.arange(n) + 1)
    d_in = pos[nxt] - pos
    d_in[last] = rng.normal(size=(int(last.sum()), 3))
    put("directionIn", unit(d_in))
    first = index == 0
    prv = np.where(first, np.arange(n), np.arange(n) - 1)
    d_out = pos[prv] - pos
    d_out[first] = rng.normal(size=(int(first.sum()), 3))
    put("directionOut", unit(d_out))
    put("normal", unit(rng.normal(size=(n, 3))))
    flags = raw.view(np.uint8)
    flags[:, PATH_SEGMENT_DTYPE.fields["volumeScatter"][1]] = rng.random(n) < 0.2
    delta = rng.random(n) < 0.15
    flags[:, PATH_SEGMENT_DTYPE.fields["isDelta"][1]] = delta
    rough = rng.random(n)
    rough[rng.random(n) < 0.1] = 0.005
    put("roughness", rough)
    eta = np.ones(n)
    eta[delta & (rng.random(n) < 0.5)] = 1.5
    put("eta", eta)
    put("pdfDirectionIn", np.exp(rng.uniform(np.log(0.01), np.log(5.0), n)))
    put("scatteringWeight", rng.uniform(0.0, 1.5, (n, 3)))
    put("transmittanceWeight", np.where(rng.random((n, 1)) < 0.7, 1.0, rng.uniform(0.3, 1.0, (n, 3))))
    emit = rng.random(n) < 0.2
    put("directContribution", rng.uniform(0.0, 4.0, (n, 3)) * emit[:, None])
    sc = rng.random(n) < 0.3
    put("scatteredContribution", rng.uniform(0.0, 2.0, (n, 3)) * sc[:, None])
    put("miWeight", rng.random(n))
    rr = np.where(index < 3, 1.0, rng.uniform(0.2, 1.0, n))
    rr[(index >= 3) & (rng.random(n) < 0.05)] = 0.0
    put("russianRouletteProbability", rr)
    return raw.view(PATH_SEGMENT_DTYPE).reshape(n), offsets
