// Open PGL 0.6.0 field files: pglFieldStoreToFile / pglDeviceNewFieldFromFile.
//
// The byte layout is the reference's stream serialisation, member by member
// (native little-endian, bool = 1 byte, size_t = 8 bytes, no padding):
//   SurfaceVolumeField::storeToFile   field/SurfaceVolumeField.h:197-213
//   SurfaceVolumeField::serialize     field/SurfaceVolumeField.h:171-185
//   Field::serialize                  field/Field.h:240-267
//   Configuration (split and merge)   directional/vmm/AdaptiveSplitandMergeFactory.h:246-275
//   Configuration (weighted EM)       directional/vmm/ParallaxAwareVonMisesFisherWeightedEMFactory.h:1335-1367
//   KDTreePartitionBuilder::Settings  spatial/kdtree/KDTreeBuilder.h:377-390
//   KDTree / KDNode                   spatial/kdtree/KDTree.h:136-146, 592-619
//   Region                            spatial/Region.h:63-80
//   ParallaxAwareVonMisesFisherMixture  directional/vmm/ParallaxAwareVonMisesFisherMixture.h:386-419
//   SufficientStatistics              directional/vmm/ParallaxAwareVonMisesFisherWeightedEMFactory.h:346-375
//   ComponentSplitStatistics          directional/vmm/VMMChiSquareComponentSplitter.h:871-901
//   Statistics (split and merge)      directional/vmm/AdaptiveSplitandMergeFactory.h:141-160
//   SampleStatistics                  data/SampleStatistics.h:157-171
//   Range                             data/Range.h:29-38
//   KNearestRegionsSearchTree         spatial/KNN.h:320-368
//   Device::newFieldFromFile          device/Device.h:217-271
//
// The per-component arrays are stored as the reference's vfloat<W> blocks:
// for block b = 0 .. 32/W - 1, lanes [bW, bW + W) of each array of the block
// in member order (a Vec3 of vfloat<W> is x[W], y[W], z[W]).  W is the SIMD
// width of the device the file is written for or read by; the file itself
// does not record it.
#include <cstring>
#include <fstream>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "field.h"

namespace pglgpu {

namespace {

const char kFieldHeader[] = "OPENPGL_0.6.0_FIELD";  // "OPENPGL_" OPENPGL_VERSION_STRING "_FIELD"
const int32_t kKdTree = 0;                         // PGL_SPATIAL_STRUCTURE_KDTREE
const int32_t kParallaxAwareVMM = 0;               // PGL_DIRECTIONAL_DISTRIBUTION_PARALLAX_AWARE_VMM
const int32_t kQuadtree = 1;                       // PGL_DIRECTIONAL_DISTRIBUTION_QUADTREE
const int32_t kVMM = 2;                            // PGL_DIRECTIONAL_DISTRIBUTION_VMM

struct Writer {
  std::ostream& os;
  template <class T> void put(const T& v) { os.write(reinterpret_cast<const char*>(&v), sizeof(T)); }
  void bytes(const void* p, size_t n) { os.write(reinterpret_cast<const char*>(p), (std::streamsize)n); }
  void flag(bool b) { put<uint8_t>(b ? 1 : 0); }
  void size(uint64_t v) { put<uint64_t>(v); }
  void lanes(const float* a, int b, int W) { bytes(a + b * W, sizeof(float) * W); }
  void zeros(int W) {
    const float z[16] = {};
    bytes(z, sizeof(float) * W);
  }
};

struct Reader {
  std::istream& is;
  void bytes(void* p, size_t n) {
    is.read(reinterpret_cast<char*>(p), (std::streamsize)n);
    if (!is) throw std::runtime_error("error: unexpected end of field file");
  }
  template <class T> T get() {
    T v;
    bytes(&v, sizeof(T));
    return v;
  }
  bool flag() { return get<uint8_t>() != 0; }
  uint64_t size() { return get<uint64_t>(); }
  void lanes(float* a, int b, int W) { bytes(a + b * W, sizeof(float) * W); }
  void skip(int W) {
    float z[16];
    bytes(z, sizeof(float) * W);
  }
};

uint32_t checked_count(uint64_t v, uint64_t limit, const char* what) {
  if (v > limit) throw std::runtime_error(std::string("error: invalid field file (") + what + ")");
  return (uint32_t)v;
}

void check_width(int W) {
  if (W != 4 && W != 8 && W != 16) throw std::runtime_error("openpgl_amd: field files need a SIMD width of 4, 8 or 16");
}

// Region::serialize (spatial/Region.h:63-71) followed by the storage's Range
void write_region(Writer& w, const RegionState& S, const KindField::RegionHost& h, int W) {
  const int B = kMaxComponents / W;
  w.flag(S.valid != 0);
  // distribution
  const VMMParams& V = S.vmm;
  for (int b = 0; b < B; b++) {
    w.lanes(V.w, b, W);
    w.lanes(V.kappa, b, W);
    w.lanes(V.mux, b, W);
    w.lanes(V.muy, b, W);
    w.lanes(V.muz, b, W);
    w.lanes(V.norm, b, W);
    w.lanes(V.em2k, b, W);
    w.lanes(V.mc, b, W);
    w.lanes(V.dist, b, W);
  }
  w.size((uint64_t)V.K);
  w.bytes(V.pivot, 12);
  // region bounds
  w.bytes(h.lo, 12);
  w.bytes(h.hi, 12);
  // training statistics: sufficient statistics
  const SuffStats& U = S.suff;
  for (int b = 0; b < B; b++) {
    w.lanes(U.swdx, b, W);
    w.lanes(U.swdy, b, W);
    w.lanes(U.swdz, b, W);
    w.lanes(U.sw, b, W);
    w.lanes(U.sdw, b, W);
  }
  w.put<float>(U.sumWeights);
  w.put<float>(U.numSamples);
  w.put<float>(U.overall);
  w.size((uint64_t)U.K);
  w.flag(U.normalized != 0);
  // training statistics: split statistics (zero-mean: the split means stay 0)
  const SplitStatsState& P = S.split;
  for (int b = 0; b < B; b++) {
    w.lanes(P.chi, b, W);
    w.zeros(W);
    w.zeros(W);
    w.lanes(P.covx, b, W);
    w.lanes(P.covy, b, W);
    w.lanes(P.covz, b, W);
    w.lanes(P.sn, b, W);
    w.lanes(P.ssw, b, W);
    w.lanes(P.ssa, b, W);
  }
  w.size((uint64_t)P.K);
  w.size(((uint64_t)S.afterSplitHi << 32) | S.afterSplitLo);
  w.size(((uint64_t)S.afterMergeHi << 32) | S.afterMergeLo);
  // sample statistics
  w.bytes(h.ss.mean, 12);
  w.bytes(h.ss.var, 12);
  w.put<float>(h.ss.n);
  w.bytes(h.ss.lo, 12);
  w.bytes(h.ss.hi, 12);
  w.flag(h.splitFlag);
  // Range: reset after every fit/update (Field.h:433, 506)
  w.size(0);
  w.size(0);
}

int32_t component_count(uint64_t v) { return (int32_t)checked_count(v, kMaxComponents, "component count"); }

void read_region(Reader& r, RegionState& S, KindField::RegionHost& h, int W) {
  const int B = kMaxComponents / W;
  memset(&S, 0, sizeof(S));
  S.valid = r.flag() ? 1 : 0;
  VMMParams& V = S.vmm;
  for (int b = 0; b < B; b++) {
    r.lanes(V.w, b, W);
    r.lanes(V.kappa, b, W);
    r.lanes(V.mux, b, W);
    r.lanes(V.muy, b, W);
    r.lanes(V.muz, b, W);
    r.lanes(V.norm, b, W);
    r.lanes(V.em2k, b, W);
    r.lanes(V.mc, b, W);
    r.lanes(V.dist, b, W);
  }
  V.K = component_count(r.size());
  r.bytes(V.pivot, 12);
  r.bytes(h.lo, 12);
  r.bytes(h.hi, 12);
  SuffStats& U = S.suff;
  for (int b = 0; b < B; b++) {
    r.lanes(U.swdx, b, W);
    r.lanes(U.swdy, b, W);
    r.lanes(U.swdz, b, W);
    r.lanes(U.sw, b, W);
    r.lanes(U.sdw, b, W);
  }
  U.sumWeights = r.get<float>();
  U.numSamples = r.get<float>();
  U.overall = r.get<float>();
  U.K = component_count(r.size());
  U.normalized = r.flag() ? 1 : 0;
  SplitStatsState& P = S.split;
  for (int b = 0; b < B; b++) {
    r.lanes(P.chi, b, W);
    r.skip(W);  // split means: always zero under the zero-mean splitter
    r.skip(W);
    r.lanes(P.covx, b, W);
    r.lanes(P.covy, b, W);
    r.lanes(P.covz, b, W);
    r.lanes(P.sn, b, W);
    r.lanes(P.ssw, b, W);
    r.lanes(P.ssa, b, W);
  }
  P.K = component_count(r.size());
  const uint64_t afterSplit = r.size(), afterMerge = r.size();
  S.afterSplitLo = (uint32_t)afterSplit;
  S.afterSplitHi = (uint32_t)(afterSplit >> 32);
  S.afterMergeLo = (uint32_t)afterMerge;
  S.afterMergeHi = (uint32_t)(afterMerge >> 32);
  r.bytes(h.ss.mean, 12);
  r.bytes(h.ss.var, 12);
  h.ss.n = r.get<float>();
  r.bytes(h.ss.lo, 12);
  r.bytes(h.ss.hi, 12);
  h.splitFlag = r.flag();
  S.splitFlag = h.splitFlag ? 1 : 0;
  h.count = 0;
  h.copyFrom = -1;
  (void)r.size();  // Range begin / end
  (void)r.size();
}

}  // namespace

void KindField::writeFile(std::ostream& os, int W, bool isSurface, hipStream_t s) const {
  check_width(W);
  syncMirror(s);
  Writer w{os};
  w.flag(isSurface);
  w.put<float>(fc_.decayOnSpatialSplit);
  w.size(iteration_);
  w.size(0);  // m_totalSPP (never incremented by the reference)
  w.flag(deterministic_);
  w.flag(fileSettings.fitRegions);
  w.flag(boundsSet_);
  w.bytes(sceneLo_, 12);
  w.bytes(sceneHi_, 12);
  w.flag(initialized_);
  w.bytes(lastTimes_, 16);
  // directional factory configuration (weighted EM, then split and merge)
  w.size((uint64_t)fc_.initK);
  w.put<float>(fc_.initKappa);
  w.size(fileSettings.maxK);
  w.size((uint64_t)fc_.maxEMIterations);
  w.put<float>(fileSettings.maxKappa);
  w.put<float>(fc_.maxMeanCosine);
  w.put<float>(fc_.convergenceThreshold);
  w.put<float>(fc_.weightPrior);
  w.put<float>(fc_.meanCosinePriorStrength);
  w.put<float>(fc_.meanCosinePrior);
  w.put<float>(fc_.splittingThreshold);
  w.put<float>(fc_.mergingThreshold);
  w.flag(fc_.partialReFit != 0);
  w.put<int32_t>(fileSettings.maxSplitItr);
  w.put<int32_t>(fc_.minSamplesForSplitting);
  w.put<int32_t>(fc_.minSamplesForMerging);
  w.put<int32_t>(fc_.minSamplesForPartialRefitting);
  // kd builder settings
  w.size(kd_.minSamples);
  w.size(kd_.maxSamples);
  w.size(kd_.maxDepth);
  // kd tree
  w.flag(initialized_);
  if (initialized_) {
    w.bytes(treeLo_, 12);
    w.bytes(treeHi_, 12);
  } else {
    const float z[6] = {};
    w.bytes(z, 24);
  }
  w.size(nodes_.size());
  for (const KDNode& n : nodes_) {
    w.put<float>(n.pivot);
    w.put<uint32_t>(n.bits);
  }
  // regions
  const uint32_t R = (uint32_t)regions_.size();
  w.size(R);
  for (uint32_t i = 0; i < R; i++) write_region(w, hostStates_[i], regions_[i], W);
  // KNN search tree over the regions' sample means
  w.flag(kd_.knnLookup);
  const bool built = kd_.knnLookup && initialized_;
  w.flag(built);
  if (built) {
    w.put<uint32_t>(R);
    for (uint32_t i = 0; i < R; i++) {
      const float p[4] = {regions_[i].ss.mean[0], regions_[i].ss.mean[1], regions_[i].ss.mean[2], 0.f};
      w.bytes(p, 16);
    }
  }
}

void KindField::readFile(std::istream& is, int W, bool parallaxAware, hipStream_t s) {
  check_width(W);
  finish();
  Reader r{is};
  (void)r.flag();  // m_isSurface: the kind is given by the position in the file
  const float decay = r.get<float>();
  const uint64_t iteration = r.size();
  (void)r.size();  // m_totalSPP
  const bool deterministic = r.flag();
  const bool fitRegions = r.flag();
  const bool boundsSet = r.flag();
  float sceneLo[3], sceneHi[3];
  r.bytes(sceneLo, 12);
  r.bytes(sceneHi, 12);
  const bool initialized = r.flag();
  float times[4];
  r.bytes(times, 16);

  FactoryConfig fc = fc_;
  FieldFileSettings fs;
  fc.initK = (int32_t)checked_count(r.size(), kMaxComponents, "initK");
  fc.initKappa = r.get<float>();
  fs.maxK = r.size();
  fc.maxEMIterations = (int32_t)checked_count(r.size(), 0x7fffffff, "maxEMIterrations");
  fs.maxKappa = r.get<float>();
  fc.maxMeanCosine = r.get<float>();
  fc.convergenceThreshold = r.get<float>();
  fc.weightPrior = r.get<float>();
  fc.meanCosinePriorStrength = r.get<float>();
  fc.meanCosinePrior = r.get<float>();
  fc.splittingThreshold = r.get<float>();
  fc.mergingThreshold = r.get<float>();
  fc.partialReFit = r.flag() ? 1 : 0;
  fs.maxSplitItr = r.get<int32_t>();
  fc.minSamplesForSplitting = r.get<int32_t>();
  fc.minSamplesForMerging = r.get<int32_t>();
  fc.minSamplesForPartialRefitting = r.get<int32_t>();
  fc.useSplitAndMerge = 1;  // not stored: a loaded Configuration keeps its default (AdaptiveSplitandMergeFactory.h:43)
  fc.decayOnSpatialSplit = decay;
  fc.vectorWidth = W;
  fc.noParallax = parallaxAware ? 0 : 1;
  if (fc.initK < 1) throw std::runtime_error("error: invalid field file (initK)");
  fs.fitRegions = fitRegions;

  KdSettings kd = kd_;
  kd.minSamples = checked_count(r.size(), 0xffffffffu, "minSamples");
  kd.maxSamples = checked_count(r.size(), 0xffffffffu, "maxSamples");
  kd.maxDepth = checked_count(r.size(), 0xffffffffu, "maxDepth");

  const bool treeInit = r.flag();
  float treeLo[3], treeHi[3];
  r.bytes(treeLo, 12);
  r.bytes(treeHi, 12);
  const uint32_t N = checked_count(r.size(), 0x3fffffffu, "node count");
  std::vector<KDNode> nodes(N);
  for (uint32_t i = 0; i < N; i++) {
    nodes[i].pivot = r.get<float>();
    nodes[i].bits = r.get<uint32_t>();
  }
  const uint32_t R = checked_count(r.size(), 0x3fffffffu, "region count");
  std::vector<RegionState> states(R);
  std::vector<RegionHost> regions(R);
  for (uint32_t i = 0; i < R; i++) read_region(r, states[i], regions[i], W);
  const bool knn = r.flag();
  const bool built = r.flag();
  std::vector<float4> knnPts;
  if (built) {
    const uint32_t P = r.get<uint32_t>();
    if (P != R) throw std::runtime_error("error: invalid field file (KNN point count)");
    knnPts.resize(P);
    for (uint32_t i = 0; i < P; i++) {
      float p[4];
      r.bytes(p, 16);
      knnPts[i] = make_float4(p[0], p[1], p[2], 0.f);
    }
  }
  // the tree must reference only existing nodes and regions
  for (uint32_t i = 0; i < N; i++) {
    const KDNode& n = nodes[i];
    if (n.isLeaf() ? n.idx() >= R : (n.idx() + 1 >= N || n.idx() <= i))
      throw std::runtime_error("error: invalid field file (kd tree)");
  }
  if (initialized && (N == 0 || R == 0 || !treeInit)) throw std::runtime_error("error: invalid field file (empty tree)");

  // replace the state
  kd_ = kd;
  fc_ = fc;
  fileSettings = fs;
  deterministic_ = deterministic;
  kd_.knnLookup = knn;
  for (int i = 0; i < 3; i++) {
    sceneLo_[i] = sceneLo[i];
    sceneHi_[i] = sceneHi[i];
    treeLo_[i] = treeLo[i];
    treeHi_[i] = treeHi[i];
  }
  boundsSet_ = boundsSet;
  initialized_ = initialized;
  iteration_ = (uint32_t)iteration;
  for (int i = 0; i < 4; i++) lastTimes_[i] = times[i];
  nodes_.swap(nodes);
  regions_.swap(regions);
  keyNodes_ = nodes_;
  splitRegion_.assign(nodes_.size(), -1);
  hintRegions_ = 0;
  info_ = UpdateInfo();
  history_.clear();

  if (R > 0) {
    ensureRegionCapacity(R, s);
    PGL_HIP_CHECK(hipMemcpyAsync(dStates_.p, states.data(), size_t(R) * sizeof(RegionState), hipMemcpyHostToDevice, s));
  }
  if (N > 0) {
    dNodes_.ensure(size_t(N) * sizeof(KDNode));
    PGL_HIP_CHECK(hipMemcpyAsync(dNodes_.p, nodes_.data(), size_t(N) * sizeof(KDNode), hipMemcpyHostToDevice, s));
  }
  if (kd_.knnLookup && R > 0) {
    // neighbourhoods of the stored search-tree points (or of the region means
    // when the file holds no search tree)
    if (knnPts.empty()) {
      knnPts.resize(R);
      for (uint32_t i = 0; i < R; i++)
        knnPts[i] = make_float4(regions_[i].ss.mean[0], regions_[i].ss.mean[1], regions_[i].ss.mean[2], 0.f);
    }
    dNbIds_.ensure(size_t(R) * 8 * 4);
    dNbPts_.ensure(size_t(R) * 8 * sizeof(float4));
    dTables_[0].ensure(size_t(R) * sizeof(float4));  // no update is pending: the tables are free
    PGL_HIP_CHECK(hipMemcpyAsync(dTables_[0].p, knnPts.data(), size_t(R) * sizeof(float4), hipMemcpyHostToDevice, s));
    launch_knn8(dTables_[0].as<float4>(), R, dNbIds_.as<uint32_t>(), dNbPts_.as<float4>(), s);
  }
  PGL_HIP_CHECK(hipStreamSynchronize(s));
  {
    std::lock_guard<std::mutex> lock(mirrorMutex_);
    mirrorDirty_ = true;
  }
}

void store_field_file(const GuidingField& gf, const char* path, int W) {
  check_width(W);
  std::filebuf fb;
  if (!fb.open(path, std::ios::out | std::ios::binary)) throw std::runtime_error("error: couldn't open file!");
  std::ostream os(&fb);
  Writer w{os};
  w.bytes(kFieldHeader, sizeof(kFieldHeader));
  w.put<int32_t>(kKdTree);
  w.put<int32_t>(gf.surface.parallaxAware() ? kParallaxAwareVMM : kVMM);
  w.size(gf.iteration);
  w.size(0);  // m_totalSPP
  gf.surface.writeFile(os, W, true, gf.stream);
  gf.volume.writeFile(os, W, false, gf.stream);
  os.flush();
  if (!os) throw std::runtime_error("error: couldn't write field file");
  fb.close();
}

void load_field_file(GuidingField& gf, const char* path, int W) {
  check_width(W);
  std::filebuf fb;
  if (!fb.open(path, std::ios::in | std::ios::binary)) throw std::runtime_error("error: couldn't open file");
  std::istream is(&fb);
  Reader r{is};
  char header[sizeof(kFieldHeader)];
  is.read(header, sizeof(header));
  if (!is) throw std::runtime_error("error: invalid file header");
  const int32_t spatial = r.get<int32_t>();
  const int32_t directional = r.get<int32_t>();
  if (spatial != kKdTree || (directional != kParallaxAwareVMM && directional != kVMM && directional != kQuadtree))
    throw std::runtime_error("error: unrecognized field type");
  if (directional == kQuadtree)
    throw std::runtime_error("openpgl_amd: quadtree field files are not supported (VMM field types only)");
  const bool parallaxAware = directional == kParallaxAwareVMM;
  const uint64_t iteration = r.size();
  (void)r.size();  // m_totalSPP
  gf.surface.readFile(is, W, parallaxAware, gf.stream);
  gf.volume.readFile(is, W, parallaxAware, gf.stream);
  gf.iteration = (uint32_t)iteration;
}

}  // namespace pglgpu
