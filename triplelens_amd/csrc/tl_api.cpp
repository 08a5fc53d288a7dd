// Host side of the IRS engine: LightCurveIRS objects (lens, critical curves,
// caustic boxes, image-plane grid), the GPU engine that runs the seed and fill
// kernels, and the C ABI declared in include/tl_lcirs.h.
//
// Set-up follows the reference constructor (src/LightCurves/lcirs.cc:6-92,
// src/CCC/ccc.cc:42-203) with the same IEEE operations, so the grid (N,
// corners, pixel size) is the reference's to the last bit.  Everything per
// source position runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "tl_engine.h"
#include "../../include/tl_lcirs.h"

using namespace tl;

namespace {

thread_local std::string g_err;

void set_err(const std::string& s) {
  g_err = s;
  if (std::getenv("TL_VERBOSE")) std::fprintf(stderr, "liblcirs: %s\n", s.c_str());
}

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_));    \
  } while (0)

TrigTable host_trig() {
  TrigTable t;
  for (int i = 0; i <= kLagMaxIt; i++) host_sincos(1.0 * i, &t.s[i], &t.c[i]);
  return t;
}
const TrigTable& trig() {
  static const TrigTable t = host_trig();
  return t;
}

// ---------------------------------------------------------------------------
// Light-curve object
// ---------------------------------------------------------------------------
struct Curve {
  double a, b, th, m2, m3, rho;
  unsigned lcLength;
  long ppr;
  Lens L;
  cx fixedPoly[11];            // ImgPoint::getCoeffsOptNoZ (the reference caches it per lens)
  double rho2, rho2pow, ldA, ldB, intensity, ampScale;
  long long N;
  double Ldbl, step, blx, bly, trx, try_;
  cx boxMin[6], boxMax[6];
  bool binImg = false, binIrs = false;  // the binary lens's image-count rule and ray map
  std::vector<cx> cc, ca;      // 6 x kCccLen, root-major
  RootChain chain;             // root-reuse state carried between calls
  std::vector<cx> slotTrace;   // recorded tempRoots[10] values per getImages call (may be empty)
  std::vector<double> lcVec;   // lcLength values (copy_lc)
  std::vector<double> full;    // lcLength + 1 values of the last IRS curve
  std::vector<uint64_t> rays;
  std::vector<uint32_t> flags;
  std::vector<int> nSeeds, seedPix;  // seedPix: seedStride pixel pairs per position
  int seedStride = kMaxSeeds;        // the seed search's largest seed count (at least 1)
  // position p's seed row as 64 pixel pairs (pairs past the stride are 0)
  // the position's seeds in kMaxSeeds slots, the slots past its count zero
  void seed_row(int p, int* out) const {
    const size_t w = (size_t)std::min(nSeeds[p], seedStride) * 2;
    std::memcpy(out, seedPix.data() + (size_t)p * seedStride * 2, w * sizeof(int));
    std::memset(out + w, 0, ((size_t)kMaxSeeds * 2 - w) * sizeof(int));
  }
  std::vector<double> pointAmp;   // point-source magnification of the last IRS curve
  cx seedStart{0.0, 0.0}, seedEnd{0.0, 0.0};  // trajectory of the last seed search
  std::vector<double> centroid;   // x, y of the weighted image centroid per position
  std::vector<double> causDist;   // signed distance to the nearest caustic sample
  std::vector<double> pairShare;  // share of the point magnification in the two brightest images
  bool ring = false;              // seed searches also compute the disk-sample magnifications
  int split = 1;                  // independent pieces of the seed search
  std::vector<double> ringAmp;    // kRingSamples per position
  std::vector<int> ringCross;     // samples per position whose image count differs
  std::vector<cx> centreRoots;    // 10 per position: the roots of the position's own solve
  bool hasSeeds = false;          // a seed search (with or without fills) has run
  bool hasIRS = false;
  bool amoebaOut = false;
  std::string amoebaPrefix, parFile;
};

// utils.h:16-28
void minmax_(const cx* v, int n, cx* mn, cx* mx) {
  for (int i = 0; i < n; i++) {
    if (v[i].re < mn->re) mn->re = v[i].re;
    if (v[i].im < mn->im) mn->im = v[i].im;
    if (v[i].re > mx->re) mx->re = v[i].re;
    if (v[i].im > mx->im) mx->im = v[i].im;
  }
}

// exp(i phi_j), phi_j = 2*3.14159265359*j/kCccLen, with the host's sincos
// (the values the reference's critical-curve loop uses, ccc.cc:57).
const CccTable& ccc_table() {
  static const CccTable t = [] {
    CccTable r;
    for (int j = 0; j < kCccLen; j++) {
      const double phi = 2.0 * 3.14159265359 * j / kCccLen;
      host_sincos(phi, &r.e[j].im, &r.e[j].re);
    }
    return r;
  }();
  return t;
}

void critical_curves(const Lens& L, std::vector<cx>& cc, std::vector<cx>& ca) {
  cc.assign(6 * kCccLen, mk(0, 0));
  ca.assign(6 * kCccLen, mk(0, 0));
  critical_curves_into(L, ccc_table(), &trig(), cc.data(), ca.data());
}

void ld_default(Curve& c) {
  c.ldA = 0.6;
  c.ldB = 0.4;
  c.intensity = 1.0 / (c.ldA + 2.0 / 3.0 * c.ldB);
}

// cc, ca: the lens's critical curves and caustics (6 * kCccLen each, as
// critical_curves makes them), or null to compute them here.
Curve* make_curve(double a, double b, double th, double m2, double m3, double rho,
                  unsigned lcLength, long ppr, const cx* cc = nullptr, const cx* ca = nullptr) {
  if (!(rho > 0.0) || !std::isfinite(rho)) throw std::invalid_argument("sourceSize must be > 0");
  if (ppr <= 0) throw std::invalid_argument("pointsPerRadius must be > 0");
  if (lcLength == 0) throw std::invalid_argument("lcLength must be > 0");
  for (double v : {a, b, th, m2, m3})
    if (!std::isfinite(v)) throw std::invalid_argument("lens parameters must be finite");
  auto* c = new Curve();
  c->a = a; c->b = b; c->th = th; c->m2 = m2; c->m3 = m3; c->rho = rho;
  c->lcLength = lcLength;
  c->ppr = ppr;
  c->L = make_lens(a, b, th, m2, m3);
  image_poly_static(c->L, c->fixedPoly);
  c->rho2 = rho * rho;
  c->rho2pow = std::pow(rho, 2.0);
  ld_default(*c);
  if (cc && ca) {
    c->cc.assign(cc, cc + 6 * kCccLen);
    c->ca.assign(ca, ca + 6 * kCccLen);
  } else {
    critical_curves(c->L, c->cc, c->ca);
  }
  // the binary lens (m3 = 0): irsBinary (lcirs.cc:29-36) and the binary
  // image-count rule, which also holds for a or b zero (imgpoint.cc:64-67);
  // the critical curves stay the triple-lens form with m3 = 0
  c->binIrs = m3 == 0.0;
  c->binImg = a == 0.0 || b == 0.0 || m3 == 0.0;
  // _getCaBoxes (lcirs.cc:39-48): cumulative, origin included
  cx mn = mk(0, 0), mx = mk(0, 0);
  for (int k = 0; k < 6; k++) {
    minmax_(c->ca.data() + k * kCccLen, kCccLen, &mn, &mx);
    c->boxMin[k] = mn;
    c->boxMax[k] = mx;
  }
  // _getImgPlanePars (lcirs.cc:51-92)
  mn = mk(0, 0);
  mx = mk(0, 0);
  for (int k = 0; k < 6; k++) minmax_(c->cc.data() + k * kCccLen, kCccLen, &mn, &mx);
  double halfEdge = mx.re - mn.re;
  if (mx.im - mn.im > halfEdge) halfEdge = mx.im - mn.im;
  halfEdge *= 3.0 / 2.0;
  const double cxr = (mx.re + mn.re) / 2.0, cxi = (mx.im + mn.im) / 2.0;
  c->trx = cxr + halfEdge;
  c->try_ = cxi + halfEdge;
  c->blx = cxr - halfEdge;
  c->bly = cxi - halfEdge;
  const double Nd = 2 * halfEdge / rho * (double)ppr;
  if (!(Nd >= 3.0) || Nd > (double)((1ll << 25) - 2)) {
    delete c;
    throw std::invalid_argument("image-plane grid size out of range (3 .. 2^25-2 pixels)");
  }
  c->N = (long long)Nd;
  c->Ldbl = 2 * halfEdge;
  c->ampScale = c->intensity / M_PI / std::pow((double)ppr, 2.0);
  c->step = c->Ldbl / (double)((double)c->N - 1.0);
  for (int k = 0; k < 11; k++) c->chain.tempRoots[k] = mk(0, 0);
  c->chain.have = 0;
  c->chain.calls = 0;
  c->lcVec.assign(lcLength, 0.0);
  return c;
}

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  void ensure(size_t k) {
    if (k <= n) return;
    if (p) HIPCHK(hipFree(p));
    p = nullptr;
    n = 0;
    size_t want = std::max<size_t>(k, 16);
    HIPCHK(hipMalloc(&p, want * sizeof(T)));
    n = want;
  }
};

struct Options {
  int fillWorkgroups = 0;  // 0: one per compute unit (16 waves), capped at the occupancy limit
  int probeLimit = 4096;
  int queueLog2 = 21;
  int scoutRows = 8;
  int scoutJump = 512;
  int scoutWidth = 768;     // runs wider than this shorten the scout jump (0: never)
  int scoutRetry = 1;       // a scout that finds nothing is sent again half as far
  int useGraph = 1;         // each batch's reset, fill and resolve launched as one hipGraph
  int scoutShare = 0;       // sweeps queue their scouts in the shared queue at once
  double shareRays = 1e7;   // ... on positions predicted to have this many rays (0: never)
  int leapEvery = 4;        // every this many scouts a chain sends a leap scout (0: never)
  int leapJumps = 8;        // ... this many jumps ahead
  int pollBackoff = 6;
  int poolNewest = 1;       // workers take the pool's newest task first
  double runPixels = 128.0;  // predicted rays per run: sizes the run table
  double tableLoad = 4.0;   // run-table slots per predicted run
  double timeLimitS = 60.0;
  int seedChunks = 3;       // the seed search is cut into this many chunks of positions
  int seedChunkMin = 64;    // ... of at least this many positions per curve
  int specPiece = 32;       // verified speculative chain: positions per piece (0: the
                            // serial chain in chunks)
  int specOverlap = 16;     // ... positions a piece's chain runs into the next piece
  int specDepth = 4096;     // ... piece boundaries a repair chain may cross (at most the
                            // batch's largest piece count is used: a repair chain can then
                            // always run to its curve's end)
};

struct Stats {
  double tasks = 0, evals = 0, ins = 0, dup = 0, seedMs = 0, fillMs = 0, fillLaunches = 0,
         totalMs = 0, busyMs = 0, insertMs = 0, idlePolls = 0, waves = 0, scouts = 0,
         scoutRuns = 0, trialMs = 0, ringMs = 0, solves = 0, specPieces = 0, specRepairs = 0,
         specFallback = 0, graphBuilds = 0;
};

// One recorded run of the amoeba print-out: row, leftmost and rightmost pixel.
struct Segment {
  long long ny, xl, xr;
};

// Amoeba::printOut (amoeba.cc:90-108) for point i of the last curve, into
// amoebaPrefix + i as getLCIRS names it (lcirs.cc:183-186): one line
// "ny,xleft,xright" per run.  The reference lists runs in hash-map order;
// here they are sorted by row, then column.
void write_amoeba_file(const Curve& c, int i, std::vector<Segment>& segs) {
  std::sort(segs.begin(), segs.end(), [](const Segment& x, const Segment& y) {
    return x.ny != y.ny ? x.ny < y.ny : x.xl < y.xl;
  });
  std::ofstream f(c.amoebaPrefix + std::to_string(i));
  for (const Segment& g : segs) f << g.ny << "," << g.xl << "," << g.xr << "\n";
}

struct Job {          // one engine run over a set of curves
  std::vector<Curve*> curves;
  std::vector<cx> starts, ends;
  bool runSeeds = true;
  bool runFill = true;   // false: the seed search only
  // when runSeeds is false (lcirs_fill_seeds): explicit positions and seeds,
  // and (may be empty) the predicted rays of each position
  std::vector<double2> pos;
  std::vector<int> nSeeds;
  std::vector<int2> seeds;
  std::vector<float> cost;
  std::vector<int> nPos;  // positions per curve (runSeeds false; empty: all pos to one curve)
};

struct Engine {
  std::mutex mu;
  bool ready = false;
  int cus = 0;
  std::string name;
  hipStream_t stream = nullptr;   // fills and copies
  hipStream_t sstream = nullptr;  // seed search (overlaps the fills)
  std::vector<hipEvent_t> seedEvents;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evTrial0 = nullptr, evTrial1 = nullptr, evRing0 = nullptr, evRing1 = nullptr;
  DBuf<TrigTable> dTrig;
  DBuf<CurveDev> dCurves;
  DBuf<cx> dCaus;
  DBuf<int2> dSeeds;
  DBuf<int> dPoolBase, dPoolCount, dPoolOwner;
  DBuf<cx> dPoolZeta, dPoolCoef, dSlots;
  DBuf<cx> dPolyScratch;  // polynomials of the disk samples / point-source positions
  DBuf<double> dPtCausD2;
  std::vector<int2> hSeeds;  // host copy of the last seed search's seed table
  DBuf<unsigned long long> dPoolCursor;
  DBuf<int> dNSeeds, dPtCurve, dBatchPts, dLParent, dLPoint;
  DBuf<double2> dPtPos;
  DBuf<float> dCost;
  DBuf<double> dPtAmp, dPtCaus, dPtPair, dRing;
  DBuf<cx> dPtRoots;
  DBuf<int> dRingCross;
  DBuf<double2> dPtCen;
  DBuf<uint32_t> dFlags;
  DBuf<int2> dSeedTasks;
  DBuf<unsigned long long> dHtab, dLCount, dQSeq, dSync, dPtCount;
  DBuf<unsigned long long> dGrp;  // run-table slices: group bases, then masks
  DBuf<unsigned long long> dLSum, dPtSum;  // split sums (hi * 2^32 + lo): 2 words each
  DBuf<Task> dQBuf;
  DBuf<RootChain> dChain;
  DBuf<double> dPointLC;
  DBuf<Lens> dLenses;
  DBuf<CccTable> dCccTab;
  DBuf<cx> dCcOut, dCaOut;
  // verified speculative chain (SpecArgs)
  DBuf<int> dSpecBase, dSpecMerged, dSpecTarget, dSpecFailed;
  DBuf<cx> dSpecOwn, dSpecExt, dSpecRep, dSpecStart;
  std::vector<int> specCalls;  // per curve: the chain's solve count after the run
  int specPieces = 0;
  // pinned host staging for the large device-to-host copies (pageable
  // copies run at a fraction of the link's rate)
  void* hstage = nullptr;
  size_t hstageCap = 0;
  void copy_out(void* dst, const void* dev, size_t bytes) {
    if (bytes < (64u << 10)) {
      HIPCHK(hipMemcpy(dst, dev, bytes, hipMemcpyDeviceToHost));
      return;
    }
    constexpr size_t kStage = 16u << 20;
    if (!hstage) {
      HIPCHK(hipHostMalloc(&hstage, kStage, hipHostMallocDefault));
      hstageCap = kStage;
    }
    for (size_t off = 0; off < bytes; off += hstageCap) {
      const size_t n = std::min(hstageCap, bytes - off);
      HIPCHK(hipMemcpy(hstage, static_cast<const char*>(dev) + off, n, hipMemcpyDeviceToHost));
      std::memcpy(static_cast<char*>(dst) + off, hstage, n);
    }
  }
  // rows of `width` bytes, `pitch` bytes apart on the device, packed into dst
  void copy_out_rows(void* dst, const void* dev, size_t rows, size_t pitch, size_t width) {
    if (width == pitch) {
      copy_out(dst, dev, rows * width);
      return;
    }
    constexpr size_t kStage = 16u << 20;
    if (!hstage) {
      HIPCHK(hipHostMalloc(&hstage, kStage, hipHostMallocDefault));
      hstageCap = kStage;
    }
    const size_t per = std::max<size_t>(1, hstageCap / width);
    for (size_t r = 0; r < rows; r += per) {
      const size_t n = std::min(per, rows - r);
      HIPCHK(hipMemcpy2D(hstage, width, static_cast<const char*>(dev) + r * pitch, pitch, width, n,
                         hipMemcpyDeviceToHost));
      std::memcpy(static_cast<char*>(dst) + r * width, hstage, n * width);
    }
  }
  // the last fill launch's graph and the arguments it was captured with
  struct FillGraphKey {
    FillArgs fa;
    int nWg, labelCap, nb;
    unsigned long long* ptCount;
    unsigned long long* ptSum;
  };
  hipGraphExec_t fillExec = nullptr;
  FillGraphKey fillKey;
  void drop_fill_graph() {
    if (fillExec) (void)hipGraphExecDestroy(fillExec);
    fillExec = nullptr;
  }
  Options opt;
  Stats stats;

  void init() {
    if (ready) return;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
      throw std::runtime_error("no HIP device available: the IRS engine needs a gfx950 GPU");
    HIPCHK(hipSetDevice(0));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, 0));
    cus = prop.multiProcessorCount;
    name = prop.gcnArchName;
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&sstream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    for (hipEvent_t* e : {&evTrial0, &evTrial1, &evRing0, &evRing1}) HIPCHK(hipEventCreate(e));
    dTrig.ensure(1);
    HIPCHK(hipMemcpy(dTrig.p, &trig(), sizeof(TrigTable), hipMemcpyHostToDevice));
    ready = true;
  }

  float elapsed() {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
    return ms;
  }

  void run(Job& job);
  bool run_spec(std::vector<CurveDev>& hc, const std::vector<int>& pointBase, SeedArgs& sa);
  void point_lc(Curve& c, cx s, cx e);
  void collect_amoebae(const Job& job, const std::vector<int>& ptCurve,
                       const std::vector<int>& batchPts,
                       const std::vector<unsigned long long>& grpBase,
                       const std::vector<unsigned long long>& grpCap, int labelCap,
                       std::vector<std::vector<Segment>>& out);
};

// Reads the run table of the last fill back and files each counted run under
// its point: a run counts when its label's class holds the point's seed label
// (runs that only scouts reached, in components without a seed, do not).
void Engine::collect_amoebae(const Job& job, const std::vector<int>& ptCurve,
                             const std::vector<int>& batchPts,
                             const std::vector<unsigned long long>& grpBase,
                             const std::vector<unsigned long long>& grpCap, int labelCap,
                             std::vector<std::vector<Segment>>& out) {
  const unsigned long long hcap = grpBase.back() + grpCap.back();
  std::vector<unsigned long long> tab(2 * hcap);
  std::vector<int> parent(labelCap);
  HIPCHK(hipMemcpy(tab.data(), dHtab.p, tab.size() * sizeof(unsigned long long),
                   hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(parent.data(), dLParent.p, parent.size() * sizeof(int), hipMemcpyDeviceToHost));
  auto root = [&](int x) {
    while (x >= 0 && x < labelCap && parent[x] != x) x = parent[x];
    return x;
  };
  const int nb = (int)batchPts.size();
  for (size_t g = 0; g < grpBase.size(); g++) {
    for (unsigned long long e = grpBase[g]; e < grpBase[g] + grpCap[g]; e++) {
      const unsigned long long w0 = tab[2 * e], w1 = tab[2 * e + 1];
      if (w0 == 0) continue;
      const int b = (int)(g * kGroupPoints + (w0 & 0x3FFF));
      const long long xl = (long long)((w0 >> 14) & 0x1FFFFFF);
      const long long row = (long long)(w0 >> 39) - 1;
      const int label = (int)(w1 >> 32);
      const long long xr = (long long)(w1 & 0xFFFFFFFFull);
      if (b >= nb) continue;
      const int gp = batchPts[b];
      if (!job.curves[ptCurve[gp]]->amoebaOut) continue;
      if (label < 0 || label >= labelCap || root(label) != root(b)) continue;
      out[gp].push_back(Segment{row, xl, xr});
    }
  }
}

Engine& engine() {
  static Engine e;
  return e;
}

CurveDev to_dev(const Curve& c, cx s, cx e, int pointBase, int causticBase) {
  CurveDev d;
  std::memset(&d, 0, sizeof(d));
  d.L = c.L;
  for (int k = 0; k < 11; k++) d.fixedPoly[k] = c.fixedPoly[k];
  d.rho = c.rho;
  d.rho2 = c.rho2;
  d.rho2pow = c.rho2pow;
  d.ldA = c.ldA;
  d.ldB = c.ldB;
  d.sumScale = sum_scale_for(c.ldA, c.ldB);
  d.blx = c.blx;
  d.bly = c.bly;
  d.step = c.step;
  d.Ldbl = c.Ldbl;
  d.Nd = (double)c.N;
  d.N = c.N;
  d.sx = s.re; d.sy = s.im; d.ex = e.re; d.ey = e.im;
  d.ppr = (double)c.ppr;
  d.lcLength = c.lcLength;
  d.pointBase = pointBase;
  d.causticBase = causticBase;
  d.chainHave = c.chain.have;
  d.chainCalls = c.chain.calls;
  d.slotBase = 0;
  d.slotCount = 0;
  d.ring = c.ring ? 1 : 0;
  d.split = c.split;
  d.binImg = c.binImg ? 1 : 0;
  d.binIrs = c.binIrs ? 1 : 0;
  for (int k = 0; k < 11; k++) d.chain[k] = c.chain.tempRoots[k];
  for (int k = 0; k < 6; k++) {
    d.boxMin[k] = c.boxMin[k];
    d.boxMax[k] = c.boxMax[k];
  }
  return d;
}

// The verified speculative chain's rounds on the seed stream (after the
// solve pool is built): cuts every curve into pieces of about
// opt.specPiece positions, runs rounds 1 and 2 and the resolution, and on
// success leaves the piece start states in `sa` for seed_kernel (and the
// pieces in the uploaded curves).  Returns false, with the curves restored
// to one piece, when some curve could not be resolved within
// opt.specDepth piece boundaries.
bool Engine::run_spec(std::vector<CurveDev>& hc, const std::vector<int>& pointBase,
                      SeedArgs& sa) {
  const int nc = (int)hc.size();
  const int nPts = pointBase[nc];
  std::vector<int> base(nc + 1, 0);
  int maxP = 1;
  for (int k = 0; k < nc; k++) {
    const int n = pointBase[k + 1] - pointBase[k];
    const int P = std::max(1, n / opt.specPiece);
    hc[k].split = P;
    base[k + 1] = base[k] + P;
    maxP = std::max(maxP, P);
  }
  const int G = base[nc];
  HIPCHK(hipMemcpyAsync(dCurves.p, hc.data(), nc * sizeof(CurveDev), hipMemcpyHostToDevice,
                        sstream));
  dSpecBase.ensure(nc + 1);
  HIPCHK(hipMemcpyAsync(dSpecBase.p, base.data(), (nc + 1) * sizeof(int), hipMemcpyHostToDevice,
                        sstream));
  dSpecOwn.ensure((size_t)nPts * 10);
  dSpecExt.ensure((size_t)nPts * 10);
  dSpecMerged.ensure(G);
  dSpecTarget.ensure(G);
  const int depth = std::max(1, std::min(opt.specDepth, maxP));
  dSpecRep.ensure((size_t)G * depth * 10);
  dSpecStart.ensure((size_t)G * 10);
  dSpecFailed.ensure(1);
  int zero = 0;
  HIPCHK(hipMemcpyAsync(dSpecFailed.p, &zero, sizeof(int), hipMemcpyHostToDevice, sstream));
  SpecArgs sp;
  sp.pieceBase = dSpecBase.p;
  sp.nPieces = G;
  sp.overlap = std::max(1, opt.specOverlap);
  sp.depth = depth;
  sp.own = dSpecOwn.p;
  sp.ext = dSpecExt.p;
  sp.merged = dSpecMerged.p;
  sp.repState = dSpecRep.p;
  sp.repTarget = dSpecTarget.p;
  sp.start = dSpecStart.p;
  sp.failed = dSpecFailed.p;
  SeedArgs a = sa;
  a.maxSplit = maxP;
  a.chunk = 0;
  a.nChunks = 1;
  a.chainIn = nullptr;
  HIPCHK((hipError_t)launch_spec(a, sp, sstream));
  int failed = 0;
  std::vector<int> merged(G);
  HIPCHK(hipMemcpyAsync(&failed, dSpecFailed.p, sizeof(int), hipMemcpyDeviceToHost, sstream));
  HIPCHK(hipMemcpyAsync(merged.data(), dSpecMerged.p, G * sizeof(int), hipMemcpyDeviceToHost,
                        sstream));
  HIPCHK(hipStreamSynchronize(sstream));
  if (failed) {
    for (auto& d : hc) d.split = 1;
    HIPCHK(hipMemcpyAsync(dCurves.p, hc.data(), nc * sizeof(CurveDev), hipMemcpyHostToDevice,
                          sstream));
    stats.specFallback = 1;
    return false;
  }
  int repairs = 0;
  for (int k = 0; k < nc; k++)
    for (int q = 1; q < hc[k].split; q++) repairs += merged[base[k] + q] ? 0 : 1;
  stats.specRepairs = repairs;
  specPieces = G;
  sa.maxSplit = maxP;
  sa.chunk = 0;
  sa.nChunks = 1;
  sa.chainIn = nullptr;
  sa.pieceBase = dSpecBase.p;
  sa.pieceStart = dSpecStart.p;
  return true;
}

void Engine::run(Job& job) {
  init();
  auto tStart = std::chrono::steady_clock::now();
  stats = Stats();
  const int nc = (int)job.curves.size();
  std::vector<CurveDev> hc(nc);
  std::vector<int> pointBase(nc + 1, 0);
  for (int k = 0; k < nc; k++) {
    const int npts = job.runSeeds ? (int)job.curves[k]->lcLength + 1
                                  : (job.nPos.empty() ? (int)job.pos.size() : job.nPos[k]);
    pointBase[k + 1] = pointBase[k] + npts;
  }
  const int nPts = pointBase[nc];
  if (nPts <= 0) return;
  std::vector<int> ptCurve(nPts);
  for (int k = 0; k < nc; k++)
    for (int p = pointBase[k]; p < pointBase[k + 1]; p++) ptCurve[p] = k;
  std::vector<cx> caus((size_t)nc * 6 * kCccLen);
  for (int k = 0; k < nc; k++) {
    hc[k] = to_dev(*job.curves[k], job.starts[k], job.ends[k], pointBase[k], k * 6 * kCccLen);
    std::copy(job.curves[k]->ca.begin(), job.curves[k]->ca.end(), caus.begin() + (size_t)k * 6 * kCccLen);
  }
  // recorded tempRoots[10] values of the curves that carry them
  std::vector<cx> slots;
  for (int k = 0; k < nc && job.runSeeds; k++) {
    const std::vector<cx>& tr = job.curves[k]->slotTrace;
    if (tr.empty()) continue;
    hc[k].slotBase = (int)slots.size();
    hc[k].slotCount = (int)tr.size();
    slots.insert(slots.end(), tr.begin(), tr.end());
  }
  if (!slots.empty()) {
    dSlots.ensure(slots.size());
    HIPCHK(hipMemcpyAsync(dSlots.p, slots.data(), slots.size() * sizeof(cx), hipMemcpyHostToDevice,
                          stream));
  }
  dCurves.ensure(nc);
  dCaus.ensure(caus.size());
  dSeeds.ensure((size_t)nPts * kMaxSeeds);
  dNSeeds.ensure(nPts);
  dPtPos.ensure(nPts);
  dCost.ensure(nPts);
  dPtAmp.ensure(nPts);
  dPtCen.ensure(nPts);
  dPtCaus.ensure(nPts);
  dPtPair.ensure(nPts);
  dPtRoots.ensure((size_t)nPts * 10);
  bool anyRing = false;
  for (auto* c : job.curves) anyRing |= c->ring && job.runSeeds;
  if (anyRing) {
    dRing.ensure((size_t)nPts * kRingSamples);
    dRingCross.ensure(nPts);
    dPolyScratch.ensure((size_t)nPts * kRingSamples * 11);
  }
  dFlags.ensure(nPts);
  dPtCurve.ensure(nPts);
  HIPCHK(hipMemcpyAsync(dCurves.p, hc.data(), nc * sizeof(CurveDev), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(dCaus.p, caus.data(), caus.size() * sizeof(cx), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(dPtCurve.p, ptCurve.data(), nPts * sizeof(int), hipMemcpyHostToDevice, stream));

  std::vector<int> nSeeds(nPts);
  std::vector<float> cost(nPts);
  std::vector<uint32_t> flags(nPts, 0);
  // Seed search and fills overlap: the seed kernel walks the positions in
  // chunks on its own stream, and each chunk's fill runs on the engine stream
  // as soon as the chunk's seeds are there.
  unsigned maxLen = 0;
  for (int k = 0; k < nc; k++)
    maxLen = std::max<unsigned>(maxLen, (unsigned)(pointBase[k + 1] - pointBase[k]));
  // chunk ch holds positions [len ch / n, len (ch+1) / n) of every curve of
  // len positions, so that every chain does the same share of its work
  int nChunks = 1;
  std::vector<hipEvent_t>& evSeed = seedEvents;
  // the verified speculative chain: the single chain's seeds, computed in
  // pieces in parallel (not with a split search, and not where the chain
  // replays recorded slot values)
  bool spec = job.runSeeds && opt.specPiece > 0 && slots.empty() &&
              (int)maxLen >= 2 * opt.specPiece;
  for (auto* c : job.curves) spec = spec && c->split == 1;
  if (job.runSeeds) {
    nChunks = std::max(1, std::min<int>(opt.seedChunks, (int)(maxLen / std::max(1, opt.seedChunkMin))));
    // a split seed search runs in parallel and needs no overlap with the
    // fills; one chunk keeps piece q at positions [n q / P, n (q+1) / P)
    for (auto* c : job.curves)
      if (c->split > 1) nChunks = 1;
    if (spec) nChunks = 1;
    while ((int)evSeed.size() < nChunks + 1) {
      hipEvent_t ev;
      HIPCHK(hipEventCreate(&ev));
      evSeed.push_back(ev);
    }
    dChain.ensure(nc);
    // the seed stream starts after the set-up copies on the engine stream
    HIPCHK(hipEventRecord(ev0, stream));
    HIPCHK(hipStreamWaitEvent(sstream, ev0, 0));
    HIPCHK(hipEventRecord(evSeed[nChunks], sstream));
    // every point's solve list and polynomials, before the chains need them
    TrialArgs ta;
    ta.curves = dCurves.p;
    ta.ptCurve = dPtCurve.p;
    ta.nPts = nPts;
    ta.caustics = dCaus.p;
    ta.ptPos = dPtPos.p;
    dPtCausD2.ensure(nPts);
    ta.ptCausD2 = dPtCausD2.p;
    dPoolBase.ensure(nPts);
    dPoolCount.ensure(nPts);
    dPoolCursor.ensure(1);
    size_t cap = std::max<size_t>(dPoolZeta.n, 2 * (size_t)nPts + 1024);
    HIPCHK(hipEventRecord(evTrial0, sstream));
    unsigned long long used = 0;
    for (int attempt = 0;; attempt++) {
      if (cap > (size_t)INT32_MAX) throw std::runtime_error("solve pool exceeds 2^31 entries");
      dPoolZeta.ensure(cap);
      dPoolOwner.ensure(cap);
      dPoolCoef.ensure(cap * 11);
      ta.pool.base = dPoolBase.p;
      ta.pool.count = dPoolCount.p;
      ta.pool.zeta = dPoolZeta.p;
      ta.pool.owner = dPoolOwner.p;
      ta.pool.coef = dPoolCoef.p;
      ta.pool.cursor = dPoolCursor.p;
      ta.pool.cap = (int)cap;
      HIPCHK((hipError_t)launch_clear_u64(dPoolCursor.p, 1, 0, sstream));
      HIPCHK((hipError_t)launch_trials(ta, sstream));
      HIPCHK(hipMemcpyAsync(&used, dPoolCursor.p, sizeof(used), hipMemcpyDeviceToHost, sstream));
      HIPCHK(hipStreamSynchronize(sstream));
      if (used <= cap) break;
      if (attempt >= 2) throw std::runtime_error("solve pool sizing did not converge");
      cap = (size_t)used + 1024;
    }
    HIPCHK((hipError_t)launch_trial_polys(ta, (int)used, sstream));
    HIPCHK(hipEventRecord(evTrial1, sstream));
    stats.solves = (double)used;
    SeedArgs sa;
    sa.curves = dCurves.p;
    sa.nCurves = nc;
    sa.caustics = dCaus.p;
    sa.trig = dTrig.p;
    sa.pool = ta.pool;
    sa.ptCausD2 = dPtCausD2.p;
    sa.slotTrace = slots.empty() ? nullptr : dSlots.p;
    sa.seeds = dSeeds.p;
    sa.nSeeds = dNSeeds.p;
    sa.cost = dCost.p;
    sa.ptAmp = dPtAmp.p;
    sa.ptCen = dPtCen.p;
    sa.ptCaus = dPtCaus.p;
    sa.ptPair = dPtPair.p;
    sa.ptRoots = dPtRoots.p;
    sa.maxSplit = 1;
    for (auto* c : job.curves) sa.maxSplit = std::max(sa.maxSplit, c->split);
    sa.flags = dFlags.p;
    sa.chainOut = dChain.p;
    sa.pieceBase = nullptr;
    sa.pieceStart = nullptr;
    if (spec && run_spec(hc, pointBase, sa)) {
      HIPCHK((hipError_t)launch_seed(sa, sstream));
      HIPCHK(hipEventRecord(evSeed[0], sstream));
      // the pieces' seed launch restarted its solve count at each piece:
      // the chain's count after the run is the start count plus every
      // solve of the curve
      std::vector<int> cnt(nPts);
      HIPCHK(hipMemcpyAsync(cnt.data(), dPoolCount.p, nPts * sizeof(int), hipMemcpyDeviceToHost,
                            sstream));
      HIPCHK(hipStreamSynchronize(sstream));
      specCalls.assign(nc, 0);
      for (int k = 0; k < nc; k++) {
        long long t = job.curves[k]->chain.calls;
        for (int p = pointBase[k]; p < pointBase[k + 1]; p++) t += cnt[p];
        specCalls[k] = (int)t;
      }
      stats.specPieces = specPieces;
    } else for (int ch = 0; ch < nChunks; ch++) {
      sa.chunk = ch;
      sa.nChunks = nChunks;
      sa.chainIn = ch == 0 ? nullptr : dChain.p;
      HIPCHK((hipError_t)launch_seed(sa, sstream));
      HIPCHK(hipEventRecord(evSeed[ch], sstream));
    }
    // disk samples of the curves that asked for them, behind the seed search
    // (and beside the fills)
    if (anyRing) {
      HIPCHK(hipEventRecord(evRing0, sstream));
      HIPCHK((hipError_t)launch_ring_amp(dCurves.p, dPtCurve.p, dPtPos.p, dPtRoots.p, nPts,
                                         dTrig.p, dRing.p, dRingCross.p, dPolyScratch.p, sstream));
      HIPCHK(hipEventRecord(evRing1, sstream));
    }
  } else {
    for (int p = 0; p < nPts; p++) {
      nSeeds[p] = std::min(job.nSeeds[p], kMaxSeeds);
      cost[p] = job.cost.empty() ? (float)nSeeds[p] : job.cost[p];
    }
    HIPCHK(hipMemcpy(dPtPos.p, job.pos.data(), nPts * sizeof(double2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dSeeds.p, job.seeds.data(), (size_t)nPts * kMaxSeeds * sizeof(int2),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dFlags.p, flags.data(), nPts * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dCost.p, cost.data(), nPts * sizeof(float), hipMemcpyHostToDevice));
  }

  std::vector<uint64_t> ptRays(nPts, 0);
  std::vector<double> ptSumAll(nPts, 0.0);  // brightness sums in fixed-point units
  // amoeba print-out: the runs of every point of the curves that asked for it
  bool wantAmoebae = false;
  for (auto* c : job.curves) wantAmoebae |= c->amoebaOut && job.runSeeds;
  std::vector<std::vector<Segment>> amoebae(wantAmoebae ? nPts : 0);
  const unsigned long long qcap = 1ull << opt.queueLog2;
  dQBuf.ensure(qcap);
  dQSeq.ensure(qcap);
  // the fill's shared counters, each on a 128-byte line of its own: waves
  // that poll one of them do not slow the atomics on the others
  constexpr int kLine = 16;  // u64 per 128 bytes
  dSync.ensure(7 * kLine);
  // one fill workgroup per compute unit (it fills the unit's registers),
  // capped at the occupancy limit; while the seed search runs beside the
  // fills, one unit is left to its wavefront (a group that cannot start
  // would only start, find nothing and leave once the seed search is done,
  // but the launch would last until then)
  const int resident = fill_resident_workgroups();
  if (resident <= 0) throw std::runtime_error("fill kernel occupancy query failed");
  int nWg = opt.fillWorkgroups > 0 ? opt.fillWorkgroups : cus;
  if (nWg > resident) nWg = resident;
  if (job.runSeeds && nChunks > 1 && opt.fillWorkgroups <= 0 && nWg > 1) nWg--;

  std::vector<int> order;
  for (int ch = 0; ch < nChunks; ch++) {
  if (job.runSeeds) {
    // this chunk's seed counts and predicted costs (the copy waits for the
    // chunk's seed launch; later chunks are still being written)
    HIPCHK(hipStreamWaitEvent(stream, evSeed[ch], 0));
    HIPCHK(hipMemcpyAsync(nSeeds.data(), dNSeeds.p, nPts * sizeof(int), hipMemcpyDeviceToHost,
                          stream));
    HIPCHK(hipMemcpyAsync(cost.data(), dCost.p, nPts * sizeof(float), hipMemcpyDeviceToHost,
                          stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  // the chunk's positions, largest predicted cost first, in launches of
  // about kRunBudget predicted runs (groups of kGroupPoints positions with a
  // run-table slice each)
  order.clear();
  for (int k = 0; k < nc; k++) {
    const long long len = pointBase[k + 1] - pointBase[k];
    const int i0 = (int)(len * ch / nChunks), i1 = (int)(len * (ch + 1) / nChunks);
    for (int i = i0; i < i1 && job.runFill; i++)
      if (nSeeds[pointBase[k] + i] > 0) order.push_back(pointBase[k] + i);
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  // A launch holds about kRunBudget predicted runs (the run table is sized
  // from the prediction, tableLoad entries per run: up to 16 GiB).
  constexpr double kRunBudget = 256.0 * 1024 * 1024;
  size_t stepPts = 0;
  for (size_t b0 = 0; b0 < order.size(); b0 += stepPts) {
    double predicted = 0;
    size_t b1 = b0;
    std::vector<double> grpPred;
    while (b1 < order.size()) {
      const double pc = std::max(256.0, (double)cost[order[b1]] / opt.runPixels);
      if (b1 > b0 && predicted + pc > kRunBudget) break;
      predicted += pc;
      if ((b1 - b0) % kGroupPoints == 0) grpPred.push_back(0.0);
      grpPred.back() += pc;
      b1++;
    }
    stepPts = b1 - b0;
    const int nb = (int)stepPts;
    const int nGrp = (int)grpPred.size();
    std::vector<int> batchPts(order.begin() + b0, order.begin() + b1);
    std::vector<int2> seedTasks;
    for (int b = 0; b < nb; b++) {
      const int gp = batchPts[b];
      for (int k = 0; k < nSeeds[gp]; k++) seedTasks.push_back(make_int2(b, k));
    }
    dBatchPts.ensure(nb);
    dSeedTasks.ensure(seedTasks.size());
    dPtCount.ensure(nb);
    dPtSum.ensure(2 * (size_t)nb);
    HIPCHK(hipMemcpyAsync(dBatchPts.p, batchPts.data(), nb * sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(dSeedTasks.p, seedTasks.data(), seedTasks.size() * sizeof(int2),
                          hipMemcpyHostToDevice, stream));
    // run-table slice of each group: a power of two of at least tableLoad
    // entries per predicted run
    std::vector<unsigned long long> grpCap(nGrp), grpBase(nGrp), grpMask(nGrp);
    for (int g = 0; g < nGrp; g++) {
      unsigned long long c = 1ull << 12;
      while ((double)c < opt.tableLoad * grpPred[g] && c < (1ull << 30)) c <<= 1;
      grpCap[g] = c;
    }
    // labels: one per seed class and one per scout run; heavy positions
    // send many scouts, so the capacity follows the predicted runs as well
    int labelCap = nb + std::max(std::max(1 << 16, 16 * nb),
                                 (int)std::min(predicted / 8.0, (double)(1 << 26)));
    for (int attempt = 0; attempt < 4; attempt++) {
      unsigned long long hcap = 0;
      for (int g = 0; g < nGrp; g++) {
        grpBase[g] = hcap;
        grpMask[g] = grpCap[g] - 1;
        hcap += grpCap[g];
      }
      dHtab.ensure(2 * hcap);
      dGrp.ensure(2 * (size_t)nGrp);
      HIPCHK(hipMemcpyAsync(dGrp.p, grpBase.data(), nGrp * sizeof(unsigned long long),
                            hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(dGrp.p + nGrp, grpMask.data(), nGrp * sizeof(unsigned long long),
                            hipMemcpyHostToDevice, stream));
      dLParent.ensure(labelCap);
      dLPoint.ensure(labelCap);
      dLCount.ensure(labelCap);
      dLSum.ensure(2 * (size_t)labelCap);
      FillArgs fa;
      std::memset(&fa, 0, sizeof(fa));
      fa.curves = dCurves.p;
      fa.ptCurve = dPtCurve.p;
      fa.ptPos = dPtPos.p;
      fa.batchPts = dBatchPts.p;
      fa.nBatch = nb;
      fa.seedTasks = dSeedTasks.p;
      fa.nSeedTasks = (int)seedTasks.size();
      fa.seeds = dSeeds.p;
      fa.htab = dHtab.p;
      fa.htabEntries = hcap;
      fa.grpBase = dGrp.p;
      fa.grpMask = dGrp.p + nGrp;
      fa.probeLimit = opt.probeLimit;
      fa.lparent = dLParent.p;
      fa.lpoint = dLPoint.p;
      fa.lcount = dLCount.p;
      fa.lsum = dLSum.p;
      fa.labelNext = reinterpret_cast<int*>(dSync.p + 4 * kLine);
      fa.labelCap = labelCap;
      fa.qbuf = dQBuf.p;
      fa.qseq = dQSeq.p;
      fa.qhead = dSync.p + 0 * kLine;
      fa.qtail = dSync.p + 1 * kLine;
      fa.qcap = qcap;
      fa.active = reinterpret_cast<int*>(dSync.p + 2 * kLine);
      fa.nGroups = nWg;
      fa.seedNext = reinterpret_cast<int*>(dSync.p + 3 * kLine);
      fa.flags = dFlags.p;
      fa.abortFlag = reinterpret_cast<int*>(dSync.p + 5 * kLine);
      fa.stats = dSync.p + 6 * kLine;
      fa.timeLimitTicks = (long long)(opt.timeLimitS * 1.0e8);
      fa.scoutRows = opt.scoutRows;
      fa.scoutJump = opt.scoutJump;
      fa.scoutWidth = opt.scoutWidth;
      fa.scoutRetry = opt.scoutRetry;
      fa.scoutShare = opt.scoutShare;
      fa.ptCost = dCost.p;
      fa.shareRays = (float)opt.shareRays;
      fa.leapEvery = opt.leapEvery;
      fa.leapJumps = opt.leapJumps;
      fa.pollBackoff = opt.pollBackoff;
      fa.poolNewest = opt.poolNewest;
      if (opt.useGraph) {
        // the launch's reset, clears, fill and resolve as one graph (kernel
        // nodes only: the resets are kernels, not memsets), captured once
        // and launched again while its arguments are the same (the next
        // step of a batch of the same sizes); the events around its launch
        // time all five
        FillGraphKey key;
        std::memset(&key, 0, sizeof(key));
        key.fa = fa;
        key.nWg = nWg;
        key.labelCap = labelCap;
        key.nb = nb;
        key.ptCount = dPtCount.p;
        key.ptSum = dPtSum.p;
        if (!fillExec || std::memcmp(&key, &fillKey, sizeof(key)) != 0) {
          drop_fill_graph();
          hipGraph_t graph = nullptr;
          HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
          int rc = launch_fill_reset(fa, stream);
          if (!rc) rc = launch_clear_u64(dPtCount.p, nb, 0, stream);
          if (!rc) rc = launch_clear_u64(dPtSum.p, 2 * (unsigned long long)nb, 0, stream);
          if (!rc) rc = launch_fill(fa, nWg, stream);
          if (!rc) rc = launch_resolve(fa, labelCap, dPtCount.p, dPtSum.p, stream);
          const hipError_t ec = hipStreamEndCapture(stream, &graph);
          if ((rc != 0 || ec != hipSuccess) && graph) (void)hipGraphDestroy(graph);
          HIPCHK((hipError_t)rc);
          HIPCHK(ec);
          const hipError_t ei = hipGraphInstantiate(&fillExec, graph, nullptr, nullptr, 0);
          (void)hipGraphDestroy(graph);
          if (ei != hipSuccess) fillExec = nullptr;
          HIPCHK(ei);
          fillKey = key;
          stats.graphBuilds += 1;
        }
        HIPCHK(hipEventRecord(ev0, stream));
        const hipError_t el = hipGraphLaunch(fillExec, stream);
        if (el == hipSuccess) HIPCHK(hipEventRecord(ev1, stream));
        const hipError_t es = el == hipSuccess ? hipStreamSynchronize(stream) : el;
        if (es != hipSuccess) drop_fill_graph();
        HIPCHK(es);
      } else {
        HIPCHK((hipError_t)launch_fill_reset(fa, stream));
        HIPCHK((hipError_t)launch_clear_u64(dPtCount.p, nb, 0, stream));
        HIPCHK((hipError_t)launch_clear_u64(dPtSum.p, 2 * (unsigned long long)nb, 0, stream));
        HIPCHK(hipEventRecord(ev0, stream));
        HIPCHK((hipError_t)launch_fill(fa, nWg, stream));
        HIPCHK(hipEventRecord(ev1, stream));
        HIPCHK((hipError_t)launch_resolve(fa, labelCap, dPtCount.p, dPtSum.p, stream));
        HIPCHK(hipStreamSynchronize(stream));
      }
      stats.fillMs += elapsed();
      stats.fillLaunches += 1;
      std::vector<unsigned long long> sync(7 * kLine);
      HIPCHK(hipMemcpy(sync.data(), dSync.p, sync.size() * sizeof(unsigned long long),
                       hipMemcpyDeviceToHost));
      const unsigned long long* st = sync.data() + 6 * kLine - 2;  // st[2..11] = stats[0..9]
      const bool aborted = *reinterpret_cast<const int*>(sync.data() + 5 * kLine) != 0;
      std::vector<uint32_t> bflags(nPts);
      HIPCHK(hipMemcpy(bflags.data(), dFlags.p, nPts * sizeof(uint32_t), hipMemcpyDeviceToHost));
      bool hashFull = false, labelFull = false;
      std::vector<char> grpFull(nGrp, 0);
      for (int b = 0; b < nb; b++) {
        const bool hf = (bflags[batchPts[b]] & kFlagHashFull) != 0;
        hashFull |= hf;
        if (hf) grpFull[b / kGroupPoints] = 1;
        labelFull |= (bflags[batchPts[b]] & kFlagLabelFull) != 0;
      }
      stats.tasks += (double)st[2];
      stats.evals += (double)st[3];
      stats.ins += (double)st[4];
      stats.dup += (double)st[5];
      stats.busyMs += (double)st[6] * 1e-5;
      stats.insertMs += (double)st[7] * 1e-5;
      stats.idlePolls += (double)st[8];
      stats.waves = (double)st[9];
      stats.scouts += (double)st[10];
      stats.scoutRuns += (double)st[11];
      if (aborted) {
        for (int b = 0; b < nb; b++) flags[batchPts[b]] |= kFlagTimeout;
        throw std::runtime_error("fill kernel hit its time limit");
      }
      if ((hashFull || labelFull) && attempt < 3) {
        // clear the overflow bits of the batch's points and run it again
        // with bigger tables (other points' flags may be in use by the seed
        // search)
        HIPCHK((hipError_t)launch_clear_flag_bits(dFlags.p, dBatchPts.p, nb,
                                                  kFlagHashFull | kFlagLabelFull, stream));
        for (int g = 0; g < nGrp; g++)
          if (grpFull[g]) grpCap[g] = std::min<unsigned long long>(grpCap[g] << 2, 1ull << 31);
        if (labelFull) labelCap *= 4;
        continue;
      }
      std::vector<unsigned long long> cnt(nb);
      std::vector<unsigned long long> sum(2 * (size_t)nb);
      HIPCHK(hipMemcpy(cnt.data(), dPtCount.p, nb * sizeof(unsigned long long), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(sum.data(), dPtSum.p, sum.size() * sizeof(unsigned long long),
                       hipMemcpyDeviceToHost));
      for (int b = 0; b < nb; b++) {
        ptRays[batchPts[b]] = cnt[b];
        // the split total: signed high word * 2^32 + unsigned low word
        ptSumAll[batchPts[b]] = (double)(long long)sum[2 * (size_t)b + 1] * 4294967296.0 +
                                (double)sum[2 * (size_t)b];
      }
      if (wantAmoebae) collect_amoebae(job, ptCurve, batchPts, grpBase, grpCap, labelCap, amoebae);
      break;
    }
  }
  }  // chunks
  if (job.runSeeds) {
    // the seed search's outputs are final once its last launch is done; the
    // disk samples behind it on the seed stream are waited for after the
    // other results are copied back
    HIPCHK(hipEventSynchronize(evSeed[nChunks - 1]));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, evTrial1, evSeed[nChunks - 1]));
    stats.seedMs = ms;
    HIPCHK(hipEventElapsedTime(&ms, evTrial0, evTrial1));
    stats.trialMs = ms;
    std::vector<RootChain> chains(nc);
    HIPCHK(hipMemcpy(chains.data(), dChain.p, nc * sizeof(RootChain), hipMemcpyDeviceToHost));
    for (int k = 0; k < nc; k++) {
      if (!specCalls.empty()) chains[k].calls = specCalls[k];
      job.curves[k]->chain = chains[k];
    }
    specCalls.clear();
    HIPCHK(hipMemcpy(nSeeds.data(), dNSeeds.p, nPts * sizeof(int), hipMemcpyDeviceToHost));
  }
  HIPCHK(hipMemcpy(flags.data(), dFlags.p, nPts * sizeof(uint32_t), hipMemcpyDeviceToHost));
  uint32_t allFlags = 0;
  for (uint32_t f : flags) allFlags |= f;

  // results back into the objects
  std::vector<int2>& seedsHost = hSeeds;  // kept: a fresh 8 MB buffer per run costs page faults
  std::vector<double> ampHost, causHost, pairHost, ringHost;
  std::vector<cx> rootsHost;
  std::vector<int> crossHost;
  std::vector<double2> cenHost;
  int stride = 1;  // seed slots kept per position: the largest seed count
  if (job.runSeeds) {
    for (int p = 0; p < nPts; p++) stride = std::max(stride, std::min(nSeeds[p], kMaxSeeds));
    seedsHost.resize((size_t)nPts * stride);
    copy_out_rows(seedsHost.data(), dSeeds.p, nPts, kMaxSeeds * sizeof(int2),
                  stride * sizeof(int2));
    ampHost.resize(nPts);
    cenHost.resize(nPts);
    copy_out(ampHost.data(), dPtAmp.p, nPts * sizeof(double));
    copy_out(cenHost.data(), dPtCen.p, nPts * sizeof(double2));
    causHost.resize(nPts);
    copy_out(causHost.data(), dPtCaus.p, nPts * sizeof(double));
    pairHost.resize(nPts);
    copy_out(pairHost.data(), dPtPair.p, nPts * sizeof(double));
  }
  for (int k = 0; k < nc; k++) {
    Curve& c = *job.curves[k];
    const int n = pointBase[k + 1] - pointBase[k];
    c.full.assign(n, 0.0);
    c.rays.assign(n, 0);
    c.flags.assign(n, 0);
    for (int i = 0; i < n; i++) {
      const int p = pointBase[k] + i;
      c.rays[i] = ptRays[p];
      c.full[i] = (ptSumAll[p] / hc[k].sumScale) * c.ampScale;
      c.flags[i] = flags[p];
    }
    if (job.runSeeds) {
      c.nSeeds.assign(nSeeds.begin() + pointBase[k], nSeeds.begin() + pointBase[k + 1]);
      const int* sp = reinterpret_cast<const int*>(seedsHost.data() + (size_t)pointBase[k] * stride);
      c.seedPix.assign(sp, sp + (size_t)n * stride * 2);
      c.seedStride = stride;
      if (job.runFill) {
        for (unsigned i = 0; i < c.lcLength && i < (unsigned)n; i++) c.lcVec[i] = c.full[i];
        c.hasIRS = true;
      }
      c.hasSeeds = true;
      c.seedStart = job.starts[k];
      c.seedEnd = job.ends[k];
      c.causDist.assign(causHost.begin() + pointBase[k], causHost.begin() + pointBase[k + 1]);
      c.pairShare.assign(pairHost.begin() + pointBase[k], pairHost.begin() + pointBase[k + 1]);
      if (!c.ring) {
        c.ringCross.clear();
        c.ringAmp.clear();
        c.centreRoots.clear();
      }
      c.pointAmp.assign(ampHost.begin() + pointBase[k], ampHost.begin() + pointBase[k + 1]);
      c.centroid.resize(2 * (size_t)n);
      for (int i = 0; i < n; i++) {
        c.centroid[2 * i] = cenHost[pointBase[k] + i].x;
        c.centroid[2 * i + 1] = cenHost[pointBase[k] + i].y;
      }
      if (wantAmoebae && c.amoebaOut)
        for (int i = 0; i < n; i++) write_amoeba_file(c, i, amoebae[pointBase[k] + i]);
    }
  }
  // the seed stream's last work: the disk samples and centre roots of the
  // curves that asked for them
  HIPCHK(hipStreamSynchronize(sstream));
  if (anyRing) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, evRing0, evRing1));
    stats.ringMs = ms;
    ringHost.resize((size_t)nPts * kRingSamples);
    crossHost.resize(nPts);
    rootsHost.resize((size_t)nPts * 10);
    for (int k = 0; k < nc; k++) {
      Curve& c = *job.curves[k];
      if (!c.ring) continue;
      const size_t p0 = pointBase[k], n = pointBase[k + 1] - pointBase[k];
      copy_out(ringHost.data() + p0 * kRingSamples, dRing.p + p0 * kRingSamples,
               n * kRingSamples * sizeof(double));
      copy_out(crossHost.data() + p0, dRingCross.p + p0, n * sizeof(int));
      copy_out(rootsHost.data() + p0 * 10, dPtRoots.p + p0 * 10, n * 10 * sizeof(cx));
      c.ringAmp.assign(ringHost.begin() + p0 * kRingSamples,
                       ringHost.begin() + (p0 + n) * kRingSamples);
      c.centreRoots.assign(rootsHost.begin() + p0 * 10, rootsHost.begin() + (p0 + n) * 10);
      c.ringCross.assign(crossHost.begin() + p0, crossHost.begin() + p0 + n);
    }
  }
  stats.totalMs =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tStart).count();
  // results are in place (with their flags); a lost task makes them wrong
  if (allFlags & kFlagQueueFull)
    throw std::runtime_error("fill task queue overflow: some light-curve points are incomplete");
  if (allFlags & (kFlagHashFull | kFlagLabelFull))
    throw std::runtime_error("fill tables still full after enlarging: results are incomplete");
}

void Engine::point_lc(Curve& c, cx s, cx e) {
  init();
  if (c.lcLength < 1) return;
  CurveDev d = to_dev(c, s, e, 0, 0);
  dCurves.ensure(1);
  dChain.ensure(1);
  dPointLC.ensure(c.lcLength);
  dPolyScratch.ensure((size_t)c.lcLength * 11);
  HIPCHK(hipMemcpyAsync(dCurves.p, &d, sizeof(CurveDev), hipMemcpyHostToDevice, stream));
  HIPCHK((hipError_t)launch_point_lc(dCurves.p, 1, (int)c.lcLength, dTrig.p, dPointLC.p, dChain.p,
                                     dPolyScratch.p, stream));
  HIPCHK(hipStreamSynchronize(stream));
  HIPCHK(hipMemcpy(c.lcVec.data(), dPointLC.p, c.lcLength * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&c.chain, dChain.p, sizeof(RootChain), hipMemcpyDeviceToHost));
}

template <class F>
int guarded(F&& f) {
  try {
    g_err.clear();
    return f();
  } catch (const std::exception& e) {
    set_err(e.what());
    return -1;
  } catch (...) {
    set_err("unknown error");
    return -1;
  }
}

void write_par_file(const Curve& c) {
  // printOutLCParameters (lcirs.cc:479-512), including its "m3" entry that
  // prints m2
  std::ofstream f(c.parFile);
  const char* in = "\t";
  f << "{" << '\n';
  f << in << "\"sourceRadius\": " << c.rho << ",\n";
  f << in << "\"pointsPerRadius\": " << c.ppr << ",\n";
  f << in << "\"imgPlaneSize\": " << c.N << ",\n";
  f << in << "\"imgPlaneSizeDouble\": " << c.Ldbl << ",\n";
  f << in << "\"bottomLeftCornerImgX\": " << c.blx << ",\n";
  f << in << "\"bottomLeftCornerImgY\": " << c.bly << ",\n";
  f << in << "\"topRightCornerImgX\": " << c.trx << ",\n";
  f << in << "\"topRightCornerImgY\": " << c.try_ << ",\n";
  f << in << "\"a\": " << c.a << ",\n";
  f << in << "\"b\": " << c.b << ",\n";
  f << in << "\"th\": " << c.th << ",\n";
  f << in << "\"m1\": " << c.L.m1 << ",\n";
  f << in << "\"m2\": " << c.m2 << ",\n";
  f << in << "\"m3\": " << c.m2 << ",\n";
  f << in << "\"z1x\": " << c.L.z1.re << ",\n";
  f << in << "\"z1y\": " << c.L.z1.im << ",\n";
  f << in << "\"z2x\": " << c.L.z2.re << ",\n";
  f << in << "\"z2y\": " << c.L.z2.im << ",\n";
  f << in << "\"z3x\": " << c.L.z3.re << ",\n";
  f << in << "\"z3y\": " << c.L.z3.im << "\n";
  f << "}";
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

void* lcirs_new(double a, double b, double th, double m2, double m3, double sourceSize,
                unsigned lcLength, long pointsPerRadius) {
  Curve* c = nullptr;
  guarded([&] {
    c = make_curve(a, b, th, m2, m3, sourceSize, lcLength, pointsPerRadius);
    return 0;
  });
  return c;
}

void* lcirs_new_ccc(double a, double b, double th, double m2, double m3, double sourceSize,
                    unsigned lcLength, long pointsPerRadius, const double* cc, const double* ca) {
  Curve* c = nullptr;
  guarded([&] {
    if (!cc || !ca) throw std::invalid_argument("null critical-curve arrays");
    c = make_curve(a, b, th, m2, m3, sourceSize, lcLength, pointsPerRadius,
                   reinterpret_cast<const cx*>(cc), reinterpret_cast<const cx*>(ca));
    return 0;
  });
  return c;
}

void lcirs_delete(void* lc) { delete static_cast<Curve*>(lc); }

void* lcirs_clone(void* lc) {
  Curve* out = nullptr;
  guarded([&] {
    if (!lc) throw std::invalid_argument("null light-curve object");
    std::unique_ptr<Curve> c(new Curve(*static_cast<Curve*>(lc)));
    // the set-up, options and root chain, without the last results
    c->lcVec.clear(); c->full.clear(); c->rays.clear(); c->flags.clear();
    c->nSeeds.clear(); c->seedPix.clear(); c->pointAmp.clear(); c->centroid.clear();
    c->causDist.clear(); c->pairShare.clear(); c->ringAmp.clear(); c->ringCross.clear();
    c->centreRoots.clear();
    c->hasSeeds = c->hasIRS = false;
    c->amoebaOut = false;
    out = c.release();
    return 0;
  });
  return out;
}

void get_lc_irs(void* lc, double iniX, double iniY, double finX, double finY) {
  int rc = guarded([&] {
    if (!lc) throw std::invalid_argument("null light-curve object");
    Curve* c = static_cast<Curve*>(lc);
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    Job job;
    job.curves = {c};
    job.starts = {mk(iniX, iniY)};
    job.ends = {mk(finX, finY)};
    E.run(job);
    if (c->amoebaOut) write_par_file(*c);
    return 0;
  });
  if (rc) std::fprintf(stderr, "get_lc_irs failed: %s\n", g_err.c_str());
}

void get_lc(void* lc, double iniX, double iniY, double finX, double finY) {
  int rc = guarded([&] {
    if (!lc) throw std::invalid_argument("null light-curve object");
    Curve* c = static_cast<Curve*>(lc);
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    E.point_lc(*c, mk(iniX, iniY), mk(finX, finY));
    return 0;
  });
  if (rc) std::fprintf(stderr, "get_lc failed: %s\n", g_err.c_str());
}

void copy_lc(void* lc, double* out) {
  g_err.clear();
  if (!lc || !out) return;
  const Curve* c = static_cast<Curve*>(lc);
  std::copy(c->lcVec.begin(), c->lcVec.end(), out);
}

void set_limb_darkening(void* lc, const char* model, double v) {
  guarded([&] {
    if (!lc) throw std::invalid_argument("null light-curve object");
    Curve* c = static_cast<Curve*>(lc);
    if (model && std::strcmp(model, "linear") == 0) {
      const double i0 = 1.0 / ((1.0 - v / 3.0));
      c->ldA = i0 * (1.0 - v);
      c->ldB = i0 * v;
      c->intensity = 1.0;
    } else {
      ld_default(*c);
    }
    c->ampScale = c->intensity / M_PI / std::pow((double)c->ppr, 2.0);
    return 0;
  });
}

void set_amoeba_printout(void* lc, const char* amoebaPrefix, const char* parFile) {
  guarded([&] {
    if (!lc) throw std::invalid_argument("null light-curve object");
    Curve* c = static_cast<Curve*>(lc);
    c->amoebaOut = true;
    c->amoebaPrefix = amoebaPrefix ? amoebaPrefix : "";
    c->parFile = parFile ? parFile : "";
    return 0;
  });
}

// LightCurveCUDA's C ABI (src/Cuda/cudalc.cc:486-542), on this engine: the
// same objects, and get_lc_cuda computes the exact FP64 extended-source curve
// that get_lc_irs computes.
void* lccuda_new(double a, double b, double th, double m2, double m3, double sourceSize,
                 unsigned lcLength, long pointsPerRadius) {
  return lcirs_new(a, b, th, m2, m3, sourceSize, lcLength, pointsPerRadius);
}

void get_lc_cuda(void* lc, double iniX, double iniY, double finX, double finY) {
  get_lc_irs(lc, iniX, iniY, finX, finY);
}

void set_limb_darkening_cuda(void* lc, const char* model, double v) {
  set_limb_darkening(lc, model, v);
}

int lcirs_copy_point_amp(void* lc, double* amp, double* centroid, double* causticDist,
                         double* pairShare) {
  return guarded([&] {
    if (!lc) throw std::invalid_argument("null light-curve object");
    Curve* c = static_cast<Curve*>(lc);
    if (!c->hasSeeds) throw std::runtime_error("no seed search has run on this object");
    if (amp) std::copy(c->pointAmp.begin(), c->pointAmp.end(), amp);
    if (centroid) std::copy(c->centroid.begin(), c->centroid.end(), centroid);
    if (causticDist) std::copy(c->causDist.begin(), c->causDist.end(), causticDist);
    if (pairShare) std::copy(c->pairShare.begin(), c->pairShare.end(), pairShare);
    return 0;
  });
}

int lcirs_set_ring(void* lc, int on) {
  return guarded([&] {
    if (!lc) throw std::invalid_argument("null light-curve object");
    static_cast<Curve*>(lc)->ring = on != 0;
    return 0;
  });
}

int lcirs_set_seed_split(void* lc, int parts) {
  return guarded([&] {
    if (!lc) throw std::invalid_argument("null light-curve object");
    if (parts < 1 || parts > 65536) throw std::invalid_argument("parts must be in [1, 65536]");
    static_cast<Curve*>(lc)->split = parts;
    return 0;
  });
}

int lcirs_copy_ring_amp(void* lc, double* ring, double* centreRoots, int* crossings) {
  return guarded([&] {
    if (!lc) throw std::invalid_argument("null light-curve object");
    Curve* c = static_cast<Curve*>(lc);
    if (!c->hasSeeds || c->ringAmp.empty())
      throw std::runtime_error("no seed search with disk samples has run on this object");
    if (ring) std::copy(c->ringAmp.begin(), c->ringAmp.end(), ring);
    if (crossings) std::copy(c->ringCross.begin(), c->ringCross.end(), crossings);
    if (centreRoots)
      for (size_t i = 0; i < c->centreRoots.size(); i++) {
        centreRoots[2 * i] = c->centreRoots[i].re;
        centreRoots[2 * i + 1] = c->centreRoots[i].im;
      }
    return 0;
  });
}

int lcirs_find_seeds(void* lc, double iniX, double iniY, double finX, double finY) {
  return guarded([&] {
    if (!lc) throw std::invalid_argument("null light-curve object");
    Curve* c = static_cast<Curve*>(lc);
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    Job job;
    job.curves = {c};
    job.starts = {mk(iniX, iniY)};
    job.ends = {mk(finX, finY)};
    job.runFill = false;
    E.run(job);
    return 0;
  });
}

int lcirs_batch_find_seeds(int n, void** lcs, const double* traj) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!lcs || !traj))) throw std::invalid_argument("bad batch arguments");
    if (n == 0) return 0;
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    Job job;
    for (int k = 0; k < n; k++) {
      if (!lcs[k]) throw std::invalid_argument("null light-curve object in batch");
      job.curves.push_back(static_cast<Curve*>(lcs[k]));
      job.starts.push_back(mk(traj[4 * k], traj[4 * k + 1]));
      job.ends.push_back(mk(traj[4 * k + 2], traj[4 * k + 3]));
    }
    job.runFill = false;
    E.run(job);
    return 0;
  });
}

int lcirs_copy_lc_full(void* lc, double* out) {
  g_err.clear();
  const Curve* c = static_cast<Curve*>(lc);
  if (!c || !c->hasIRS) return -1;
  std::copy(c->full.begin(), c->full.end(), out);
  return 0;
}

int lcirs_copy_ray_counts(void* lc, uint64_t* out) {
  g_err.clear();
  const Curve* c = static_cast<Curve*>(lc);
  if (!c || !c->hasIRS) return -1;
  std::copy(c->rays.begin(), c->rays.end(), out);
  return 0;
}

int lcirs_copy_flags(void* lc, uint32_t* out) {
  g_err.clear();
  const Curve* c = static_cast<Curve*>(lc);
  if (!c || !c->hasSeeds) return -1;
  std::copy(c->flags.begin(), c->flags.end(), out);
  return 0;
}

int lcirs_copy_seeds(void* lc, int* nSeeds, int* pix) {
  g_err.clear();
  const Curve* c = static_cast<Curve*>(lc);
  if (!c || !c->hasSeeds) return -1;
  std::copy(c->nSeeds.begin(), c->nSeeds.end(), nSeeds);
  for (int p = 0; p < (int)c->nSeeds.size(); p++) c->seed_row(p, pix + (size_t)p * kMaxSeeds * 2);
  return 0;
}

int lcirs_copy_seeds_at(void* lc, int n, const int* idx, int* nSeeds, int* pix) {
  return guarded([&] {
    const Curve* c = static_cast<Curve*>(lc);
    if (!c) throw std::invalid_argument("null light-curve object");
    if (!c->hasSeeds) throw std::runtime_error("no seed search has run on this object");
    if (n < 0 || (n > 0 && (!idx || !nSeeds || !pix))) throw std::invalid_argument("bad arguments");
    const int np = (int)c->nSeeds.size();
    for (int i = 0; i < n; i++)
      if (idx[i] < 0 || idx[i] >= np) throw std::invalid_argument("position index out of range");
    const size_t row = (size_t)kMaxSeeds * 2;
    for (int i = 0; i < n; i++) {
      nSeeds[i] = c->nSeeds[idx[i]];
      c->seed_row(idx[i], pix + i * row);
    }
    return 0;
  });
}

int lcirs_copy_seeds_union(void* lc, int nOther, void** others, int n, const int* idx,
                           int* nSeeds, int* pix) {
  return guarded([&] {
    const Curve* c = static_cast<Curve*>(lc);
    if (!c) throw std::invalid_argument("null light-curve object");
    if (nOther < 0 || (nOther > 0 && !others) || n < 0 || (n > 0 && (!idx || !nSeeds || !pix)))
      throw std::invalid_argument("bad arguments");
    std::vector<const Curve*> src{c};
    for (int k = 0; k < nOther; k++) {
      if (!others[k]) throw std::invalid_argument("null light-curve object");
      src.push_back(static_cast<const Curve*>(others[k]));
    }
    const size_t np = c->nSeeds.size();
    for (const Curve* s : src) {
      if (!s->hasSeeds) throw std::runtime_error("no seed search has run on this object");
      if (s->nSeeds.size() != np) throw std::invalid_argument("objects of different lengths");
    }
    for (int i = 0; i < n; i++)
      if (idx[i] < 0 || (size_t)idx[i] >= np) throw std::invalid_argument("position index out of range");
    const size_t row = (size_t)kMaxSeeds * 2;
    for (int i = 0; i < n; i++) {
      const int p = idx[i];
      int* out = pix + i * row;
      int cnt = 0;
      for (const Curve* s : src) {
        const int* sp = s->seedPix.data() + (size_t)p * s->seedStride * 2;
        const int m = std::min(s->nSeeds[p], s->seedStride);
        for (int k = 0; k < m && cnt < kMaxSeeds; k++) {
          const int x = sp[2 * k], y = sp[2 * k + 1];
          bool seen = false;
          for (int j = 0; j < cnt && !seen; j++) seen = out[2 * j] == x && out[2 * j + 1] == y;
          if (seen) continue;
          out[2 * cnt] = x;
          out[2 * cnt + 1] = y;
          cnt++;
        }
      }
      std::memset(out + 2 * cnt, 0, (row - 2 * (size_t)cnt) * sizeof(int));
      nSeeds[i] = cnt;
    }
    return 0;
  });
}

int lcirs_get_grid(void* lc, double* out) {
  g_err.clear();
  const Curve* c = static_cast<Curve*>(lc);
  if (!c) return -1;
  out[0] = (double)c->N; out[1] = c->Ldbl; out[2] = c->step;
  out[3] = c->blx; out[4] = c->bly; out[5] = c->trx; out[6] = c->try_;
  out[7] = c->ampScale;
  for (int k = 0; k < 6; k++) {
    out[8 + 4 * k] = c->boxMin[k].re; out[9 + 4 * k] = c->boxMin[k].im;
    out[10 + 4 * k] = c->boxMax[k].re; out[11 + 4 * k] = c->boxMax[k].im;
  }
  out[32] = c->ldA; out[33] = c->ldB; out[34] = c->rho2; out[35] = c->rho2pow;
  return 0;
}

int lcirs_copy_cc_ca(void* lc, double* cc, double* ca) {
  g_err.clear();
  const Curve* c = static_cast<Curve*>(lc);
  if (!c) return -1;
  std::memcpy(cc, c->cc.data(), c->cc.size() * sizeof(cx));
  std::memcpy(ca, c->ca.data(), c->ca.size() * sizeof(cx));
  return 0;
}

int lcirs_batch_irs(int n, void** lcs, const double* traj) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!lcs || !traj))) throw std::invalid_argument("bad batch arguments");
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    Job job;
    for (int k = 0; k < n; k++) {
      if (!lcs[k]) throw std::invalid_argument("null light-curve object in batch");
      job.curves.push_back(static_cast<Curve*>(lcs[k]));
      job.starts.push_back(mk(traj[4 * k], traj[4 * k + 1]));
      job.ends.push_back(mk(traj[4 * k + 2], traj[4 * k + 3]));
    }
    E.run(job);
    for (auto* c : job.curves)
      if (c->amoebaOut) write_par_file(*c);
    return 0;
  });
}

int lcirs_fill_seeds(void* lc, int n, const double* pos, const int* nSeeds, const int* pix,
                     uint64_t* rays, double* amp, uint32_t* flags) {
  return guarded([&] {
    if (!lc || n <= 0 || !pos || !nSeeds || !pix) throw std::invalid_argument("bad arguments");
    for (int i = 0; i < n; i++)
      if (nSeeds[i] < 0 || nSeeds[i] > kMaxSeeds)
        throw std::invalid_argument("seed count out of range (0 .. 64 per position)");
    Curve* c = static_cast<Curve*>(lc);
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    Job job;
    job.curves = {c};
    job.starts = {mk(0, 0)};
    job.ends = {mk(0, 0)};
    job.runSeeds = false;
    job.pos.resize(n);
    job.nSeeds.assign(nSeeds, nSeeds + n);
    job.seeds.resize((size_t)n * kMaxSeeds);
    for (int i = 0; i < n; i++) {
      job.pos[i] = make_double2(pos[2 * i], pos[2 * i + 1]);
      for (int k = 0; k < kMaxSeeds; k++)
        job.seeds[(size_t)i * kMaxSeeds + k] =
            make_int2(pix[2 * ((size_t)i * kMaxSeeds + k)], pix[2 * ((size_t)i * kMaxSeeds + k) + 1]);
    }
    E.run(job);
    for (int i = 0; i < n; i++) {
      if (rays) rays[i] = c->rays[i];
      if (amp) amp[i] = c->full[i];
      if (flags) flags[i] = c->flags[i];
    }
    return 0;
  });
}

int lcirs_fill_positions(void* lc, int n, const int* idx, uint64_t* rays, double* amp,
                         uint32_t* flags) {
  return lcirs_fill_positions_seeds(lc, n, idx, nullptr, nullptr, rays, amp, flags);
}

int lcirs_batch_fill_positions(int n, void** lcs, const int* counts, const int* idx,
                               const int* nSeeds, const int* pix, uint64_t* rays, double* amp,
                               uint32_t* flags) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!lcs || !counts)) || (!nSeeds) != (!pix))
      throw std::invalid_argument("bad batch arguments");
    long long total = 0;
    for (int k = 0; k < n; k++) {
      if (!lcs[k]) throw std::invalid_argument("null light-curve object in batch");
      if (counts[k] < 0) throw std::invalid_argument("negative position count");
      const Curve* c = static_cast<Curve*>(lcs[k]);
      if (counts[k] > 0 && !c->hasSeeds)
        throw std::runtime_error("no seed search has run on an object of the batch");
      for (long long i = total; i < total + counts[k]; i++)
        if (idx[i] < 0 || idx[i] >= (int)c->nSeeds.size())
          throw std::invalid_argument("position index out of range (0 .. lcLength)");
      total += counts[k];
    }
    if (total == 0) return 0;
    if (nSeeds)
      for (long long i = 0; i < total; i++)
        if (nSeeds[i] < 0 || nSeeds[i] > kMaxSeeds)
          throw std::invalid_argument("seed count out of range (0 .. 64 per position)");
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    Job job;
    job.runSeeds = false;
    job.pos.resize(total);
    job.nSeeds.resize(total);
    job.seeds.resize((size_t)total * kMaxSeeds);
    job.cost.resize(total);
    long long off = 0;
    for (int k = 0; k < n; k++) {
      if (counts[k] == 0) continue;
      Curve* c = static_cast<Curve*>(lcs[k]);
      job.curves.push_back(c);
      job.starts.push_back(mk(0, 0));
      job.ends.push_back(mk(0, 0));
      job.nPos.push_back(counts[k]);
      const double area = 3.141592653589793 * (double)c->ppr * (double)c->ppr;
      for (long long i = off; i < off + counts[k]; i++) {
        const int p = idx[i];
        const cx x = traj_pos(c->seedStart, c->seedEnd, (unsigned)p, c->lcLength);
        job.pos[i] = make_double2(x.re, x.im);
        job.nSeeds[i] = nSeeds ? nSeeds[i] : c->nSeeds[p];
        if (nSeeds)
          std::memcpy(&job.seeds[(size_t)i * kMaxSeeds], &pix[(size_t)i * kMaxSeeds * 2],
                      kMaxSeeds * sizeof(int2));
        else
          c->seed_row(p, reinterpret_cast<int*>(&job.seeds[(size_t)i * kMaxSeeds]));
        job.cost[i] = (float)(c->pointAmp[p] * area);
      }
      off += counts[k];
    }
    E.run(job);
    off = 0;
    for (int k = 0; k < n; k++) {
      if (counts[k] == 0) continue;
      const Curve* c = static_cast<Curve*>(lcs[k]);
      for (int i = 0; i < counts[k]; i++) {
        if (rays) rays[off + i] = c->rays[i];
        if (amp) amp[off + i] = c->full[i];
        if (flags) flags[off + i] = c->flags[i];
      }
      off += counts[k];
    }
    return 0;
  });
}

int lcirs_fill_positions_seeds(void* lc, int n, const int* idx, const int* nSeeds, const int* pix,
                               uint64_t* rays, double* amp, uint32_t* flags) {
  return guarded([&] {
    if (!lc || n <= 0 || !idx || (!nSeeds) != (!pix)) throw std::invalid_argument("bad arguments");
    if (nSeeds)
      for (int i = 0; i < n; i++)
        if (nSeeds[i] < 0 || nSeeds[i] > kMaxSeeds)
          throw std::invalid_argument("seed count out of range (0 .. 64 per position)");
    Curve* c = static_cast<Curve*>(lc);
    if (!c->hasSeeds) throw std::runtime_error("no seed search has run on this object");
    const int npos = (int)c->nSeeds.size();
    for (int i = 0; i < n; i++)
      if (idx[i] < 0 || idx[i] >= npos)
        throw std::invalid_argument("position index out of range (0 .. lcLength)");
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    Job job;
    job.curves = {c};
    job.starts = {mk(0, 0)};
    job.ends = {mk(0, 0)};
    job.runSeeds = false;
    job.pos.resize(n);
    job.nSeeds.resize(n);
    job.seeds.resize((size_t)n * kMaxSeeds);
    job.cost.resize(n);
    const double area = 3.141592653589793 * (double)c->ppr * (double)c->ppr;
    for (int i = 0; i < n; i++) {
      const int p = idx[i];
      const cx x = traj_pos(c->seedStart, c->seedEnd, (unsigned)p, c->lcLength);
      job.pos[i] = make_double2(x.re, x.im);
      if (nSeeds) {
        job.nSeeds[i] = nSeeds[i];
        std::memcpy(&job.seeds[(size_t)i * kMaxSeeds], &pix[(size_t)i * kMaxSeeds * 2],
                    kMaxSeeds * sizeof(int2));
      } else {
        job.nSeeds[i] = c->nSeeds[p];
        c->seed_row(p, reinterpret_cast<int*>(&job.seeds[(size_t)i * kMaxSeeds]));
      }
      // the seed search's prediction, as get_lc_irs orders and sizes by it
      job.cost[i] = (float)(c->pointAmp[p] * area);
    }
    // a run without seed search leaves the object's seeds and point data
    E.run(job);
    for (int i = 0; i < n; i++) {
      if (rays) rays[i] = c->rays[i];
      if (amp) amp[i] = c->full[i];
      if (flags) flags[i] = c->flags[i];
    }
    return 0;
  });
}

int tl_ccc_batch(int n, const double* params, double* cc, double* ca) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!params || !cc || !ca))) throw std::invalid_argument("bad arguments");
    if (n == 0) return 0;
    std::vector<Lens> lenses(n);
    for (int g = 0; g < n; g++) {
      const double* q = params + 5 * (size_t)g;
      for (int k = 0; k < 5; k++)
        if (!std::isfinite(q[k])) throw std::invalid_argument("lens parameters must be finite");
      lenses[g] = make_lens(q[0], q[1], q[2], q[3], q[4]);
    }
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    E.init();
    E.dLenses.ensure(n);
    E.dCccTab.ensure(1);
    E.dCcOut.ensure((size_t)n * 6 * kCccLen);
    E.dCaOut.ensure((size_t)n * 6 * kCccLen);
    HIPCHK(hipMemcpyAsync(E.dLenses.p, lenses.data(), n * sizeof(Lens), hipMemcpyHostToDevice,
                          E.stream));
    HIPCHK(hipMemcpyAsync(E.dCccTab.p, &ccc_table(), sizeof(CccTable), hipMemcpyHostToDevice,
                          E.stream));
    HIPCHK((hipError_t)launch_ccc(E.dLenses.p, n, E.dCccTab.p, E.dTrig.p, E.dCcOut.p, E.dCaOut.p,
                                  E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    HIPCHK(hipMemcpy(cc, E.dCcOut.p, (size_t)n * 6 * kCccLen * sizeof(cx), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ca, E.dCaOut.p, (size_t)n * 6 * kCccLen * sizeof(cx), hipMemcpyDeviceToHost));
    return 0;
  });
}

int lcirs_set_slot_trace(void* lc, const double* xy, long n) {
  return guarded([&] {
    if (!lc) throw std::invalid_argument("null light-curve object");
    if (n < 0 || (n > 0 && !xy)) throw std::invalid_argument("bad slot trace");
    if (n > INT32_MAX) throw std::invalid_argument("slot trace too long");
    Curve* c = static_cast<Curve*>(lc);
    c->slotTrace.resize((size_t)n);
    for (long k = 0; k < n; k++) c->slotTrace[k] = mk(xy[2 * k], xy[2 * k + 1]);
    c->chain.calls = 0;
    return 0;
  });
}

int tl_image_poly(double a, double b, double th, double m2, double m3, int n, const double* zeta,
                  double* out, int onDevice) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!zeta || !out))) throw std::invalid_argument("bad arguments");
    const Lens L = make_lens(a, b, th, m2, m3);
    cx fixed[11];
    image_poly_static(L, fixed);
    if (!onDevice) {
      for (int g = 0; g < n; g++) {
        cx c[11];
        image_poly(L, fixed, mk(zeta[2 * g], zeta[2 * g + 1]), c);
        for (int k = 0; k < 11; k++) {
          out[22 * (size_t)g + 2 * k] = c[k].re;
          out[22 * (size_t)g + 2 * k + 1] = c[k].im;
        }
      }
      return 0;
    }
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    E.init();
    if (n == 0) return 0;
    cx *dz = nullptr, *dc = nullptr, *df = nullptr;
    auto release = [&] {
      if (dz) hipFree(dz);
      if (dc) hipFree(dc);
      if (df) hipFree(df);
    };
    try {
      HIPCHK(hipMalloc(&dz, (size_t)n * sizeof(cx)));
      HIPCHK(hipMalloc(&dc, (size_t)n * 11 * sizeof(cx)));
      HIPCHK(hipMalloc(&df, 11 * sizeof(cx)));
      HIPCHK(hipMemcpy(dz, zeta, (size_t)n * sizeof(cx), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(df, fixed, 11 * sizeof(cx), hipMemcpyHostToDevice));
      HIPCHK((hipError_t)launch_image_poly(L, df, dz, n, dc, E.stream));
      HIPCHK(hipStreamSynchronize(E.stream));
      HIPCHK(hipMemcpy(out, dc, (size_t)n * 11 * sizeof(cx), hipMemcpyDeviceToHost));
    } catch (...) {
      release();
      throw;
    }
    release();
    return 0;
  });
}

int tl_div_selftest(long long n, unsigned long long seed) {
  return guarded([&] {
    if (n < 0) throw std::invalid_argument("bad arguments");
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    E.init();
    unsigned long long* d = nullptr;
    unsigned long long bad = 0;
    HIPCHK(hipMalloc(&d, sizeof bad));
    try {
      HIPCHK(hipMemsetAsync(d, 0, sizeof bad, E.stream));
      HIPCHK((hipError_t)launch_div_check(n, seed, d, E.stream));
      HIPCHK(hipMemcpyAsync(&bad, d, sizeof bad, hipMemcpyDeviceToHost, E.stream));
      HIPCHK(hipStreamSynchronize(E.stream));
    } catch (...) {
      hipFree(d);
      throw;
    }
    hipFree(d);
    return bad > 0x7fffffffull ? 0x7fffffff : (int)bad;
  });
}

int tl_math_selftest(int n, unsigned long long seed) {
  // the branch-free complex helpers against the general ones, on the host
  unsigned long long st = seed ? seed : 1;
  auto next = [&]() {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
  };
  auto val = [&]() -> double {
    const unsigned long long r = next();
    switch (r & 15) {
      case 0: return 0.0;
      case 1: return -0.0;
      case 2: return (r & 16) ? (double)INFINITY : -(double)INFINITY;
      case 3: return (double)NAN;
      case 4: return ldexp((double)(next() >> 11) * 0x1p-53, (int)(next() % 2100) - 1075);
      default: {
        const double m = (double)(next() >> 11) * 0x1p-53 + 0.5;
        const double v = ldexp(m, (int)(next() % 41) - 20);
        return (r & 32) ? -v : v;
      }
    }
  };
  auto same = [](double x, double y) {
    return (std::isnan(x) && std::isnan(y)) || std::memcmp(&x, &y, sizeof x) == 0;
  };
  int bad = 0;
  for (int i = 0; i < n; i++) {
    const cx u = mk(val(), val()), v = mk(val(), val());
    const double h0 = hyp(u.re, u.im), h1 = hyp_sel(u.re, u.im);
    const cx d0 = cdiv(u, v), d1 = cdiv_sel(u, v);
    const cx q0 = csqrt_(u), q1 = csqrt_sel(u);
    if (!same(h0, h1) || !same(d0.re, d1.re) || !same(d0.im, d1.im) || !same(q0.re, q1.re) ||
        !same(q0.im, q1.im))
      bad++;
  }
  return bad;
}

int tl_last_stats(double* out) {
  g_err.clear();
  const Stats& s = engine().stats;
  out[0] = s.tasks; out[1] = s.evals; out[2] = s.ins; out[3] = s.dup;
  out[4] = s.seedMs; out[5] = s.fillMs; out[6] = s.fillLaunches; out[7] = s.totalMs;
  out[8] = s.busyMs; out[9] = s.insertMs; out[10] = s.idlePolls; out[11] = s.waves;
  out[12] = s.scouts; out[13] = s.scoutRuns; out[14] = s.trialMs; out[15] = s.ringMs;
  out[16] = s.solves; out[17] = s.specPieces; out[18] = s.specRepairs;
  out[19] = s.specFallback; out[20] = s.graphBuilds;
  return 0;
}

int tl_set_option(const char* name, double value) {
  g_err.clear();
  if (!name) return -1;
  Options& o = engine().opt;
  const std::string n(name);
  if (n == "fill_workgroups") o.fillWorkgroups = (int)value;
  else if (n == "probe_limit") o.probeLimit = std::max(64, (int)value);
  else if (n == "queue_log2") o.queueLog2 = std::min(28, std::max(10, (int)value));
  else if (n == "scout_rows") o.scoutRows = std::max(0, (int)value);
  else if (n == "scout_jump") o.scoutJump = std::min(32767, std::max(1, (int)value));
  else if (n == "scout_width") o.scoutWidth = std::max(0, (int)value);
  else if (n == "scout_retry") o.scoutRetry = value != 0.0 ? 1 : 0;
  else if (n == "leap_every") o.leapEvery = std::max(0, (int)value);
  else if (n == "scout_share") o.scoutShare = value != 0.0 ? 1 : 0;
  else if (n == "use_graph") o.useGraph = value != 0.0 ? 1 : 0;
  else if (n == "share_rays") o.shareRays = std::max(0.0, value);
  else if (n == "leap_jumps") o.leapJumps = std::min(64, std::max(2, (int)value));
  else if (n == "poll_backoff") o.pollBackoff = std::min(6, std::max(0, (int)value));
  else if (n == "pool_newest") o.poolNewest = value != 0.0 ? 1 : 0;
  else if (n == "run_pixels") o.runPixels = std::min(65536.0, std::max(1.0, value));
  else if (n == "table_load") o.tableLoad = std::min(64.0, std::max(1.5, value));
  else if (n == "time_limit_s") o.timeLimitS = std::max(0.1, value);
  else if (n == "seed_chunks") o.seedChunks = std::max(1, (int)value);
  else if (n == "seed_chunk_min") o.seedChunkMin = std::max(1, (int)value);
  else if (n == "spec_piece") o.specPiece = std::max(0, (int)value);
  else if (n == "spec_overlap") o.specOverlap = std::min(4096, std::max(1, (int)value));
  else if (n == "spec_depth") o.specDepth = std::min(4096, std::max(1, (int)value));
  else return -1;
  return 0;
}

const char* tl_last_error(void) { return g_err.c_str(); }

int tl_device_info(char* name, int len, int* cus) {
  return guarded([&] {
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    E.init();
    if (name && len > 0) {
      std::strncpy(name, E.name.c_str(), len - 1);
      name[len - 1] = 0;
    }
    if (cus) *cus = E.cus;
    return 0;
  });
}

}  // extern "C"
