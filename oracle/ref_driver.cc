// Test-infrastructure driver around the reference LightCurveIRS, compiled
// together with the reference's own sources (see oracle/Makefile) into
// oracle/_ref/libtlref.so.  Nothing in the product links this file: it exists
// so that the golden fixtures and the CPU baseline come from the reference
// itself.
//
// What it adds on top of the reference:
//   * per-point ray counts: LightCurveIRS only prints them
//     (src/LightCurves/lcirs.cc:176-177); a streambuf parses that line;
//   * the lcLength+1'th magnification: getLCIRS writes lcVec[lcLength] one past
//     the end (lcirs.cc:117,181), so the vector is grown before the call;
//   * a large stack: the recursive lineFloodFill (lcirs.cc:451,460) needs one
//     frame per image row, so every call runs on a thread with a 2 GiB stack;
//   * read access to the grid parameters (protected members, lcirs.h:113-136).
#include <lcirs.h>
#include <lcbase.h>
#include <ccc.h>
#include <imgpoint.h>
#include <laguerre.h>

#include <pthread.h>
#include <cstdint>
#include <cstring>
#include <sstream>
#include <streambuf>
#include <string>
#include <vector>
#include <functional>

namespace {

// Collects "the count N" values from the lines getLCIRS prints per point.
class CountParser : public std::streambuf {
 public:
  std::vector<uint64_t> counts;
 protected:
  int overflow(int ch) override {
    if (ch == '\n') { parse(); line_.clear(); }
    else if (ch != EOF) line_.push_back(static_cast<char>(ch));
    return ch;
  }
  std::streamsize xsputn(const char* s, std::streamsize n) override {
    for (std::streamsize i = 0; i < n; ++i) overflow(s[i]);
    return n;
  }
 private:
  void parse() {
    static const char key[] = "and the count ";
    size_t p = line_.find(key);
    if (p == std::string::npos) return;
    counts.push_back(std::stoull(line_.substr(p + sizeof(key) - 1)));
  }
  std::string line_;
};

// Runs fn on a thread with a big stack and with std::cout captured.
void run_big_stack(const std::function<void()>& fn) {
  pthread_attr_t attr;
  pthread_attr_init(&attr);
  pthread_attr_setstacksize(&attr, size_t(2) << 30);
  pthread_t th;
  auto* task = new std::function<void()>(fn);
  pthread_create(&th, &attr, [](void* p) -> void* {
    auto* f = static_cast<std::function<void()>*>(p);
    (*f)();
    delete f;
    return nullptr;
  }, task);
  pthread_join(th, nullptr);
  pthread_attr_destroy(&attr);
}

struct Probe : public LightCurveIRS {
  using LightCurveIRS::LightCurveIRS;
  long N() const { return _imgPlaneSize; }
  double L() const { return _imgPlaneSizeDouble; }
  double step() const { return _imgGridPointSize; }
  double blx() const { return _bottomLeftCornerImg.real(); }
  double bly() const { return _bottomLeftCornerImg.imag(); }
  double trx() const { return _topRightCornerImg.real(); }
  double try_() const { return _topRightCornerImg.imag(); }
  double ampScale() const { return _ampScale; }
  const vector<pair<complex<double>>>& boxes() const { return caBoxes; }
  ImgPoint& images() { return _pointImages; }
  double rho() const { return _sourceRadius; }
};

// The slot recorder.  Solving from the previous roots, the reference reads
// initRoots[10], one element past the ten roots of its previous solve
// (laguerre.cc:111-128 via imgpoint.cc:187).  The roots vector grows by
// push_back to a capacity of 16, so that element lies inside the buffer but
// was never written: it holds whatever the allocator's recycled chunk held
// before.  The library is linked with --wrap on ImgPoint::getImages(complex),
// so that the calls getLCIRS makes (lcirs.cc:123,154) pass through
// __wrap_..., which records that element before the real call.  The
// recorder allocates nothing, so the heap history it observes is the
// reference's own.
//
// _tempRoots is private; a pointer to it may still be named in an explicit
// instantiation, which is how MemberPtr reaches it.
struct TempRootsTag {
  using type = vector<complex<double>> ImgPoint::*;
  friend type member_of(TempRootsTag);
};
template <typename Tag, typename Tag::type M>
struct MemberPtr {
  friend typename Tag::type member_of(Tag) { return M; }
};
template struct MemberPtr<TempRootsTag, &ImgPoint::_tempRoots>;

double* g_slot_out = nullptr;
long g_slot_cap = 0;
long g_slot_n = 0;

void set_ld(Probe& lc, int model, double v) {
  if (model == 1) lc.setLimbDarkeningModel(LimbDarkeningModel(v));
  else lc.setLimbDarkeningModel(LimbDarkeningModel());
}

}  // namespace

extern "C" vector<complex<double>> __real__ZN8ImgPoint9getImagesESt7complexIdE(ImgPoint* self,
                                                                              complex<double> pos);
extern "C" vector<complex<double>> __wrap__ZN8ImgPoint9getImagesESt7complexIdE(ImgPoint* self,
                                                                              complex<double> pos) {
  if (g_slot_out) {
    const vector<complex<double>>& tr = self->*member_of(TempRootsTag());
    if (g_slot_n < g_slot_cap) {
      double* o = g_slot_out + 2 * g_slot_n;
      if (tr.capacity() > 10) {
        std::memcpy(o, tr.data() + 10, sizeof(complex<double>));
      } else {
        o[0] = o[1] = __builtin_nan("");
      }
    }
    g_slot_n++;
  }
  return __real__ZN8ImgPoint9getImagesESt7complexIdE(self, pos);
}

extern "C" {

// Grid parameters: out[0..9] = N, L, step, BLx, BLy, TRx, TRy, ampScale,
// followed by the 6 cumulative caustic boxes (minx,miny,maxx,maxy) at out[8..31].
int tlref_grid(double a, double b, double th, double m2, double m3, double rho,
               unsigned lcLength, long ppr, int ldModel, double v, double* out) {
  run_big_stack([&] {
    CountParser sink;
    std::streambuf* old = std::cout.rdbuf(&sink);
    Probe lc(a, b, th, m2, m3, rho, lcLength, ppr);
    set_ld(lc, ldModel, v);
    std::cout.rdbuf(old);
    out[0] = double(lc.N()); out[1] = lc.L(); out[2] = lc.step();
    out[3] = lc.blx(); out[4] = lc.bly(); out[5] = lc.trx(); out[6] = lc.try_();
    out[7] = lc.ampScale();
    for (int k = 0; k < 6; ++k) {
      out[8 + 4 * k + 0] = lc.boxes()[k].first.real();
      out[8 + 4 * k + 1] = lc.boxes()[k].first.imag();
      out[8 + 4 * k + 2] = lc.boxes()[k].second.real();
      out[8 + 4 * k + 3] = lc.boxes()[k].second.imag();
    }
  });
  return 0;
}

// Extended-source light curve through the reference getLCIRS.
// ampOut and raysOut hold lcLength+1 values (the reference evaluates
// t = i/lcLength for i = 0..lcLength).  Returns the number of parsed counts.
int tlref_lc_irs(double a, double b, double th, double m2, double m3, double rho,
                 unsigned lcLength, long ppr, int ldModel, double v,
                 double sx, double sy, double ex, double ey,
                 double* ampOut, uint64_t* raysOut) {
  int parsed = 0;
  run_big_stack([&] {
    CountParser sink;
    std::streambuf* old = std::cout.rdbuf(&sink);
    Probe lc(a, b, th, m2, m3, rho, lcLength, ppr);
    set_ld(lc, ldModel, v);
    lc.lcVec.resize(lcLength + 1);
    lc.getLCIRS({sx, sy}, {ex, ey});
    std::cout.rdbuf(old);
    for (unsigned i = 0; i <= lcLength; ++i) ampOut[i] = lc.lcVec[i];
    parsed = int(sink.counts.size());
    for (int i = 0; i < parsed && i <= int(lcLength); ++i) raysOut[i] = sink.counts[i];
  });
  return parsed;
}

// tlref_lc_irs that also records, for every getImages call of getLCIRS (the
// centre and caustic-trial solves, in order), the value of the slot the
// chained solve reads past the end of the previous roots: slots[2k], [2k+1]
// for call k (NaN while the vector holds no such element).  Returns the
// number of calls (also when larger than maxSlots, whose excess is not
// stored), or -1 when the ray counts could not all be parsed.
long tlref_lc_irs_slots(double a, double b, double th, double m2, double m3, double rho,
                        unsigned lcLength, long ppr, int ldModel, double v,
                        double sx, double sy, double ex, double ey,
                        double* ampOut, uint64_t* raysOut, double* slots, long maxSlots) {
  int parsed = 0;
  long calls = 0;
  run_big_stack([&] {
    CountParser sink;
    std::streambuf* old = std::cout.rdbuf(&sink);
    Probe lc(a, b, th, m2, m3, rho, lcLength, ppr);
    set_ld(lc, ldModel, v);
    lc.lcVec.resize(lcLength + 1);
    g_slot_out = slots;
    g_slot_cap = maxSlots;
    g_slot_n = 0;
    lc.getLCIRS({sx, sy}, {ex, ey});
    calls = g_slot_n;
    g_slot_out = nullptr;
    std::cout.rdbuf(old);
    for (unsigned i = 0; i <= lcLength; ++i) ampOut[i] = lc.lcVec[i];
    parsed = int(sink.counts.size());
    for (int i = 0; i < parsed && i <= int(lcLength); ++i) raysOut[i] = sink.counts[i];
  });
  return parsed == int(lcLength) + 1 ? calls : -1;
}

// getLCIRS with the amoeba and parameter print-out switched on
// (setAmoebaPrintOut, lcirs.cc:469-477): the reference writes parFile and one
// amoeba file per point, amoebaPrefix + i.
int tlref_lc_irs_print(double a, double b, double th, double m2, double m3, double rho,
                       unsigned lcLength, long ppr, int ldModel, double v,
                       double sx, double sy, double ex, double ey,
                       const char* amoebaPrefix, const char* parFile) {
  run_big_stack([&] {
    CountParser sink;
    std::streambuf* old = std::cout.rdbuf(&sink);
    Probe lc(a, b, th, m2, m3, rho, lcLength, ppr);
    set_ld(lc, ldModel, v);
    lc.lcVec.resize(lcLength + 1);
    lc.setAmoebaPrintOut(true, amoebaPrefix, parFile);
    lc.getLCIRS({sx, sy}, {ex, ey});
    std::cout.rdbuf(old);
  });
  return 0;
}

// Seeds of every point, in the order getLCIRS visits them: the same calls on
// the same ImgPoint object (so the root-reuse chain is the reference's own),
// following the loop of lcirs.cc:116-165.  For point i, seeds are written to
// seedXY[2*(i*maxSeeds + k) + {0,1}], nSeeds[i] is their count (capped at
// maxSeeds; the true count is still reported).  nImg[i] = point-image count.
int tlref_seeds(double a, double b, double th, double m2, double m3, double rho,
                unsigned lcLength, long ppr,
                double sx, double sy, double ex, double ey,
                int maxSeeds, double* seedXY, int* nSeeds, int* nImg) {
  run_big_stack([&] {
    CountParser sink;
    std::streambuf* old = std::cout.rdbuf(&sink);
    Probe lc(a, b, th, m2, m3, rho, lcLength, ppr);
    complex<double> s{sx, sy}, e{ex, ey};
    for (unsigned i = 0; i <= lcLength; ++i) {
      complex<double> pos = s + (e - s) * (double(i) / double(lcLength));
      vector<complex<double>> imgPos = lc.images().getImages(pos);
      nImg[i] = int(imgPos.size());
      complex<double> trial;
      bool taken = false;
      for (unsigned r = 0; r < 6; ++r) {
        const auto& box = lc.boxes()[r];
        if (pos.real() + lc.rho() >= box.first.real() &&
            pos.real() - lc.rho() <= box.second.real() &&
            pos.imag() + lc.rho() >= box.first.imag() &&
            pos.imag() - lc.rho() <= box.second.imag()) {
          taken = false;
          for (unsigned k = 0; k + 1 < lc.ccc.caVec[r].size(); ++k) {
            if (lc.intersectionCheck(pos, lc.ccc.caVec[r][k], lc.ccc.caVec[r][k + 1], trial)) {
              if (!taken) {
                for (auto img : lc.images().getImages(trial)) imgPos.push_back(img);
                taken = true;
              }
            } else {
              taken = false;
            }
          }
        }
      }
      nSeeds[i] = int(imgPos.size());
      for (int k = 0; k < int(imgPos.size()) && k < maxSeeds; ++k) {
        seedXY[2 * (size_t(i) * maxSeeds + k) + 0] = imgPos[k].real();
        seedXY[2 * (size_t(i) * maxSeeds + k) + 1] = imgPos[k].imag();
      }
    }
    std::cout.rdbuf(old);
  });
  return 0;
}

// Critical curves and caustics as the LightCurveIRS constructor builds them
// (ccc(lensPar, 500) then getCa, lcirs.cc:16,23).  cc and ca hold 6*len
// complex values (root-major), as copy_cc_ca lays them out (ccc.cc:280-294).
int tlref_ccc(double a, double b, double th, double m2, double m3, int len,
              double* cc, double* ca) {
  CountParser sink;
  std::streambuf* old = std::cout.rdbuf(&sink);
  CriticalCurveCaustic c(LensPar(a, b, th, m2, m3), unsigned(len));
  c.getCa();
  std::cout.rdbuf(old);
  for (size_t r = 0; r < c.ccVec.size(); ++r)
    for (int j = 0; j < len; ++j) {
      size_t idx = r * len + j;
      cc[2 * idx] = c.ccVec[r][j].real(); cc[2 * idx + 1] = c.ccVec[r][j].imag();
      ca[2 * idx] = c.caVec[r][j].real(); ca[2 * idx + 1] = c.caVec[r][j].imag();
    }
  return int(c.ccVec.size());
}

// Point images along a sequence of source positions, one ImgPoint object
// (root reuse chain), as LightCurveBase::getPointAmp reaches them.
// roots: 10 complex per position; isImg: 10 flags per position.
int tlref_images(double a, double b, double th, double m2, double m3, int n,
                 const double* pos, double* roots, int* isImg, int* nImg) {
  CountParser sink;
  std::streambuf* old = std::cout.rdbuf(&sink);
  ImgPoint ip(LensPar(a, b, th, m2, m3));
  for (int i = 0; i < n; ++i) {
    auto imgs = ip.getImages(complex<double>{pos[2 * i], pos[2 * i + 1]});
    nImg[i] = int(imgs.size());
    for (int k = 0; k < 10; ++k) {
      roots[20 * i + 2 * k] = k < int(ip.roots.size()) ? ip.roots[k].real() : 0.0;
      roots[20 * i + 2 * k + 1] = k < int(ip.roots.size()) ? ip.roots[k].imag() : 0.0;
      isImg[10 * i + k] = k < int(ip.isImg.size()) ? int(ip.isImg[k]) : 0;
    }
  }
  std::cout.rdbuf(old);
  return 0;
}

// Point-source light curve (LightCurveBase::getLC, grid i/(lcLength-1)).
int tlref_point_lc(double a, double b, double th, double m2, double m3,
                   unsigned lcLength, double sx, double sy, double ex, double ey,
                   double* out) {
  CountParser sink;
  std::streambuf* old = std::cout.rdbuf(&sink);
  LightCurveBase lc(a, b, th, m2, m3, lcLength);
  lc.getLC({sx, sy}, {ex, ey});
  std::cout.rdbuf(old);
  for (unsigned i = 0; i < lcLength; ++i) out[i] = lc.lcVec[i];
  return 0;
}

// Point-source magnification at arbitrary positions, one chain.
int tlref_point_amp(double a, double b, double th, double m2, double m3, int n,
                    const double* pos, double* out) {
  CountParser sink;
  std::streambuf* old = std::cout.rdbuf(&sink);
  LightCurveBase lc(a, b, th, m2, m3, 1);
  for (int i = 0; i < n; ++i) out[i] = lc.getPointAmp({pos[2 * i], pos[2 * i + 1]});
  std::cout.rdbuf(old);
  return 0;
}

// Laguerre::solveRoots on a polynomial of degree deg (coefficients low to
// high, interleaved re/im), with optional initial roots (init == nullptr for
// none).  solveRoots(init) reads initRoots[j] for j = deg..1
// (laguerre.cc:118,126), i.e. one element past the end of a deg-long vector;
// init therefore holds deg+1 values and the last one is placed in that slot
// (the vector is built with deg+1 elements and shrunk, which keeps the value
// in its storage).  roots receives deg complex values.
int tlref_laguerre_solve(int deg, const double* coeffs, const double* init, double* roots) {
  CountParser sink;
  std::streambuf* old = std::cout.rdbuf(&sink);
  vector<complex<double>> p(deg + 1);
  for (int i = 0; i <= deg; ++i) p[i] = {coeffs[2 * i], coeffs[2 * i + 1]};
  Laguerre lag(p);
  vector<complex<double>> r;
  if (init) {
    vector<complex<double>> in(deg + 1);
    for (int i = 0; i <= deg; ++i) in[i] = {init[2 * i], init[2 * i + 1]};
    in.resize(deg);
    r = lag.solveRoots(in);
  } else {
    r = lag.solveRoots();
  }
  std::cout.rdbuf(old);
  for (int i = 0; i < deg && i < int(r.size()); ++i) {
    roots[2 * i] = r[i].real(); roots[2 * i + 1] = r[i].imag();
  }
  return int(r.size());
}

}  // extern "C"

extern "C" {
// ImgPoint coefficients as getRootsPrecalculated sums them
// (imgpoint.cc:163-176): 11 complex values, lowest order first.
int tlref_coeffs(double a, double b, double th, double m2, double m3,
                 double zr, double zi, double* out) {
  ImgPoint ip(LensPar(a, b, th, m2, m3));
  ip.setPos(zr, zi);
  vector<complex<double>> zf = ip.getCoeffsOptNoZ();
  vector<complex<double>> zd = ip.getCoeffsOptJustZ();
  for (int i = 0; i < 11; ++i) {
    complex<double> c = zf[i] + zd[i];
    out[2 * i] = c.real(); out[2 * i + 1] = c.imag();
  }
  return 0;
}
}
