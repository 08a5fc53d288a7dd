// Device-side data layout of the IRS engine and the host entry points that
// launch its kernels.  Shared by tl_kernels.hip and tl_api.cpp.
#pragma once

#include <stdint.h>
#include "tl_math.h"
#include "tl_coeffs.h"

namespace tl {

constexpr int kCccLen = 500;       // critical-curve samples per branch (lcirs.cc:16)
constexpr int kMaxSeeds = 64;      // seeds kept per source position
constexpr int kGroupPoints = 16384;  // source positions per run-table group (14-bit key
                                     // field); a fill launch holds any number of groups
constexpr int kMaxCover = 16;      // known runs read per row lookup
constexpr int kFillWaves = 16;     // waves per fill workgroup: 1 manager + 15 workers, four
                                   // waves per SIMD
constexpr int kFillThreads = 64 * kFillWaves;
constexpr int kLocalCap = 64;      // tasks in each worker's own ring (LDS)
constexpr int kPoolCap = 512;      // tasks in the workgroup's shared pool (LDS)

// Per-point status bits (OR-ed into flags[p]).
enum : uint32_t {
  kFlagEdge = 1u,          // a counted pixel lies in column <= 1 or >= N-2, or a seed
                           // fell outside the columns [0, N-1]: the reference's
                           // scan-bound behaviour there is not reproduced
  kFlagSeedOverflow = 2u,  // more than kMaxSeeds seeds; the extra ones were dropped
  kFlagHashFull = 4u,      // run table probe limit reached
  kFlagLabelFull = 8u,     // label table full
  kFlagQueueFull = 16u,    // task queue full
  kFlagTimeout = 32u,      // the fill launch hit its time limit
  kFlagRowOverflow = 64u,  // more than kMaxCover runs in one row (work repeated, result exact)
  kFlagKickTable = 128u,   // a Laguerre iteration of the point's solves reached step 64, where
                           // the reference reads its 8-entry fractional-step table past the
                           // end (laguerre.cc:22,92): heap memory, so the reference's result
                           // there is undefined; the engine wraps the index modulo 8
};

// One light curve (a LightCurveIRS object) as the kernels see it.
struct CurveDev {
  Lens L;
  cx fixedPoly[11];            // ImgPoint::getCoeffsOptNoZ of the lens (image_poly_static)
  double rho, rho2, rho2pow;   // rho, rho*rho (ray test), pow(rho, 2.0) (intersection test)
  double ldA, ldB;             // limb-darkening profile A + B sqrt(1 - r^2)
  double sumScale;             // fixed-point scale of the brightness sum (sum_scale_for)
  double blx, bly, step, Ldbl, Nd;
  long long N;
  double sx, sy, ex, ey;       // trajectory
  double ppr;
  unsigned lcLength;           // positions t = i/lcLength, i = 0..lcLength
  int pointBase;               // first global point index
  int causticBase;             // first caustic sample (6*kCccLen per curve)
  int chainHave;               // root-reuse state at the start of the run
  int ring;                    // compute the disk-sample magnifications (ring_lane_kernel)
  int split;                   // independent pieces of the seed search (1: one chain)
  int binImg;                  // binary image-count rule (a, b or m3 zero)
  int binIrs;                  // binary ray map (m3 = 0)
  int chainCalls;              // getImages calls made by the chain before this run
  int slotBase, slotCount;     // recorded values of tempRoots[10] (SeedArgs::slotTrace):
                               // getImages call k of the chain (k < slotCount) starts its
                               // chained solve with slotTrace[slotBase + k]; slotCount 0: none
  cx chain[11];
  cx boxMin[6], boxMax[6];
};

// A row task of batch point `b`: find the maximal hit runs of row `row` that
// meet [xa, xb].  kind: 0 = sweep step (arriving in direction dir from the
// parent run [xa, xb] of row row-dir; f1, f2 = sweep anchor column and rows
// since the anchor), 1 = seed pixel, 2 = scout (speculative interval, new
// label per run), 3 = look-back (row row, [xa, xb] = the child run's extent,
// pixels [f1, f2] known: the parent run and its two bounding misses).
struct alignas(16) Task {
  int b, row, xa, xb, label, kd, f1, f2;  // kd = kind | (dir < 0 ? 4 : 0)
};

// The source positions each light-curve point solves for, in the order the
// reference solves them (lcirs.cc:123-165): its centre, then the caustic
// trial points.  Written by trial_kernel for every point at once (they do
// not depend on the root chain), with the image polynomial of every solve.
struct SolvePool {
  int* base;          // [point] first solve in the pool
  int* count;         // [point] solves (1 + caustic trials)
  cx* zeta;           // [solve] source position
  int* owner;         // [solve] the point it belongs to
  cx* coef;           // [solve * 11 + k] coefficient of z^k (image_poly)
  unsigned long long* cursor;  // solves handed out (may exceed cap: then rerun bigger)
  int cap;            // pool capacity in solves
};

struct TrialArgs {
  const CurveDev* curves;
  const int* ptCurve;  // global point -> curve
  int nPts;
  const cx* caustics;
  double2* ptPos;      // source position of every point
  double* ptCausD2;    // squared distance to the nearest caustic sample
  SolvePool pool;
};

struct SeedArgs {
  const CurveDev* curves;
  int nCurves;
  const cx* caustics;
  const TrigTable* trig;
  SolvePool pool;      // filled by trial_kernel
  const double* ptCausD2;
  const cx* slotTrace; // recorded tempRoots[10] values (CurveDev::slotBase), may be null
  int2* seeds;        // [point][kMaxSeeds] pixel (nx, ny)
  int* nSeeds;
  float* cost;        // predicted rays (point-source magnification x pi ppr^2)
  double* ptAmp;      // point-source magnification at every position
  double2* ptCen;     // magnification-weighted centroid of the point images
  double* ptCaus;     // distance to the nearest caustic sample, negative when the
                      // position has more than 4 images (inside a caustic)
  double* ptPair;     // share of the point magnification in its two brightest images
  cx* ptRoots;        // the 10 roots of each position's own image solve
  uint32_t* flags;
  RootChain* chainOut; // per curve: root-reuse state after the last position
  const RootChain* chainIn;  // per curve: state to start from (nullptr: the curve's own)
  int chunk, nChunks;  // positions [n chunk / nChunks, n (chunk+1) / nChunks) of every
                       // curve of n positions
  int maxSplit;        // workgroups per curve: the largest CurveDev::split of the launch
  // Pieces with known start states (the verified speculative chain): piece
  // q > 0 of curve c starts from the roots pieceStart[(pieceBase[c] + q) * 10
  // + k] (the true chain's state before its first position) instead of a
  // fresh solve.  Null: fresh starts.
  const int* pieceBase;
  const cx* pieceStart;
};

// The verified speculative chain.  Every curve of n positions is cut into
// P = CurveDev::split pieces [n q / P, n (q+1) / P).  Round 1 runs every piece
// from a fresh solve (piece 0 from the curve's own chain state) and on into
// the next piece for `overlap` positions, recording the state after every
// position.  Where the next piece's own state equals the predecessor's
// continuation at one of its first `overlap` positions, both chains agree
// from there on.  Where none does, round 2 runs a repair chain from the
// predecessor's end state on through the following pieces until it agrees
// with one of them.  spec_resolve_kernel then walks each curve's pieces from
// its first (whose start is known) and picks the true state at every piece
// start; seed_kernel recomputes all pieces from those states in parallel.
// A chain that has not agreed after `depth` further pieces sets `failed`,
// and the host falls back to the serial chain.
struct SpecArgs {
  const int* pieceBase;  // [curve] first global piece
  int nPieces;           // global pieces
  int overlap;           // positions of the next piece a chain runs into (<= its length)
  int depth;             // piece boundaries a repair chain may cross
  cx* own;               // [point * 10 + k] state after the point, by the piece that owns it
  cx* ext;               // [point * 10 + k] ... by the preceding piece's continuation
  int* merged;           // [piece] round 1: the piece agrees with its predecessor
  cx* repState;          // [(piece * depth + j) * 10 + k] repair chain of piece q: state
                         // before the first position of piece q + 1 + j
  int* repTarget;        // [piece] the piece its repair chain agreed with; -1: ran to the
                         // curve's end; -2: crossed `depth` boundaries without agreeing
  cx* start;             // [piece * 10 + k] resolved start states
  int* failed;           // set when some curve could not be resolved
};

struct FillArgs {
  const CurveDev* curves;
  const int* ptCurve;         // global point -> curve
  const double2* ptPos;       // global point -> source position
  const int* batchPts;        // batch point -> global point
  int nBatch;
  const int2* seedTasks;      // (batch point, seed slot) in processing order
  int nSeedTasks;
  const int2* seeds;
  // run table: one slice per group of kGroupPoints batch points (b >> 14),
  // slice g holding grpMask[g] + 1 entries from entry grpBase[g] on
  unsigned long long* htab;   // 2 words per entry
  unsigned long long htabEntries;  // all slices
  const unsigned long long* grpBase;
  const unsigned long long* grpMask;
  int probeLimit;
  // labels
  int* lparent;
  int* lpoint;
  unsigned long long* lcount;
  unsigned long long* lsum;   // 2 words per label: brightness total hi * 2^32 + lo (add_split)
  int* labelNext;
  int labelCap;
  // shared queue (tasks handed between waves) and termination
  Task* qbuf;
  unsigned long long* qseq;
  unsigned long long* qhead;
  unsigned long long* qtail;
  unsigned long long qcap;
  int* active;                // started workgroups holding work (the fill ends when 0 and no
                              // task is queued)
  int nGroups;                // workgroups of the launch
  int* seedNext;
  // results and status
  uint32_t* flags;            // per global point
  int* abortFlag;
  unsigned long long* stats;  // [0] tasks, [1] evaluated pixels, [2] runs inserted, [3] duplicates,
                              // [4] busy ticks, [5] insert ticks, [6] idle polls, [7] waves,
                              // [8] scout tasks, [9] runs found by scouts
  long long timeLimitTicks;   // s_memrealtime ticks (100 MHz)
  int scoutRows;              // rows swept before a scout is sent (0 = no scouts)
  int scoutJump;              // rows a scout jumps ahead
  int scoutWidth;             // runs wider than this shorten the jump (0: never)
  int scoutRetry;             // a scout that finds nothing is sent again half as far
  int scoutShare;             // sweeps hand their scouts to the workgroup's pool at once
  const float* ptCost;        // predicted rays per position
  float shareRays;            // ... from which its sweeps' scouts go to the pool at once (0: never)
  int leapEvery;              // every this many scouts a chain sends a leap scout (0: never)
  int leapJumps;              // ... this many jumps ahead
  int pollBackoff;            // idle polls of the global queue sleep up to 2^pollBackoff x 1024
  int poolNewest;             // a worker takes the pool's newest task (else its oldest)
                              // cycles
};

// Host-side launchers (tl_kernels.hip).  All return hipError_t as int.
int launch_trials(const TrialArgs& a, hipStream_t s);
// The polynomials of the first nSolves solves of the pool (after launch_trials).
int launch_trial_polys(const TrialArgs& a, int nSolves, hipStream_t s);
int launch_seed(const SeedArgs& a, hipStream_t s);
// Rounds of the verified speculative chain (SpecArgs), with a.maxSplit and
// the curves' split set to the pieces: round 1, the repair chains, and the
// resolution of the piece start states.
int launch_spec(const SeedArgs& a, const SpecArgs& sp, hipStream_t s);
// The fill's in-range division against IEEE division on n generated operand
// pairs (test access): adds the number of differing results to *bad (device).
int launch_div_check(long long n, unsigned long long seed, unsigned long long* bad,
                     hipStream_t s);
// image_poly for n source positions of one lens on the GPU (test access):
// zeta[n] in, coef[11 n] out (device pointers).
int launch_image_poly(const Lens& L, const cx* fixedPoly, const cx* zeta, int n, cx* coef,
                      hipStream_t s);
// Point magnification at the kRingSamples disk samples around each position
// of the curves with `ring` set: ring[kRingSamples p + s], and cross[p], the
// number of samples whose count of significant images (at least 1e-4 of the
// magnification each) differs from the centre's.
constexpr int kRingSamples = 12;
// coef: nPts * kRingSamples * 11 values of scratch for the samples' polynomials.
int launch_ring_amp(const CurveDev* curves, const int* ptCurve, const double2* ptPos,
                    const cx* ptRoots, int nPts, const TrigTable* trig, double* ring, int* cross,
                    cx* coef, hipStream_t s);
int launch_ccc(const Lens* lenses, int n, const CccTable* tab, const TrigTable* trig, cx* cc,
               cx* ca, hipStream_t s);
int launch_fill(const FillArgs& a, int nWorkgroups, hipStream_t s);
int fill_resident_workgroups();
int launch_clear_flag_bits(uint32_t* flags, const int* pts, int n, uint32_t bits, hipStream_t s);
int launch_clear_u64(unsigned long long* p, unsigned long long n, unsigned long long v, hipStream_t s);
int launch_fill_reset(const FillArgs& a, hipStream_t s);
// ptSum: 2 words per batch point, the brightness total as hi * 2^32 + lo
// (lo unsigned, hi signed).
int launch_resolve(const FillArgs& a, int nLabels, unsigned long long* ptCount,
                   unsigned long long* ptSum, hipStream_t s);
// Point-source light curves (LightCurveBase::getLC): curve k writes
// lcLength values at out[curves[k].pointBase ...].
// nTotal: the largest pointBase + lcLength of the curves; coef: nTotal * 11
// values of scratch for the positions' polynomials.
int launch_point_lc(const CurveDev* curves, int nCurves, int nTotal, const TrigTable* trig,
                    double* out, RootChain* chainOut, cx* coef, hipStream_t s);

// Fixed-point brightness sum: each pixel's A + B sqrt(1 - r^2) is rounded to
// a multiple of 1/scale and summed as an integer, which makes the sum
// independent of the order in which pixels are visited.  The scale is 2^34
// (rounding below 3e-11 of a pixel's value) for profiles whose brightness
// stays within 2 (every linear profile with 0 <= u <= 1, and the default),
// halved per doubling beyond, so that one pixel never exceeds 2^35 units and
// one run (at most 2^25 pixels) 2^60.  Wave-local totals are flushed before
// they pass 2^61; label and point totals are kept in two 64-bit words, the
// low 32 bits of every addition and the rest, so that they do not wrap.
constexpr double kSumScale = 17179869184.0;  // 2^34
inline double sum_scale_for(double ldA, double ldB) {
  const double peak = fabs(ldA) + fabs(ldB);
  double s = kSumScale;
  while (peak * s > 2.0 * kSumScale && s > 1.0) s *= 0.5;
  return s;
}

}  // namespace tl
