/* C ABI of the MI355X IRS light-curve engine (liblcirs.so).
 *
 * The first group of entry points is the reference's own LightCurveIRS ABI,
 * with the same names, argument types and meaning, so that the reference's
 * ctypes class LC_irs (Examples/ctypes_classes.py:128-193) drives this
 * library unchanged.  Every function is safe to call from C or ctypes: no C++
 * exception crosses the boundary.  Where the reference has no error channel
 * (its functions return void) failures are recorded and can be read with
 * tl_last_error().
 *
 * Threading: one object must not be used from two threads at once (as in the
 * reference); different objects may be used concurrently; the GPU work of all
 * objects goes through one engine per process and is serialised there.
 */
#ifndef TL_LCIRS_H
#define TL_LCIRS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference ABI (src/LightCurves/lcirs.cc:515-609) -------------------- */

/* LightCurveIRS constructor (lcirs.cc:517-536).  Lens a, b, th, m2, m3 in the
 * reference frame (z1 = 0, z2 = a, z3 = b e^{i th}), source radius, lcLength
 * and pointsPerRadius.  Builds the critical curves, caustic boxes and
 * image-plane grid on the host.  m3 = 0 is the reference's binary lens
 * (lcirs.cc:29-36): irsBinary and the binary image-count rule.  Returns NULL
 * on invalid input (the reference has no such check). */
void* lcirs_new(double a, double b, double th, double m2, double m3, double sourceSize,
                unsigned lcLength, long pointsPerRadius);

/* Point-source light curve (lcirs.cc:539-549 -> LightCurveBase::getLC,
 * lcbase.cc:40-52): lcLength positions at t = i/(lcLength-1). */
void get_lc(void* lc, double iniX, double iniY, double finX, double finY);

/* Extended-source light curve by inverse ray shooting (lcirs.cc:552-562 ->
 * LightCurveIRS::getLCIRS, lcirs.cc:103-191): lcLength+1 positions at
 * t = i/lcLength, computed on the GPU. */
void get_lc_irs(void* lc, double iniX, double iniY, double finX, double finY);

/* Copies the first lcLength values of the last light curve (lcirs.cc:566-576). */
void copy_lc(void* lc, double* lcArray);

/* Limb darkening (lcirs.cc:579-598): "linear" selects the normalised linear
 * profile with parameter v, anything else the default A = 0.6, B = 0.4. */
void set_limb_darkening(void* lc, const char* model, double v);

/* Amoeba and parameter print-out (lcirs.cc:601-607): after each get_lc_irs,
 * point i's runs are written to "<amoebaPrefix><i>" as "ny,xleft,xright"
 * lines (amoeba.cc:90-108) and the grid parameters to parFile as JSON
 * (lcirs.cc:479-512). */
void set_amoeba_printout(void* lc, const char* amoebaPrefix, const char* parFile);

/* ---- reference ABI of LightCurveCUDA (src/Cuda/cudalc.cc:486-542) -------- */

/* lccuda_new (cudalc.cc:488-507): the same object as lcirs_new. */
void* lccuda_new(double a, double b, double th, double m2, double m3, double sourceSize,
                 unsigned lcLength, long pointsPerRadius);
/* get_lc_cuda (cudalc.cc:510-520): the extended-source curve.  The
 * reference's CUDA path is an FP32 approximation; here it is the exact FP64
 * curve of get_lc_irs (lcLength+1 positions, copy_lc copies lcLength). */
void get_lc_cuda(void* lc, double iniX, double iniY, double finX, double finY);
/* set_limb_darkening_cuda (cudalc.cc:523-540): as set_limb_darkening. */
void set_limb_darkening_cuda(void* lc, const char* model, double v);

/* ---- additions (no counterpart in the reference) ------------------------ */

/* Frees an object (the reference never frees them). */
void lcirs_delete(void* lc);
/* A new object with the same geometry, grid, critical curves, options and
 * root chain, without the last results (no set-up is recomputed).  NULL on
 * error.  No counterpart in the reference. */
void* lcirs_clone(void* lc);
/* lcirs_new with the lens's critical curves and caustics given (cc, ca: 6 *
 * 500 complex values each, branch-major as lcirs_copy_cc_ca and tl_ccc_batch
 * give them) instead of computed on the host: for callers that build many
 * geometries at once and compute their critical curves with tl_ccc_batch
 * (a Metropolis step's proposals).  The object equals lcirs_new's when the
 * curves are that geometry's.  NULL on error.  No counterpart in the
 * reference. */
void* lcirs_new_ccc(double a, double b, double th, double m2, double m3, double sourceSize,
                    unsigned lcLength, long pointsPerRadius, const double* cc,
                    const double* ca);

/* Per-point results of the last get_lc_irs, lcLength+1 values each:
 * magnifications (all of them, copy_lc stops at lcLength), ray counts and
 * status flags (bit values in TL_FLAG_*).  Return 0, or -1 if there is no
 * light curve yet. */
int lcirs_copy_lc_full(void* lc, double* out);
/* Point-source magnification (LightCurveBase::getPointAmp, lcbase.cc:17-38)
 * at each of the lcLength+1 positions of the last seed search (get_lc_irs or
 * lcirs_find_seeds), from the images it found; the magnification-weighted
 * centroid of those images (x, y per position; 2*(lcLength+1) values); and
 * the distance from the position to the nearest caustic sample, negative
 * where the position has more than 4 images; and the share of the point
 * magnification carried by its two brightest images.  Any pointer may be
 * NULL. */
int lcirs_copy_point_amp(void* lc, double* amp, double* centroid, double* causticDist,
                         double* pairShare);

/* Disk samples.  With `on`, every later seed search of the object (get_lc_irs,
 * lcirs_find_seeds, lcirs_batch_irs) also computes the point magnification
 * at 12 points around each position, solved from that position's own image
 * roots: offsets rho/2 on the axes (+x, +y, -x, -y), rho on the axes, and
 * rho on the diagonals (45, 135, 225, 315 degrees).  These are the samples
 * of Gould's (2008) hexadecapole approximation of the finite-source
 * magnification.  No counterpart in the reference. */
int lcirs_set_ring(void* lc, int on);
/* Split seed search.  With parts = P > 1, the seed search of the object
 * (every call that runs one) cuts its positions into P contiguous pieces that
 * run in parallel; the first piece continues the object's root chain, the
 * others start from a fresh solve.  Faster for long curves (the chain is
 * serial), but the roots, and at rare ill-conditioned positions the images
 * and ray counts, no longer follow the reference's single chain
 * (lcirs.cc:116-130, imgpoint.cc:160-205).  parts = 1 (the default) is the
 * reference's behaviour.  Piece q holds positions [n q / P, n (q+1) / P) of
 * the n = lcLength+1 positions.  No counterpart in the reference. */
int lcirs_set_seed_split(void* lc, int parts);
/* The disk-sample magnifications of the last seed search, (lcLength+1) x 12;
 * (may be NULL) the 10 roots of each position's own solve, (lcLength+1) x 10
 * complex as re, im pairs; and (may be NULL) per position the number of
 * samples whose count of significant images (each at least 1e-4 of the
 * magnification) differs from the centre's: samples across a caustic from
 * the centre. */
int lcirs_copy_ring_amp(void* lc, double* ring, double* centreRoots, int* crossings);
/* The seed search of get_lc_irs alone (lcirs.cc:116-165, no fill): seeds
 * (lcirs_copy_seeds), point magnifications, centroids and caustic distances
 * (lcirs_copy_point_amp), with the root chain advanced as get_lc_irs
 * advances it.  For callers that fill only some positions
 * (lcirs_fill_seeds). */
int lcirs_find_seeds(void* lc, double iniX, double iniY, double finX, double finY);
/* lcirs_find_seeds of n objects in one engine run (one seed launch):
 * objects lcs[k] with trajectories traj[4k..4k+3].  No counterpart in the
 * reference. */
int lcirs_batch_find_seeds(int n, void** lcs, const double* traj);
int lcirs_copy_ray_counts(void* lc, uint64_t* out);
int lcirs_copy_flags(void* lc, uint32_t* out);

/* Grid parameters: out[0..7] = N, L, step, BLx, BLy, TRx, TRy, ampScale, the
 * 6 cumulative caustic boxes at out[8..31] (minx, miny, maxx, maxy), then
 * out[32] = A, out[33] = B of the limb-darkening profile, out[34] = rho^2,
 * out[35] = pow(rho, 2). */
int lcirs_get_grid(void* lc, double* out);

/* Critical curves and caustics of the object, 6*500 complex values each,
 * root-major (as the reference's copy_cc_ca lays them out, ccc.cc:280-294). */
int lcirs_copy_cc_ca(void* lc, double* cc, double* ca);

/* Many light curves in one engine run: objects lcs[k] with trajectories
 * traj[4k..4k+3] = iniX, iniY, finX, finY.  Afterwards each object holds its
 * light curve as after get_lc_irs.  Returns 0 or a negative error. */
int lcirs_batch_irs(int n, void** lcs, const double* traj);

/* Seeds of the last light curve's positions as the GPU seed kernel found
 * them: for point i, nSeeds[i] pixel pairs at pix[2*(i*64 + k) + {0,1}].
 * Returns 0 or -1. */
int lcirs_copy_seeds(void* lc, int* nSeeds, int* pix);

/* lcirs_copy_seeds for the n positions idx[0..n-1] only: nSeeds[i] and the
 * 64 pixel pairs of row i at pix[2*(i*64 + k) + {0,1}].  Returns 0 or a
 * negative error (bad index, no seed search).  No counterpart in the
 * reference: it saves callers that need a few positions' seeds the copy of
 * the whole table. */
int lcirs_copy_seeds_at(void* lc, int n, const int* idx, int* nSeeds, int* pix);

/* The union of the seeds of lc and of the nOther objects others[] (same
 * trajectory length, each after a seed search) at the n positions idx: per
 * position, lc's seeds in order, then each other object's seeds not seen
 * before, in order, at most 64; rows as in lcirs_copy_seeds_at.  Returns 0
 * or a negative error.  No counterpart in the reference: the surrogate's
 * fallback positions are filled from the union of several split seed
 * searches. */
int lcirs_copy_seeds_union(void* lc, int nOther, void** others, int n, const int* idx,
                           int* nSeeds, int* pix);

/* Fills given seeds on the GPU (no seed search): n source positions pos[2i..],
 * seeds as pixel pairs pix[2*(i*64+k)+{0,1}] for k < nSeeds[i].  Writes ray
 * counts, magnifications and flags per position.  For tests of the fill. */
int lcirs_fill_seeds(void* lc, int n, const double* pos, const int* nSeeds, const int* pix,
                     uint64_t* rays, double* amp, uint32_t* flags);

/* Ray-shoot positions idx[0..n) (indices 0..lcLength) of the object's last
 * seed search with that search's seeds, the fill ordered and its tables
 * sized by the search's predicted rays (point magnification x pi ppr^2) as
 * in get_lc_irs.  Per position: rays, magnification, status flags.  For
 * callers that shoot a subset of a curve (a surrogate's fallback).  No
 * counterpart in the reference. */
int lcirs_fill_positions(void* lc, int n, const int* idx, uint64_t* rays, double* amp,
                         uint32_t* flags);
/* As lcirs_fill_positions, with the caller's seeds in place of the seed
 * search's: position idx[i] gets nSeeds[i] pixel pairs at
 * pix[2*(i*64+k)+{0,1}].  For seeds merged from several seed searches (a
 * surrogate's fallback takes the union of two split searches).  No
 * counterpart in the reference. */
int lcirs_fill_positions_seeds(void* lc, int n, const int* idx, const int* nSeeds, const int* pix,
                               uint64_t* rays, double* amp, uint32_t* flags);
/* lcirs_fill_positions(_seeds) of n objects in one engine run: object k
 * fills counts[k] positions, taken in order from the flat arrays idx (and,
 * unless nSeeds and pix are NULL, its seeds there); the outputs are flat
 * in the same order.  No counterpart in the reference. */
int lcirs_batch_fill_positions(int n, void** lcs, const int* counts, const int* idx,
                               const int* nSeeds, const int* pix, uint64_t* rays, double* amp,
                               uint32_t* flags);

/* Engine counters of the last run: out[0] fill tasks, out[1] evaluated
 * pixels, out[2] runs inserted, out[3] duplicate runs, out[4] seed-kernel
 * milliseconds, out[5] fill milliseconds (sum over launches), out[6] fill
 * launches, out[7] total milliseconds of the last run, out[8] summed busy
 * time of the fill's wavefronts (ms), out[9] of that, time in run-table
 * inserts (ms), out[10] idle polls, out[11] wavefronts of the last launch,
 * out[12] scout tasks, out[13] runs found by scouts, out[14] milliseconds
 * of the solve lists and their polynomials (trial kernels), out[15] of the
 * disk samples, out[16] solves (centre and caustic-trial positions),
 * out[17] pieces of the verified speculative chain, out[18] of them not in
 * agreement with their predecessor (repair chains run), out[19] 1 when the
 * speculative chain could not be resolved and the serial chain ran, out[20]
 * fill graphs captured (0 when every fill reused the previous launch's
 * graph).  out must hold 21 values. */
int tl_last_stats(double* out);

/* Critical curves and caustics of n geometries on the GPU (one lane each):
 * params[5*g..5*g+4] = a, b, th, m2, m3; cc and ca receive 6*500 complex
 * values per geometry (branch-major, as lcirs_copy_cc_ca), equal to the host
 * set-up's (CriticalCurveCaustic::getCC/getCa, ccc.cc:42-98,160-203) to the
 * bit.  For callers that need many geometries at once (a Metropolis chain
 * makes a new one per proposal).  Returns 0, or -1 with tl_last_error(). */
int tl_ccc_batch(int n, const double* params, double* cc, double* ca);

/* Values for the slot the reference's chained solve reads one element past
 * the previous roots (initRoots[10]; laguerre.cc:111-128 via
 * imgpoint.cc:187).  The reference never writes that element: it holds
 * whatever its allocator left there, so at the few positions where it
 * steers a root search the reference's own ray counts depend on its heap
 * history.  The engine uses 0 there (what the reference reads from zeroed
 * memory).  With a trace of n values set, getImages call k of this object's
 * chain (centre and caustic-trial solves in order, counted from the next
 * light curve on, k < n) starts its chained solve with (xy[2k], xy[2k+1])
 * instead: a recorded trace makes the engine reproduce a given run of the
 * reference bit for bit.  n = 0 clears it.  No counterpart in the
 * reference.  Returns 0, or -1 with tl_last_error(). */
int lcirs_set_slot_trace(void* lc, const double* xy, long n);

/* The degree-10 image polynomial at n source positions of one lens, as the
 * reference's getRootsPrecalculated forms it (imgpoint.cc:163-176:
 * getCoeffsOptNoZ + getCoeffsOptJustZ, imgpointcoeff.cc:223-447): zeta[2g],
 * zeta[2g+1] in, out[22g + 2k], out[22g + 2k + 1] = coefficient of z^k.
 * onDevice = 0 evaluates on the host, 1 with the GPU kernel the engine
 * uses; both give the reference's bits.  No counterpart in the reference
 * (test access).  Returns 0, or -1 with tl_last_error(). */
int tl_image_poly(double a, double b, double th, double m2, double m3, int n, const double* zeta,
                  double* out, int onDevice);

/* Host self-test of the branch-free complex routines the kernels use (modulus,
 * division, square root) against the general ones on n random operand pairs,
 * normal, tiny, huge, zero, infinite and NaN; returns the number of pairs
 * whose results differ in any bit.  No counterpart in the reference. */
int tl_math_selftest(int n, unsigned long long seed);

/* GPU self-test of the fill's in-range division (the IEEE double division
 * sequence without operand pre-scaling, used for the ray test's quotients
 * m_k / |z - z_k|^2 when both lie in its range) against IEEE division on n
 * generated operand pairs; returns the number whose results differ in any
 * bit, or -1 on error.  No counterpart in the reference. */
int tl_div_selftest(long long n, unsigned long long seed);

/* Engine settings: name/value pairs ("fill_workgroups", "probe_limit",
 * "queue_log2", "scout_rows", "scout_jump", "scout_retry", "scout_width",
 * "scout_share", "share_rays", "use_graph", "leap_every", "leap_jumps", "poll_backoff",
 * "run_pixels", "table_load", "seed_chunks", "spec_piece", "spec_overlap", "spec_depth",
 * "seed_chunk_min", "pool_newest", "time_limit_s").  None changes a result.  Returns 0 or
 * -1 for an unknown name.  No counterpart in the reference. */
int tl_set_option(const char* name, double value);

/* Description of the last error on this thread ("" if none). */
const char* tl_last_error(void);

/* Device name and compute-unit count of the engine's GPU; 0 or -1. */
int tl_device_info(char* name, int len, int* cus);

#define TL_FLAG_EDGE 1u
#define TL_FLAG_SEED_OVERFLOW 2u
#define TL_FLAG_HASH_FULL 4u
#define TL_FLAG_LABEL_FULL 8u
#define TL_FLAG_QUEUE_FULL 16u
#define TL_FLAG_TIMEOUT 32u
#define TL_FLAG_ROW_OVERFLOW 64u
/* Informational: a Laguerre iteration of the point's solves reached step 64,
 * where the reference reads its 8-entry fractional-step table past the end
 * (laguerre.cc:22,92), i.e. heap memory: the reference's result at this point
 * is undefined.  The engine wraps the table index modulo 8. */
#define TL_FLAG_KICK_TABLE 128u

#ifdef __cplusplus
}
#endif
#endif
