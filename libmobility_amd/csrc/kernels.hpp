// Launch interface of the gfx950 kernels (nbody_kernels.hip, lanczos_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lm {

// One unit of pair work: targets [tgt0, tgt0 + ntgt) against sources
// [src0, src1) of the same system; the partial result goes to slab `chunk`.
struct PairItem {
  int64_t tgt0;
  int64_t src0;
  int64_t src1;
  int32_t ntgt;
  int32_t chunk;
  int32_t wall;  // kernel of the item's system: 0 open RPY, 1 bottom wall
  int32_t sys;   // system index (items of inactive systems are skipped)
};

struct PairArgs {
  int wall;       // 0 open RPY, 1 bottom wall, 2 per item (mixed ensembles)
  int persistent; // 1: grid of `grid` workgroups fetching items through `counter`
  int grid;
  unsigned int* counter;  // zero before the launch (persistent mode)
  int fp32;       // 1: FP32 pair terms, FP64 per-tile accumulation
  int blocks;     // BlockBits mask
  int tpl;        // targets per lane (1 or 2)
  const void* pos;      // double4 or float4 records, N entries
  const double* F;      // 3N AoS, nullptr when absent
  const double* T;      // 3N AoS, nullptr when absent
  double* U;            // output slab base, 3N per slab
  double* W;            // output slab base, 3N per slab (nullptr if no angular)
  int64_t slab;         // elements between slabs (3N)
  const PairItem* items;
  int nitems;
  const int* active;    // per-system flags; items of systems with 0 are skipped (nullptr: all)
  // Non-persistent launches: a workgroup that finds *stop != 0 when it starts
  // does nothing (products a finished Lanczos solve no longer needs; nullptr:
  // never).  Its outputs are then undefined.
  const int* stop;
  double t0, rt0, r0;
};

constexpr int kPairThreads = 256;  // threads per workgroup = sources per LDS tile
constexpr int kChunkGrain = 64;    // source chunks of the ordered kernel: multiples of 64

void launch_pair(const PairArgs& a, hipStream_t s);

// Translational product with several force vectors in one pass over the pairs:
// U_r = t0 * M F_r, r < nrhs.  F holds nrhs consecutive 3N vectors; the
// partial of vector r and chunk c goes to U + (r * nch + c) * n3.
struct PairMultiArgs {
  int kind;   // 0 open RPY, 1 bottom wall, 2 per item
  int fp32;   // 1: FP32 pair terms, FP64 per-tile accumulation
  int nrhs;   // vectors per launch: 2 (two targets per lane) or 4 (one)
  const void* pos;
  const double* F;
  double* U;
  int64_t n3;
  int nch;
  const PairItem* items;
  int nitems;
  double t0;
};
constexpr int multi_tpl(int nrhs) { return nrhs >= 4 ? 1 : 2; }

// Symmetric evaluation (any block set, open or bottom-wall systems, one system
// or an ensemble): the grand mobility matrix is symmetric (M_ji = M_ij^T
// blockwise), so each unordered pair is evaluated once and applied to both
// particles.  Each system is cut into tiles of 256 * tpl particles; tile I
// meets tiles I + o (mod ntiles) for o = 1 .. hmax, hmax = ntiles / 2 (offset
// ntiles / 2 of an even count only from I < ntiles / 2), so every unordered
// tile pair of the system is met once.  Target-side sums of an item go to
// target slab `chunk`, source-side sums of offset o to source slab o - obase
// (each row of it written by exactly one item); the diagonal tile (o0 = 0) is
// an ordinary item into target slab `nchunks`.
struct SymSys {
  int64_t off;     // first particle
  int32_t n;       // particles
  int32_t ntiles;  // tiles of 256 * tpl
  int32_t hmax;    // ring offsets, ntiles / 2
  int32_t wall;    // 0 open RPY, 1 bottom wall
};
struct SymItem {
  int32_t sys;     // system
  int32_t tile;    // target tile I within the system
  int32_t o0, o1;  // partner offsets [o0, o1); o0 = 0: the diagonal tile
  int32_t chunk;   // target slab
};
// At most 256 particles of one tile of one system, for the reduction.
struct SymRows {
  int64_t p0;
  int32_t np;
  int32_t sys;
};
struct SymArgs {
  int fp32;
  int tpl;
  int wall;          // 0 open RPY, 1 bottom wall, 2 per system
  int blocks;        // BlockBits mask
  const void* pos;   // float4 / double4 records
  const double* F;   // 3N AoS (nullptr when absent)
  const double* T;   // 3N AoS (nullptr when absent)
  double* slab_t;    // nchunks + 1 target slabs of n3 (linear velocities)
  double* slab_s;    // source slabs of n3, one per ring offset of the pass
  double* slab_tw;   // as slab_t for angular velocities (nullptr without them)
  double* slab_sw;   // as slab_s for angular velocities
  int64_t n3;
  int obase;         // first ring offset of this pass: source slab index = o - obase
  const SymSys* sys;
  const SymItem* items;
  int nitems;
  const int* active;  // per-system flags; items of systems with 0 return at once (nullptr: all)
  // a workgroup that finds *stop != 0 when it starts does nothing (products a
  // finished Lanczos solve no longer needs; nullptr: never)
  const int* stop;
  double t0, rt0, r0;
};
void launch_pair_sym(const SymArgs& a, hipStream_t s);
// Per row block: out (= or +=, accumulate) the system's target slabs of the
// pass (its chunks of offsets [o0, o1), at most nch of them, then the
// diagonal slab nch when `diag`), then its source slabs of offsets o0 ..
// min(o1, hmax + 1) - 1 (offset ntiles/2 of an even count skipped for tiles
// below ntiles/2), in that fixed order.  Blocks of systems with active == 0
// are skipped (active nullptr: none).
struct SymReduceArgs {
  const double* slab_t;
  const double* slab_s;
  int64_t n3;
  int nch, o0, o1, diag, accumulate, tpl;
  const SymSys* sys;
  const SymRows* rows;
  int nrows;
  const int* active;
  double* out;
};
void launch_sym_reduce(const SymReduceArgs& r, hipStream_t s);
void launch_pair_multi(const PairMultiArgs& a, hipStream_t s);
void launch_chunk_reduce(const double* part, int nchunks, int64_t n, double* out, hipStream_t s);

// setPositions: per-system bounding boxes (6 doubles each: min xyz, max xyz)
// of the wall-shifted positions.
// kinds[s] (0 open, 1 wall) selects per system; nullptr: all systems `wall`.
void launch_system_bbox(const double* pos, const int64_t* sys_off, int nsys, double wall_shift,
                        const int* kinds, int wall, double* bbox, hipStream_t s);
// Records for the pair kernels; `shifted` receives the wall-shifted AoS copy.
void launch_make_records(const double* pos, int64_t N, const int64_t* sys_off, int nsys,
                         const double* bbox, const int* kinds, int wall, double wall_shift,
                         double a, double* shifted, void* rec64, void* rec32, hipStream_t s);
// v[3i..3i+2] = 0 for the particles of open systems (kinds[s] == 0)
void launch_zero_open_systems(double* v, int64_t N, const int64_t* sys_off, int nsys,
                              const int* kinds, hipStream_t s);

// ---- Lanczos / vector kernels (FP64) ---------------------------------------
// Normal(0,1) samples from Philox4x32-10 + Box-Muller: out[k], k < n.
void launch_normal_fill(double* out, int64_t n, uint64_t seed, uint64_t stream, hipStream_t s);

// Scratch for the single-pass reductions: partial sums + a self-resetting
// arrival counter.  kReduceBlocks partials per reduction slot (measured: a
// wider grid costs more in contention on the arrival counter than it gains
// in memory parallelism at n ~ 4e5).
constexpr int kReduceBlocks = 256;
struct ReduceScratch {
  double* partial;       // kReduceBlocks * 4 doubles
  unsigned int* counter; // 1 counter, zero at rest
};

// Largest Krylov dimension the device-side Lanczos state can hold.
constexpr int kLanczosMaxIterations = 1024;

// Convergence checks evaluated by one launch: the square roots of T_m for
// consecutive m (one workgroup each) and their estimates in one pass over V.
constexpr int kCheckBatch = 4;

// Convergence state of a Lanczos solve, decided on the device by the estimate
// kernel; the kernel also writes it to a host-mapped mirror.
struct LanczosStatus {
  int conv;       // 0 iterating, 1 converged, 2 NaN estimate
  int conv_iter;  // iteration that converged (-1 before)
  int old;        // estimate slot holding the previous estimate (oldBz)
  int result;     // slot holding the converged estimate
  int first;      // first iteration of the last batch of checks
  int nchk;       // checks of that batch that ran (up to the deciding one)
  int pad[2];
  double err[kCheckBatch];  // their error estimates, in order
};

// Lanczos scalar state resident on the device.
struct LanczosScalars {
  double normz;    // ||z||
  double alpha[kLanczosMaxIterations];  // hdiag
  double beta[kLanczosMaxIterations];   // hsup
  double err;      // last error estimate
  double errnum, errden;
  double rscale;   // reorthogonalisation: 1 / norm of the reorthogonalised v_{i+1}
  LanczosStatus status;
};

// ||z|| and v0 = z / ||z||; resets the status (previous estimate in slot `old`)
void launch_lanczos_start(const double* z, double* v0, int64_t n, ReduceScratch rs,
                          LanczosScalars* sc, int old, hipStream_t s);
// One Lanczos step after the product, in one launch (n <= kLanczosFusedMaxN):
// w = M v_i (from `w`, or summed over the nch source-chunk slabs of the pair
// kernel: entries k < n3 from slab_u, the rest from slab_w); w -= beta[i-1]
// v_prev; alpha[i] = w . v_i; w -= alpha[i] v_i; beta[i] = ||w|| with the
// breakdown rule; v_next = w / beta[i] or e1.  w_out (n doubles) holds w in
// between; partial / counter: reduction scratch (counter zero at rest).
constexpr int kLanczosFusedThreads = 1024;
constexpr int kLanczosFusedPerThread = 16;
constexpr int64_t kLanczosFusedMaxN = (int64_t)kLanczosFusedThreads * kLanczosFusedPerThread;
struct LanczosStepArgs {
  const double* w;
  const double* slab_u;
  const double* slab_w;
  int nch;
  int64_t n3;
  const double* v_prev;
  const double* v_i;
  double* v_next;
  double* w_out;
  double* partial;
  unsigned int* counter;
  // a workgroup that finds *stop != 0 when it starts skips the arithmetic
  // (the step is past convergence; its outputs are then undefined) but still
  // takes part in the arrival count (nullptr: never)
  const int* stop;
  int64_t n;
  int i;
  LanczosScalars* sc;
};
void launch_lanczos_step(const LanczosStepArgs& a, hipStream_t s);
// Full reorthogonalisation of v_{i+1} (column i + 1 of V) against columns
// 0..i, classical Gram-Schmidt in FP64 (an extension; the reference's
// recurrence has none): c = V^T v, v -= V c, s = ||v||, beta[i] *= s (the
// breakdown rule applied to the result), v /= s, or e1 after a breakdown.
// Skipped when beta[i] is already 0.  partial: kLanczosMaxIterations *
// kReduceBlocks doubles; c: kLanczosMaxIterations doubles; counter zero at
// rest.
void launch_lanczos_reorth(const double* V, int64_t ldv, int i, int64_t n, double* partial,
                           unsigned int* counter, double* c, LanczosScalars* sc, hipStream_t s);
// w -= beta[i-1] v_prev (if i > 0); alpha[i] = w . v_i
void launch_lanczos_alpha(double* w, const double* v_i, const double* v_prev, int64_t n, int i,
                          ReduceScratch rs, LanczosScalars* sc, hipStream_t s);
// w -= alpha[i] v_i; beta[i] = ||w|| with the breakdown rule
void launch_lanczos_beta(double* w, const double* v_i, int64_t n, int i, ReduceScratch rs,
                         LanczosScalars* sc, hipStream_t s);
// v_next = w / beta[i], or e1 after breakdown
void launch_lanczos_next(const double* w, double* v_next, int64_t n, int i,
                         const LanczosScalars* sc, hipStream_t s);
// s[0..m) = T_m^{1/2} e1 (FP64, one workgroup); work has tsqrt_scratch(m) doubles
constexpr int kTsqrtNodes = 512;
constexpr int64_t tsqrt_scratch(int m) { return (int64_t)m * (kTsqrtNodes + kTsqrtNodes / 64); }
// The same kernel on host arrays alpha (m) and beta (m - 1): out = T^{1/2} e1
// (synchronous; temporary device buffers).  For tests.
void tridiag_sqrt_selftest(const double* alpha, const double* beta, int m, double* out);
// Checks of iterations i0 .. i0+c-1 (c <= kCheckBatch), all skipped once the
// status says converged.  Launch 1: coef row j = T_{m_j}^{1/2} e1 with
// m_j = i0 + 1 + j, zero-padded to i0 + c entries (one workgroup per row,
// `work` holds c slices of tsqrt_scratch(i0 + c) doubles).  Launch 2: the
// estimates y_j = normz V[:, 0:m_j] coef_j in one pass over V, each compared
// with the one before it (the first with slot `old`): err_j = ||y_{j-1} -
// y_j|| / ||y_{j-1}|| (LanczosAlgorithm.cu:158-167, 202-222).  From i = 1 on
// the first err_j <= tol converges (y_j is the result and, as in the
// reference, the previous estimate becomes y_{j-1} - y_j), NaN fails;
// otherwise y_{c-1} becomes the previous estimate (cu:196).  Estimates live
// in a ring of nslots >= c + 1 slots of `ystride` doubles.  The status is also
// written to `mirror` (host-mapped, may be null).  Launch 3 forms that
// difference when this batch converged.
struct CheckArgs {
  const double* V;
  int64_t ldv;
  int i0, c;
  double* coef;  // kCheckBatch rows of kLanczosMaxIterations
  double* work;
  double* Y;
  int64_t ystride;
  int nslots;
  int64_t n;
  double tol;
  ReduceScratch rs;  // partials: 2 * kCheckBatch slots
  LanczosScalars* sc;
  LanczosStatus* mirror;
};
void launch_lanczos_checks(const CheckArgs& a, hipStream_t s);
// out[k] += pref * y[k]
void launch_axpy_accumulate(double* out, const double* y, double pref, int64_t n, hipStream_t s);
// dst = src + c * w
void launch_axpy_to(double* dst, const double* src, const double* w, double c, int64_t n,
                    hipStream_t s);
// out += pref * (a - b)
void launch_diff_accumulate(double* out, const double* a, const double* b, double pref,
                            int64_t n, hipStream_t s);
void launch_copy(double* dst, const double* src, int64_t n, hipStream_t s);
// zero fills by kernel (no memset nodes in code a caller may capture in a graph)
void launch_zero(double* p, int64_t n, hipStream_t s);
void launch_zero_u32(unsigned int* p, int64_t n, hipStream_t s);

// ---- independent Lanczos recurrences, one per system ---------------------------
// A block of the segmented vector kernels: particles [p0, p1) of system `sys`,
// i.e. vector entries [3 p0, 3 p1) and, for 6N vectors (torques after the
// forces), [3N + 3 p0, 3N + 3 p1).
struct SegBlock {
  int64_t p0, p1;
  int32_t sys;
  int32_t pad;
};
constexpr int kSegParticles = 512;  // particles per block
constexpr int kSegTsqrtGroups = 256; // concurrent tridiagonal solves (scratch slices)

struct SegLayout {
  const SegBlock* blocks;
  int nblocks;
  const int* first;       // [nsys + 1]: blocks of system s are first[s] .. first[s+1]-1
  int nsys;
  int64_t N3;             // 3N
  int angular;            // 1: vectors hold 6N entries
  double* partial;        // [nblocks * 2]
};

// Per-system state; alpha/beta rows have L entries (iteration capacity).
struct SegScalars {
  double* normz;   // [nsys]
  double* alpha;   // [nsys * L]
  double* beta;    // [nsys * L]
  double* err;     // [nsys]
  const int* active;  // [nsys] 1 while the system iterates
  const int* check;   // [nsys] 1: estimate this iteration, 2: store y as the previous
                      // estimate, 3: zero result (the system's noise is zero)
  int L;
};

void launch_seg_start(const double* z, double* v0, const SegLayout& g, const SegScalars& sc,
                      hipStream_t s);
void launch_seg_alpha(double* w, const double* vi, const double* vp, int i, const SegLayout& g,
                      const SegScalars& sc, hipStream_t s);
void launch_seg_beta(double* w, const double* vi, int i, const SegLayout& g, const SegScalars& sc,
                     hipStream_t s);
void launch_seg_next(const double* w, double* vn, int i, const SegLayout& g, const SegScalars& sc,
                     hipStream_t s);
// coef[s * L + k] = (T_m^{1/2} e1)_k for the listed systems; work holds
// kSegTsqrtGroups * tsqrt_scratch(m) doubles
void launch_seg_tridiag_sqrt(const int* list, int nlist, int m, const SegScalars& sc, double* coef,
                             double* work, hipStream_t s);
// for systems with check != 0: y = normz V[:, :m] coef_s ; old <- old - y ; err_s
void launch_seg_estimate(const double* V, int64_t ldv, int m, const double* coef, double* y,
                         double* old, const SegLayout& g, const SegScalars& sc, hipStream_t s);
// old <- y for systems with check == 2
void launch_seg_store(const double* y, double* old, const SegLayout& g, const SegScalars& sc,
                      hipStream_t s);
// y <- 0 for systems with check == 3
void launch_seg_clear(double* y, const SegLayout& g, const SegScalars& sc, hipStream_t s);
// launch_chunk_reduce over the particle blocks of systems with active != 0
// only (same summation order); rows of other systems are left as they are.
void launch_seg_chunk_reduce(const double* part, int nchunks, int64_t n, double* out,
                             const SegBlock* blocks, int nblocks, const int* active,
                             hipStream_t s);


}  // namespace lm
