// Host-side NBody solver: the MI355X counterpart of libmobility::Mobility +
// NBody (include/MobilityInterface/MobilityInterface.h:56-224,
// solvers/NBody/mobility.h:11-237).  One instance = one solver handle of the
// C ABI; calls on one instance are serialised by the caller (as with the
// reference, which is not thread-safe either).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <random>
#include <tuple>
#include <vector>

#include "coarse.hpp"
#include "kernels.hpp"
#include "lm_common.hpp"
#include "pair_math.hpp"
#include "precond.hpp"

namespace lm {

// Same numbering as libmobility::periodicity_mode (MobilityInterface.h:27-33).
enum Periodicity : int { kSingleWall = 0, kTwoWalls = 1, kOpen = 2, kPeriodic = 3, kUnspecified = 4 };
// Same numbering as nbody_rpy::algorithm (solvers/NBody/extra/interface.h:10).
enum Algorithm : int { kFast = 0, kNaive = 1, kBlock = 2, kAdvise = 3 };
enum PairPrecision : int { kPairFP64 = 0, kPairFP32 = 1 };

using LanczosCallback = void (*)(int iteration, double residual, void* user);

// Device memory of the allocation-free Mdot (one static array per GPU).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;

class NBodySolver {
 public:
  NBodySolver(int px, int py, int pz);
  ~NBodySolver();

  void setParameters(int algorithm, int nbatch, int nperbatch, bool has_wall_height,
                     double wall_height, double delta);
  void setPairPrecision(int p);
  // Pair arithmetic of the thermal drift's two products (default FP64
  // whatever the pair precision: the drift is a difference divided by the
  // step, and FP32 positions far from the system centre cannot resolve a
  // step of 1e-4 a).
  void setDriftPrecision(int p);
  void setStream(hipStream_t s) { stream_ = s; }
  hipStream_t stream() const { return stream_; }
  void initialize(double viscosity, double radius, bool include_angular, double tolerance,
                  uint64_t seed, LanczosCallback cb, void* cb_user);

  void setPositions(const double* pos, int64_t n3, bool dev);
  // Ragged systems (extension): offsets[0..nsys], offsets[nsys] = N; kinds[s]
  // (0 open, 1 wall) per system, nullptr = the solver's periodicity.
  void setSystems(const int64_t* offsets, int nsys, const int* kinds);
  // Persistent pair launches (a fixed grid draining a work queue).
  void setPersistent(int mode) { persistent_mode_ = mode; }
  // Lanczos iterations before "[Lanczos] Could not converge" (the public
  // iterationHardLimit member of lanczos::Solver, LanczosAlgorithm.h:49).
  void setLanczosIterationLimit(int limit);
  // 0: the reference's plain recurrence; 1: full FP64 reorthogonalisation of
  // every new basis vector (single-recurrence solves)
  void setLanczosReorthogonalization(int mode);
  int lanczosReorthogonalization() const { return reorth_; }
  int lanczosIterationLimit() const { return iter_limit_; }
  // Ensembles: one Lanczos recurrence per system (each with its own tolerance
  // test, check schedule, previous estimate and iteration count, as separate
  // solvers per system would have) instead of one over the whole
  // block-diagonal operator.
  void setLanczosPerSystem(bool on) { seg_mode_ = on; }
  bool lanczosPerSystem() const { return seg_mode_; }
  int numSystems() const { return (int)sys_off_host_.size() - 1; }
  // iterations of the last solve per system (per-system mode), else -1s
  void lastLanczosIterationsPerSystem(int* out, int nsys) const;

  // Several force vectors in one pass over the pairs (translational block):
  // U_r = M F_r, r < nrhs; F and U hold nrhs consecutive 3N vectors.
  void MdotMulti(const double* F, int nrhs, double* U, bool dev);
  void mdot_multi_dev(const double* F, int nrhs, double* U);
  // nrhs independent Lanczos square roots sharing every pass over the pairs:
  // U_r += prefactor M^{1/2} W_r.  noise: nrhs consecutive 3N vectors, or
  // nullptr to draw them as nrhs successive sqrtMdotW calls would.
  void sqrtMdotWMulti(double* U, int nrhs, double prefactor, bool dev, const double* noise);
  // iterations of the last multi-vector solve per vector (largest over systems)
  void lastLanczosIterationsMulti(int* out, int nrhs) const;
  int64_t numParticles() const { return N_; }
  bool includeAngular() const { return angular_; }

  void Mdot(const double* F, int64_t nF, const double* T, int64_t nT, double* U, double* W,
            bool dev);
  void sqrtMdotW(double* U, double* W, double prefactor, bool dev, const double* noise);
  // noise: the 3N random displacement W (nullptr: drawn as the reference does)
  void divM(double* U, double* W, double prefactor, bool dev, const double* noise = nullptr);
  void langevin(double dt, double kbt, const double* F, int64_t nF, const double* T, int64_t nT,
                double* U, double* W, bool dev);
  void clean();
  int lastLanczosIterations() const { return last_iters_; }
  // Pair-kernel tiling: targets per lane (1, 2, 4) and source chunks; 0 = auto.
  void setPairTiling(int tpl, int nch);
  void setAutotune(bool on) { autotune_ = on; }
  // The symmetric path reports its targets per lane and offset chunks.
  void lastPairTiling(int* tpl, int* nch) const {
    *tpl = last_sym_ ? sym_.tpl : items_tpl_;
    *nch = last_sym_ ? sym_.nchunks : items_nch_;
  }
  // Symmetric evaluation of one open or single-wall system, any block set
  // (-1 auto: from N = kSymAutoN; 0 off; 1 on).
  void setPairSymmetric(int mode);
  // Particles (in total) from which the automatic mode evaluates each
  // unordered pair once: from 16384 it was 1.10-1.37x faster than the ordered
  // evaluation (open UF and wall 6x6, FP32 and FP64); at 8192 equal or slower.
  static constexpr int64_t kSymAutoN = 16384;
  // Memory for the symmetric path's partial sums (0: the default, 4 GiB);
  // ring offsets are split into passes that fit it.
  void setPairSymmetricMemory(int64_t bytes);
  bool lastPairSymmetric() const { return last_sym_; }
  int checkConvergenceSteps() const { return check_steps_; }

  // Learned block preconditioner for sqrtMdotW (nullptr: plain Lanczos).  The
  // caller keeps the object alive while it is attached.
  void setPreconditioner(LearnedPreconditioner* p) { precond_ = p; pc_layout_ = 0; }
  // Training pairs from the current positions: X (ncenters x 48) and the FP64
  // local targets (ncenters x 9), host arrays.
  // kind 0: 3x3 targets (T: ncenters x 9); 1: 6x6 targets (ncenters x 36).
  // status (may be null): 0 where the local iteration converged; without it
  // a center that did not converge is an error.
  void precondDataset(int kind, int64_t ncenters, uint64_t seed, float* X, float* T, int* status);
  // The blocks G (d*d per particle, host) the attached network gives.
  void precondBlocks(double* G);
  // The K nearest neighbours of every particle within its system (N x K,
  // host; -1 past the system's other particles), nearest first.
  void precondNeighbours(int* nbr);
  // Coarse level of the sqrtMdotW preconditioner (per system): cells of about
  // `particles_per_cell` particles, Lanczos on G M G^T with G = D Gc (D the
  // learned blocks when attached, Gc the coarse whitening), result
  // G^{-1} (G M G^T)^{1/2} W.  0 switches it off.
  void setCoarseLevel(int particles_per_cell);
  int coarseLevel() const { return coarse_ppc_; }
  // Largest coarse dimension 3C per system (default CoarseLevel::kMaxDim);
  // denser systems get bigger cells.  The build holds three dense (3C)^2 FP64
  // matrices per system.
  void setCoarseMaxDim(int64_t max_dim);
  // Pair arithmetic of the coarse matrix Q^T M Q: FP32 by default (FP64 when
  // the FP32 matrix is indefinite), FP64 always with fp64.
  void setCoarseFP64(bool fp64) { coarse_fp64_ = fp64; }
  bool coarseUsedFP64() const { return coarse_.fp64_pairs(); }
  // Cells of the last coarse build (cell_of: N entries, host) and the grid;
  // returns the number of cells.
  int coarseCells(int* cell_of, int* grid) const;
  // L and L^{-1} of one system's coarse block (3C_s x 3C_s each, host).
  void coarseFactors(int system, double* L, double* X) const;
  int coarseSystemCells(int system) const;
  double coarseBuildMs() const { return coarse_.build_ms; }
  int64_t coarseBuilds() const { return coarse_builds_; }
  // Keep the preconditioner (learned blocks and coarse level) for `interval`
  // setPositions calls of the same systems before rebuilding it (default 1:
  // rebuilt for every new positions).  Results stay exact square roots of the
  // current M; stale parts only cost iterations.
  void setPreconditionerRefresh(int interval);

  // Device-pointer cores (no staging, no synchronisation beyond what the
  // algorithm needs).
  void mdot_dev(const double* F, const double* T, double* U, double* W);
  // Mdot that allocates no device memory (device pointers only, no wait):
  // the ordered evaluation with its work items and chunk slabs in a static
  // device array of kNoAllocScratchBytes per GPU (noalloc.hip).  One system
  // or equal batches; calls overlapping on one GPU race for that array.
  void mdotNoAlloc(const double* F, const double* T, double* U, double* W);
  void sqrtMdotW_dev(double* U, double* W, double prefactor, const double* noise);
  void divM_dev(double* U, double* W, double prefactor, const double* noise = nullptr);

 private:
  // State of a set of independent recurrences over a stacked layout: the
  // systems of an ensemble, or (vector, system) pairs of a multi-vector solve.
  struct SegSolve {
    DevBuf blocks, first, partial, state, flags, coef, work, old, y;
    HostBuf host;  // active, check (2 slots), list (int), err (double): one each per recurrence
    int nblocks = 0, L = 0, angular = -1;
    int64_t n = 0;
    size_t work_cols = 0;
    std::vector<int> check_steps, iters;
    std::vector<int64_t> off;  // particle offsets of the recurrences in the stacked layout
  };
  // A pair work list with the key it was built for.
  struct ItemList {
    DevBuf buf;
    int tpl = -1, nch = 0, count = 0;
    uint64_t structure = 0;
    int64_t N = -1;
  };
  void require_positions(const char* msg) const;
  void ensure_records();
  void build_systems(int64_t N, std::vector<int64_t>& off, bool& valid) const;
  void ensure_items(int tpl, int nch_hint);
  void choose_tiling(int blocks, bool fp32, const double* F, const double* T, double* U,
                     double* W, const void* rec, int& tpl, int& nch);
  void launch_pair_items(const PairItem* list, int count, int tpl, int nch, int blocks, bool fp32,
                         const void* rec, const double* F, const double* T, double* U, double* W);
  void launch_main_items(int blocks, bool fp32, const void* rec, const double* F, const double* T,
                         double* U, double* W) {
    launch_pair_items(items_.as<PairItem>(), items_count_, items_tpl_, items_nch_, blocks, fp32,
                      rec, F, T, U, W);
  }
  // prec: kPairFP64 / kPairFP32, or -1 for the solver's pair precision
  void pair_product(const void* rec64, const void* rec32, const double* F, const double* T,
                    double* U, double* W, int prec = -1);
  int lanczos_run(const double* z, double** result);
  // The convergence checks run on a second stream while the main stream goes
  // on with the next iterations; systems up to this many particles check
  // kCheckBatch iterations per launch, larger ones every iteration on its own
  // (pair workgroups launched after the device has decided return at once).
  static constexpr int64_t kLanczosLookaheadN = 32768;
  // Batches of checks in flight before the host waits for the oldest.
  static constexpr int kLanczosMaxLag = 1;
  int lanczos_run_seg(SegSolve& S, const std::vector<int64_t>& off, bool angular,
                      bool skip_converged_pairs,
                      const std::function<void(const double*, double*)>& product, const double* z,
                      double** result);
  void ensure_seg_layout(SegSolve& S, const std::vector<int64_t>& off, int angular, int64_t n);
  void ensure_basis(int64_t n, int64_t cols);
  // w = M v (or G M G^T v with the preconditioner) for the Lanczos recurrences
  void lanczos_product(const double* v, double* w);
  // G = D Gc, either factor optional: t = G^T v; w = G u (u overwritten);
  // out = G^{-1} y.  Vectors of length n >= 3N.
  bool pre_on() const { return precond_ != nullptr || coarse_ppc_ > 0; }
  void ensure_pre();
  bool pre_current(uint64_t layout, uint64_t structure, double a) const;
  void pre_forward(const double* v, double* t, int64_t n);
  void pre_back(double* u, double* w, int64_t n);
  void pre_inverse(const double* y, double* out, int64_t n);

  // configuration
  bool wall_ = false;
  int algorithm_ = kAdvise;
  int nbatch_ = -1, nperbatch_ = -1;
  double wall_height_ = 0.0;
  double delta_ = 1e-3;
  int pair_prec_ = kPairFP64;
  int drift_prec_ = kPairFP64;
  hipStream_t stream_ = nullptr;

  // initialize()
  bool initialized_ = false;
  double eta_ = 1.0, a_ = 1.0, t0_ = 0.0, rt0_ = 0.0, r0_ = 0.0;
  bool angular_ = false;
  double tol_ = 1e-4;
  LanczosCallback cb_ = nullptr;
  void* cb_user_ = nullptr;
  std::mt19937 rng_nbody_;    // mobility.h:34,110
  std::mt19937 rng_base_;     // MobilityInterface.h:64,108
  uint64_t lanczos_seed_ = 0;
  bool lanczos_engine_live_ = false;
  std::mt19937 lanczos_engine_;  // lanczos.h:32,37

  // positions
  int64_t N_ = 0;
  std::vector<int64_t> sys_off_host_;
  std::vector<int> sys_kind_host_;  // per system of the current layout
  std::vector<int> explicit_kinds_;
  bool explicit_systems_ = false;
  bool any_wall_ = false, any_open_ = true, kinds_dev_valid_ = false;
  DevBuf sys_kind_, work_counter_;
  int persistent_mode_ = -1;  // -1 auto (mixed ensembles), 0 off, 1 on
  int kind_code() const { return (any_wall_ && any_open_) ? 2 : (any_wall_ ? 1 : 0); }
  const int* kinds_dev() const { return kinds_dev_valid_ ? sys_kind_.as<int>() : nullptr; }
  DevBuf sys_off_, bbox_, shifted_, rec64_, rec32_;
  double rec_a_ = 0.0;  // radius the records were scaled with
  // pair work lists
  DevBuf items_;
  int items_tpl_ = -1, items_nch_ = 0, items_count_ = 0;
  bool items_persist_ = false;
  bool use_persistent() const {
    return persistent_mode_ == 1 || (persistent_mode_ == -1 && kind_code() == 2);
  }
  int64_t items_N_ = -1;
  // layout_id_: positions (and radius) changed; struct_id_: N, system
  // offsets or kinds changed (work lists depend on these only)
  uint64_t items_struct_ = 0, layout_id_ = 0, struct_id_ = 0;
  DevBuf part_u_, part_w_;
  // With defer_chunk_sum_ set, an ordered product over nch > 1 source chunks
  // leaves its partial sums in part_u_ / part_w_ (deferred_nch_ = nch) instead
  // of summing them into the outputs; deferred_nch_ = 0 when the outputs hold
  // the product.
  bool defer_chunk_sum_ = false;
  int deferred_nch_ = 0;
  // device flag that stops non-persistent pair workgroups (PairArgs::stop)
  const int* pair_stop_ = nullptr;
  int tile_tpl_ = 0, tile_nch_ = 0;  // user override (0 = auto)
  bool autotune_ = false;
  std::map<std::tuple<int64_t, int, int, int, int>, std::pair<int, int>> tuned_;
  // staging for host arrays
  DevBuf st_pos_, st_f_, st_t_, st_u_, st_w_;
  // Lanczos
  DevBuf V_, w_, coef_, noise_, red_partial_, red_counter_, scalars_, tsq_work_;
  int64_t V_cols_ = 0, V_n_ = 0;
  // Estimates of the convergence checks: a ring of ring_slots_ vectors of
  // ring_stride_ doubles; old_host_ names the slot holding the previous
  // estimate (oldBz), as the device status last reported it.
  DevBuf est_ring_;
  int ring_slots_ = 0;
  int64_t ring_stride_ = 0;
  int old_host_ = 0;
  bool est_valid_ = false;
  int64_t old_n_ = 0;
  // Small systems run the checks (tridiagonal square roots and estimates) on
  // a second stream of default priority, overlapped with the next
  // iterations; the estimate kernel writes each batch's status into a
  // host-mapped slot, read once the batch's event has completed.  The checks
  // have their own reduction scratch.
  static constexpr int kLzSlots = 8;
  HostBuf lz_status_;       // host-mapped LanczosStatus, one per slot
  hipEvent_t lz_ev_[kLzSlots] = {};
  hipEvent_t lz_vec_ev_ = nullptr;  // main stream: iteration's vectors written
  hipStream_t chk_stream_ = nullptr;
  DevBuf chk_partial_, chk_counter_;
  int check_steps_ = 3;          // LanczosAlgorithm.cu:171
  int iter_limit_ = 200;         // LanczosAlgorithm.h:49
  int reorth_ = 0;
  DevBuf reorth_partial_, reorth_counter_, reorth_c_;
  int last_iters_ = 0;
  // per-system Lanczos
  bool seg_mode_ = false;
  const int* pair_active_ = nullptr;  // per-system flags seen by the pair kernel
  // Pair work of the systems still iterating (rebuilt when a system
  // converges; count 0: the full list) and the particle blocks the chunk sum
  // is restricted to.
  ItemList seg_items_;
  HostBuf seg_items_host_;
  const SegBlock* seg_blocks_ = nullptr;
  int seg_nblocks_ = 0;
  void build_active_items(const int* act);
  // symmetric evaluation (open or wall systems, any block set, ensembles)
  struct SymPass {
    int item0 = 0, nitems = 0, oa = 1, ob = 1, nch = 0, diag = 0;
  };
  struct SymPlan {
    DevBuf items, sys, rows, active_items;
    std::vector<SymItem> host_items;  // all passes, in launch order
    int tpl = 0, nchunks = 0, hmax = 0, nitems = 0, nrows = 0, ang = -1;
    int64_t N = -1, budget = 0;
    uint64_t structure = 0;
    std::vector<SymPass> passes, active_passes;
    size_t max_slabs = 0;
    bool active_valid = false;  // active_items / active_passes hold the iterating systems
  };
  static constexpr int64_t kDefaultSymBudget = (int64_t)4 << 30;
  int64_t sym_budget_ = kDefaultSymBudget;
  SymPlan sym_;
  HostBuf sym_active_host_;
  int sym_mode_ = -1;
  bool last_sym_ = false;
  // (N, fp32, kind + 4 * systems, blocks) -> (targets per lane, offset chunks)
  std::map<std::tuple<int64_t, int, int, int>, std::pair<int, int>> sym_tuned_;
  void sym_shape(bool special, int tpl_req, int nch_req, int& tpl, int& nch, int& ntiles,
                 int& hmax) const;
  bool use_sym(int blocks) const;
  void ensure_sym_plan(int tpl, int nch, int h, bool ang);
  void build_sym_active(const int* act);
  void run_sym(int tpl, int nch, int h, int blocks, bool fp32, const void* rec, const double* F,
               const double* T, double* U, double* W);
  void launch_sym(int blocks, bool fp32, const void* rec, const double* F, const double* T,
                  double* U, double* W);
  std::map<std::tuple<int, int, int>, std::vector<PairItem>> host_items_;
  uint64_t host_items_struct_ = 0;
  int64_t host_items_N_ = -1;
  const std::vector<PairItem>& host_items(int tpl, int nch);
  SegSolve seg_sys_, seg_multi_;
  DevBuf seg_w_;
  // multi-vector products
  ItemList multi_items_[2];  // 4 vectors (one target per lane), 2 vectors (two)
  DevBuf part_multi_, multi_f_, multi_u_, multi_noise_, multi_t_, multi_w_;
  int multi_nrhs_ = 0, multi_per_ = 1;
  // act (host, per system; nullptr: all): only systems with act[s] != 0
  int resolve_nch(int tpl, int nch_hint, const int* act = nullptr) const;
  void build_items(int tpl, int nch, bool persist, std::vector<PairItem>& items,
                   const int* act = nullptr) const;
  void ensure_multi_items(ItemList& L, int tpl);
  void multi_product(const double* v, int nrhs, double* w);
  // learned preconditioner
  void ensure_knn();
  void ensure_precond_blocks();
  LearnedPreconditioner* precond_ = nullptr;
  DevBuf pc_feat_, pc_nbr_, pc_out_, pc_G_, pc_Gi_, pc_t_, pc_u_;
  // cell grid of the neighbour search
  std::vector<double> knn_xyz_;
  std::vector<KnnSys> knn_sys_host_;
  std::vector<int> knn_cell_host_, knn_perm_host_;
  std::vector<int64_t> knn_start_host_;
  DevBuf knn_sys_, knn_cell_, knn_perm_, knn_start_;
  uint64_t knn_layout_ = 0, pc_layout_ = 0, pc_struct_ = 0, pc_version_ = 0;
  double pc_a_ = 0.0;
  int pc_refresh_ = 1;
  // coarse level
  void ensure_coarse();
  CoarseLevel coarse_;
  int coarse_ppc_ = 0;
  int64_t coarse_max_dim_ = CoarseLevel::kMaxDim;
  bool coarse_fp64_ = false;
  int64_t coarse_builds_ = 0;
  // RFD
  DevBuf rfd_noise_, rfd_pos_, rfd_rec64_, rfd_rec32_, rfd_mp_, rfd_mm_, rfd_wp_, rfd_wm_;
};

}  // namespace lm
