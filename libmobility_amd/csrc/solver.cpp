// Host orchestration of the NBody solver.  Reference behaviour followed (paths
// relative to the reference root):
//   construction            solvers/NBody/mobility.h:37-47
//   setParameters           solvers/NBody/mobility.h:78-97
//   initialize + seeds      solvers/NBody/mobility.h:99-114, MobilityInterface.h:102-114
//   setPositions            solvers/NBody/mobility.h:116-162
//   Mdot + block selection  solvers/NBody/mobility.h:166-182, extra/NbodyRPY.cu:254-287
//   sqrtMdotW               MobilityInterface.h:132-188, lanczos.h:43-58,
//                           include/lanczos/LanczosAlgorithm.cu:173-232
//   LangevinVelocities      MobilityInterface.h:191-206
//   divM (RFD)              solvers/NBody/mobility.h:184-236,
//                           include/MobilityInterface/random_finite_differences.h:30-63
#include "solver.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>

namespace lm {


NBodySolver::NBodySolver(int px, int py, int pz) {
  if (px < 0 || px > 4 || py < 0 || py > 4 || pz < 0 || pz > 4)
    throw Error("[libMobility] Invalid periodicity");
  if (px != kOpen || py != kOpen) throw Error("[Mobility] NBody must be open in the plane");
  if (pz == kOpen)
    wall_ = false;
  else if (pz == kSingleWall)
    wall_ = true;
  else
    throw Error("[Mobility] Invalid periodicity");
}

NBodySolver::~NBodySolver() {
  if (chk_stream_) (void)hipStreamSynchronize(chk_stream_);
  for (auto& e : lz_ev_)
    if (e) (void)hipEventDestroy(e);
  if (lz_vec_ev_) (void)hipEventDestroy(lz_vec_ev_);
  if (chk_stream_) (void)hipStreamDestroy(chk_stream_);
}

void NBodySolver::setParameters(int algorithm, int nbatch, int nperbatch, bool has_wall_height,
                                double wall_height, double delta) {
  if (algorithm < kFast || algorithm > kAdvise) throw Error("Invalid algorithm selected");
  algorithm_ = algorithm;
  nbatch_ = nbatch;
  nperbatch_ = nperbatch;
  if (wall_) {
    if (has_wall_height)
      wall_height_ = wall_height;
    else
      throw Error(
          "[Mobility] Wall height parameter is required for a bottom wall. If you want to use "
          "a wall, set the wallHeight parameter.");
  } else if (has_wall_height) {
    throw Error(
        "[Mobility] Wall height parameter is only valid for bottom wall. If you want to use a "
        "wall, set periodicityZ to single_wall in the configuration.");
  }
  delta_ = delta;
  items_tpl_ = -1;  // algorithm may change the work decomposition
}

void NBodySolver::setPairPrecision(int p) {
  if (p != kPairFP64 && p != kPairFP32) throw Error("[libmobility_amd] Invalid pair precision");
  pair_prec_ = p;
  items_tpl_ = -1;
}

void NBodySolver::setDriftPrecision(int p) {
  if (p != kPairFP64 && p != kPairFP32) throw Error("[libmobility_amd] Invalid pair precision");
  drift_prec_ = p;
}

void NBodySolver::initialize(double viscosity, double radius, bool include_angular,
                             double tolerance, uint64_t seed, LanczosCallback cb, void* cb_user) {
  if (initialized_) clean();  // MobilityInterface.h:104-105
  eta_ = viscosity;
  a_ = radius;
  t0_ = 1.0 / (6.0 * M_PI * viscosity * radius);
  rt0_ = 1.0 / (8.0 * M_PI * viscosity * radius * radius);
  r0_ = 1.0 / (8.0 * M_PI * viscosity * radius * radius * radius);
  // Seed chain of mobility.h:107-113 and MobilityInterface.h:106-110.
  if (seed == 0) seed = std::random_device()();
  rng_nbody_ = std::mt19937(seed);
  uint64_t base_seed = rng_nbody_();
  if (base_seed == 0) base_seed = std::random_device()();
  rng_base_ = std::mt19937(base_seed);
  lanczos_seed_ = rng_base_();
  tol_ = tolerance;
  cb_ = cb;
  cb_user_ = cb_user;
  angular_ = include_angular;
  initialized_ = true;
  items_tpl_ = -1;
}

void NBodySolver::clean() {
  // MobilityInterface.h:220-223: drop the Lanczos solver and its workspace.
  lanczos_engine_live_ = false;
  check_steps_ = 3;
  old_n_ = 0;
  V_.release();
  w_.release();
  est_ring_.release();
  ring_slots_ = 0;
  ring_stride_ = 0;
  est_valid_ = false;
  old_host_ = 0;
  V_cols_ = V_n_ = 0;
}

void NBodySolver::require_positions(const char* msg) const {
  if (N_ <= 0) throw Error(msg);
}

void NBodySolver::build_systems(int64_t N, std::vector<int64_t>& off, bool& valid) const {
  off.clear();
  if (explicit_systems_ && !sys_off_host_.empty() && sys_off_host_.back() == N) {
    off = sys_off_host_;
    valid = true;
    return;
  }
  const int64_t nb = nbatch_ < 0 ? 1 : nbatch_;
  const int64_t npb = nperbatch_ < 0 ? N : nperbatch_;
  valid = (nb * npb == N) && nb > 0;
  if (!valid || N == 0) {
    off = {0, N};
    return;
  }
  off.resize(nb + 1);
  for (int64_t b = 0; b <= nb; ++b) off[b] = b * npb;
}

void NBodySolver::setSystems(const int64_t* offsets, int nsys, const int* kinds) {
  if (nsys < 1) throw Error("[libmobility_amd] need at least one system");
  if (offsets[0] != 0) throw Error("[libmobility_amd] system offsets must start at 0");
  for (int s = 0; s < nsys; ++s)
    if (offsets[s + 1] < offsets[s]) throw Error("[libmobility_amd] system offsets must be non-decreasing");
  sys_off_host_.assign(offsets, offsets + nsys + 1);
  explicit_kinds_.clear();
  if (kinds) {
    for (int s = 0; s < nsys; ++s)
      if (kinds[s] != 0 && kinds[s] != 1) throw Error("[libmobility_amd] system kind must be 0 (open) or 1 (wall)");
    explicit_kinds_.assign(kinds, kinds + nsys);
  }
  explicit_systems_ = true;
  ++struct_id_;
  items_tpl_ = -1;
}

void NBodySolver::setPositions(const double* pos, int64_t n3, bool dev) {
  if (n3 % 3 != 0)
    throw Error(
        "[libMobility] The input arrays for positions, forces, and torques must each have "
        "total size 3*N.");
  const int64_t N = n3 / 3;
  std::vector<int64_t> off;
  bool valid = true;
  build_systems(N, off, valid);
  if (explicit_systems_ && off.back() != N)
    throw Error("[libmobility_amd] system offsets do not cover the positions");
  const int nsys = (int)off.size() - 1;
  std::vector<int> kinds(nsys, wall_ ? 1 : 0);
  const bool use_kinds = explicit_systems_ && !explicit_kinds_.empty() &&
                         (int)explicit_kinds_.size() == nsys && valid;
  if (use_kinds) kinds = explicit_kinds_;
  const int* kinds_d = nullptr;
  if (N > 0 && use_kinds) {
    sys_kind_.reserve(sizeof(int) * nsys);
    LM_HIP(hipMemcpyAsync(sys_kind_.as<int>(), kinds.data(), sizeof(int) * nsys,
                          hipMemcpyHostToDevice, stream_));
    kinds_d = sys_kind_.as<int>();
  }
  bool any_wall = false, any_open = false;
  for (int k : kinds) (k ? any_wall : any_open) = true;
  if (N > 0) {
    const double* src = pos;
    if (!dev) {
      st_pos_.reserve(sizeof(double) * n3);
      LM_HIP(hipMemcpyAsync(st_pos_.as<double>(), pos, sizeof(double) * n3,
                            hipMemcpyHostToDevice, stream_));
      src = st_pos_.as<double>();
    }
    sys_off_.reserve(sizeof(int64_t) * off.size());
    LM_HIP(hipMemcpyAsync(sys_off_.as<int64_t>(), off.data(), sizeof(int64_t) * off.size(),
                          hipMemcpyHostToDevice, stream_));
    bbox_.reserve(sizeof(double) * 6 * nsys);
    const double shift = wall_height_;
    launch_system_bbox(src, sys_off_.as<int64_t>(), nsys, shift, kinds_d, wall_ ? 1 : 0,
                       bbox_.as<double>(), stream_);
    if (any_wall) {
      std::vector<double> bb(6 * (size_t)nsys);
      LM_HIP(hipMemcpyAsync(bb.data(), bbox_.as<double>(), sizeof(double) * bb.size(),
                            hipMemcpyDeviceToHost, stream_));
      LM_HIP(hipStreamSynchronize(stream_));
      double minz = 1e300;
      for (int s = 0; s < nsys; ++s)
        if (off[s + 1] > off[s] && kinds[s]) minz = std::min(minz, bb[6 * s + 2]);
      if (minz < 0)
        throw Error(
            "[Mobility] The center of a particle has fallen below the wall. Check your "
            "configuration and forces to ensure particle centers stay above the wall.");
    }
    shifted_.reserve(sizeof(double) * n3);
    rec64_.reserve(sizeof(double) * 4 * N);
    rec32_.reserve(sizeof(float) * 4 * N);
    launch_make_records(src, N, sys_off_.as<int64_t>(), nsys, bbox_.as<double>(), kinds_d,
                        wall_ ? 1 : 0, shift, a_, shifted_.as<double>(), rec64_.as<void>(),
                        rec32_.as<void>(), stream_);
    rec_a_ = a_;
  }
  // Work lists depend on N, the system offsets and kinds only: a new
  // configuration of the same systems (every step of a time loop) keeps them.
  const bool same_structure =
      N == N_ && kinds == sys_kind_host_ && (explicit_systems_ || off == sys_off_host_);
  N_ = N;
  sys_kind_host_ = kinds;
  kinds_dev_valid_ = kinds_d != nullptr;
  any_wall_ = any_wall;
  any_open_ = any_open;
  if (!explicit_systems_) sys_off_host_ = off;
  ++layout_id_;
  if (!same_structure) {
    ++struct_id_;
    items_tpl_ = -1;
  }
  if (!valid)
    throw Error(
        "[Mobility] Invalid batch parameters for NBody. If in doubt, use the defaults.");
  if (!dev || wall_) LM_HIP(hipStreamSynchronize(stream_));
}

// Work decomposition: target tiles of 256*tpl particles x source chunks.
// Source chunks per target block: the hint (auto: about 8192 items, ~32
// workgroups per CU for a short tail on every XCD), clamped to [1, 32] and to
// one chunk per 64 sources of the largest system (a small system still
// spreads over several workgroups per target tile).
int NBodySolver::resolve_nch(int tpl, int nch_hint, const int* act) const {
  const std::vector<int64_t>& off = sys_off_host_;
  const int nsys = (int)off.size() - 1;
  const int64_t bt = (int64_t)kPairThreads * tpl;
  int64_t tiles = 0, maxn = 0;
  for (int s = 0; s < nsys; ++s) {
    if (act && !act[s]) continue;
    const int64_t ns = off[s + 1] - off[s];
    tiles += (ns + bt - 1) / bt;
    maxn = std::max(maxn, ns);
  }
  int nch = nch_hint;
  if (nch <= 0) {
    const int64_t want = 8192;
    nch = (int)std::max<int64_t>(1, (want + std::max<int64_t>(tiles, 1) - 1) / std::max<int64_t>(tiles, 1));
  }
  const int maxch = (int)std::max<int64_t>(1, (maxn + kChunkGrain - 1) / kChunkGrain);
  return std::max(1, std::min({nch, 32, maxch}));
}

// Work items: per system, target blocks of 256 * tpl targets x nch source chunks.
void NBodySolver::build_items(int tpl, int nch, bool persist, std::vector<PairItem>& items,
                              const int* act) const {
  const std::vector<int64_t>& off = sys_off_host_;
  const int nsys = (int)off.size() - 1;
  const int64_t bt = (int64_t)kPairThreads * tpl;
  items.clear();
  for (int s = 0; s < nsys; ++s) {
    if (act && !act[s]) continue;
    const int64_t b = off[s], e = off[s + 1], ns = e - b;
    if (ns <= 0) continue;
    const int kind = s < (int)sys_kind_host_.size() ? sys_kind_host_[s] : (wall_ ? 1 : 0);
    int64_t cs = (ns + nch - 1) / nch;
    cs = ((cs + kChunkGrain - 1) / kChunkGrain) * kChunkGrain;
    for (int64_t t0 = b; t0 < e; t0 += bt) {
      for (int c = 0; c < nch; ++c) {
        PairItem it;
        it.tgt0 = t0;
        it.ntgt = (int32_t)std::min<int64_t>(bt, e - t0);
        it.src0 = std::min<int64_t>(b + (int64_t)c * cs, e);
        it.src1 = std::min<int64_t>(b + (int64_t)(c + 1) * cs, e);
        it.chunk = c;
        it.wall = kind;
        it.sys = s;
        items.push_back(it);
      }
    }
  }
  if (persist) {
    // work queue: longest items first (LPT order) for a balanced drain
    auto cost = [](const PairItem& q) {
      return (double)q.ntgt * (double)(q.src1 - q.src0) * (q.wall ? 5.0 : 1.0);
    };
    std::stable_sort(items.begin(), items.end(),
                     [&](const PairItem& x, const PairItem& y) { return cost(x) > cost(y); });
  }
}

void NBodySolver::ensure_items(int tpl, int nch_hint) {
  const int nch = resolve_nch(tpl, nch_hint);
  if (items_tpl_ == tpl && items_nch_ == nch && items_N_ == N_ && items_struct_ == struct_id_ &&
      items_persist_ == use_persistent())
    return;
  std::vector<PairItem> items;
  build_items(tpl, nch, use_persistent(), items);
  items_.reserve(sizeof(PairItem) * std::max<size_t>(items.size(), 1));
  LM_HIP(hipMemcpyAsync(items_.as<PairItem>(), items.data(), sizeof(PairItem) * items.size(),
                        hipMemcpyHostToDevice, stream_));
  LM_HIP(hipStreamSynchronize(stream_));  // the host vector goes out of scope
  items_tpl_ = tpl;
  items_nch_ = nch;
  items_N_ = N_;
  items_struct_ = struct_id_;
  items_persist_ = use_persistent();
  items_count_ = (int)items.size();
}

void NBodySolver::ensure_multi_items(ItemList& L, int tpl) {
  if (L.tpl == tpl && L.N == N_ && L.structure == struct_id_) return;
  std::vector<PairItem> items;
  const int nch = resolve_nch(tpl, 0);
  build_items(tpl, nch, false, items);
  L.buf.reserve(sizeof(PairItem) * std::max<size_t>(items.size(), 1));
  LM_HIP(hipMemcpyAsync(L.buf.as<PairItem>(), items.data(), sizeof(PairItem) * items.size(),
                        hipMemcpyHostToDevice, stream_));
  LM_HIP(hipStreamSynchronize(stream_));
  L.tpl = tpl;
  L.nch = nch;
  L.count = (int)items.size();
  L.N = N_;
  L.structure = struct_id_;
}

void NBodySolver::setPairSymmetric(int mode) {
  if (mode < -1 || mode > 1)
    throw Error("[libmobility_amd] symmetric pair mode must be -1 (auto), 0 (off) or 1 (on)");
  sym_mode_ = mode;
}

void NBodySolver::setPairSymmetricMemory(int64_t bytes) {
  if (bytes < 0) throw Error("[libmobility_amd] the symmetric slab budget must be >= 0");
  sym_budget_ = bytes == 0 ? kDefaultSymBudget : bytes;
}

// Shape of a symmetric tiling request (0 = default rule): targets per lane
// (open translational block of one system: 2 from N = 65536, else 1; other
// cases 1, at most 2) and offset chunks per tile (about 8192 items), at most
// the longest ring.  ntiles is the total over systems, hmax the longest ring.
void NBodySolver::sym_shape(bool special, int tpl_req, int nch_req, int& tpl, int& nch,
                            int& ntiles, int& hmax) const {
  const int nsys = (int)sys_off_host_.size() - 1;
  tpl = tpl_req != 0 ? tpl_req : (special && nsys == 1 && N_ >= 65536 ? 2 : 1);
  if (!(special && nsys == 1)) tpl = std::min(tpl, 2);  // 4 only for one open system
  const int64_t bt = (int64_t)kPairThreads * tpl;
  ntiles = 0;
  hmax = 0;
  for (int s = 0; s < nsys; ++s) {
    const int nt = (int)((sys_off_host_[s + 1] - sys_off_host_[s] + bt - 1) / bt);
    ntiles += nt;
    hmax = std::max(hmax, nt / 2);
  }
  nch = nch_req != 0 ? nch_req : (8192 + std::max(ntiles, 1) - 1) / std::max(ntiles, 1);
  nch = std::min(nch, hmax);
  if (hmax > 0) nch = std::max(nch, 1);
}

bool NBodySolver::use_sym(int blocks) const {
  if (sym_mode_ == 0 || N_ <= 0) return false;
  if (blocks != kUF && blocks != (kUF | kWF) && blocks != (kUT | kWT) &&
      blocks != (kUF | kUT | kWT | kWF))
    return false;
  return sym_mode_ == 1 || N_ >= kSymAutoN;
}

// Items: per pass, for every system and tile, the system's chunks of the
// pass's offsets; the diagonal tiles in the first pass.  Ensembles order each
// pass longest first (wall pairs weigh 5).  Ring offsets are split into
// passes whose partial-sum slabs fit the memory budget: per pass the target
// slabs (one per chunk, plus the diagonal in the first) and one source slab
// per offset, reduced into the output in that order, the first pass
// assigning and later ones adding.
void NBodySolver::ensure_sym_plan(int tpl, int nch, int h, bool ang) {
  if (sym_.N == N_ && sym_.structure == struct_id_ && sym_.tpl == tpl && sym_.nchunks == nch &&
      sym_.ang == (ang ? 1 : 0) && sym_.budget == sym_budget_)
    return;
  const int nsys = (int)sys_off_host_.size() - 1;
  const int64_t bt = (int64_t)kPairThreads * tpl;
  std::vector<SymSys> systems(nsys);
  std::vector<SymRows> rows;
  for (int s = 0; s < nsys; ++s) {
    const int64_t b = sys_off_host_[s], e = sys_off_host_[s + 1];
    SymSys& S = systems[s];
    S.off = b;
    S.n = (int32_t)(e - b);
    S.ntiles = (int32_t)((e - b + bt - 1) / bt);
    S.hmax = S.ntiles / 2;
    S.wall = s < (int)sys_kind_host_.size() ? sys_kind_host_[s] : (wall_ ? 1 : 0);
    for (int64_t p0 = b; p0 < e; p0 += kPairThreads)
      rows.push_back(SymRows{p0, (int32_t)std::min<int64_t>(kPairThreads, e - p0), s});
  }
  const double unit = 8.0 * 3.0 * (double)N_ * (ang ? 2.0 : 1.0);  // bytes per slab (U and W)
  const double per_offset = 1.0 + (h > 0 ? (double)nch / h : 0.0);
  const int span = std::max(1, (int)(((double)sym_budget_ / unit - 2.0) / per_offset));
  std::vector<SymItem> items;
  sym_.passes.clear();
  sym_.max_slabs = 0;
  int oa = 1;
  do {
    const int ob = std::min(h + 1, oa + span);
    const bool first = oa == 1;
    const int width = ob - oa;
    SymPass ps;
    ps.item0 = (int)items.size();
    ps.oa = oa;
    ps.ob = ob;
    ps.nch = width > 0 ? std::max(1, std::min(width, (int)(((int64_t)nch * width + h - 1) / h))) : 0;
    ps.diag = first ? 1 : 0;
    std::vector<std::pair<double, SymItem>> pass_items;
    for (int s = 0; s < nsys; ++s) {
      const SymSys& S = systems[s];
      const double wgt = S.wall ? 5.0 : 1.0;
      const int ws = std::min(ob, S.hmax + 1) - oa;  // this system's offsets in the pass
      const int cs = ws > 0 ? std::min(ps.nch, ws) : 0;
      for (int I = 0; I < S.ntiles; ++I)
        for (int c = 0; c < cs; ++c) {
          SymItem it;
          it.sys = s;
          it.tile = I;
          it.o0 = oa + (int)((int64_t)c * ws / cs);
          it.o1 = oa + (int)((int64_t)(c + 1) * ws / cs);
          it.chunk = c;
          pass_items.emplace_back(wgt * (it.o1 - it.o0), it);
        }
      if (first)
        for (int I = 0; I < S.ntiles; ++I) {
          const double nt = (double)std::min<int64_t>(bt, S.n - (int64_t)I * bt) / bt;
          pass_items.emplace_back(0.5 * wgt * nt * nt, SymItem{s, I, 0, 0, ps.nch});
        }
    }
    if (nsys > 1)
      std::stable_sort(pass_items.begin(), pass_items.end(),
                       [](const std::pair<double, SymItem>& x, const std::pair<double, SymItem>& y) {
                         return x.first > y.first;
                       });
    for (const auto& pi : pass_items) items.push_back(pi.second);
    ps.nitems = (int)items.size() - ps.item0;
    sym_.max_slabs = std::max(sym_.max_slabs, (size_t)(ps.nch + ps.diag + width));
    sym_.passes.push_back(ps);
    oa = ob;
  } while (oa <= h);
  sym_.items.reserve(sizeof(SymItem) * std::max<size_t>(items.size(), 1));
  sym_.sys.reserve(sizeof(SymSys) * systems.size());
  sym_.rows.reserve(sizeof(SymRows) * rows.size());
  LM_HIP(hipMemcpyAsync(sym_.items.as<void>(), items.data(), sizeof(SymItem) * items.size(),
                        hipMemcpyHostToDevice, stream_));
  LM_HIP(hipMemcpyAsync(sym_.sys.as<void>(), systems.data(), sizeof(SymSys) * systems.size(),
                        hipMemcpyHostToDevice, stream_));
  LM_HIP(hipMemcpyAsync(sym_.rows.as<void>(), rows.data(), sizeof(SymRows) * rows.size(),
                        hipMemcpyHostToDevice, stream_));
  LM_HIP(hipStreamSynchronize(stream_));  // the host vectors go out of scope (items are kept)
  sym_.host_items = std::move(items);
  sym_.N = N_;
  sym_.structure = struct_id_;
  sym_.tpl = tpl;
  sym_.nchunks = nch;
  sym_.hmax = h;
  sym_.ang = ang ? 1 : 0;
  sym_.budget = sym_budget_;
  sym_.nitems = (int)sym_.host_items.size();
  sym_.nrows = (int)rows.size();
  sym_.active_valid = false;
}

// Per-system Lanczos: the items of the systems still iterating, pass by pass
// (filtered from the plan's host list, already in launch order).
void NBodySolver::build_sym_active(const int* act) {
  sym_.active_valid = false;
  if (sym_.N != N_ || sym_.structure != struct_id_ || sym_.host_items.empty()) return;
  sym_.active_passes = sym_.passes;
  sym_active_host_.reserve(sizeof(SymItem) * sym_.host_items.size());
  SymItem* dst = sym_active_host_.as<SymItem>();
  int count = 0;
  for (size_t p = 0; p < sym_.passes.size(); ++p) {
    const SymPass& ps = sym_.passes[p];
    sym_.active_passes[p].item0 = count;
    for (int k = ps.item0; k < ps.item0 + ps.nitems; ++k)
      if (act[sym_.host_items[k].sys]) dst[count++] = sym_.host_items[k];
    sym_.active_passes[p].nitems = count - sym_.active_passes[p].item0;
  }
  // the previous upload from the pinned buffer has completed: the caller
  // synchronised the stream after it
  sym_.active_items.reserve(sizeof(SymItem) * std::max(count, 1));
  if (count > 0)
    LM_HIP(hipMemcpyAsync(sym_.active_items.as<void>(), dst, sizeof(SymItem) * count,
                          hipMemcpyHostToDevice, stream_));
  sym_.active_valid = true;
}

void NBodySolver::run_sym(int tpl, int nch, int h, int blocks, bool fp32, const void* rec,
                          const double* F, const double* T, double* U, double* W) {
  const bool ang = (blocks & (kWT | kWF)) != 0;
  ensure_sym_plan(tpl, nch, h, ang);
  const bool compact = pair_active_ != nullptr && sym_.active_valid;
  const std::vector<SymPass>& passes = compact ? sym_.active_passes : sym_.passes;
  const SymItem* items = compact ? sym_.active_items.as<SymItem>() : sym_.items.as<SymItem>();
  const int64_t n3 = 3 * N_;
  part_u_.reserve(sizeof(double) * n3 * sym_.max_slabs);
  if (ang) part_w_.reserve(sizeof(double) * n3 * sym_.max_slabs);
  const int nsys = (int)sys_off_host_.size() - 1;
  for (size_t p = 0; p < passes.size(); ++p) {
    const SymPass& ps = passes[p];
    const int nslab_t = ps.nch + ps.diag;
    double* slab_t = part_u_.as<double>();
    double* slab_s = slab_t + (int64_t)nslab_t * n3;
    double* slab_tw = ang ? part_w_.as<double>() : nullptr;
    double* slab_sw = ang ? slab_tw + (int64_t)nslab_t * n3 : nullptr;
    SymArgs a{};
    a.fp32 = fp32 ? 1 : 0;
    a.tpl = tpl;
    a.wall = nsys > 1 ? 2 : kind_code();
    a.blocks = blocks;
    a.pos = rec;
    a.F = F;
    a.T = T;
    a.slab_t = slab_t;
    a.slab_s = slab_s;
    a.slab_tw = slab_tw;
    a.slab_sw = slab_sw;
    a.n3 = n3;
    a.obase = ps.oa;
    a.sys = sym_.sys.as<SymSys>();
    a.items = items + ps.item0;
    a.nitems = ps.nitems;
    a.active = pair_active_;
    a.stop = pair_stop_;
    a.t0 = t0_;
    a.rt0 = rt0_;
    a.r0 = r0_;
    launch_pair_sym(a, stream_);
    SymReduceArgs r{};
    r.slab_t = slab_t;
    r.slab_s = slab_s;
    r.n3 = n3;
    r.nch = ps.nch;
    r.o0 = ps.oa;
    r.o1 = ps.ob;
    r.diag = ps.diag;
    r.accumulate = p > 0 ? 1 : 0;
    r.tpl = tpl;
    r.sys = sym_.sys.as<SymSys>();
    r.rows = sym_.rows.as<SymRows>();
    r.nrows = sym_.nrows;
    r.active = pair_active_;
    r.out = U;
    launch_sym_reduce(r, stream_);
    if (ang) {
      r.slab_t = slab_tw;
      r.slab_s = slab_sw;
      r.out = W;
      launch_sym_reduce(r, stream_);
    }
  }
}

void NBodySolver::launch_sym(int blocks, bool fp32, const void* rec, const double* F,
                             const double* T, double* U, double* W) {
  const int nsys = (int)sys_off_host_.size() - 1;
  const bool special = kind_code() == 0 && blocks == kUF;
  const auto key = std::make_tuple(N_, fp32 ? 1 : 0, kind_code() + 4 * nsys, blocks);
  double pairs = 0.0;
  for (int s = 0; s < nsys; ++s) {
    const double n = (double)(sys_off_host_[s + 1] - sys_off_host_[s]);
    pairs += n * n;
  }
  // time every candidate on the live inputs (outputs are overwritten by the
  // real launch that follows); not where one product takes seconds
  if (autotune_ && tile_tpl_ == 0 && tile_nch_ == 0 && pairs <= 1e11 &&
      sym_tuned_.find(key) == sym_tuned_.end()) {
    hipEvent_t e0, e1;
    LM_HIP(hipEventCreate(&e0));
    LM_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    std::pair<int, int> pick{0, 0};
    std::vector<std::tuple<int, int, int>> seen;
    for (int tr : {1, 2, 4})
      for (int cr : {0, 16, 32, 64}) {
        if (tr == 4 && !(special && nsys == 1)) continue;
        int tpl, nch, nt, h;
        sym_shape(special, tr, cr, tpl, nch, nt, h);
        if (std::find(seen.begin(), seen.end(), std::make_tuple(tpl, nch, nt)) != seen.end())
          continue;
        seen.emplace_back(tpl, nch, nt);
        run_sym(tpl, nch, h, blocks, fp32, rec, F, T, U, W);
        LM_HIP(hipEventRecord(e0, stream_));
        run_sym(tpl, nch, h, blocks, fp32, rec, F, T, U, W);
        run_sym(tpl, nch, h, blocks, fp32, rec, F, T, U, W);
        LM_HIP(hipEventRecord(e1, stream_));
        LM_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        LM_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
          best = ms;
          pick = {tpl, nch};
        }
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    sym_tuned_[key] = pick;
  }
  int tr = tile_tpl_, cr = tile_nch_;
  if (tr == 0 && cr == 0) {
    auto it = sym_tuned_.find(key);
    if (it != sym_tuned_.end()) {
      tr = it->second.first;
      cr = it->second.second;
    }
  }
  int tpl, nch, nt, h;
  sym_shape(special, tr, cr, tpl, nch, nt, h);
  run_sym(tpl, nch, h, blocks, fp32, rec, F, T, U, W);
}

void NBodySolver::setPairTiling(int tpl, int nch) {
  if (tpl != 0 && tpl != 1 && tpl != 2 && tpl != 4)
    throw Error("[libmobility_amd] targets per lane must be 0 (auto), 1, 2 or 4");
  if (nch < 0) throw Error("[libmobility_amd] source chunks must be >= 0");
  tile_tpl_ = tpl;
  tile_nch_ = nch;
  items_tpl_ = -1;
}

// Launch the pair kernel over a work list of `count` items (tpl targets per
// lane, nch source chunks).
void NBodySolver::launch_pair_items(const PairItem* list, int count, int tpl, int nch, int blocks,
                                    bool fp32, const void* rec, const double* F, const double* T,
                                    double* U, double* W) {
  const int64_t n3 = 3 * N_;
  PairArgs a{};
  a.wall = kind_code();
  a.persistent = use_persistent() ? 1 : 0;
  if (a.persistent) {
    work_counter_.reserve(sizeof(unsigned int) * 64);
    a.counter = work_counter_.as<unsigned int>();
    a.grid = std::min(count, 1024);  // 4 workgroups per CU drain the queue
  }
  a.fp32 = fp32 ? 1 : 0;
  a.blocks = blocks;
  a.tpl = tpl;
  a.pos = rec;
  a.F = F;
  a.T = (blocks & (kUT | kWT)) ? T : nullptr;
  a.items = list;
  a.nitems = count;
  a.active = pair_active_;
  a.stop = a.persistent ? nullptr : pair_stop_;
  a.t0 = t0_;
  a.rt0 = rt0_;
  a.r0 = r0_;
  const bool ang = (blocks & (kWT | kWF)) != 0;
  if (nch == 1) {
    a.U = U;
    a.W = ang ? W : nullptr;
    a.slab = 0;
    launch_pair(a, stream_);
  } else {
    part_u_.reserve(sizeof(double) * n3 * nch);
    if (ang) part_w_.reserve(sizeof(double) * n3 * nch);
    a.U = part_u_.as<double>();
    a.W = ang ? part_w_.as<double>() : nullptr;
    a.slab = n3;
    launch_pair(a, stream_);
    if (defer_chunk_sum_ && !(seg_blocks_ && pair_active_)) {
      // the caller sums the slabs itself (the fused Lanczos step)
      deferred_nch_ = nch;
      return;
    }
    auto reduce = [&](const double* part, double* out) {
      if (seg_blocks_ && pair_active_)
        launch_seg_chunk_reduce(part, nch, n3, out, seg_blocks_, seg_nblocks_, pair_active_,
                                stream_);
      else
        launch_chunk_reduce(part, nch, n3, out, stream_);
    };
    reduce(part_u_.as<double>(), U);
    if (ang) reduce(part_w_.as<double>(), W);
  }
}

// Work list of the systems with act[s] != 0: the full list's tiling while it
// still gives every CU several items (filtered from the complete host list,
// cached per tiling and already in work-queue order, so a rebuild is one pass
// over the items), else one target per lane and enough source chunks for
// about 8192 items, built for the few systems left.
const std::vector<PairItem>& NBodySolver::host_items(int tpl, int nch) {
  if (host_items_struct_ != struct_id_ || host_items_N_ != N_) {
    host_items_.clear();
    host_items_struct_ = struct_id_;
    host_items_N_ = N_;
  }
  auto& v = host_items_[{tpl, nch, use_persistent() ? 1 : 0}];
  if (v.empty()) build_items(tpl, nch, use_persistent(), v);
  return v;
}

void NBodySolver::build_active_items(const int* act) {
  int tpl = items_tpl_ > 0 ? items_tpl_ : 1;
  int nch = items_nch_ > 0 ? items_nch_ : resolve_nch(tpl, 0, act);
  auto count_active = [&](const std::vector<PairItem>& all) {
    size_t c = 0;
    for (const PairItem& it : all) c += act[it.sys] ? 1 : 0;
    return c;
  };
  const std::vector<PairItem>* all = &host_items(tpl, nch);
  std::vector<PairItem> fine;
  if (count_active(*all) < 2048) {
    tpl = 1;
    nch = resolve_nch(1, 0, act);
    build_items(tpl, nch, use_persistent(), fine, act);
    all = &fine;
  }
  // the previous upload from the pinned buffer has completed: the caller
  // synchronised the stream after it
  seg_items_host_.reserve(sizeof(PairItem) * std::max<size_t>(all->size(), 1));
  PairItem* dst = seg_items_host_.as<PairItem>();
  int count = 0;
  for (const PairItem& it : *all)
    if (act[it.sys]) dst[count++] = it;
  seg_items_.count = 0;
  if (count == 0) return;
  const size_t bytes = sizeof(PairItem) * (size_t)count;
  seg_items_.buf.reserve(bytes);
  LM_HIP(hipMemcpyAsync(seg_items_.buf.as<void>(), dst, bytes, hipMemcpyHostToDevice, stream_));
  seg_items_.tpl = tpl;
  seg_items_.nch = nch;
  seg_items_.count = count;
}

// Tiling: user override, else the autotuned choice, else the default rule.
void NBodySolver::choose_tiling(int blocks, bool fp32, const double* F, const double* T,
                                double* U, double* W, const void* rec, int& tpl, int& nch) {
  const bool tpl4_ok = kind_code() == 0 && blocks == kUF;
  if (tile_tpl_ != 0 || tile_nch_ != 0) {
    tpl = tile_tpl_ != 0 ? tile_tpl_ : 1;
    if (tpl == 4 && !tpl4_ok) tpl = 2;
    nch = tile_nch_;
    return;
  }
  tpl = 1;
  nch = 0;
  if (algorithm_ == kNaive) {
    nch = 1;
    return;
  }
  if (algorithm_ == kBlock) {
    nch = 32;
    return;
  }
  // Measured on MI355X: two targets per lane pay off once the system is
  // large enough to keep every CU busy with 512-target tiles.
  int64_t maxn = 0;
  for (size_t s = 0; s + 1 < sys_off_host_.size(); ++s)
    maxn = std::max(maxn, sys_off_host_[s + 1] - sys_off_host_[s]);
  if (maxn >= 16384) tpl = 2;
  if (!autotune_) return;
  const int nsys = (int)sys_off_host_.size() - 1;
  const auto key = std::make_tuple(N_, nsys, kind_code() + 4 * (use_persistent() ? 1 : 0),
                                   fp32 ? 1 : 0, blocks);
  auto it = tuned_.find(key);
  if (it != tuned_.end()) {
    tpl = it->second.first;
    nch = it->second.second;
    return;
  }
  // Time every candidate on the live inputs (outputs are overwritten by the
  // real launch that follows).
  hipEvent_t e0, e1;
  LM_HIP(hipEventCreate(&e0));
  LM_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int best_tpl = tpl, best_nch = 0;
  const int tpls[3] = {1, 2, 4};
  const int nchs[6] = {0, 1, 2, 4, 8, 16};
  for (int ti = 0; ti < 3; ++ti) {
    if (tpls[ti] == 4 && !tpl4_ok) continue;
    for (int ci = 0; ci < 6; ++ci) {
      ensure_items(tpls[ti], nchs[ci]);
      if (nchs[ci] != 0 && items_nch_ != nchs[ci]) continue;  // clamped duplicate
      // per-launch time of `reps` back-to-back launches after one untimed
      auto time_reps = [&](int reps) {
        launch_main_items(blocks, fp32, rec, F, T, U, W);
        LM_HIP(hipEventRecord(e0, stream_));
        for (int r = 0; r < reps; ++r) launch_main_items(blocks, fp32, rec, F, T, U, W);
        LM_HIP(hipEventRecord(e1, stream_));
        LM_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        LM_HIP(hipEventElapsedTime(&t, e0, e1));
        return t / reps;
      };
      float ms = time_reps(2);
      // short products: two launches are within the clock's noise
      if (ms < 0.1f) ms = time_reps(16);
      if (ms < best) {
        best = ms;
        best_tpl = items_tpl_;
        best_nch = items_nch_;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  tuned_[key] = {best_tpl, best_nch};
  tpl = best_tpl;
  nch = best_nch;
}

// U (3N) and, when angular, W (3N) are overwritten (NbodyRPY.cu:91-96).
void NBodySolver::pair_product(const void* rec64, const void* rec32, const double* F,
                               const double* T, double* U, double* W, int prec) {
  int blocks = 0;
  // Block selection of NbodyRPY.cu:264-275.
  if (F) blocks |= kUF | (angular_ ? kWF : 0);
  if (T && angular_) blocks |= kUT | kWT;
  const int64_t n3 = 3 * N_;
  if (!U) throw Error("[libmobility_amd] Mdot needs a linear output array");
  if (blocks == 0) {
    launch_zero(U, n3, stream_);
    if (angular_ && W) launch_zero(W, n3, stream_);
    return;
  }
  const bool ang = (blocks & (kWT | kWF)) != 0;
  if (ang && !W)
    throw Error("[libmobility_amd] the solver includes angular velocities: an angular output array is required");
  const bool fp32 = (prec < 0 ? pair_prec_ : prec) == kPairFP32;
  const void* rec = fp32 ? rec32 : rec64;
  last_sym_ = false;
  deferred_nch_ = 0;
  if (use_sym(blocks)) {
    launch_sym(blocks, fp32, rec, F, T, U, W);
    last_sym_ = true;
  } else if (seg_items_.count > 0) {
    // per-system Lanczos after some systems converged: their work is dropped
    launch_pair_items(seg_items_.buf.as<PairItem>(), seg_items_.count, seg_items_.tpl,
                      seg_items_.nch, blocks, fp32, rec, F, T, U, W);
  } else {
    int tpl = 1, nch = 0;
    choose_tiling(blocks, fp32, F, T, U, W, rec, tpl, nch);
    ensure_items(tpl, nch);
    if (items_count_ <= 0) throw Error("[libmobility_amd] empty pair work list");
    launch_main_items(blocks, fp32, rec, F, T, U, W);
  }
  if (angular_ && W && !ang) launch_zero(W, n3, stream_);
}

// Records are scaled by 1/a; rebuild them if initialize() changed a after
// setPositions() (the reference reads the radius only when it multiplies).
void NBodySolver::ensure_records() {
  if (N_ <= 0 || rec_a_ == a_) return;
  const int nsys = (int)sys_off_host_.size() - 1;
  launch_system_bbox(shifted_.as<double>(), sys_off_.as<int64_t>(), nsys, 0.0, kinds_dev(),
                     wall_ ? 1 : 0, bbox_.as<double>(), stream_);
  launch_make_records(shifted_.as<double>(), N_, sys_off_.as<int64_t>(), nsys, bbox_.as<double>(),
                      kinds_dev(), wall_ ? 1 : 0, 0.0, a_, nullptr, rec64_.as<void>(),
                      rec32_.as<void>(), stream_);
  rec_a_ = a_;
  ++layout_id_;  // features are in units of a
}

void NBodySolver::mdot_dev(const double* F, const double* T, double* U, double* W) {
  require_positions("[Mobility] Positions have 0 particles. Did you call setPositions?");
  ensure_records();
  pair_product(rec64_.as<void>(), rec32_.as<void>(), F, T, U, W);
}

void NBodySolver::Mdot(const double* F, int64_t nF, const double* T, int64_t nT, double* U,
                       double* W, bool dev) {
  // pythonify.h:109-149 checks, then NBody::Mdot (mobility.h:166-182)
  const int64_t n3 = 3 * N_;
  if (nF > 0 && nF != n3) throw Error("The forces array must have size 3*N.");
  if (nT > 0 && nT != n3) throw Error("The torques array must have size 3*N.");
  if (!angular_ && nT > 0)
    throw Error(
        "The solver was configured without including angular velocities. Set includeAngular "
        "to true when initializing if you want to use torques");
  require_positions("[Mobility] Positions have 0 particles. Did you call setPositions?");
  const double* f = nF > 0 ? F : nullptr;
  const double* t = nT > 0 ? T : nullptr;
  if (dev) {
    mdot_dev(f, t, U, angular_ ? W : nullptr);
    LM_HIP(hipStreamSynchronize(stream_));
    return;
  }
  double *df = nullptr, *dt = nullptr;
  if (f) {
    st_f_.reserve(sizeof(double) * n3);
    df = st_f_.as<double>();
    LM_HIP(hipMemcpyAsync(df, f, sizeof(double) * n3, hipMemcpyHostToDevice, stream_));
  }
  if (t) {
    st_t_.reserve(sizeof(double) * n3);
    dt = st_t_.as<double>();
    LM_HIP(hipMemcpyAsync(dt, t, sizeof(double) * n3, hipMemcpyHostToDevice, stream_));
  }
  st_u_.reserve(sizeof(double) * n3);
  st_w_.reserve(sizeof(double) * n3);
  mdot_dev(df, dt, st_u_.as<double>(), angular_ ? st_w_.as<double>() : nullptr);
  LM_HIP(hipMemcpyAsync(U, st_u_.as<double>(), sizeof(double) * n3, hipMemcpyDeviceToHost, stream_));
  if (angular_ && W)
    LM_HIP(hipMemcpyAsync(W, st_w_.as<double>(), sizeof(double) * n3, hipMemcpyDeviceToHost, stream_));
  LM_HIP(hipStreamSynchronize(stream_));
}

// ---- Lanczos ---------------------------------------------------------------------

// Krylov basis, column-major n x cols, grown geometrically up to the
// iteration limit + 1 columns, keeping the columns already computed.
void NBodySolver::ensure_basis(int64_t n, int64_t cols) {
  if (V_n_ != n) {
    V_.release();
    V_cols_ = 0;
    V_n_ = n;
  }
  if (cols <= V_cols_) return;
  int64_t want = std::max<int64_t>(cols, std::min<int64_t>(iter_limit_ + 1,
                                                          std::max<int64_t>(32, 2 * V_cols_)));
  V_.grow_keep(sizeof(double) * n * want, sizeof(double) * n * V_cols_, stream_);
  V_cols_ = want;
}

void NBodySolver::setLanczosReorthogonalization(int mode) {
  if (mode != 0 && mode != 1)
    throw Error("[libmobility_amd] Lanczos reorthogonalisation mode must be 0 (off) or 1 (full)");
  reorth_ = mode;
}

void NBodySolver::setLanczosIterationLimit(int limit) {
  if (limit < 3 || limit > kLanczosMaxIterations)
    throw Error("[libmobility_amd] Lanczos iteration limit must be in [3, " +
                std::to_string(kLanczosMaxIterations) + "]");
  iter_limit_ = limit;
}

int NBodySolver::lanczos_run(const double* z, double** result) {
  const int64_t n = (angular_ ? 6 : 3) * N_;
  ensure_records();
  // The checks run on a second stream, overlapped with the next iterations:
  // kCheckBatch at a time for small systems, one at a time for large ones
  // (where a product costs milliseconds, a check of every iteration decides
  // before the next product is far along).
  const int batch = N_ <= kLanczosLookaheadN ? kCheckBatch : 1;
  const bool fused = n <= kLanczosFusedMaxN;
  // workspace
  red_partial_.reserve(sizeof(double) * 4 * kReduceBlocks);
  if (red_counter_.capacity() == 0) {
    red_counter_.reserve(sizeof(unsigned int) * 64);
    launch_zero_u32(red_counter_.as<unsigned int>(), 64, stream_);
  }
  scalars_.reserve(sizeof(LanczosScalars));
  tsq_work_.reserve(sizeof(double) * kCheckBatch * tsqrt_scratch(iter_limit_));
  coef_.reserve(sizeof(double) * kCheckBatch * kLanczosMaxIterations);
  w_.reserve(sizeof(double) * n);
  lz_status_.reserve(kLzSlots * sizeof(LanczosStatus), hipHostMallocMapped | hipHostMallocCoherent);
  for (auto& e : lz_ev_)
    if (!e) LM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (!chk_stream_) LM_HIP(hipStreamCreateWithFlags(&chk_stream_, hipStreamNonBlocking));
  if (!lz_vec_ev_) LM_HIP(hipEventCreateWithFlags(&lz_vec_ev_, hipEventDisableTiming));
  chk_partial_.reserve(sizeof(double) * 2 * kCheckBatch * kReduceBlocks);
  if (chk_counter_.capacity() == 0) {
    chk_counter_.reserve(sizeof(unsigned int) * 64);
    launch_zero_u32(chk_counter_.as<unsigned int>(), 64, stream_);
  }
  ReduceScratch rs{red_partial_.as<double>(), red_counter_.as<unsigned int>()};
  ReduceScratch rs_chk{chk_partial_.as<double>(), chk_counter_.as<unsigned int>()};
  hipStream_t cs = chk_stream_;
  LanczosScalars* sc = scalars_.as<LanczosScalars>();
  // Estimate ring; the previous estimate (oldBz) keeps its prefix across
  // calls (LanczosAlgorithm.cu:175).
  const int nslots = batch + 1;
  const int64_t stride = (n + 31) / 32 * 32;
  if (old_n_ != n || !est_valid_ || ring_slots_ != nslots) {
    const int64_t keep = est_valid_ ? std::min(old_n_, n) : 0;
    std::vector<double> host_prev(keep);
    if (keep > 0) {
      LM_HIP(hipMemcpyAsync(host_prev.data(), est_ring_.as<double>() + (int64_t)old_host_ * ring_stride_,
                            sizeof(double) * keep, hipMemcpyDeviceToHost, stream_));
      LM_HIP(hipStreamSynchronize(stream_));
    }
    est_ring_.release();
    est_ring_.reserve(sizeof(double) * nslots * stride);
    launch_zero(est_ring_.as<double>(), nslots * stride, stream_);
    if (keep > 0) {
      LM_HIP(hipMemcpyAsync(est_ring_.as<double>(), host_prev.data(), sizeof(double) * keep,
                            hipMemcpyHostToDevice, stream_));
      LM_HIP(hipStreamSynchronize(stream_));
    }
    old_host_ = 0;
    old_n_ = n;
    ring_slots_ = nslots;
    ring_stride_ = stride;
    est_valid_ = true;
  }
  // A basis that grows is copied and the old one freed: the check stream
  // must be done with it first.
  auto ensure_cols = [&](int64_t cols) {
    if (cols > V_cols_) LM_HIP(hipStreamSynchronize(chk_stream_));
    ensure_basis(n, cols);
  };
  if (pre_on()) {
    ensure_pre();
    pc_t_.reserve(sizeof(double) * n);
    pc_u_.reserve(sizeof(double) * n);
  }
  if (reorth_) {
    reorth_partial_.reserve(sizeof(double) * kLanczosMaxIterations * kReduceBlocks);
    reorth_c_.reserve(sizeof(double) * kLanczosMaxIterations);
    if (reorth_counter_.capacity() == 0) {
      reorth_counter_.reserve(sizeof(unsigned int) * 64);
      launch_zero_u32(reorth_counter_.as<unsigned int>(), 64, stream_);
    }
  }
  // the fused step sums the pair kernel's chunk slabs itself
  const bool defer = fused && !pre_on();
  // Products launched after the solve has converged on the device (the main
  // stream runs ahead of the checks) skip their pair work.
  struct StopGuard {
    const int*& p;
    ~StopGuard() { p = nullptr; }
  } stop_guard{pair_stop_};
  ensure_cols(2);
  double* V = V_.as<double>();
  double* w = w_.as<double>();
  launch_lanczos_start(z, V, n, rs, sc, old_host_, stream_);  // LanczosAlgorithm.cu:110-118
  pair_stop_ = &sc->status.conv;
  const int check = std::min(check_steps_, iter_limit_ - 2);
  LanczosStatus* mirror = lz_status_.as<LanczosStatus>();
  LanczosStatus* mirror_dev = lz_status_.device_as<LanczosStatus>();
  // On every exit the check stream has finished (the next solve resets the
  // status it reads).
  auto drain = [&]() {
    LM_HIP(hipStreamSynchronize(chk_stream_));
  };
  // Reads the status of the batch of checks starting at iteration `first`;
  // returns true once the solve is finished (the tolerance met).
  auto settle = [&](int first, int slot) -> bool {
    LM_HIP(hipEventSynchronize(lz_ev_[slot]));
    const LanczosStatus st = mirror[slot];
    if (st.first != first || st.nchk < 1 || st.nchk > kCheckBatch) {
      drain();
      throw Error("[libmobility_amd] Lanczos convergence status out of order");
    }
    for (int j = 0; j < st.nchk; ++j) {
      if (st.conv == 2 && j == st.nchk - 1) {
        drain();
        throw Error("[Lanczos] Unknown error (found NaN in result guess)");
      }
      if (first + j > 0 && cb_) cb_(first + j, st.err[j], cb_user_);
    }
    old_host_ = st.old;
    if (st.conv != 1) return false;
    const int i = st.conv_iter;
    // registerRequiredStepsForConverge, LanczosAlgorithm.cu:224-232
    if (i - 2 > check_steps_)
      check_steps_ += 1;
    else
      check_steps_ = std::max(1, check_steps_ - 2);
    last_iters_ = i;
    *result = est_ring_.as<double>() + (int64_t)st.result * ring_stride_;
    drain();
    return true;
  };
  // batches in flight, oldest first: (first iteration, slot)
  int pend_f[kLzSlots], pend_s[kLzSlots];
  int npend = 0, nbatches = 0, bfirst = -1;
  auto settle_oldest = [&]() -> bool {
    const bool done = settle(pend_f[0], pend_s[0]);
    for (int k = 1; k < npend; ++k) {
      pend_f[k - 1] = pend_f[k];
      pend_s[k - 1] = pend_s[k];
    }
    --npend;
    return done;
  };
  auto finish = [&]() {
    LM_HIP(hipStreamWaitEvent(stream_, lz_ev_[(nbatches - 1) % kLzSlots], 0));
    return last_iters_;
  };
  for (int i = 0; i < iter_limit_; ++i) {
    ensure_cols(i + 2);
    V = V_.as<double>();
    const double* vi = V + (int64_t)i * n;
    const double* vp = i > 0 ? V + (int64_t)(i - 1) * n : nullptr;
    double* vn = V + (int64_t)(i + 1) * n;
    if (fused) {
      defer_chunk_sum_ = defer;
      try {
        lanczos_product(vi, w);
      } catch (...) {
        defer_chunk_sum_ = false;
        drain();
        throw;
      }
      defer_chunk_sum_ = false;
      LanczosStepArgs sa{};
      if (deferred_nch_ > 0) {
        sa.slab_u = part_u_.as<double>();
        sa.slab_w = angular_ ? part_w_.as<double>() : nullptr;
        sa.nch = deferred_nch_;
      } else {
        sa.w = w;
      }
      sa.n3 = 3 * N_;
      sa.w_out = w;
      sa.partial = rs.partial;
      sa.counter = rs.counter;
      sa.stop = pair_stop_;
      sa.v_prev = vp;
      sa.v_i = vi;
      sa.v_next = vn;
      sa.n = n;
      sa.i = i;
      sa.sc = sc;
      launch_lanczos_step(sa, stream_);  // LanczosAlgorithm.cu:120-153
    } else {
      lanczos_product(vi, w);
      launch_lanczos_alpha(w, vi, vp, n, i, rs, sc, stream_);
      launch_lanczos_beta(w, vi, n, i, rs, sc, stream_);
      launch_lanczos_next(w, vn, n, i, sc, stream_);
    }
    if (reorth_)
      launch_lanczos_reorth(V, n, i, n, reorth_partial_.as<double>(),
                            reorth_counter_.as<unsigned int>(), reorth_c_.as<double>(), sc,
                            stream_);
    if (i < check && cb_ == nullptr) continue;
    // checks from iteration `check` on (every one with a callback), in batches
    if (bfirst < 0) bfirst = i;
    if (i - bfirst + 1 < batch && i < iter_limit_ - 1) continue;
    LM_HIP(hipEventRecord(lz_vec_ev_, stream_));
    LM_HIP(hipStreamWaitEvent(chk_stream_, lz_vec_ev_, 0));
    const int slot = nbatches++ % kLzSlots;
    mirror[slot].first = -2;
    CheckArgs ca{};
    ca.V = V;
    ca.ldv = n;
    ca.i0 = bfirst;
    ca.c = i - bfirst + 1;
    ca.coef = coef_.as<double>();
    ca.work = tsq_work_.as<double>();
    ca.Y = est_ring_.as<double>();
    ca.ystride = ring_stride_;
    ca.nslots = ring_slots_;
    ca.n = n;
    ca.tol = tol_;
    ca.rs = rs_chk;
    ca.sc = sc;
    ca.mirror = mirror_dev + slot;
    launch_lanczos_checks(ca, cs);
    LM_HIP(hipEventRecord(lz_ev_[slot], cs));
    const int first = bfirst;
    bfirst = -1;
    pend_f[npend] = first;
    pend_s[npend] = slot;
    ++npend;
    // read every finished batch in order; wait when too many are in flight
    while (npend > 0) {
      if (npend <= kLanczosMaxLag) {
        const hipError_t q = hipEventQuery(lz_ev_[pend_s[0]]);
        if (q == hipErrorNotReady) break;
        if (q != hipSuccess) LM_HIP(q);
      }
      if (settle_oldest()) return finish();
    }
  }
  while (npend > 0)
    if (settle_oldest()) return finish();
  drain();
  throw Error("[Lanczos] Could not converge");
}

// w = M v ; torques are stored after the forces (MobilityInterface.h:164-176).
// With the preconditioner: w = G M G^T v, G = D Gc.
void NBodySolver::lanczos_product(const double* v, double* w) {
  const int64_t n = (angular_ ? 6 : 3) * N_;
  const int64_t N3 = 3 * N_;
  if (pre_on()) {
    double* t = pc_t_.as<double>();
    double* u = pc_u_.as<double>();
    pre_forward(v, t, n);
    pair_product(rec64_.as<void>(), rec32_.as<void>(), t, angular_ ? t + N3 : nullptr, u,
                 angular_ ? u + N3 : nullptr);
    pre_back(u, w, n);
  } else {
    pair_product(rec64_.as<void>(), rec32_.as<void>(), v, angular_ ? v + N3 : nullptr, w,
                 angular_ ? w + N3 : nullptr);
  }
}

// ---- one Lanczos recurrence per system ------------------------------------------
// Each system runs lanczos::Solver::run (LanczosAlgorithm.cu:171-201) on its
// own block of the block-diagonal operator: its own ||z||, alpha/beta,
// breakdown test, check schedule (check_convergence_steps, adapted by
// registerRequiredStepsForConverge, cu:224-232), previous estimate (oldBz)
// and convergence test.  All systems advance together; a converged system
// leaves the pair work list (its items are skipped) and the vector kernels.

void NBodySolver::ensure_seg_layout(SegSolve& S, const std::vector<int64_t>& off, int angular,
                                    int64_t n) {
  const int nrec = (int)off.size() - 1;
  if (S.off != off) {
    std::vector<SegBlock> blocks;
    std::vector<int> first(nrec + 1, 0);
    for (int s = 0; s < nrec; ++s) {
      first[s] = (int)blocks.size();
      for (int64_t p0 = off[s]; p0 < off[s + 1]; p0 += kSegParticles) {
        SegBlock b;
        b.p0 = p0;
        b.p1 = std::min<int64_t>(p0 + kSegParticles, off[s + 1]);
        b.sys = s;
        b.pad = 0;
        blocks.push_back(b);
      }
    }
    first[nrec] = (int)blocks.size();
    if (blocks.empty()) throw Error("[libmobility_amd] no particles in any system");
    S.blocks.reserve(sizeof(SegBlock) * blocks.size());
    S.first.reserve(sizeof(int) * first.size());
    LM_HIP(hipMemcpyAsync(S.blocks.as<void>(), blocks.data(), sizeof(SegBlock) * blocks.size(),
                          hipMemcpyHostToDevice, stream_));
    LM_HIP(hipMemcpyAsync(S.first.as<void>(), first.data(), sizeof(int) * first.size(),
                          hipMemcpyHostToDevice, stream_));
    LM_HIP(hipStreamSynchronize(stream_));  // host vectors go out of scope
    S.nblocks = (int)blocks.size();
    S.partial.reserve(sizeof(double) * 2 * blocks.size());
    S.flags.reserve(sizeof(int) * 3 * nrec);
    S.host.reserve(sizeof(int) * 4 * nrec + sizeof(double) * nrec);
    S.check_steps.assign(nrec, 3);  // LanczosAlgorithm.cu:171, one solver per recurrence
    S.iters.assign(nrec, 0);
    S.off = off;
    S.angular = angular;
    S.n = 0;
    S.L = 0;
  }
  if (S.n != n || S.angular != angular) {
    // previous estimates of every recurrence start at zero (cu:175)
    S.old.reserve(sizeof(double) * n);
    S.y.reserve(sizeof(double) * n);
    launch_zero(S.old.as<double>(), n, stream_);
    launch_zero(S.y.as<double>(), n, stream_);
    S.n = n;
    S.angular = angular;
  }
  if (S.L < iter_limit_) {
    S.state.reserve(sizeof(double) * ((size_t)2 * nrec + (size_t)2 * nrec * iter_limit_));
    S.coef.reserve(sizeof(double) * (size_t)nrec * iter_limit_);
    S.L = iter_limit_;
  }
}

void NBodySolver::lastLanczosIterationsPerSystem(int* out, int nsys) const {
  const int have = (int)seg_sys_.iters.size();
  for (int s = 0; s < nsys; ++s) out[s] = (seg_mode_ && s < have) ? seg_sys_.iters[s] : -1;
}

int NBodySolver::lanczos_run_seg(SegSolve& S, const std::vector<int64_t>& off, bool angular,
                                 bool skip_converged_pairs,
                                 const std::function<void(const double*, double*)>& product,
                                 const double* z, double** result) {
  const int nrec = (int)off.size() - 1;
  const int64_t nvirt = off.back();              // particles of the stacked layout
  const int64_t n = (angular ? 6 : 3) * nvirt;  // entries of the stacked vector
  ensure_records();
  ensure_seg_layout(S, off, angular ? 1 : 0, n);
  const int L = S.L;
  seg_w_.reserve(sizeof(double) * n);
  double* st = S.state.as<double>();
  int* d_active = S.flags.as<int>();
  int* d_check = d_active + nrec;
  int* d_list = d_active + 2 * nrec;
  SegScalars sc{st, st + 2 * nrec, st + 2 * nrec + (int64_t)nrec * L, st + nrec, d_active,
                d_check, L};
  SegLayout g{S.blocks.as<SegBlock>(), S.nblocks, S.first.as<int>(), nrec, 3 * nvirt,
              angular ? 1 : 0, S.partial.as<double>()};
  // pinned host mirrors: the "check" flags alternate between two slots so a
  // slot is only rewritten after the stream has consumed its last upload
  int* h_active = S.host.as<int>();
  int* h_check_est = h_active + nrec;
  int* h_check_store = h_active + 2 * nrec;
  int* h_list = h_active + 3 * nrec;
  double* h_err = reinterpret_cast<double*>(h_active + 4 * nrec);
  double* old = S.old.as<double>();
  double* y = S.y.as<double>();
  double* coef = S.coef.as<double>();

  int nact = 0;
  std::vector<int> check_at(nrec);
  for (int s = 0; s < nrec; ++s) {
    h_active[s] = off[s + 1] > off[s] ? 1 : 0;
    nact += h_active[s];
    S.iters[s] = 0;
    check_at[s] = std::min(S.check_steps[s], iter_limit_ - 2);  // cu:178-179
  }
  LM_HIP(hipMemcpyAsync(d_active, h_active, sizeof(int) * nrec, hipMemcpyHostToDevice, stream_));
  ensure_basis(n, 2);
  launch_seg_start(z, V_.as<double>(), g, sc, stream_);  // cu:110-118 per recurrence
  // A recurrence whose noise is exactly zero has the result zero; it does not
  // iterate (its ||z|| = 0 would make every estimate NaN).
  LM_HIP(hipMemcpyAsync(h_err, sc.normz, sizeof(double) * nrec, hipMemcpyDeviceToHost, stream_));
  LM_HIP(hipStreamSynchronize(stream_));
  {
    bool any_zero = false;
    for (int s = 0; s < nrec; ++s) {
      h_check_store[s] = 0;
      if (h_active[s] && h_err[s] == 0.0) {
        h_active[s] = 0;
        h_check_store[s] = 3;
        --nact;
        any_zero = true;
      }
    }
    if (any_zero) {
      LM_HIP(hipMemcpyAsync(d_check, h_check_store, sizeof(int) * nrec, hipMemcpyHostToDevice,
                            stream_));
      launch_seg_clear(y, g, sc, stream_);
      LM_HIP(hipMemcpyAsync(d_active, h_active, sizeof(int) * nrec, hipMemcpyHostToDevice, stream_));
    }
  }
  struct ActiveGuard {
    NBodySolver& s;
    ~ActiveGuard() {
      s.pair_active_ = nullptr;
      s.seg_blocks_ = nullptr;
      s.seg_items_.count = 0;
      s.sym_.active_valid = false;
    }
  } guard{*this};
  if (skip_converged_pairs) {
    pair_active_ = d_active;
    seg_blocks_ = g.blocks;
    seg_nblocks_ = g.nblocks;
  }
  int last = 0;
  for (int i = 0; i < iter_limit_ && nact > 0; ++i) {
    ensure_basis(n, i + 2);
    double* V = V_.as<double>();
    const double* vi = V + (int64_t)i * n;
    double* w = seg_w_.as<double>();
    product(vi, w);
    launch_seg_alpha(w, vi, i > 0 ? V + (int64_t)(i - 1) * n : nullptr, i, g, sc, stream_);
    launch_seg_beta(w, vi, i, g, sc, stream_);
    launch_seg_next(w, V + (int64_t)(i + 1) * n, i, g, sc, stream_);
    int ncheck = 0;
    for (int s = 0; s < nrec; ++s) {
      const bool c = h_active[s] && (i >= check_at[s] || cb_ != nullptr);
      h_check_est[s] = c ? 1 : 0;
      if (c) h_list[ncheck++] = s;
    }
    if (ncheck == 0) continue;
    const int m = i + 1;
    if (S.work_cols < (size_t)m) {
      // grown geometrically with the Krylov dimension (slices of m x nodes)
      const int cols = std::min(L, std::max({m, 2 * (int)S.work_cols, 16}));
      S.work.reserve(sizeof(double) * kSegTsqrtGroups * tsqrt_scratch(cols));
      S.work_cols = (size_t)cols;
    }
    LM_HIP(hipMemcpyAsync(d_check, h_check_est, sizeof(int) * nrec, hipMemcpyHostToDevice, stream_));
    LM_HIP(hipMemcpyAsync(d_list, h_list, sizeof(int) * ncheck, hipMemcpyHostToDevice, stream_));
    launch_seg_tridiag_sqrt(d_list, ncheck, m, sc, coef, S.work.as<double>(), stream_);
    launch_seg_estimate(V, n, m, coef, y, old, g, sc, stream_);
    LM_HIP(hipMemcpyAsync(h_err, sc.err, sizeof(double) * nrec, hipMemcpyDeviceToHost, stream_));
    LM_HIP(hipStreamSynchronize(stream_));
    bool changed = false;
    double worst = 0.0;
    for (int s = 0; s < nrec; ++s) h_check_store[s] = 0;
    for (int j = 0; j < ncheck; ++j) {
      const int s = h_list[j];
      h_check_store[s] = 2;  // oldBz <- Bz unless converged (cu:196)
      if (i == 0) continue;  // the error is only tested from the second step (cu:185)
      const double err = h_err[s];
      if (std::isnan(err)) throw Error("[Lanczos] Unknown error (found NaN in result guess)");
      worst = std::max(worst, err);
      if (err <= tol_) {
        // registerRequiredStepsForConverge, cu:224-232
        if (i - 2 > S.check_steps[s])
          S.check_steps[s] += 1;
        else
          S.check_steps[s] = std::max(1, S.check_steps[s] - 2);
        S.iters[s] = i;
        last = std::max(last, i);
        h_active[s] = 0;
        h_check_store[s] = 0;
        --nact;
        changed = true;
      }
    }
    if (cb_ && i > 0) cb_(i, worst, cb_user_);
    LM_HIP(hipMemcpyAsync(d_check, h_check_store, sizeof(int) * nrec, hipMemcpyHostToDevice,
                          stream_));
    launch_seg_store(y, old, g, sc, stream_);
    if (changed) {
      LM_HIP(hipMemcpyAsync(d_active, h_active, sizeof(int) * nrec, hipMemcpyHostToDevice, stream_));
      if (skip_converged_pairs && nact > 0) {
        build_active_items(h_active);
        build_sym_active(h_active);
      }
    }
  }
  if (nact > 0) throw Error("[Lanczos] Could not converge");
  last_iters_ = last;
  *result = y;
  return last;
}

void NBodySolver::sqrtMdotW_dev(double* U, double* W, double prefactor, const double* noise) {
  const int64_t n = (angular_ ? 6 : 3) * N_;
  const double* z = noise;
  if (!z) {
    // LanczosStochasticVelocities: per-call seeds from an mt19937 (lanczos.h:35-52)
    if (!lanczos_engine_live_) {
      lanczos_engine_ = std::mt19937(lanczos_seed_);
      lanczos_engine_live_ = true;
    }
    const uint32_t s1 = std::uniform_int_distribution<uint32_t>(0, UINT32_MAX)(lanczos_engine_);
    const uint32_t s2 = std::uniform_int_distribution<uint32_t>(0, UINT32_MAX)(lanczos_engine_);
    noise_.reserve(sizeof(double) * n);
    launch_normal_fill(noise_.as<double>(), n, ((uint64_t)s2 << 32) | s1, 0, stream_);
    z = noise_.as<double>();
  }
  double* y = nullptr;
  if (seg_mode_ && numSystems() > 1) {
    if (pre_on()) {
      ensure_pre();
      pc_t_.reserve(sizeof(double) * n);
      pc_u_.reserve(sizeof(double) * n);
    }
    lanczos_run_seg(seg_sys_, sys_off_host_, angular_, true,
                    [this](const double* v, double* w) { lanczos_product(v, w); }, z, &y);
  } else {
    lanczos_run(z, &y);
  }
  if (pre_on()) {
    // B W = G^{-1} (G M G^T)^{1/2} W
    pre_inverse(y, pc_t_.as<double>(), n);
    y = pc_t_.as<double>();
  }
  // out += prefactor * y (MobilityInterface.h:179-187)
  launch_axpy_accumulate(U, y, prefactor, 3 * N_, stream_);
  if (angular_ && W) launch_axpy_accumulate(W, y + 3 * N_, prefactor, 3 * N_, stream_);
}

void NBodySolver::sqrtMdotW(double* U, double* W, double prefactor, bool dev,
                            const double* noise) {
  if (prefactor == 0) return;  // MobilityInterface.h:134-135
  if (!initialized_)
    throw Error(
        "[libMobility] You must initialize the base class in order to use the default "
        "stochastic displacement computation");
  require_positions(
      "[libMobility] The number of particles is not set. Did you forget to call setPositions?");
  if (!U) throw Error("[libMobility] This solver requires linear velocities");
  const int64_t n3 = 3 * N_;
  const int64_t n = (angular_ ? 6 : 3) * N_;
  if (dev) {
    sqrtMdotW_dev(U, W, prefactor, noise);
    LM_HIP(hipStreamSynchronize(stream_));
    return;
  }
  st_u_.reserve(sizeof(double) * n3);
  st_w_.reserve(sizeof(double) * n3);
  LM_HIP(hipMemcpyAsync(st_u_.as<double>(), U, sizeof(double) * n3, hipMemcpyHostToDevice, stream_));
  if (angular_ && W)
    LM_HIP(hipMemcpyAsync(st_w_.as<double>(), W, sizeof(double) * n3, hipMemcpyHostToDevice, stream_));
  const double* dnoise = nullptr;
  if (noise) {
    st_f_.reserve(sizeof(double) * n);
    LM_HIP(hipMemcpyAsync(st_f_.as<double>(), noise, sizeof(double) * n, hipMemcpyHostToDevice, stream_));
    dnoise = st_f_.as<double>();
  }
  sqrtMdotW_dev(st_u_.as<double>(), (angular_ && W) ? st_w_.as<double>() : nullptr, prefactor, dnoise);
  LM_HIP(hipMemcpyAsync(U, st_u_.as<double>(), sizeof(double) * n3, hipMemcpyDeviceToHost, stream_));
  if (angular_ && W)
    LM_HIP(hipMemcpyAsync(W, st_w_.as<double>(), sizeof(double) * n3, hipMemcpyDeviceToHost, stream_));
  LM_HIP(hipStreamSynchronize(stream_));
}

// ---- several vectors per pass over the pairs ---------------------------------------

void NBodySolver::mdot_multi_dev(const double* F, int nrhs, double* U) {
  require_positions("[Mobility] Positions have 0 particles. Did you call setPositions?");
  if (angular_)
    throw Error(
        "[libmobility_amd] the multi-vector product is translational only: initialize the solver "
        "without includeAngular");
  ensure_records();
  const int64_t n3 = 3 * N_;
  const bool fp32 = pair_prec_ == kPairFP32;
  const void* rec = fp32 ? rec32_.as<void>() : rec64_.as<void>();
  int r0 = 0;
  while (r0 < nrhs) {
    const int left = nrhs - r0;
    const int nr = left >= 4 ? 4 : (left >= 2 ? 2 : 1);
    if (nr == 1) {
      pair_product(rec64_.as<void>(), rec32_.as<void>(), F + (int64_t)r0 * n3, nullptr,
                   U + (int64_t)r0 * n3, nullptr);
      r0 += 1;
      continue;
    }
    const int tpl = multi_tpl(nr);
    ItemList& L = multi_items_[tpl == 1 ? 0 : 1];
    ensure_multi_items(L, tpl);
    PairMultiArgs a{};
    a.kind = kind_code();
    a.fp32 = fp32 ? 1 : 0;
    a.nrhs = nr;
    a.pos = rec;
    a.F = F + (int64_t)r0 * n3;
    a.n3 = n3;
    a.nch = L.nch;
    a.items = L.buf.as<PairItem>();
    a.nitems = L.count;
    a.t0 = t0_;
    if (L.nch == 1) {
      a.U = U + (int64_t)r0 * n3;
      launch_pair_multi(a, stream_);
    } else {
      part_multi_.reserve(sizeof(double) * (size_t)nr * L.nch * n3);
      a.U = part_multi_.as<double>();
      launch_pair_multi(a, stream_);
      for (int r = 0; r < nr; ++r)
        launch_chunk_reduce(part_multi_.as<double>() + (int64_t)r * L.nch * n3, L.nch, n3,
                            U + (int64_t)(r0 + r) * n3, stream_);
    }
    r0 += nr;
  }
}

void NBodySolver::MdotMulti(const double* F, int nrhs, double* U, bool dev) {
  if (nrhs < 1) throw Error("[libmobility_amd] at least one force vector is required");
  if (!F || !U) throw Error("[libmobility_amd] null force or output array");
  require_positions("[Mobility] Positions have 0 particles. Did you call setPositions?");
  const int64_t total = (int64_t)nrhs * 3 * N_;
  if (dev) {
    mdot_multi_dev(F, nrhs, U);
    LM_HIP(hipStreamSynchronize(stream_));
    return;
  }
  multi_f_.reserve(sizeof(double) * total);
  multi_u_.reserve(sizeof(double) * total);
  LM_HIP(hipMemcpyAsync(multi_f_.as<double>(), F, sizeof(double) * total, hipMemcpyHostToDevice,
                        stream_));
  mdot_multi_dev(multi_f_.as<double>(), nrhs, multi_u_.as<double>());
  LM_HIP(hipMemcpyAsync(U, multi_u_.as<double>(), sizeof(double) * total, hipMemcpyDeviceToHost,
                        stream_));
  LM_HIP(hipStreamSynchronize(stream_));
}

// w = M v (or G M G^T v) for nrhs stacked vectors
void NBodySolver::multi_product(const double* v, int nrhs, double* w) {
  const int64_t n3 = 3 * N_;
  if (pre_on()) {
    double* t = multi_t_.as<double>();
    double* u = multi_w_.as<double>();
    for (int r = 0; r < nrhs; ++r) pre_forward(v + r * n3, t + r * n3, n3);
    mdot_multi_dev(t, nrhs, u);
    for (int r = 0; r < nrhs; ++r) pre_back(u + r * n3, w + r * n3, n3);
  } else {
    mdot_multi_dev(v, nrhs, w);
  }
}

void NBodySolver::sqrtMdotWMulti(double* U, int nrhs, double prefactor, bool dev,
                                 const double* noise) {
  if (nrhs < 1) throw Error("[libmobility_amd] at least one noise vector is required");
  if (prefactor == 0) return;  // MobilityInterface.h:134-135
  if (!initialized_)
    throw Error(
        "[libMobility] You must initialize the base class in order to use the default "
        "stochastic displacement computation");
  require_positions(
      "[libMobility] The number of particles is not set. Did you forget to call setPositions?");
  if (!U) throw Error("[libMobility] This solver requires linear velocities");
  if (angular_)
    throw Error(
        "[libmobility_amd] the multi-vector square root is translational only: initialize the "
        "solver without includeAngular");
  const int64_t n3 = 3 * N_;
  const int64_t total = (int64_t)nrhs * n3;
  double* Ud = U;
  const double* z = noise;
  if (!dev) {
    multi_u_.reserve(sizeof(double) * total);
    LM_HIP(hipMemcpyAsync(multi_u_.as<double>(), U, sizeof(double) * total, hipMemcpyHostToDevice,
                          stream_));
    Ud = multi_u_.as<double>();
    if (noise) {
      multi_f_.reserve(sizeof(double) * total);
      LM_HIP(hipMemcpyAsync(multi_f_.as<double>(), noise, sizeof(double) * total,
                            hipMemcpyHostToDevice, stream_));
      z = multi_f_.as<double>();
    }
  }
  if (!z) {
    // the seeds nrhs successive sqrtMdotW calls would draw (lanczos.h:35-52)
    if (!lanczos_engine_live_) {
      lanczos_engine_ = std::mt19937(lanczos_seed_);
      lanczos_engine_live_ = true;
    }
    multi_noise_.reserve(sizeof(double) * total);
    for (int r = 0; r < nrhs; ++r) {
      const uint32_t s1 = std::uniform_int_distribution<uint32_t>(0, UINT32_MAX)(lanczos_engine_);
      const uint32_t s2 = std::uniform_int_distribution<uint32_t>(0, UINT32_MAX)(lanczos_engine_);
      launch_normal_fill(multi_noise_.as<double>() + r * n3, n3, ((uint64_t)s2 << 32) | s1, 0,
                         stream_);
    }
    z = multi_noise_.as<double>();
  }
  // recurrences: one per vector, or one per (vector, system) in per-system mode
  const int nsys = numSystems();
  const bool per_sys = seg_mode_ && nsys > 1;
  std::vector<int64_t> off(1, 0);
  for (int r = 0; r < nrhs; ++r) {
    if (per_sys)
      for (int s = 1; s <= nsys; ++s) off.push_back((int64_t)r * N_ + sys_off_host_[s]);
    else
      off.push_back((int64_t)(r + 1) * N_);
  }
  if (pre_on()) {
    ensure_pre();
    multi_t_.reserve(sizeof(double) * total);
    multi_w_.reserve(sizeof(double) * total);
  }
  double* y = nullptr;
  lanczos_run_seg(seg_multi_, off, false, false,
                  [this, nrhs](const double* v, double* w) { multi_product(v, nrhs, w); }, z, &y);
  multi_nrhs_ = nrhs;
  multi_per_ = per_sys ? nsys : 1;
  if (pre_on()) {
    double* t = multi_t_.as<double>();
    for (int r = 0; r < nrhs; ++r) pre_inverse(y + r * n3, t + r * n3, n3);
    y = t;
  }
  launch_axpy_accumulate(Ud, y, prefactor, total, stream_);  // out += prefactor * y
  if (!dev)
    LM_HIP(hipMemcpyAsync(U, Ud, sizeof(double) * total, hipMemcpyDeviceToHost, stream_));
  LM_HIP(hipStreamSynchronize(stream_));
}

void NBodySolver::lastLanczosIterationsMulti(int* out, int nrhs) const {
  for (int r = 0; r < nrhs; ++r) {
    int best = -1;
    if (r < multi_nrhs_)
      for (int s = 0; s < multi_per_; ++s)
        best = std::max(best, seg_multi_.iters[(size_t)r * multi_per_ + s]);
    out[r] = best;
  }
}

// ---- learned preconditioner ------------------------------------------------------

void NBodySolver::ensure_knn() {
  require_positions("[Mobility] Positions have 0 particles. Did you call setPositions?");
  ensure_records();
  if (knn_layout_ == layout_id_ && knn_layout_ != 0) return;
  pc_feat_.reserve(sizeof(float) * N_ * kPrecondIn);
  pc_nbr_.reserve(sizeof(int) * N_ * kPrecondK);
  const int nsys = (int)sys_off_host_.size() - 1;
  // cell grid on the host (deterministic), neighbour search on the device
  knn_xyz_.resize((size_t)4 * N_);
  LM_HIP(hipMemcpyAsync(knn_xyz_.data(), rec64_.as<void>(), sizeof(double) * 4 * N_,
                        hipMemcpyDeviceToHost, stream_));
  LM_HIP(hipStreamSynchronize(stream_));
  knn_grid(knn_xyz_.data(), sys_off_host_.data(), nsys, knn_sys_host_, knn_cell_host_,
           knn_perm_host_, knn_start_host_);
  auto upload = [&](DevBuf& b, const void* src, size_t bytes) {
    b.reserve(bytes);
    LM_HIP(hipMemcpyAsync(b.as<void>(), src, bytes, hipMemcpyHostToDevice, stream_));
  };
  upload(knn_sys_, knn_sys_host_.data(), sizeof(KnnSys) * knn_sys_host_.size());
  upload(knn_cell_, knn_cell_host_.data(), sizeof(int) * N_);
  upload(knn_perm_, knn_perm_host_.data(), sizeof(int) * N_);
  upload(knn_start_, knn_start_host_.data(), sizeof(int64_t) * knn_start_host_.size());
  launch_knn_cells(rec64_.as<double4>(), sys_off_.as<int64_t>(), nsys, knn_sys_.as<KnnSys>(),
                   knn_cell_.as<int>(), knn_perm_.as<int>(), knn_start_.as<int64_t>(), N_,
                   kinds_dev(), wall_ ? 1 : 0, pc_feat_.as<float>(), pc_nbr_.as<int>(), stream_);
  knn_layout_ = layout_id_;
}

void NBodySolver::precondNeighbours(int* nbr) {
  ensure_knn();
  LM_HIP(hipMemcpyAsync(nbr, pc_nbr_.as<int>(), sizeof(int) * kPrecondK * N_,
                        hipMemcpyDeviceToHost, stream_));
  LM_HIP(hipStreamSynchronize(stream_));
}

// A preconditioner part built for earlier positions of the same systems (same
// N, offsets, kinds and radius) is still exact: G keeps its exact inverse, so
// B B^T = M holds for the new positions and only the iteration count can
// drift.  It is kept for pc_refresh_ setPositions calls.
bool NBodySolver::pre_current(uint64_t layout, uint64_t structure, double a) const {
  return layout != 0 && structure == struct_id_ && a == a_ && layout_id_ >= layout &&
         layout_id_ - layout < (uint64_t)pc_refresh_;
}

void NBodySolver::setPreconditionerRefresh(int interval) {
  if (interval < 1) throw Error("[libmobility_amd] the refresh interval must be >= 1");
  pc_refresh_ = interval;
}

void NBodySolver::ensure_precond_blocks() {
  if (!precond_) throw Error("[libmobility_amd] no preconditioner attached");
  ensure_records();
  if (pc_version_ == precond_->version() && pre_current(pc_layout_, pc_struct_, pc_a_)) return;
  const int kind = precond_->kind();
  if (kind == 1 && !angular_)
    throw Error("[libmobility_amd] a preconditioner of 6x6 blocks needs a solver initialised with "
                "includeAngular");
  ensure_knn();
  const int d = precond_->blockDim();
  pc_out_.reserve(sizeof(float) * N_ * precond_->outputs());
  pc_G_.reserve(sizeof(double) * d * d * N_);
  pc_Gi_.reserve(sizeof(double) * d * d * N_);
  precond_->forward(pc_feat_.as<float>(), (int)N_, pc_out_.as<float>(), stream_);
  launch_make_blocks(pc_out_.as<float>(), pc_nbr_.as<int>(), N_, kind, a_, pc_G_.as<double>(),
                     pc_Gi_.as<double>(), stream_);
  pc_layout_ = layout_id_;
  pc_struct_ = struct_id_;
  pc_a_ = a_;
  pc_version_ = precond_->version();
}

void NBodySolver::setCoarseLevel(int particles_per_cell) {
  if (particles_per_cell < 0)
    throw Error("[libmobility_amd] particles per cell must be >= 0 (0: no coarse level)");
  coarse_ppc_ = particles_per_cell;
}

void NBodySolver::setCoarseMaxDim(int64_t max_dim) {
  if (max_dim < 3 || max_dim > CoarseLevel::kMaxDimLimit)
    throw Error("[libmobility_amd] coarse dimension cap must be in [3, " +
                std::to_string(CoarseLevel::kMaxDimLimit) + "]");
  coarse_max_dim_ = max_dim;
}

void NBodySolver::ensure_coarse() {
  if (coarse_ppc_ <= 0) return;
  ensure_records();
  if (coarse_.ppc == coarse_ppc_ && coarse_.max_dim == coarse_max_dim_ &&
      coarse_.fp64_requested == coarse_fp64_ && coarse_.cells() > 0 &&
      pre_current(coarse_.layout, coarse_.structure, coarse_.a))
    return;
  const int nsys = numSystems();
  std::vector<int> wall(nsys);
  for (int q = 0; q < nsys; ++q)
    wall[q] = q < (int)sys_kind_host_.size() ? (sys_kind_host_[q] != 0) : (wall_ ? 1 : 0);
  coarse_.build(rec64_.as<void>(), rec32_.as<void>(), sys_off_host_, wall, coarse_ppc_,
                coarse_max_dim_, coarse_fp64_, stream_);
  coarse_.layout = layout_id_;
  coarse_.structure = struct_id_;
  coarse_.a = a_;
  coarse_.ppc = coarse_ppc_;
  coarse_.max_dim = coarse_max_dim_;
  coarse_.fp64_requested = coarse_fp64_;
  ++coarse_builds_;
}

void NBodySolver::ensure_pre() {
  if (precond_) ensure_precond_blocks();
  ensure_coarse();
}

// G = D Gc: t = G^T v = Gc^T D^T v
void NBodySolver::pre_forward(const double* v, double* t, int64_t n) {
  if (precond_) {
    launch_apply_blocks(pc_G_.as<double>(), v, N_, n, precond_->kind(), 1, t, stream_);
    if (coarse_ppc_ > 0) coarse_.apply(CoarseLevel::kGt, t, t, n, stream_);
  } else {
    coarse_.apply(CoarseLevel::kGt, v, t, n, stream_);
  }
}

// w = G u = D Gc u
void NBodySolver::pre_back(double* u, double* w, int64_t n) {
  if (precond_) {
    if (coarse_ppc_ > 0) coarse_.apply(CoarseLevel::kG, u, u, n, stream_);
    launch_apply_blocks(pc_G_.as<double>(), u, N_, n, precond_->kind(), 0, w, stream_);
  } else {
    coarse_.apply(CoarseLevel::kG, u, w, n, stream_);
  }
}

// out = G^{-1} y = Gc^{-1} D^{-1} y
void NBodySolver::pre_inverse(const double* y, double* out, int64_t n) {
  if (precond_) {
    launch_apply_blocks(pc_Gi_.as<double>(), y, N_, n, precond_->kind(), 0, out, stream_);
    if (coarse_ppc_ > 0) coarse_.apply(CoarseLevel::kGinv, out, out, n, stream_);
  } else {
    coarse_.apply(CoarseLevel::kGinv, y, out, n, stream_);
  }
}

int NBodySolver::coarseCells(int* cell_of, int* grid) const {
  const CoarseCells& c = coarse_.host_cells();
  if (c.C == 0) throw Error("[libmobility_amd] no coarse level has been built");
  if (cell_of) std::copy(c.cell_of.begin(), c.cell_of.end(), cell_of);
  if (grid)
    for (int k = 0; k < 3; ++k) grid[k] = c.g[k];
  return c.C;
}

void NBodySolver::coarseFactors(int system, double* L, double* X) const {
  coarse_.factors(system, L, X, stream_);
}

int NBodySolver::coarseSystemCells(int system) const {
  if (coarse_.cells() == 0) throw Error("[libmobility_amd] no coarse level has been built");
  if (system < 0 || system >= coarse_.systems())
    throw Error("[libmobility_amd] coarse level: no such system");
  return coarse_.system_cells(system);
}

void NBodySolver::precondDataset(int kind, int64_t ncenters, uint64_t seed, float* X, float* T,
                                 int* status_out) {
  ensure_knn();
  if (ncenters <= 0 || ncenters > N_) throw Error("[libmobility_amd] invalid number of centers");
  const int d = precond_dim(kind);
  // distinct random centers
  std::vector<int64_t> idx(N_);
  for (int64_t i = 0; i < N_; ++i) idx[i] = i;
  std::mt19937_64 rng(seed);
  for (int64_t i = 0; i < ncenters; ++i) {
    std::uniform_int_distribution<int64_t> dist(i, N_ - 1);
    std::swap(idx[i], idx[dist(rng)]);
  }
  // the mobility of each center's system (open or bottom wall)
  std::vector<int> wall(ncenters);
  for (int64_t c = 0; c < ncenters; ++c) {
    const int64_t i = idx[c];
    const int sys = (int)(std::upper_bound(sys_off_host_.begin(), sys_off_host_.end(), i) -
                          sys_off_host_.begin()) - 1;
    wall[c] = sys < (int)sys_kind_host_.size() ? (sys_kind_host_[sys] != 0) : (wall_ ? 1 : 0);
  }
  DevBuf centers, walls, tg, st, xf;
  centers.reserve(sizeof(int64_t) * ncenters);
  walls.reserve(sizeof(int) * ncenters);
  tg.reserve(sizeof(float) * d * d * ncenters);
  st.reserve(sizeof(int) * ncenters);
  xf.reserve(sizeof(float) * kPrecondIn * ncenters);
  LM_HIP(hipMemcpyAsync(centers.as<int64_t>(), idx.data(), sizeof(int64_t) * ncenters,
                        hipMemcpyHostToDevice, stream_));
  LM_HIP(hipMemcpyAsync(walls.as<int>(), wall.data(), sizeof(int) * ncenters,
                        hipMemcpyHostToDevice, stream_));
  launch_local_targets(rec64_.as<double4>(), pc_nbr_.as<int>(), centers.as<int64_t>(),
                       walls.as<int>(), ncenters, kind, tg.as<float>(), st.as<int>(), stream_);
  launch_gather_rows(pc_feat_.as<float>(), centers.as<int64_t>(), (int)ncenters, kPrecondIn,
                     xf.as<float>(), stream_);
  std::vector<int> status(ncenters);
  LM_HIP(hipMemcpyAsync(status.data(), st.as<int>(), sizeof(int) * ncenters, hipMemcpyDeviceToHost, stream_));
  LM_HIP(hipMemcpyAsync(T, tg.as<float>(), sizeof(float) * d * d * ncenters, hipMemcpyDeviceToHost, stream_));
  LM_HIP(hipMemcpyAsync(X, xf.as<float>(), sizeof(float) * kPrecondIn * ncenters, hipMemcpyDeviceToHost, stream_));
  LM_HIP(hipStreamSynchronize(stream_));
  if (status_out) {
    std::copy(status.begin(), status.end(), status_out);
    return;
  }
  for (int64_t c = 0; c < ncenters; ++c)
    if (status[c] != 0) throw Error("[libmobility_amd] local inverse square root did not converge");
}

void NBodySolver::precondBlocks(double* G) {
  ensure_precond_blocks();
  const int d = precond_->blockDim();
  LM_HIP(hipMemcpyAsync(G, pc_G_.as<double>(), sizeof(double) * d * d * N_, hipMemcpyDeviceToHost, stream_));
  LM_HIP(hipStreamSynchronize(stream_));
}

// ---- random finite differences -------------------------------------------------------

void NBodySolver::divM_dev(double* U, double* W, double prefactor, const double* given) {
  const int64_t N = N_, n3 = 3 * N;
  ensure_records();
  const double delta = delta_ * a_;  // mobility.h:224
  rfd_noise_.reserve(sizeof(double) * n3);
  rfd_pos_.reserve(sizeof(double) * n3);
  rfd_rec64_.reserve(sizeof(double) * 4 * N);
  rfd_rec32_.reserve(sizeof(float) * 4 * N);
  rfd_mp_.reserve(sizeof(double) * n3);
  rfd_mm_.reserve(sizeof(double) * n3);
  if (angular_) {
    rfd_wp_.reserve(sizeof(double) * n3);
    rfd_wm_.reserve(sizeof(double) * n3);
  }
  double* noise = rfd_noise_.as<double>();
  if (given) {
    LM_HIP(hipMemcpyAsync(noise, given, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_));
  } else {
    const uint32_t seed = rng_nbody_();  // mobility.h:218
    launch_normal_fill(noise, n3, seed, 1, stream_);
  }
  const int nsys = (int)sys_off_host_.size() - 1;
  // open systems of a mixed ensemble have no drift (mobility.h:189-190):
  // no displacement and no force there
  if (kind_code() == 2)
    launch_zero_open_systems(noise, N, sys_off_.as<int64_t>(), nsys, kinds_dev(), stream_);
  // q + (delta/2) W, then M(q + delta/2 W) W  (random_finite_differences.h:47-49)
  launch_axpy_to(rfd_pos_.as<double>(), shifted_.as<double>(), noise, 0.5 * delta, n3, stream_);
  launch_system_bbox(rfd_pos_.as<double>(), sys_off_.as<int64_t>(), nsys, 0.0, kinds_dev(),
                     wall_ ? 1 : 0, bbox_.as<double>(), stream_);
  launch_make_records(rfd_pos_.as<double>(), N, sys_off_.as<int64_t>(), nsys, bbox_.as<double>(),
                      kinds_dev(), wall_ ? 1 : 0, 0.0, a_, nullptr, rfd_rec64_.as<void>(),
                      rfd_rec32_.as<void>(), stream_);
  pair_product(rfd_rec64_.as<void>(), rfd_rec32_.as<void>(), noise, nullptr, rfd_mp_.as<double>(),
               angular_ ? rfd_wp_.as<double>() : nullptr, drift_prec_);
  // q - (delta/2) W  (:50-52)
  launch_axpy_to(rfd_pos_.as<double>(), shifted_.as<double>(), noise, -0.5 * delta, n3, stream_);
  launch_system_bbox(rfd_pos_.as<double>(), sys_off_.as<int64_t>(), nsys, 0.0, kinds_dev(),
                     wall_ ? 1 : 0, bbox_.as<double>(), stream_);
  launch_make_records(rfd_pos_.as<double>(), N, sys_off_.as<int64_t>(), nsys, bbox_.as<double>(),
                      kinds_dev(), wall_ ? 1 : 0, 0.0, a_, nullptr, rfd_rec64_.as<void>(),
                      rfd_rec32_.as<void>(), stream_);
  pair_product(rfd_rec64_.as<void>(), rfd_rec32_.as<void>(), noise, nullptr, rfd_mm_.as<double>(),
               angular_ ? rfd_wm_.as<double>() : nullptr, drift_prec_);
  // out += (prefactor/delta) (M+ W - M- W)  (:54-62, mobility.h:227-235)
  launch_diff_accumulate(U, rfd_mp_.as<double>(), rfd_mm_.as<double>(), prefactor / delta, n3, stream_);
  if (angular_ && W)
    launch_diff_accumulate(W, rfd_wp_.as<double>(), rfd_wm_.as<double>(), prefactor / delta, n3, stream_);
}

void NBodySolver::divM(double* U, double* W, double prefactor, bool dev, const double* noise) {
  if (prefactor == 0) return;  // mobility.h:186-188
  if (!any_wall_) return;      // open: no thermal drift (mobility.h:189-190)
  require_positions(
      "[libMobility] The number of particles is not set. Did you forget to call setPositions?");
  if (!U) throw Error("[libMobility] The number of forces does not match the number of particles");
  const int64_t n3 = 3 * N_;
  if (dev) {
    divM_dev(U, W, prefactor, noise);
    LM_HIP(hipStreamSynchronize(stream_));
    return;
  }
  st_u_.reserve(sizeof(double) * n3);
  st_w_.reserve(sizeof(double) * n3);
  LM_HIP(hipMemcpyAsync(st_u_.as<double>(), U, sizeof(double) * n3, hipMemcpyHostToDevice, stream_));
  if (angular_ && W)
    LM_HIP(hipMemcpyAsync(st_w_.as<double>(), W, sizeof(double) * n3, hipMemcpyHostToDevice, stream_));
  const double* dnoise = nullptr;
  if (noise) {
    st_f_.reserve(sizeof(double) * n3);
    LM_HIP(hipMemcpyAsync(st_f_.as<double>(), noise, sizeof(double) * n3, hipMemcpyHostToDevice, stream_));
    dnoise = st_f_.as<double>();
  }
  divM_dev(st_u_.as<double>(), (angular_ && W) ? st_w_.as<double>() : nullptr, prefactor, dnoise);
  LM_HIP(hipMemcpyAsync(U, st_u_.as<double>(), sizeof(double) * n3, hipMemcpyDeviceToHost, stream_));
  if (angular_ && W)
    LM_HIP(hipMemcpyAsync(W, st_w_.as<double>(), sizeof(double) * n3, hipMemcpyDeviceToHost, stream_));
  LM_HIP(hipStreamSynchronize(stream_));
}

void NBodySolver::langevin(double dt, double kbt, const double* F, int64_t nF, const double* T,
                           int64_t nT, double* U, double* W, bool dev) {
  // MobilityInterface.h:191-206
  const int64_t n3 = 3 * N_;
  if (nF > 0 && nF != n3) throw Error("The forces array must have size 3*N.");
  if (nT > 0 && nT != n3) throw Error("The torques array must have size 3*N.");
  if (!angular_ && nT > 0)
    throw Error(
        "The solver was configured without including angular velocities. Set includeAngular "
        "to true when initializing if you want to use torques");
  const bool have_ft = nF > 0 || nT > 0;
  if (!have_ft && kbt == 0) return;
  require_positions("[Mobility] Positions have 0 particles. Did you call setPositions?");
  if (kbt != 0) {
    if (!initialized_)
      throw Error(
          "[libMobility] You must initialize the base class in order to use the default "
          "stochastic displacement computation");
  }
  double *dU = U, *dW = angular_ ? W : nullptr;
  const double *dF = nF > 0 ? F : nullptr, *dT = nT > 0 ? T : nullptr;
  if (!dev) {
    st_u_.reserve(sizeof(double) * n3);
    st_w_.reserve(sizeof(double) * n3);
    dU = st_u_.as<double>();
    dW = angular_ ? st_w_.as<double>() : nullptr;
    LM_HIP(hipMemcpyAsync(dU, U, sizeof(double) * n3, hipMemcpyHostToDevice, stream_));
    if (dW && W) LM_HIP(hipMemcpyAsync(dW, W, sizeof(double) * n3, hipMemcpyHostToDevice, stream_));
    if (dF) {
      st_f_.reserve(sizeof(double) * n3);
      LM_HIP(hipMemcpyAsync(st_f_.as<double>(), F, sizeof(double) * n3, hipMemcpyHostToDevice, stream_));
      dF = st_f_.as<double>();
    }
    if (dT) {
      st_t_.reserve(sizeof(double) * n3);
      LM_HIP(hipMemcpyAsync(st_t_.as<double>(), T, sizeof(double) * n3, hipMemcpyHostToDevice, stream_));
      dT = st_t_.as<double>();
    }
  }
  if (have_ft) mdot_dev(dF, dT, dU, dW);
  if (kbt != 0) {
    sqrtMdotW_dev(dU, dW, std::sqrt(2 * kbt / dt), nullptr);
    if (any_wall_) divM_dev(dU, dW, kbt);
  }
  if (!dev) {
    LM_HIP(hipMemcpyAsync(U, dU, sizeof(double) * n3, hipMemcpyDeviceToHost, stream_));
    if (dW && W) LM_HIP(hipMemcpyAsync(W, dW, sizeof(double) * n3, hipMemcpyDeviceToHost, stream_));
  }
  LM_HIP(hipStreamSynchronize(stream_));
}

}  // namespace lm
