// C ABI (include/libmobility_amd.h) over lm::NBodySolver.
#include "../../include/libmobility_amd.h"

#include <algorithm>
#include <exception>
#include <string>

#include "coarse.hpp"
#include "pair_math.hpp"
#include "philox.hpp"
#include "self_solver.hpp"
#include "solver.hpp"

struct lm_self {
  lm::SelfSolver solver;
  lm_self(int px, int py, int pz) : solver(px, py, pz) {}
};

struct lm_nbody {
  lm::NBodySolver solver;
  lm_nbody(int px, int py, int pz) : solver(px, py, pz) {}
};

struct lm_precond {
  lm::LearnedPreconditioner net;
  explicit lm_precond(int kind) : net(kind) {}
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
  try {
    g_last_error.clear();
    f();
    return LM_OK;
  } catch (const std::exception& e) {
    g_last_error = e.what();
  } catch (...) {
    g_last_error = "[libmobility_amd] unknown error";
  }
  return LM_ERROR;
}

lm::SelfSolver& SS(lm_self* h) {
  if (!h) throw lm::Error("[libmobility_amd] null solver handle");
  return h->solver;
}

lm::NBodySolver& S(lm_nbody* h) {
  if (!h) throw lm::Error("[libmobility_amd] null solver handle");
  return h->solver;
}

bool is_dev(int mem) {
  if (mem != LM_MEM_HOST && mem != LM_MEM_DEVICE) throw lm::Error("[libmobility_amd] invalid memory space");
  return mem == LM_MEM_DEVICE;
}
}  // namespace

extern "C" {

const char* lm_last_error(void) { return g_last_error.c_str(); }
const char* lm_version(void) { return "1.1"; }
const char* lm_precision(void) { return "double"; }

int lm_nbody_create(int px, int py, int pz, lm_nbody** out) {
  return guarded([&] {
    if (!out) throw lm::Error("[libmobility_amd] null output handle");
    *out = nullptr;
    *out = new lm_nbody(px, py, pz);
  });
}

int lm_nbody_destroy(lm_nbody* h) {
  return guarded([&] { delete h; });
}

int lm_nbody_set_parameters(lm_nbody* h, int algorithm, int nbatch, int nperbatch,
                            int has_wall_height, double wall_height, double delta) {
  return guarded([&] {
    S(h).setParameters(algorithm, nbatch, nperbatch, has_wall_height != 0, wall_height, delta);
  });
}

int lm_nbody_set_pair_precision(lm_nbody* h, int precision) {
  return guarded([&] { S(h).setPairPrecision(precision); });
}

int lm_nbody_set_drift_precision(lm_nbody* h, int precision) {
  return guarded([&] { S(h).setDriftPrecision(precision); });
}

int lm_nbody_set_pair_tiling(lm_nbody* h, int tpl, int nch) {
  return guarded([&] { S(h).setPairTiling(tpl, nch); });
}

int lm_nbody_set_pair_symmetric(lm_nbody* h, int mode) {
  return guarded([&] { S(h).setPairSymmetric(mode); });
}

int lm_nbody_set_pair_symmetric_memory(lm_nbody* h, int64_t bytes) {
  return guarded([&] { S(h).setPairSymmetricMemory(bytes); });
}

int lm_nbody_last_pair_symmetric(lm_nbody* h, int* used) {
  return guarded([&] {
    if (!used) throw lm::Error("[libmobility_amd] null output");
    *used = S(h).lastPairSymmetric() ? 1 : 0;
  });
}

int lm_nbody_set_autotune(lm_nbody* h, int enable) {
  return guarded([&] { S(h).setAutotune(enable != 0); });
}

int lm_nbody_get_pair_tiling(lm_nbody* h, int* tpl, int* nch) {
  return guarded([&] {
    if (!tpl || !nch) throw lm::Error("[libmobility_amd] null output");
    S(h).lastPairTiling(tpl, nch);
  });
}

int lm_nbody_set_stream(lm_nbody* h, void* stream) {
  return guarded([&] { S(h).setStream(static_cast<hipStream_t>(stream)); });
}

int lm_nbody_initialize(lm_nbody* h, double viscosity, double radius, int include_angular,
                        double tolerance, uint64_t seed, lm_lanczos_callback cb, void* user) {
  return guarded([&] {
    S(h).initialize(viscosity, radius, include_angular != 0, tolerance, seed, cb, user);
  });
}

int lm_nbody_set_positions(lm_nbody* h, const double* pos, int64_t n3, int mem) {
  return guarded([&] { S(h).setPositions(pos, n3, is_dev(mem)); });
}

int lm_nbody_set_systems(lm_nbody* h, const int64_t* offsets, int nsys, const int* kinds) {
  return guarded([&] {
    if (!offsets) throw lm::Error("[libmobility_amd] null offsets");
    S(h).setSystems(offsets, nsys, kinds);
  });
}

int lm_nbody_set_persistent(lm_nbody* h, int mode) {
  return guarded([&] {
    if (mode < -1 || mode > 1) throw lm::Error("[libmobility_amd] persistent mode must be -1, 0 or 1");
    S(h).setPersistent(mode);
  });
}

int lm_nbody_set_lanczos_iteration_limit(lm_nbody* h, int limit) {
  return guarded([&] { S(h).setLanczosIterationLimit(limit); });
}

int lm_nbody_set_lanczos_reorthogonalization(lm_nbody* h, int mode) {
  return guarded([&] { S(h).setLanczosReorthogonalization(mode); });
}

int lm_nbody_set_lanczos_per_system(lm_nbody* h, int on) {
  return guarded([&] { S(h).setLanczosPerSystem(on != 0); });
}

int lm_nbody_last_lanczos_iterations_per_system(lm_nbody* h, int* out, int nsys) {
  return guarded([&] {
    if (!out || nsys != S(h).numSystems())
      throw lm::Error("[libmobility_amd] output must hold one entry per system");
    S(h).lastLanczosIterationsPerSystem(out, nsys);
  });
}

int lm_nbody_mdot_multi(lm_nbody* h, const double* forces, int nrhs, double* linear, int mem) {
  return guarded([&] { S(h).MdotMulti(forces, nrhs, linear, is_dev(mem)); });
}

int lm_nbody_sqrt_mdot_w_multi(lm_nbody* h, const double* noise, int nrhs, double* linear,
                               double prefactor, int mem) {
  return guarded([&] { S(h).sqrtMdotWMulti(linear, nrhs, prefactor, is_dev(mem), noise); });
}

int lm_nbody_last_lanczos_iterations_multi(lm_nbody* h, int* out, int nrhs) {
  return guarded([&] {
    if (!out || nrhs < 1) throw lm::Error("[libmobility_amd] invalid output for iteration counts");
    S(h).lastLanczosIterationsMulti(out, nrhs);
  });
}

int lm_nbody_num_systems(lm_nbody* h) { return h ? h->solver.numSystems() : -1; }

int64_t lm_nbody_num_particles(lm_nbody* h) { return h ? h->solver.numParticles() : -1; }

int lm_nbody_mdot(lm_nbody* h, const double* F, int64_t nF, const double* T, int64_t nT,
                  double* U, double* W, int mem) {
  return guarded([&] { S(h).Mdot(F, F ? nF : 0, T, T ? nT : 0, U, W, is_dev(mem)); });
}

int lm_nbody_mdot_noalloc(lm_nbody* h, const double* F, const double* T, double* U, double* W) {
  return guarded([&] { S(h).mdotNoAlloc(F, T, U, W); });
}

int lm_nbody_sqrt_mdot_w(lm_nbody* h, double* U, double* W, double prefactor, int mem) {
  return guarded([&] { S(h).sqrtMdotW(U, W, prefactor, is_dev(mem), nullptr); });
}

int lm_nbody_sqrt_mdot_w_noise(lm_nbody* h, const double* noise, double* U, double* W,
                               double prefactor, int mem) {
  return guarded([&] {
    if (!noise) throw lm::Error("[libmobility_amd] null noise");
    S(h).sqrtMdotW(U, W, prefactor, is_dev(mem), noise);
  });
}

int lm_nbody_div_m(lm_nbody* h, double* U, double* W, double prefactor, int mem) {
  return guarded([&] { S(h).divM(U, W, prefactor, is_dev(mem)); });
}

int lm_nbody_div_m_noise(lm_nbody* h, const double* noise, double* U, double* W,
                         double prefactor, int mem) {
  return guarded([&] {
    if (!noise) throw lm::Error("[libmobility_amd] null noise");
    S(h).divM(U, W, prefactor, is_dev(mem), noise);
  });
}

int lm_nbody_langevin_velocities(lm_nbody* h, double dt, double kbt, const double* F, int64_t nF,
                                 const double* T, int64_t nT, double* U, double* W, int mem) {
  return guarded(
      [&] { S(h).langevin(dt, kbt, F, F ? nF : 0, T, T ? nT : 0, U, W, is_dev(mem)); });
}

int lm_nbody_hydrodynamic_velocities(lm_nbody* h, const double* F, int64_t nF, const double* T,
                                     int64_t nT, double* U, double* W, int mem) {
  return lm_nbody_langevin_velocities(h, 2.0, 1.0, F, nF, T, nT, U, W, mem);
}

int lm_nbody_clean(lm_nbody* h) {
  return guarded([&] { S(h).clean(); });
}

int lm_nbody_last_lanczos_iterations(lm_nbody* h) { return h ? h->solver.lastLanczosIterations() : -1; }

int lm_nbody_synchronize(lm_nbody* h) {
  return guarded([&] { LM_HIP(hipStreamSynchronize(S(h).stream())); });
}

int lm_precond_create(lm_precond** out) { return lm_precond_create_blocks(3, out); }

int lm_precond_create_blocks(int block_dim, lm_precond** out) {
  return guarded([&] {
    if (!out) throw lm::Error("[libmobility_amd] null output handle");
    *out = nullptr;
    if (block_dim != 3 && block_dim != 6)
      throw lm::Error("[libmobility_amd] block size must be 3 or 6");
    *out = new lm_precond(block_dim == 6 ? 1 : 0);
  });
}

int lm_precond_destroy(lm_precond* p) {
  return guarded([&] { delete p; });
}

int64_t lm_precond_num_weights(void) { return lm::precond_layout(0).total; }

static lm::LearnedPreconditioner& P(lm_precond* p) {
  if (!p) throw lm::Error("[libmobility_amd] null preconditioner handle");
  return p->net;
}

int lm_precond_describe(lm_precond* p, int* block_dim, int64_t* num_weights, int* inputs,
                        int* outputs) {
  return guarded([&] {
    if (block_dim) *block_dim = P(p).blockDim();
    if (num_weights) *num_weights = P(p).numWeights();
    if (inputs) *inputs = lm::kPrecondIn;
    if (outputs) *outputs = P(p).outputs();
  });
}

int lm_precond_init_weights(lm_precond* p, uint64_t seed) {
  return guarded([&] { P(p).initWeights(seed); });
}

int lm_precond_set_weights(lm_precond* p, const float* w, int64_t n) {
  return guarded([&] { P(p).setWeights(w, n); });
}

int lm_precond_get_weights(lm_precond* p, float* w, int64_t n) {
  return guarded([&] { P(p).getWeights(w, n); });
}

int lm_precond_train(lm_precond* p, const float* X, const float* T9, int64_t n, int epochs,
                     int batch, float lr, uint64_t seed, float* loss_history) {
  return guarded([&] { P(p).train(X, T9, n, epochs, batch, lr, seed, loss_history, nullptr); });
}

int lm_precond_forward(lm_precond* p, const float* X, int64_t n, float* out) {
  return guarded([&] {
    const int no = P(p).outputs();
    lm::DevBuf x, o;
    x.reserve(sizeof(float) * n * lm::kPrecondIn);
    o.reserve(sizeof(float) * n * no);
    LM_HIP(hipMemcpy(x.as<float>(), X, sizeof(float) * n * lm::kPrecondIn, hipMemcpyHostToDevice));
    P(p).forward(x.as<float>(), (int)n, o.as<float>(), nullptr);
    LM_HIP(hipMemcpy(out, o.as<float>(), sizeof(float) * n * no, hipMemcpyDeviceToHost));
  });
}

int lm_precond_loss_grad(lm_precond* p, const float* X, const float* T, int64_t n, float* grad,
                         float* loss) {
  return guarded([&] {
    const int td = P(p).blockDim() * P(p).blockDim();
    lm::DevBuf x, t;
    x.reserve(sizeof(float) * n * lm::kPrecondIn);
    t.reserve(sizeof(float) * n * td);
    LM_HIP(hipMemcpy(x.as<float>(), X, sizeof(float) * n * lm::kPrecondIn, hipMemcpyHostToDevice));
    LM_HIP(hipMemcpy(t.as<float>(), T, sizeof(float) * n * td, hipMemcpyHostToDevice));
    const float l = P(p).lossGrad(x.as<float>(), t.as<float>(), (int)n, nullptr);
    if (loss) *loss = l;
    LM_HIP(hipMemcpy(grad, P(p).grad(), sizeof(float) * P(p).numWeights(), hipMemcpyDeviceToHost));
  });
}

int lm_nbody_precond_dataset(lm_nbody* h, int block_dim, int64_t ncenters, uint64_t seed,
                             float* X, float* T, int* status) {
  return guarded([&] {
    if (!X || !T) throw lm::Error("[libmobility_amd] null output");
    if (block_dim != 3 && block_dim != 6)
      throw lm::Error("[libmobility_amd] block size must be 3 or 6");
    S(h).precondDataset(block_dim == 6 ? 1 : 0, ncenters, seed, X, T, status);
  });
}

int lm_nbody_set_preconditioner(lm_nbody* h, lm_precond* p) {
  return guarded([&] { S(h).setPreconditioner(p ? &p->net : nullptr); });
}

int lm_nbody_precond_blocks(lm_nbody* h, double* G9) {
  return guarded([&] {
    if (!G9) throw lm::Error("[libmobility_amd] null output");
    S(h).precondBlocks(G9);
  });
}

int lm_nbody_precond_neighbours(lm_nbody* h, int* nbr) {
  return guarded([&] {
    if (!nbr) throw lm::Error("[libmobility_amd] null output");
    S(h).precondNeighbours(nbr);
  });
}

int lm_nbody_set_coarse_level(lm_nbody* h, int particles_per_cell) {
  return guarded([&] { S(h).setCoarseLevel(particles_per_cell); });
}

int lm_nbody_set_coarse_pair_precision(lm_nbody* h, int precision) {
  return guarded([&] {
    if (precision != LM_PAIR_FP64 && precision != LM_PAIR_FP32)
      throw lm::Error("[libmobility_amd] pair precision must be LM_PAIR_FP64 or LM_PAIR_FP32");
    S(h).setCoarseFP64(precision == LM_PAIR_FP64);
  });
}

int lm_nbody_coarse_used_fp64(lm_nbody* h, int* used) {
  return guarded([&] { *used = S(h).coarseUsedFP64() ? 1 : 0; });
}

int lm_nbody_set_coarse_max_dim(lm_nbody* h, int64_t max_dim) {
  return guarded([&] { S(h).setCoarseMaxDim(max_dim); });
}

int lm_nbody_coarse_cells(lm_nbody* h, int* ncells, int* grid, int* cell_of) {
  return guarded([&] {
    if (!ncells || !grid) throw lm::Error("[libmobility_amd] null output");
    *ncells = S(h).coarseCells(cell_of, grid);
  });
}

int lm_nbody_coarse_factors(lm_nbody* h, int system, double* L, double* Linv) {
  return guarded([&] {
    if (!L || !Linv) throw lm::Error("[libmobility_amd] null output");
    S(h).coarseFactors(system, L, Linv);
  });
}

int lm_nbody_coarse_system_cells(lm_nbody* h, int system, int* ncells) {
  return guarded([&] {
    if (!ncells) throw lm::Error("[libmobility_amd] null output");
    *ncells = S(h).coarseSystemCells(system);
  });
}

int lm_nbody_coarse_build_ms(lm_nbody* h, double* ms) {
  return guarded([&] {
    if (!ms) throw lm::Error("[libmobility_amd] null output");
    *ms = S(h).coarseBuildMs();
  });
}

int lm_nbody_coarse_builds(lm_nbody* h, int64_t* count) {
  return guarded([&] {
    if (!count) throw lm::Error("[libmobility_amd] null output");
    *count = S(h).coarseBuilds();
  });
}

int lm_nbody_set_preconditioner_refresh(lm_nbody* h, int interval) {
  return guarded([&] { S(h).setPreconditionerRefresh(interval); });
}

int lm_selftest_pair_block_host(int wall, int fp32, const double* pi, const double* pj, double a,
                                double* blk) {
  return guarded([&] {
    auto run = [&](auto zero) {
      using R = decltype(zero);
      R ri[4], rj[4];
      const double* p[2] = {pi, pj};
      R* r[2] = {ri, rj};
      for (int q = 0; q < 2; ++q) {
        const double x = p[q][0] / a, y = p[q][1] / a, z = p[q][2];
        if (wall) {
          r[q][0] = R(x);
          r[q][1] = R(y);
          r[q][2] = R((z > a ? z : a) / a);
          const double b = z / a;
          r[q][3] = R(b < 0 ? 0 : (b > 1 ? 1 : b));
        } else {
          r[q][0] = R(x);
          r[q][1] = R(y);
          r[q][2] = R(z / a);
          r[q][3] = R(1);
        }
      }
      for (int col = 0; col < 6; ++col) {
        R v[6] = {0, 0, 0, 0, 0, 0};
        v[col] = R(1);
        lm::PairAcc<R> acc;
        lm::acc_zero(acc);
        constexpr int all = lm::kUF | lm::kUT | lm::kWT | lm::kWF;
        if (wall)
          lm::wall_pair<R, all>(ri[0], ri[1], ri[2], ri[3], rj[0], rj[1], rj[2], rj[3], v[0], v[1],
                                v[2], v[3], v[4], v[5], acc);
        else
          lm::open_pair<R, all>(ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2], v[0], v[1], v[2], v[3],
                                v[4], v[5], acc);
        const double u[6] = {double(acc.uf.x + acc.ut.x), double(acc.uf.y + acc.ut.y),
                             double(acc.uf.z + acc.ut.z), double(acc.wt.x + acc.wf.x),
                             double(acc.wt.y + acc.wf.y), double(acc.wt.z + acc.wf.z)};
        for (int row = 0; row < 6; ++row) blk[row * 6 + col] = u[row];
      }
    };
    if (fp32)
      run(0.0f);
    else
      run(0.0);
  });
}

int lm_selftest_pair_both_host(int wall, const double* pi, const double* pj, double a,
                               double* fwd, double* rev) {
  return guarded([&] {
    if (!pi || !pj || !fwd || !rev) throw lm::Error("[libmobility_amd] null argument");
    double r[2][4];
    const double* p[2] = {pi, pj};
    for (int q = 0; q < 2; ++q) {
      const double z = p[q][2];
      const double b = z / a;
      r[q][0] = p[q][0] / a;
      r[q][1] = p[q][1] / a;
      r[q][2] = wall ? (z > a ? z : a) / a : z / a;
      r[q][3] = wall ? (b < 0 ? 0 : (b > 1 ? 1 : b)) : 1.0;
    }
    constexpr int all = lm::kUF | lm::kUT | lm::kWT | lm::kWF;
    const lm::Vec3<double> zero = {0, 0, 0};
    for (int col = 0; col < 6; ++col) {
      double v[6] = {0, 0, 0, 0, 0, 0};
      v[col] = 1.0;
      const lm::Vec3<double> f = {v[0], v[1], v[2]}, t = {v[3], v[4], v[5]};
      // forward: unit force/torque on j, velocities of i; reverse: unit
      // force/torque on i, what the same pair gives j (transposed blocks)
      lm::PairAcc<double> acc;
      lm::acc_zero(acc);
      lm::Vec3<double> su = zero, sw = zero, du = zero, dw = zero;
      lm::PairAcc<double> dummy;
      lm::acc_zero(dummy);
      if (wall) {
        lm::pair_both<double, true, all>(r[0][0], r[0][1], r[0][2], r[0][3], r[1][0], r[1][1],
                                         r[1][2], r[1][3], f, t, zero, zero, 1.0, 1.0, acc, du, dw);
        lm::pair_both<double, true, all>(r[0][0], r[0][1], r[0][2], r[0][3], r[1][0], r[1][1],
                                         r[1][2], r[1][3], zero, zero, f, t, 1.0, 1.0, dummy, su, sw);
      } else {
        lm::pair_both<double, false, all>(r[0][0], r[0][1], r[0][2], r[0][3], r[1][0], r[1][1],
                                          r[1][2], r[1][3], f, t, zero, zero, 1.0, 1.0, acc, du, dw);
        lm::pair_both<double, false, all>(r[0][0], r[0][1], r[0][2], r[0][3], r[1][0], r[1][1],
                                          r[1][2], r[1][3], zero, zero, f, t, 1.0, 1.0, dummy, su,
                                          sw);
      }
      const double u[6] = {acc.uf.x + acc.ut.x, acc.uf.y + acc.ut.y, acc.uf.z + acc.ut.z,
                           acc.wt.x + acc.wf.x, acc.wt.y + acc.wf.y, acc.wt.z + acc.wf.z};
      const double s6[6] = {su.x, su.y, su.z, sw.x, sw.y, sw.z};
      for (int row = 0; row < 6; ++row) {
        fwd[row * 6 + col] = u[row];
        rev[row * 6 + col] = s6[row];
      }
    }
  });
}

// ---- SelfMobility ------------------------------------------------------------

int lm_self_create(int px, int py, int pz, lm_self** out) {
  return guarded([&] {
    if (!out) throw lm::Error("[libmobility_amd] null output");
    *out = nullptr;
    *out = new lm_self(px, py, pz);
  });
}

int lm_self_destroy(lm_self* h) {
  return guarded([&] { delete h; });
}

int lm_self_set_stream(lm_self* h, void* stream) {
  return guarded([&] { SS(h).setStream(static_cast<hipStream_t>(stream)); });
}

int lm_self_initialize(lm_self* h, double viscosity, double radius, int include_angular,
                       double tolerance, uint64_t seed) {
  return guarded(
      [&] { SS(h).initialize(viscosity, radius, include_angular != 0, tolerance, seed); });
}

int lm_self_set_positions(lm_self* h, const double* positions, int64_t n, int mem) {
  (void)positions;
  (void)mem;
  return guarded([&] { SS(h).setPositions(n); });
}

int64_t lm_self_num_particles(lm_self* h) { return h ? h->solver.numParticles() : 0; }

int lm_self_mdot(lm_self* h, const double* F, int64_t nF, const double* T, int64_t nT, double* U,
                 double* W, int mem) {
  return guarded([&] { SS(h).Mdot(F, F ? nF : 0, T, T ? nT : 0, U, W, is_dev(mem)); });
}

int lm_self_sqrt_mdot_w(lm_self* h, double* U, double* W, double prefactor, int mem) {
  return guarded([&] {
    if (prefactor == 0) return;
    SS(h).sqrtMdotW(U, W, prefactor, is_dev(mem));
  });
}

int lm_self_div_m(lm_self* h, double* U, double* W, double prefactor, int mem) {
  (void)U;
  (void)W;
  (void)prefactor;
  (void)mem;
  return guarded([&] { SS(h); });
}

int lm_self_langevin_velocities(lm_self* h, double dt, double kbt, const double* F, int64_t nF,
                                const double* T, int64_t nT, double* U, double* W, int mem) {
  return guarded(
      [&] { SS(h).langevin(dt, kbt, F, F ? nF : 0, T, T ? nT : 0, U, W, is_dev(mem)); });
}

int lm_self_clean(lm_self* h) {
  return guarded([&] { SS(h).clean(); });
}

int lm_selftest_tridiag_sqrt(const double* alpha, const double* beta, int m, double* out) {
  return guarded([&] {
    if (!alpha || !out || (m > 1 && !beta)) throw lm::Error("[libmobility_amd] null array");
    lm::tridiag_sqrt_selftest(alpha, beta, m, out);
  });
}

int lm_selftest_philox_host(const uint32_t* ctr4, const uint32_t* key2, uint32_t* out4) {
  return guarded([&] {
    lm::U32x4 c;
    for (int i = 0; i < 4; ++i) c.v[i] = ctr4[i];
    const lm::U32x4 r = lm::philox4x32_10(c, key2[0], key2[1]);
    for (int i = 0; i < 4; ++i) out4[i] = r.v[i];
  });
}

int lm_selftest_coarse_cells_host(const double* xyz, int64_t N, int particles_per_cell,
                                  int* ncells, int* grid, int* cell_of) {
  return guarded([&] {
    if (!xyz || !ncells || !grid || !cell_of) throw lm::Error("[libmobility_amd] null argument");
    lm::CoarseCells c;
    lm::coarse_cells(xyz, 3, N, particles_per_cell, c, lm::CoarseLevel::kMaxDim);
    *ncells = c.C;
    for (int k = 0; k < 3; ++k) grid[k] = c.g[k];
    std::copy(c.cell_of.begin(), c.cell_of.end(), cell_of);
  });
}

}  // extern "C"
