// SelfMobility on the device (solvers/SelfMobility/mobility.h:17-111).
#include "self_solver.hpp"

#include <cmath>

#include "kernels.hpp"

namespace lm {

SelfSolver::SelfSolver(int px, int py, int pz) {
  // periodicity_mode::open == 2 (MobilityInterface.h:27-33)
  if (px != 2 || py != 2 || pz != 2) throw Error("[Mobility] This is an open boundary solver");
}

void SelfSolver::initialize(double viscosity, double radius, bool angular, double tolerance,
                            uint64_t seed) {
  (void)tolerance;  // no Lanczos: the square root of a diagonal is exact
  if (seed == 0) seed = std::random_device()();
  engine_ = std::mt19937((uint32_t)seed);
  lin_ = 1.0 / (6.0 * M_PI * viscosity * radius);
  ang_ = 1.0 / (8.0 * M_PI * viscosity * radius * radius * radius);
  angular_ = angular;
  initialized_ = true;
}

void SelfSolver::setPositions(int64_t entries) {
  if (entries % 3 != 0)
    throw Error("[libMobility] The input arrays for positions, forces, and torques must each have "
                "total size 3*N.");
  N_ = entries / 3;
}

void SelfSolver::require_particles() const {
  if (N_ <= 0)
    throw Error(
        "[libMobility] The number of particles is not set. Did you forget to call setPositions?");
}

void SelfSolver::Mdot(const double* F, int64_t nF, const double* T, int64_t nT, double* U,
                      double* W, bool dev) {
  const int64_t n3 = 3 * N_;
  if (nF > 0 && nF != n3) throw Error("The forces array must have size 3*N.");
  if (nT > 0 && nT != n3) throw Error("The torques array must have size 3*N.");
  if (!angular_ && nT > 0)
    throw Error(
        "The solver was configured without including angular velocities. Set includeAngular "
        "to true when initializing if you want to use torques");
  auto scale = [&](const double* x, double* y, double m, DevBuf& sx, DevBuf& sy) {
    if (dev) {
      launch_zero(y, n3, stream_);
      launch_axpy_accumulate(y, x, m, n3, stream_);
      return;
    }
    sx.reserve(sizeof(double) * n3);
    sy.reserve(sizeof(double) * n3);
    LM_HIP(hipMemcpyAsync(sx.as<void>(), x, sizeof(double) * n3, hipMemcpyHostToDevice, stream_));
    launch_zero(sy.as<double>(), n3, stream_);
    launch_axpy_accumulate(sy.as<double>(), sx.as<double>(), m, n3, stream_);
    LM_HIP(hipMemcpyAsync(y, sy.as<void>(), sizeof(double) * n3, hipMemcpyDeviceToHost, stream_));
  };
  if (nF > 0) {
    if (!U) throw Error("[libmobility_amd] a linear output array is required");
    scale(F, U, lin_, st_f_, st_u_);
  }
  if (nT > 0) {
    if (!W) throw Error("[libmobility_amd] an angular output array is required");
    scale(T, W, ang_, st_t_, st_w_);
  }
  LM_HIP(hipStreamSynchronize(stream_));
}

// Per call two 32-bit seeds from the solver's mt19937 select a Philox stream.
void SelfSolver::noise_into(double* out, int64_t n, double scale) {
  const uint32_t s1 = std::uniform_int_distribution<uint32_t>(0, UINT32_MAX)(engine_);
  const uint32_t s2 = std::uniform_int_distribution<uint32_t>(0, UINT32_MAX)(engine_);
  noise_.reserve(sizeof(double) * n);
  launch_normal_fill(noise_.as<double>(), n, ((uint64_t)s2 << 32) | s1, 0, stream_);
  launch_axpy_accumulate(out, noise_.as<double>(), scale, n, stream_);
}

void SelfSolver::sqrtMdotW(double* U, double* W, double prefactor, bool dev) {
  if (!initialized_) throw Error("[libMobility] You must initialize the solver first");
  require_particles();
  const int64_t n3 = 3 * N_;
  auto one = [&](double* y, double m, DevBuf& sy) {
    if (!y) return;
    if (dev) {
      noise_into(y, n3, prefactor * std::sqrt(m));
      return;
    }
    sy.reserve(sizeof(double) * n3);
    LM_HIP(hipMemcpyAsync(sy.as<void>(), y, sizeof(double) * n3, hipMemcpyHostToDevice, stream_));
    noise_into(sy.as<double>(), n3, prefactor * std::sqrt(m));
    LM_HIP(hipMemcpyAsync(y, sy.as<void>(), sizeof(double) * n3, hipMemcpyDeviceToHost, stream_));
  };
  one(U, lin_, st_u_);
  one(angular_ ? W : nullptr, ang_, st_w_);
  LM_HIP(hipStreamSynchronize(stream_));
}

void SelfSolver::langevin(double dt, double kbt, const double* F, int64_t nF, const double* T,
                          int64_t nT, double* U, double* W, bool dev) {
  if (nF > 0 || nT > 0) Mdot(F, nF, T, nT, U, W, dev);
  if (kbt != 0) sqrtMdotW(U, W, std::sqrt(2.0 * kbt / dt), dev);
}

void SelfSolver::clean() {
  st_f_.release();
  st_t_.release();
  st_u_.release();
  st_w_.release();
  noise_.release();
}

}  // namespace lm
