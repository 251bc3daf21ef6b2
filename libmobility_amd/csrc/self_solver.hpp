// SelfMobility (solvers/SelfMobility/mobility.h:17-111): no hydrodynamic
// interactions.  The mobility is diagonal, 1/(6 pi eta a) for linear and
// 1/(8 pi eta a^3) for angular velocities; the square root is applied to
// Gaussian noise drawn on the device.  Host arrays are staged through device
// buffers; device arrays are used in place.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <random>

#include "lm_common.hpp"

namespace lm {

class SelfSolver {
 public:
  // Open boundaries only (mobility.h:31-36).
  SelfSolver(int px, int py, int pz);
  void setStream(hipStream_t s) { stream_ = s; }
  // mobility.h:38-50 and Mobility::initialize (MobilityInterface.h:102-114).
  void initialize(double viscosity, double radius, bool angular, double tolerance, uint64_t seed);
  // Only the particle count is kept (mobility.h:56-59).
  void setPositions(int64_t entries);
  int64_t numParticles() const { return N_; }
  // linear = F / (6 pi eta a) when forces are given, angular = T / (8 pi eta a^3)
  // when torques are given (mobility.h:63-86).
  void Mdot(const double* F, int64_t nF, const double* T, int64_t nT, double* U, double* W,
            bool dev);
  // linear (and angular, when given) += prefactor * sqrt(mobility) * N(0,1)
  // (mobility.h:90-110).
  void sqrtMdotW(double* U, double* W, double prefactor, bool dev);
  // Mdot, then with kbt != 0 the noise with prefactor sqrt(2 kbt / dt); the
  // drift of this solver is zero (MobilityInterface.h:191-216).
  void langevin(double dt, double kbt, const double* F, int64_t nF, const double* T, int64_t nT,
                double* U, double* W, bool dev);
  void clean();

 private:
  void require_particles() const;
  void noise_into(double* out, int64_t n, double scale);  // device: out += scale * N(0,1)
  hipStream_t stream_ = nullptr;
  bool initialized_ = false, angular_ = false;
  double lin_ = 0.0, ang_ = 0.0;
  int64_t N_ = 0;
  std::mt19937 engine_;
  DevBuf st_f_, st_t_, st_u_, st_w_, noise_;
};

}  // namespace lm
