// Learned block preconditioner: constants, kernel launchers, host class.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "lm_common.hpp"

namespace lm {

constexpr int kPrecondK = 16;               // neighbours in the features
constexpr int kPrecondGeom = 3 * kPrecondK;  // 48 relative coordinates
// Features per particle: the 48 coordinates, then the wall features of the
// particle (u = a / max(z, a), the damping b = clamp(z / a, 0, 1), u^3; for an
// open system u = 0, b = 1) and zeros up to a multiple of the MFMA K slice.
constexpr int kPrecondIn = 64;
constexpr int kPrecondH = 128;               // hidden width
constexpr int kPrecondMaxOut = 21;           // Cholesky entries of a 6x6 block
constexpr float kPrecondEps = 1e-3f;         // floor of the factor's diagonal
// Block kinds: 3x3 blocks on the linear entries of a particle (solvers
// without angular velocities), or 6x6 blocks coupling its linear and angular
// entries (solvers with them).  Targets: the center block of the local
// (M / t0)^{-1/2}, over the particle and kPrecondK (3x3) or
// kPrecondK6 (6x6) neighbours.
constexpr int kPrecondK6 = 8;
constexpr int precond_dim(int kind) { return kind ? 6 : 3; }
constexpr int precond_out(int kind) { return kind ? 21 : 6; }
// FP32 weights, one flat array: W1[64][128] b1[128] W2[128][128] b2[128]
// W3[128][out] b3[out]
struct PrecondLayout {
  int64_t W1, B1, W2, B2, W3, B3, total;
};
constexpr PrecondLayout precond_layout(int kind) {
  const int64_t w1 = 0, b1 = w1 + (int64_t)kPrecondIn * kPrecondH, w2 = b1 + kPrecondH,
                b2 = w2 + (int64_t)kPrecondH * kPrecondH, w3 = b2 + kPrecondH,
                b3 = w3 + (int64_t)kPrecondH * precond_out(kind);
  return PrecondLayout{w1, b1, w2, b2, w3, b3, b3 + precond_out(kind)};
}

// Uniform grid of one system for the cell-list kNN: cell (qx, qy, qz) covers
// [lo + q edge, lo + (q + 1) edge) per axis, local id (qx g1 + qy) g2 + qz,
// global id cell0 + local id.
struct KnnSys {
  double lo[3], edge[3];
  int g[3];
  int64_t cell0;
};
// Host side: grids of all systems (about 8 particles per cell), each
// particle's local cell id, the particles sorted by (system, cell, index)
// and the cells' first positions in that order (total cells + 1 entries).
// xyz: 4 doubles per particle (the FP64 records).
void knn_grid(const double* xyz4, const int64_t* sys_off, int nsys, std::vector<KnnSys>& sys,
              std::vector<int>& cell_of, std::vector<int>& perm, std::vector<int64_t>& start);
// The K nearest neighbours of each particle within its system (K smallest
// FP32 squared distances, ties by index; -1 past the system's other
// particles), searching shells of grid cells until the K-th is closer than
// every unsearched cell; feat receives kPrecondIn features per particle.
// kinds[s] (0 open, 1 wall) per system, nullptr: all `wall`.
void launch_knn_cells(const double4* rec, const int64_t* sys_off, int nsys, const KnnSys* sys,
                      const int* cell_of, const int* perm, const int64_t* start, int64_t N,
                      const int* kinds, int wall, float* feat, int* nbr, hipStream_t s);
// Training targets: for each center, the center d x d block (d = 3 or 6 by
// `kind`) of (M_local / t0)^{-1/2} in FP64 (wall[c] != 0: bottom-wall
// mobility), status[c] = 0 when the Newton-Schulz iteration converged.
void launch_local_targets(const double4* rec, const int* nbr, const int64_t* centers,
                          const int* wall, int64_t ncenters, int kind, float* target,
                          int* status, hipStream_t s);
void launch_gemm_bf16(int M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B,
                      int64_t sbk, int64_t sbn, float* C, int64_t ldc, const float* bias, int gelu,
                      float* Zout, int accumulate, hipStream_t s);
// Loss mean_b sum_ij (G_b - T_b)_ij^2 of the d x d blocks G = L L^T and its
// gradient with respect to the network outputs.
void launch_head_loss(const float* out, const float* tgt, int Bn, int kind, float* dout,
                      float* loss_part, hipStream_t s);
void launch_gelu_back(const float* dH, const float* Z, int64_t n, float* dZ, hipStream_t s);
void launch_bias_grad(const float* dZ, int Bn, int Nc, float* db, hipStream_t s);
void launch_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, int step,
                 hipStream_t s);
void launch_gather_rows(const float* src, const int64_t* idx, int rows, int cols, float* dst,
                        hipStream_t s);
// Per particle: 3x3 kind, G = L L^T and G^{-1} (9 doubles each); 6x6 kind,
// G = L L^T S and G^{-1} = S^{-1} (L L^T)^{-1} with S = diag(1, 1, 1, c, c, c),
// c = 2a / sqrt(3), which maps the dimensionless local problem (linear and
// angular entries scaled by sqrt(t0) and sqrt(r0)) back to physical units
// (36 doubles each).  Particles without a full neighbourhood get G = I.
void launch_make_blocks(const float* out, const int* nbr, int64_t N, int kind, double a,
                        double* G, double* Gi, hipStream_t s);
// y = B x per particle (transpose: B^T x); 3x3 kind on the linear entries
// [3i, 3i + 3) (entries past 3N copied), 6x6 kind on [3i, 3i + 3) and
// [3N + 3i, 3N + 3i + 3) of a 6N vector.
void launch_apply_blocks(const double* B, const double* x, int64_t N, int64_t n, int kind,
                         int transpose, double* y, hipStream_t s);

// Network weights + training state on the device.
class LearnedPreconditioner {
 public:
  explicit LearnedPreconditioner(int kind = 0);
  int kind() const { return kind_; }
  int blockDim() const { return precond_dim(kind_); }
  int outputs() const { return precond_out(kind_); }
  int64_t numWeights() const { return layout_.total; }
  void initWeights(uint64_t seed);
  void setWeights(const float* w_host, int64_t n);
  void getWeights(float* w_host, int64_t n);
  // Forward pass for Bn samples (device arrays): out = net(X).
  void forward(const float* X, int Bn, float* out, hipStream_t s);
  // Loss and parameter gradient for Bn samples (device arrays, targets of d*d
  // entries each); grad stays in g_.
  float lossGrad(const float* X, const float* T, int Bn, hipStream_t s);
  const float* grad() const { return g_.as<float>(); }
  // Mini-batch Adam over host arrays; loss_history[e] = mean loss of epoch e.
  void train(const float* X_host, const float* T_host, int64_t n, int epochs, int batch, float lr,
             uint64_t seed, float* loss_history, hipStream_t s);
  const float* weights() const { return w_.as<float>(); }
  // Incremented whenever the weights change (solvers cache their blocks).
  uint64_t version() const { return version_; }

 private:
  int kind_;
  PrecondLayout layout_;
  uint64_t version_ = 1;
  void ensure_batch(int Bn);
  DevBuf w_, g_, m_, v_;
  DevBuf z1_, h1_, z2_, h2_, out_, dout_, dh_, dz_, loss_part_;
  int cap_ = 0;
  int step_ = 0;
};

}  // namespace lm
