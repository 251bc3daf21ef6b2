// Learned block preconditioner: constants, kernel launchers, host class.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "lm_common.hpp"

namespace lm {

constexpr int kPrecondK = 16;             // neighbours
constexpr int kPrecondIn = 3 * kPrecondK; // 48 features
constexpr int kPrecondH = 128;            // hidden width
constexpr int kPrecondOut = 6;            // Cholesky factor entries
constexpr float kPrecondEps = 1e-3f;      // floor of the factor's diagonal
// FP32 weights, one flat array: W1[48][128] b1[128] W2[128][128] b2[128] W3[128][6] b3[6]
constexpr int64_t kOffW1 = 0;
constexpr int64_t kOffB1 = kOffW1 + (int64_t)kPrecondIn * kPrecondH;
constexpr int64_t kOffW2 = kOffB1 + kPrecondH;
constexpr int64_t kOffB2 = kOffW2 + (int64_t)kPrecondH * kPrecondH;
constexpr int64_t kOffW3 = kOffB2 + kPrecondH;
constexpr int64_t kOffB3 = kOffW3 + (int64_t)kPrecondH * kPrecondOut;
constexpr int64_t kPrecondNumWeights = kOffB3 + kPrecondOut;

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
// every unsearched cell.
void launch_knn_cells(const double4* rec, const int64_t* sys_off, int nsys, const KnnSys* sys,
                      const int* cell_of, const int* perm, const int64_t* start, int64_t N,
                      float* feat, int* nbr, hipStream_t s);
void launch_local_targets(const double4* rec, const int* nbr, const int64_t* centers,
                          int64_t ncenters, float* target9, int* status, hipStream_t s);
void launch_gemm_bf16(int M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B,
                      int64_t sbk, int64_t sbn, float* C, int64_t ldc, const float* bias, int gelu,
                      float* Zout, int accumulate, hipStream_t s);
void launch_head_loss(const float* out, const float* tgt9, int Bn, float* dout, float* loss_part,
                      hipStream_t s);
void launch_gelu_back(const float* dH, const float* Z, int64_t n, float* dZ, hipStream_t s);
void launch_bias_grad(const float* dZ, int Bn, int Nc, float* db, hipStream_t s);
void launch_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, int step,
                 hipStream_t s);
void launch_gather_rows(const float* src, const int64_t* idx, int rows, int cols, float* dst,
                        hipStream_t s);
void launch_make_blocks(const float* out, const int* nbr, int64_t N, double* G9, double* Gi9,
                        hipStream_t s);
void launch_apply_blocks(const double* B9, const double* x, int64_t N, int64_t n, double* y,
                         hipStream_t s);

// Network weights + training state on the device.
class LearnedPreconditioner {
 public:
  LearnedPreconditioner();
  void initWeights(uint64_t seed);
  void setWeights(const float* w_host, int64_t n);
  void getWeights(float* w_host, int64_t n);
  // Forward pass for Bn samples (device arrays): out = net(X).
  void forward(const float* X, int Bn, float* out, hipStream_t s);
  // Loss and parameter gradient for Bn samples (device arrays); grad stays in g_.
  float lossGrad(const float* X, const float* T9, int Bn, hipStream_t s);
  const float* grad() const { return g_.as<float>(); }
  // Mini-batch Adam over host arrays; loss_history[e] = mean loss of epoch e.
  void train(const float* X_host, const float* T_host, int64_t n, int epochs, int batch, float lr,
             uint64_t seed, float* loss_history, hipStream_t s);
  const float* weights() const { return w_.as<float>(); }
  // Incremented whenever the weights change (solvers cache their blocks).
  uint64_t version() const { return version_; }

 private:
  uint64_t version_ = 1;
  void ensure_batch(int Bn);
  DevBuf w_, g_, m_, v_;
  DevBuf z1_, h1_, z2_, h2_, out_, dout_, dh_, dz_, loss_part_;
  int cap_ = 0;
  int step_ = 0;
};

}  // namespace lm
