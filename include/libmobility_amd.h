/*
 * libmobility_amd -- C ABI of the MI355X NBody mobility solver.
 *
 * Drop-in boundary for libMobility's NBody solver and its default Lanczos
 * fluctuation path.  Every entry point below replaces one member of the
 * reference's C++ interface (paths relative to the reference repository);
 * the Python class libmobility_amd.NBody binds exactly these symbols.
 *
 * Conventions
 *   - All arrays are FP64, interleaved [x1,y1,z1,x2,...] (3N values), caller
 *     owned.  `mem` says where they live: LM_MEM_HOST (pageable host memory,
 *     staged through the device) or LM_MEM_DEVICE (HIP device memory).
 *   - Mdot overwrites its outputs; sqrtMdotW and divM accumulate (+=), as
 *     in MobilityInterface.h:179-187 and mobility.h:227-235.
 *   - Every function returns LM_OK or LM_ERROR; lm_last_error() then returns
 *     the message of the reference's std::runtime_error for the same
 *     condition (thread-local).
 *   - A handle is not thread-safe; all work is ordered on one HIP stream
 *     (default: the null stream, see lm_nbody_set_stream).
 */
#ifndef LIBMOBILITY_AMD_H
#define LIBMOBILITY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LM_OK 0
#define LM_ERROR (-1)

#define LM_MEM_HOST 0
#define LM_MEM_DEVICE 1

/* periodicity_mode, include/MobilityInterface/MobilityInterface.h:27-33 */
#define LM_PERIODICITY_SINGLE_WALL 0
#define LM_PERIODICITY_TWO_WALLS 1
#define LM_PERIODICITY_OPEN 2
#define LM_PERIODICITY_PERIODIC 3
#define LM_PERIODICITY_UNSPECIFIED 4

/* nbody_rpy::algorithm, solvers/NBody/extra/interface.h:10 */
#define LM_ALGORITHM_FAST 0
#define LM_ALGORITHM_NAIVE 1
#define LM_ALGORITHM_BLOCK 2
#define LM_ALGORITHM_ADVISE 3

/* Pair arithmetic (extension; the reference fixes it at compile time,
 * include/MobilityInterface/defines.h:5-13). */
#define LM_PAIR_FP64 0
#define LM_PAIR_FP32 1 /* FP32 pair terms, FP64 per-tile accumulation */

typedef struct lm_nbody lm_nbody;

/* Lanczos progress hook, Parameters::lanczosCallback
 * (MobilityInterface.h:42, LanczosAlgorithm.cu:187-189). */
typedef void (*lm_lanczos_callback)(int iteration, double residual, void* user);

/* Message of the last failed call on this thread ("" if none). */
const char* lm_last_error(void);

/* Interface version, LIBMOBILITYVERSION (include/MobilityInterface/defines.h:4). */
const char* lm_version(void);

/* I/O precision, Mobility::precision (MobilityInterface.h:73-77): "double". */
const char* lm_precision(void);

/* NBody(Configuration), solvers/NBody/mobility.h:37-47. */
int lm_nbody_create(int periodicity_x, int periodicity_y, int periodicity_z, lm_nbody** out);

/* Destroys the handle and frees its device memory. */
int lm_nbody_destroy(lm_nbody* h);

/* NBody::setParametersNBody, solvers/NBody/mobility.h:78-97
 * (Python: solvers/NBody/python_wrapper.cu:50-61).  delta is in units of the
 * hydrodynamic radius. */
int lm_nbody_set_parameters(lm_nbody* h, int algorithm, int nbatch, int nperbatch,
                            int has_wall_height, double wall_height, double delta);

/* Extension: pair arithmetic (LM_PAIR_FP64 default, LM_PAIR_FP32). */
int lm_nbody_set_pair_precision(lm_nbody* h, int precision);

/* Extension: pair arithmetic of the thermal drift's two products
 * (NBody::divM, mobility.h:184-236).  Default LM_PAIR_FP64 whatever
 * lm_nbody_set_pair_precision says: the random finite difference divides a
 * difference of two products by delta*a (random_finite_differences.h:54-62),
 * so FP32 pair terms amplify their ~1e-7 rounding by 1/delta, and FP32
 * positions far from the system centre cannot represent a displacement of
 * delta*a/2 at all. */
int lm_nbody_set_drift_precision(lm_nbody* h, int precision);

/* Extension: pair-kernel tiling.  targets_per_lane in {0 (auto), 1, 2, 4};
 * source_chunks >= 0 (0 = auto).  Fixes the work decomposition (the
 * reference's blockDim choice, NbodyRPY.cu:105-108). */
int lm_nbody_set_pair_tiling(lm_nbody* h, int targets_per_lane, int source_chunks);

/* Extension: symmetric evaluation of the product of a single open or
 * single-wall system, every block set (-1 auto, 0 off, 1 on).  The grand
 * mobility matrix is symmetric (M_ji = M_ij^T blockwise), so each unordered
 * pair is evaluated once and applied to both particles; the reference
 * evaluates every ordered pair (NbodyRPY.cu:27-97, hydrodynamicKernels.cuh).
 * Auto uses it from N = 16384.  Results agree with the ordered evaluation to
 * rounding and are bitwise reproducible. */
int lm_nbody_set_pair_symmetric(lm_nbody* h, int mode);

/* Extension: device memory for the symmetric evaluation's partial sums
 * (0 = 4 GiB).  The ring of tile offsets is split into passes that fit. */
int lm_nbody_set_pair_symmetric_memory(lm_nbody* h, int64_t bytes);

/* Extension: 1 if the last pair product used the symmetric evaluation. */
int lm_nbody_last_pair_symmetric(lm_nbody* h, int* used);

/* Extension: when enabled, the first product for a new (N, systems, kernel,
 * precision, blocks) combination times every tiling candidate on the device
 * and keeps the fastest for later calls. */
int lm_nbody_set_autotune(lm_nbody* h, int enable);

/* Extension: tiling used by the last pair-kernel launch. */
int lm_nbody_get_pair_tiling(lm_nbody* h, int* targets_per_lane, int* source_chunks);

/* Extension: order all work of this handle on `stream` (a hipStream_t). */
int lm_nbody_set_stream(lm_nbody* h, void* stream);

/* NBody::initialize + Mobility::initialize, mobility.h:99-114 and
 * MobilityInterface.h:102-114.  seed 0 draws from std::random_device. */
int lm_nbody_initialize(lm_nbody* h, double viscosity, double hydrodynamic_radius,
                        int include_angular, double tolerance, uint64_t seed,
                        lm_lanczos_callback callback, void* callback_user);

/* NBody::setPositions, mobility.h:116-162.  n3 = 3N. */
int lm_nbody_set_positions(lm_nbody* h, const double* positions, int64_t n3, int mem);

/* Extension: ragged independent systems.  offsets[0] = 0 <= ... <=
 * offsets[nsys] = N; pairs interact only inside a system (generalises the
 * equal-size batches of mobility.h:56-77).  kinds[s] = 0 (open RPY) or 1
 * (bottom wall at wallHeight) per system, or NULL for the solver's
 * periodicity; mixed ensembles run in one launch.  Applies to the next
 * lm_nbody_set_positions. */
int lm_nbody_set_systems(lm_nbody* h, const int64_t* offsets, int nsys, const int* kinds);

/* Extension: persistent pair launches (a fixed grid draining a work queue
 * of items sorted longest first).  mode: -1 auto (mixed ensembles), 0 off,
 * 1 on. */
int lm_nbody_set_persistent(lm_nbody* h, int mode);

/* lanczos::Solver::iterationHardLimit, LanczosAlgorithm.h:49 (a public member
 * there, default 200): iterations before "[Lanczos] Could not converge".
 * Range [3, 1024]; the default stays 200. */
int lm_nbody_set_lanczos_iteration_limit(lm_nbody* h, int limit);

/* Extension, no counterpart in the reference (its recurrence,
 * LanczosAlgorithm.cu:120-153, never reorthogonalises): mode 1 orthogonalises
 * every new basis vector against all earlier ones in FP64 (classical
 * Gram-Schmidt) before it is normalised; the recurrence's beta becomes the norm
 * of the reorthogonalised vector.  It removes the loss of orthogonality that
 * lets FP32 products and the reference's rounding-sensitive stopping rule
 * move the converged iteration, at the cost of one pass over the basis per
 * iteration.  Mode 0 (default) is the reference's recurrence.  Applies to
 * single-recurrence solves (not to per-system or multi-vector recurrences). */
int lm_nbody_set_lanczos_reorthogonalization(lm_nbody* h, int mode);

/* Ragged ensembles: one lanczos::Solver per system (LanczosAlgorithm.cu:171-232
 * applied to each diagonal block, as separate per-system NBody solvers would
 * run it) instead of one recurrence over the whole block-diagonal operator
 * (the reference's batched behaviour, the default).  Each system keeps its own
 * check schedule and previous estimate across calls; all systems advance in
 * one pair launch per iteration and converged systems drop out of it.  The
 * Lanczos callback receives the largest error of the systems tested at that
 * step.  A system whose noise is exactly zero gets a zero result without
 * iterating. */
int lm_nbody_set_lanczos_per_system(lm_nbody* h, int on);
/* Several force vectors in one pass over the pairs (no single counterpart in
 * the reference; it equals nrhs calls of NBody::Mdot, mobility.h:166-182, for
 * the translational block): linear_r = M forces_r, r < nrhs.  forces and
 * linear hold nrhs consecutive 3N vectors.  The pair geometry and
 * coefficients are computed once per pair for all vectors.  Solvers
 * initialised with includeAngular are refused. */
int lm_nbody_mdot_multi(lm_nbody* h, const double* forces, int nrhs, double* linear, int mem);
/* nrhs independent square roots, one lanczos::Solver each
 * (LanczosAlgorithm.cu:171-232; one per (vector, system) in per-system mode),
 * sharing every pass over the pairs: linear_r += prefactor M^{1/2} W_r
 * (Mobility::sqrtMdotW, MobilityInterface.h:132-188).  noise: nrhs
 * consecutive 3N vectors, or NULL to draw the W_r that nrhs successive
 * sqrtMdotW calls would draw.  Translational solvers only. */
int lm_nbody_sqrt_mdot_w_multi(lm_nbody* h, const double* noise, int nrhs, double* linear,
                               double prefactor, int mem);
/* Iterations of each vector of the last multi-vector solve (largest over its
 * systems in per-system mode). */
int lm_nbody_last_lanczos_iterations_multi(lm_nbody* h, int* out, int nrhs);

/* Number of independent systems of the current positions (Nbatch of
 * mobility.h:149-160, or the ragged systems of lm_nbody_set_systems). */
int lm_nbody_num_systems(lm_nbody* h);
/* Iterations each system took in the last per-system solve (-1 otherwise). */
int lm_nbody_last_lanczos_iterations_per_system(lm_nbody* h, int* out, int nsys);

/* NBody::getNumberParticles, mobility.h:164. */
int64_t lm_nbody_num_particles(lm_nbody* h);

/* NBody::Mdot, mobility.h:166-182 (checks of pythonify.h:109-149).
 * forces/torques may be NULL (n = 0); angular may be NULL when the solver was
 * initialised without includeAngular. */
int lm_nbody_mdot(lm_nbody* h, const double* forces, int64_t n_forces, const double* torques,
                  int64_t n_torques, double* linear, double* angular, int mem);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): lm_nbody_mdot on device pointers (forces, torques may
 * be NULL) that calls no hipMalloc, hipFree or allocating routine.  It runs
 * the ordered pair evaluation with its work items and chunk partial sums in
 * a static device array of 64 MiB per GPU (held until the process ends), so
 * the slabs are cut to fit (one chunk needs none) and it refuses a layout
 * whose work items alone exceed 64 MiB (about 4e8 particles).  One system or
 * equal batches only (not lm_nbody_set_systems).  The work is queued on the
 * handle's stream and the call returns without waiting.  All calls on one GPU
 * share that GPU's array: calls that overlap in time on one GPU race.  Call
 * lm_nbody_set_positions (which allocates) beforehand. */
int lm_nbody_mdot_noalloc(lm_nbody* h, const double* forces, const double* torques,
                          double* linear, double* angular);

/* Mobility::sqrtMdotW, MobilityInterface.h:132-188 (Lanczos, lanczos.h:43-58):
 * linear (and angular) += prefactor * M^{1/2} W, W ~ N(0,1) drawn on the device. */
int lm_nbody_sqrt_mdot_w(lm_nbody* h, double* linear, double* angular, double prefactor, int mem);

/* Extension for parity tests: as lm_nbody_sqrt_mdot_w with the noise W given
 * (length 3N, or 6N with includeAngular; torque part after the forces). */
int lm_nbody_sqrt_mdot_w_noise(lm_nbody* h, const double* noise, double* linear, double* angular,
                               double prefactor, int mem);

/* NBody::divM (random finite differences), mobility.h:184-236:
 * outputs += prefactor * div M; no-op for open boundaries. */
int lm_nbody_div_m(lm_nbody* h, double* linear, double* angular, double prefactor, int mem);

/* Extension for parity tests: as lm_nbody_div_m with the random displacement
 * W (3N values, random_finite_differences.h:46-52) given instead of drawn;
 * W is used both as displacement and as force, the entries of open systems
 * of a mixed ensemble are ignored.  The solver's generator is not advanced. */
int lm_nbody_div_m_noise(lm_nbody* h, const double* noise, double* linear, double* angular,
                         double prefactor, int mem);

/* Mobility::LangevinVelocities, MobilityInterface.h:191-206:
 * Mdot (when forces/torques given) + sqrt(2 kbt/dt) M^{1/2} W + kbt div M. */
int lm_nbody_langevin_velocities(lm_nbody* h, double dt, double kbt, const double* forces,
                                 int64_t n_forces, const double* torques, int64_t n_torques,
                                 double* linear, double* angular, int mem);

/* Alias with the name used by earlier libMobility versions
 * (include/MobilityInterface/README.md:16-17): LangevinVelocities with
 * dt = 2, kbt = 1 (unit noise and drift prefactors). */
int lm_nbody_hydrodynamic_velocities(lm_nbody* h, const double* forces, int64_t n_forces,
                                     const double* torques, int64_t n_torques, double* linear,
                                     double* angular, int mem);

/* Mobility::clean, MobilityInterface.h:220-223. */
int lm_nbody_clean(lm_nbody* h);

/* Extension: iteration index at which the last Lanczos call converged
 * (the return value of lanczos::Solver::run, LanczosAlgorithm.cu:192). */
int lm_nbody_last_lanczos_iterations(lm_nbody* h);

/* Extension: wait for all work of the handle's stream. */
int lm_nbody_synchronize(lm_nbody* h);

/* ---- learned block preconditioner (no counterpart in the reference) ------
 * A per-particle network maps the relative positions of a particle's 16
 * nearest neighbours and its height above a wall (64 floats: 48 coordinates
 * in units of a, then u = a/max(z, a), b = clamp(z/a, 0, 1), u^3 -- 0, 1, 0 in
 * an open system -- and zeros) through 64-128-128-out (GELU, BF16 MFMA, FP32
 * accumulation) to the Cholesky factor of an SPD block: 3x3 on the particle's
 * linear entries (6 outputs), or 6x6 on its linear and angular entries (21
 * outputs) for solvers initialised with includeAngular.  With a
 * preconditioner attached, sqrtMdotW runs Lanczos on G M G^T and returns
 * G^{-1} (G M G^T)^{1/2} W, another square root of M (with the coarse level
 * too, G = D Gc; see lm_nbody_set_coarse_level).
 * All arrays of this group are host arrays. */
typedef struct lm_precond lm_precond;

/* A network of 3x3 blocks (block_dim 3) or 6x6 blocks (block_dim 6). */
int lm_precond_create_blocks(int block_dim, lm_precond** out);
/* lm_precond_create_blocks(3, out). */
int lm_precond_create(lm_precond** out);
int lm_precond_destroy(lm_precond* p);
/* Number of FP32 weights of a 3x3-block network:
 * W1[64][128] b1[128] W2[128][128] b2[128] W3[128][out] b3[out]. */
int64_t lm_precond_num_weights(void);
/* Block size, number of weights, inputs and outputs of a network (any may be NULL). */
int lm_precond_describe(lm_precond* p, int* block_dim, int64_t* num_weights, int* inputs,
                        int* outputs);
int lm_precond_init_weights(lm_precond* p, uint64_t seed);
int lm_precond_set_weights(lm_precond* p, const float* weights, int64_t n);
int lm_precond_get_weights(lm_precond* p, float* weights, int64_t n);
/* Mini-batch Adam on (X: n x 64, T: n x d*d) with the GPU forward/backward
 * kernels; loss_history (epochs entries, may be NULL) gets each epoch's mean
 * loss sum_ij (G - T)_ij^2. */
int lm_precond_train(lm_precond* p, const float* X, const float* T, int64_t n, int epochs,
                     int batch, float learning_rate, uint64_t seed, float* loss_history);
/* Network outputs (n x outputs) for inputs X (n x 64). */
int lm_precond_forward(lm_precond* p, const float* X, int64_t n, float* out);
/* Mean loss and its gradient w.r.t. all weights for one batch (T: n x d*d). */
int lm_precond_loss_grad(lm_precond* p, const float* X, const float* T, int64_t n, float* grad,
                         float* loss);
/* Training pairs from the solver's current positions: features of ncenters
 * random particles and the center d x d block (block_dim 3 or 6) of the local
 * (M/t0)^{-1/2} of each (open or bottom-wall mobility by the particle's
 * system; the particle and its 16 nearest neighbours for 3x3 blocks, 8 for
 * 6x6 blocks, whose linear and angular entries are scaled by sqrt(t0) and
 * sqrt(r0); FP64 on the device).  status (ncenters ints, may be NULL) gets 0
 * where the local iteration converged; without it a center that did not
 * converge is an error. */
int lm_nbody_precond_dataset(lm_nbody* h, int block_dim, int64_t ncenters, uint64_t seed,
                             float* X, float* T, int* status);
/* Attach (p != NULL) or detach (p == NULL) a preconditioner; p must outlive
 * its use by h.  6x6 blocks need a solver initialised with includeAngular. */
int lm_nbody_set_preconditioner(lm_nbody* h, lm_precond* p);
/* The blocks G of the attached network for the positions: d*d doubles per
 * particle, row-major (6x6 blocks: G = L L^T diag(1, 1, 1, c, c, c) with
 * c = 2a/sqrt(3), acting on (linear, angular) entries). */
int lm_nbody_precond_blocks(lm_nbody* h, double* G);
/* The 16 nearest neighbours of every particle within its system (16N ints,
 * nearest first, ties by index; -1 past the system's other particles), the
 * network's input geometry. */
int lm_nbody_precond_neighbours(lm_nbody* h, int* nbr);

/* Extension, no counterpart in the reference: coarse level of the sqrtMdotW
 * preconditioner.  The particles of every system are grouped into spatial
 * cells of about particles_per_cell particles (coarse dimension 3C_s at most
 * 24576 per system by default, see lm_nbody_set_coarse_max_dim); with Q the
 * orthonormal cell columns and Q^T M Q = L L^T (block diagonal over the
 * systems), Lanczos runs on G M G^T with
 * G = D (I + Q (L^{-1} - I) Q^T) (D the learned blocks when a network is
 * attached, else I) and the result is G^{-1} (G M G^T)^{1/2} W.  0 switches
 * it off.  Built on the first sqrtMdotW after each setPositions. */
int lm_nbody_set_coarse_level(lm_nbody* h, int particles_per_cell);
/* Extension: the largest coarse dimension 3C_s of one system (default 24576,
 * at most 98304); denser requests get bigger cells.  The build holds three
 * dense (3C_s)^2 FP64 matrices per system (14 GB at 24576, 55 GB at 49152). */
int lm_nbody_set_coarse_max_dim(lm_nbody* h, int64_t max_dim);
/* Extension: pair arithmetic of the coarse matrix Q^T M Q.  LM_PAIR_FP32
 * (default) falls back to FP64 by itself when the FP32 matrix turns out
 * indefinite (a non-positive Cholesky pivot); LM_PAIR_FP64 always uses FP64
 * (about twice the build's pair work).  lm_nbody_coarse_used_fp64 reports
 * whether the last build used FP64. */
int lm_nbody_set_coarse_pair_precision(lm_nbody* h, int precision);
int lm_nbody_coarse_used_fp64(lm_nbody* h, int* used);
/* Cells of the last coarse build, over all systems: *ncells, grid[3] (cells
 * per axis of the first system, before empty ones are dropped) and, if
 * cell_of != NULL, the cell of each particle (N ints). */
int lm_nbody_coarse_cells(lm_nbody* h, int* ncells, int* grid, int* cell_of);
/* L and L^{-1} of one system's coarse block, 3C_s x 3C_s row-major doubles
 * each (C_s from lm_nbody_coarse_system_cells). */
int lm_nbody_coarse_factors(lm_nbody* h, int system, double* L, double* Linv);
/* Number of cells C_s of one system in the last coarse build (0 for an empty
 * system). */
int lm_nbody_coarse_system_cells(lm_nbody* h, int system, int* ncells);
/* Wall-clock milliseconds of the last coarse build (cells, Q^T M Q, factors). */
int lm_nbody_coarse_build_ms(lm_nbody* h, double* ms);
/* Number of coarse builds so far on this handle. */
int lm_nbody_coarse_builds(lm_nbody* h, int64_t* count);
/* Extension, no counterpart in the reference: keep the preconditioner (the
 * learned blocks and the coarse level) for `interval` set_positions calls of
 * the same systems (same N, offsets, kinds and radius) before rebuilding it;
 * default 1 (rebuilt for every new positions).  G = D Gc from earlier
 * positions keeps its exact inverse, so the result is still a square root of
 * the current M; only the iteration count can drift.  interval >= 1. */
int lm_nbody_set_preconditioner_refresh(lm_nbody* h, int interval);

/* ---- self-checks of the kernel arithmetic, run on the host CPU -----------
 * Not used by any solver path.  They evaluate the same pair and generator
 * code that the kernels are compiled from, so CPU tests can check it
 * against the oracle without a GPU. */

/* 6x6 dimensionless coupling block (rows U,W of i; columns F,T of j) as the
 * pair kernels compute it; fp32 != 0 runs the FP32 arithmetic. */
int lm_selftest_pair_block_host(int wall, int fp32, const double* pi, const double* pj, double a,
                                double* block36);

/* ---- SelfMobility (solvers/SelfMobility/mobility.h:17-111) --------------------
 * No hydrodynamic interactions: M = 1/(6 pi eta a) I for linear and
 * 1/(8 pi eta a^3) I for angular velocities.  Memory modes as for NBody. */
typedef struct lm_self lm_self;

/* SelfMobility(Configuration), mobility.h:31-36: open boundaries only
 * ("[Mobility] This is an open boundary solver"). */
int lm_self_create(int periodicity_x, int periodicity_y, int periodicity_z, lm_self** out);
int lm_self_destroy(lm_self* h);
int lm_self_set_stream(lm_self* h, void* stream);
/* SelfMobility::initialize, mobility.h:38-50 (the tolerance is unused). */
int lm_self_initialize(lm_self* h, double viscosity, double radius, int include_angular,
                       double tolerance, uint64_t seed);
/* SelfMobility::setPositions, mobility.h:56-59: only the particle count. */
int lm_self_set_positions(lm_self* h, const double* positions, int64_t n, int mem);
int64_t lm_self_num_particles(lm_self* h);
/* SelfMobility::Mdot, mobility.h:63-86: linear = F/(6 pi eta a) when forces
 * are given, angular = T/(8 pi eta a^3) when torques are given. */
int lm_self_mdot(lm_self* h, const double* forces, int64_t n_forces, const double* torques,
                 int64_t n_torques, double* linear, double* angular, int mem);
/* SelfMobility::sqrtMdotW, mobility.h:90-110: linear (and angular) +=
 * prefactor sqrt(mobility) N(0,1), noise drawn on the device. */
int lm_self_sqrt_mdot_w(lm_self* h, double* linear, double* angular, double prefactor, int mem);
/* Mobility::divM default, MobilityInterface.h:211-216: the drift is zero
 * (outputs are left as they are). */
int lm_self_div_m(lm_self* h, double* linear, double* angular, double prefactor, int mem);
/* Mobility::LangevinVelocities, MobilityInterface.h:191-206. */
int lm_self_langevin_velocities(lm_self* h, double dt, double kbt, const double* forces,
                                int64_t n_forces, const double* torques, int64_t n_torques,
                                double* linear, double* angular, int mem);
/* Mobility::clean, MobilityInterface.h:220-223. */
int lm_self_clean(lm_self* h);

/* Host self-check of the symmetric evaluation's pair arithmetic (FP64, all
 * blocks): fwd = the 6x6 block of target i from unit forces/torques on j, rev
 * = the block the same pair evaluation gives j from unit forces/torques on i
 * (the transposed blocks).  They must equal the reference blocks for (i, j)
 * and (j, i) (hydrodynamicKernels.cuh:20-442). */
int lm_selftest_pair_both_host(int wall, const double* pi, const double* pj, double a,
                               double* fwd36, double* rev36);

/* Self-check of the Lanczos tridiagonal square root, run on the GPU (replaces
 * LAPACKE ?steqr + cblas ?gemv of LanczosAlgorithm.cu:48-85): out = T^{1/2} e1
 * for T = tridiag(beta, alpha, beta), alpha: m host values, beta: m - 1,
 * 1 <= m <= 1024; the same device code as the solver's.  NaN if T has a
 * negative eigenvalue. */
int lm_selftest_tridiag_sqrt(const double* alpha, const double* beta, int m, double* out);

/* Philox4x32-10 of the noise kernels: out = philox(ctr, key). */
int lm_selftest_philox_host(const uint32_t* ctr4, const uint32_t* key2, uint32_t* out4);

/* Cell assignment of the coarse level for the N positions of one system, xyz
 * (3N doubles, units of a): *ncells, grid[3] and cell_of (N ints), as lm_nbody_coarse_cells
 * reports after a build. */
int lm_selftest_coarse_cells_host(const double* xyz, int64_t N, int particles_per_cell,
                                  int* ncells, int* grid, int* cell_of);

#ifdef __cplusplus
}
#endif
#endif /* LIBMOBILITY_AMD_H */
