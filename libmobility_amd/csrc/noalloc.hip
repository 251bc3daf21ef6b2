// Allocation-free Mdot: the ordered pair evaluation of NBodySolver::Mdot with
// every buffer beyond the caller's inputs and outputs taken from one
// file-scope device array.  For callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  It calls no hipMalloc, hipFree or allocating library routine.
//
// Scratch layout per call: [work items, 256-byte aligned][U chunk slabs]
// [W chunk slabs].  Every byte the call reads is written by the same call
// first (items by k_uniform_items, every slab row by exactly one item), so
// nothing needs clearing between calls.  All calls on one GPU share its copy
// of the array: calls that overlap in time on one GPU race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "kernels.hpp"
#include "solver.hpp"

namespace lm {

namespace {

constexpr int kMaxDevices = 64;

// Device memory of the allocation-free entry point; each GPU holds its own
// copy until the process ends.  Host code uses the per-device address from
// hipGetSymbolAddress, never this name.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
unsigned char* scratch_current_device() {
  int dev = -1;
  LM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    throw Error("[libmobility_amd] device index out of range for the allocation-free scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    LM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) throw Error("[libmobility_amd] no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  return static_cast<unsigned char*>(g_scratch_ptr[dev]);
}

// Work items of `nb` equal systems of `ns` particles: system-major, then
// target tiles of bt particles, then nch source chunks of cs particles (the
// order and fields of NBodySolver::build_items).
__global__ void k_uniform_items(PairItem* __restrict__ out, int nitems, int64_t ns, int tiles,
                                int nch, int64_t cs, int64_t bt, int kind) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nitems) return;
  const int per_sys = tiles * nch;
  const int s = i / per_sys;
  const int r = i - s * per_sys;
  const int t = r / nch;
  const int c = r - t * nch;
  const int64_t b = (int64_t)s * ns, e = b + ns;
  PairItem it;
  it.tgt0 = b + (int64_t)t * bt;
  it.ntgt = (int32_t)min(bt, e - it.tgt0);
  it.src0 = min(b + (int64_t)c * cs, e);
  it.src1 = min(b + (int64_t)(c + 1) * cs, e);
  it.chunk = c;
  it.wall = kind;
  it.sys = s;
  out[i] = it;
}

}  // namespace

void NBodySolver::mdotNoAlloc(const double* F, const double* T, double* U, double* W) {
  require_positions("[Mobility] Positions have 0 particles. Did you call setPositions?");
  if (explicit_systems_)
    throw Error(
        "[libmobility_amd] the allocation-free Mdot takes one system or equal batches "
        "(Nbatch/NperBatch), not ragged systems");
  if (!U) throw Error("[libmobility_amd] Mdot needs a linear output array");
  int blocks = 0;
  if (F) blocks |= kUF | (angular_ ? kWF : 0);
  if (T && angular_) blocks |= kUT | kWT;
  if (T && !angular_)
    throw Error(
        "The solver was configured without including angular velocities. Set includeAngular "
        "to true when initializing if you want to use torques");
  const bool ang = (blocks & (kWT | kWF)) != 0;
  if (angular_ && !W)
    throw Error("[libmobility_amd] the solver includes angular velocities: an angular output array is required");
  unsigned char* base = scratch_current_device();
  ensure_records();
  const int64_t n3 = 3 * N_;
  if (blocks == 0) {
    launch_zero(U, n3, stream_);
    if (angular_) launch_zero(W, n3, stream_);
    return;
  }
  const int nb = numSystems();
  const int64_t ns = N_ / nb;
  // tiling: the solver's override, else its default rule (no autotuning here)
  int tpl = tile_tpl_ != 0 ? tile_tpl_ : (ns >= 16384 ? 2 : 1);
  if (tpl == 4 && !(kind_code() == 0 && blocks == kUF)) tpl = 2;
  const int64_t bt = (int64_t)kPairThreads * tpl;
  const int64_t tiles64 = (ns + bt - 1) / bt;
  const size_t slab_bytes = sizeof(double) * (size_t)n3 * (ang ? 2 : 1);
  auto item_bytes = [&](int nch) {
    const size_t b = sizeof(PairItem) * (size_t)(tiles64 * nb * nch);
    return (b + 255) / 256 * 256;
  };
  int nch = resolve_nch(tpl, algorithm_ == kNaive ? 1 : tile_nch_);
  // one chunk writes the outputs directly; more need a slab each
  while (nch > 1 && item_bytes(nch) + slab_bytes * nch > kNoAllocScratchBytes) --nch;
  if (item_bytes(nch) > kNoAllocScratchBytes || tiles64 * nb * nch > INT32_MAX)
    throw Error("[libmobility_amd] the allocation-free Mdot needs " +
                std::to_string(item_bytes(nch)) + " bytes of work items for " +
                std::to_string(N_) + " particles, more than its scratch of " +
                std::to_string(kNoAllocScratchBytes) + " bytes");
  const int nitems = (int)(tiles64 * nb * nch);
  int64_t cs = (ns + nch - 1) / nch;
  cs = ((cs + kChunkGrain - 1) / kChunkGrain) * kChunkGrain;
  PairItem* items = reinterpret_cast<PairItem*>(base);
  hipLaunchKernelGGL(k_uniform_items, dim3((unsigned)((nitems + 255) / 256)), dim3(256), 0,
                     stream_, items, nitems, ns, (int)tiles64, nch, cs, bt, wall_ ? 1 : 0);
  LM_HIP(hipGetLastError());
  const bool fp32 = pair_prec_ == kPairFP32;
  PairArgs a{};
  a.wall = wall_ ? 1 : 0;
  a.persistent = 0;
  a.fp32 = fp32 ? 1 : 0;
  a.blocks = blocks;
  a.tpl = tpl;
  a.pos = fp32 ? rec32_.as<void>() : rec64_.as<void>();
  a.F = F;
  a.T = (blocks & (kUT | kWT)) ? T : nullptr;
  a.items = items;
  a.nitems = nitems;
  a.active = nullptr;
  a.t0 = t0_;
  a.rt0 = rt0_;
  a.r0 = r0_;
  if (nch == 1) {
    a.U = U;
    a.W = ang ? W : nullptr;
    a.slab = 0;
    launch_pair(a, stream_);
  } else {
    double* su = reinterpret_cast<double*>(base + item_bytes(nch));
    double* sw = ang ? su + (size_t)nch * n3 : nullptr;
    a.U = su;
    a.W = sw;
    a.slab = n3;
    launch_pair(a, stream_);
    launch_chunk_reduce(su, nch, n3, U, stream_);
    if (ang) launch_chunk_reduce(sw, nch, n3, W, stream_);
  }
  if (angular_ && !ang) launch_zero(W, n3, stream_);
}

}  // namespace lm
