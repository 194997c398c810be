// common.hpp -- shared device/host helpers for the minim_amd HIP library.
//
// Conventions used by every kernel in this library:
//  * vectors are flat arrays of T (double or float) resident in HBM; the
//    PhaseField layout is AoS, index = site * nFluid + fluid, site = (x*Ny+y)*Nz+z;
//  * every reduction accumulates in FP64, per thread, then per wavefront
//    (64 lanes, shuffles), then per block (LDS), then across blocks in a fixed
//    order by the last block to finish -- results are bitwise reproducible;
//  * kernels are compiled with -ffp-contract=off: every product is rounded
//    before it is added, as in the reference's element-wise vector operators.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace minim_amd {

constexpr int kBlock = 256;          // threads per block (4 wavefronts of 64)
constexpr int kWave = 64;
constexpr int kMaxRedBlocks = 8192;  // upper bound on blocks of any reducing launch
constexpr int kMaxDots = 96;         // upper bound on values reduced by one launch
// Result slot of a deferred validity flag (Potential::gradientDeferred): the
// last one, written by that launch and read by the caller's next fetch.
constexpr int kDeferSlot = kMaxDots - 1;

enum Prec : int { F64 = 0, F32 = 1 };

inline size_t prec_bytes(int p) { return p == F64 ? 8 : 4; }

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};

#define MINIM_HIP_CHECK(expr)                                                                  \
  do {                                                                                         \
    hipError_t _e = (expr);                                                                    \
    if (_e != hipSuccess)                                                                      \
      throw ::minim_amd::HipError(std::string("HIP error ") + hipGetErrorString(_e) + " at " + \
                                  __FILE__ + ":" + std::to_string(__LINE__) + " (" #expr ")"); \
  } while (0)

// ---------------------------------------------------------------- reductions
// Sum K doubles over the 64 lanes of a wavefront.
template <int K>
__device__ inline void wave_sum(double (&v)[K]) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < K; k++) v[k] += __shfl_xor(v[k], off, kWave);
  }
}

// Block-wide sum of K doubles; the result is valid in thread 0 only.
// lds must hold (kBlock/kWave)*K doubles.
template <int K>
__device__ inline void block_sum(double (&v)[K], double* lds) {
  wave_sum<K>(v);
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) lds[wid * K + k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      double s = lds[k];
      for (int w = 1; w < kBlock / kWave; w++) s += lds[w * K + k];
      v[k] = s;
    }
  }
}

// Scratch for cross-block reductions: K rows of per-block partials and the
// final result (device).
//
// Two stages, no fences and no atomics: each block of the producing kernel
// writes its K partial sums (grid_sum_store), then a small kernel with one
// block per value sums the partials in block order (launch_finalize).  A
// "last block finishes the sum" scheme would need a device-scope fence in
// every block, which on CDNA3/4 writes back the whole L2 each time.
struct RedScratch {
  double* partials;  // [kMaxDots][kMaxRedBlocks], value k of block b at k * gridDim + b
  double* result;    // [kMaxDots]
};

// All threads of the (1-D) grid must call this; v is each thread's contribution.
template <int K>
__device__ inline void grid_sum_store(double (&v)[K], RedScratch rs) {
  __shared__ double lds[(kBlock / kWave) * (K > 0 ? K : 1)];
  block_sum<K>(v, lds);
  if (threadIdx.x == 0) {
    const size_t nb = gridDim.x;
#pragma unroll
    for (int k = 0; k < K; k++) rs.partials[(size_t)k * nb + blockIdx.x] = v[k];
  }
}

// Second stage: result[k] = sum over the nb partials of value k, in a fixed
// order (stream-ordered after the producing kernel).  With dupDst, value dupK
// is also written there (a caller's device slot: no copy launch after).
void launch_finalize(RedScratch rs, int K, int nb, hipStream_t s, int dupK = -1, double* dupDst = nullptr);

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with a private L2.  Remap the hardware block id so that every XCD receives a
// contiguous range of logical blocks: neighbouring tiles (which share halo
// rows) then run on the same XCD and hit the same L2.  A bijection for any nb.
constexpr int kXcd = 8;
__device__ inline int xcd_logical_block(int b, int nb) {
  const int q = nb / kXcd, r = nb % kXcd;
  const int k = b % kXcd, i = b / kXcd;
  return (k < r) ? k * (q + 1) + i : r * (q + 1) + (k - r) * q + i;
}

// Number of blocks for a grid-stride launch over n work items.
inline int grid_blocks(size_t n, int per_thread = 1, int max_blocks = kMaxRedBlocks) {
  size_t threads = (n + per_thread - 1) / per_thread;
  size_t b = (threads + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > (size_t)max_blocks) b = max_blocks;
  return (int)b;
}

}  // namespace minim_amd
