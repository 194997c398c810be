// batch.hip -- many independent L-BFGS minimisations of one-fluid PhaseField
// fields, one launch per phase over every instance still running.
//
// The reference runs independent minimisations as States on disjoint rank
// subsets (include/State.h:23, src/Communicator.cpp:219-240), each with its
// own Lbfgs (src/minimisers/Lbfgs.cpp:41-133) and backtracking line search
// (src/minimisers/linesearch.cpp:8-29).  Here all instances' vectors are
// concatenated in HBM and every phase of an iteration is one launch over the
// blocks of the instances that take part in it:
//
//   direction   d = c_g g + sum_s c_D[s] D_s + c_Y[s] Y_s, with d.d, d.g, d.Y_s
//   trials      energies at x + tau 2^-k d, k < 4 (per-instance tau): the next
//               four points of the backtracking sequence in one pass
//   step        x' = x + tau d into the other coordinate buffer, energy and
//               gradient at x', y = g_new - g and every dot the next
//               recursion needs
//
// A block covers 256 BR consecutive elements of one instance (BR elements per
// thread, a launch-shape parameter); it finds its instance by binary search
// over the prefix sums of the block counts of the instances in the launch,
// and writes its partial sums to its own row.  The rows of each instance are
// summed in a fixed order, so every per-instance result is the same whatever
// other instances share the launch.  The direction and step passes are
// compiled for each number of stored pairs (a block-uniform switch), so all
// their loads are unconditional.  Each instance's (2m+1)^2 Gram data and its
// two-loop recursion live either on the device (device control: one thread
// of a block per instance takes the decisions, and an iteration is a fixed
// launch sequence, launched as it is or replayed as a HIP graph; resident
// control: one workgroup runs an instance's whole minimisation) or on the
// host (host control: one host synchronisation per phase serves all
// instances).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <deque>
#include <stdexcept>
#include <vector>

#include "minim/Minimiser.h"
#include "minim/PhaseField.h"
#include "minim/State.h"
#include "autotune.hpp"
#include "runtime.hpp"

namespace minim_amd {
namespace {

constexpr int kBS = 11;            // history slots (m + 1 with m <= 10)
// Elements per thread when none is set or tuned: 8, or 4 for batches below
// kSmallBatch elements in all, whose passes have too few blocks to fill the
// device and are bound by latency (a 128^2 field alone: 62.6 -> 56.5 us per
// iteration).
constexpr int kBlockDefault = 8;
constexpr int kBlockSmall = 4;
constexpr size_t kSmallBatch = size_t(1) << 20;
constexpr int kKP = 2;             // stencil partials: energy, g.g
constexpr int kNT = 4;             // line-search trials evaluated per pass: tau, tau/2, tau/4, tau/8
constexpr int kKD = 2 + kBS;       // direction: d.d, d.g, d.Y_s
constexpr int kKG = 4 * kBS + 3;   // Gram (layout in bstep_body)
constexpr int kKS = 1 + kKG;       // step pass: energy, then the Gram block
constexpr int kKF = kKP + 2 * kBS;  // fresh gradient: energy, g.g, g.D_i, g.Y_i

struct BInst {
  size_t off;  // first element in the concatenated vectors
  int n;       // elements (sites)
  int X, Y, Z;  // canonical grid, Z contiguous
  double fb, fg, fbg, fg2;  // kappa/16 v, 0.25 kappa' v / res^2, 4 fb, 2 fg
  float fb_f, fg_f, fbg_f, fg2_f;  // the same rounded to FP32
};

template <typename T>
struct BC;
template <>
struct BC<double> {
  __device__ static double fb(const BInst& I) { return I.fb; }
  __device__ static double fg(const BInst& I) { return I.fg; }
  __device__ static double fbg(const BInst& I) { return I.fbg; }
  __device__ static double fg2(const BInst& I) { return I.fg2; }
};
template <>
struct BC<float> {
  __device__ static float fb(const BInst& I) { return I.fb_f; }
  __device__ static float fg(const BInst& I) { return I.fg_f; }
  __device__ static float fbg(const BInst& I) { return I.fbg_f; }
  __device__ static float fg2(const BInst& I) { return I.fg2_f; }
};

struct BDyn {
  int j;           // instance
  int dslot;       // slot of the direction and of the new y
  int gcur;        // G buffer that holds g
  int gout;        // G buffer a gradient pass writes
  int np;          // stored pairs (oldest first) in the direction and Gram passes
  int lp;          // 1: FP32 coordinates and arithmetic (first phase of the mixed schedule)
  int slot[kBS];   // their history slots
  double tau;      // trial or step factor
  double cg;       // coefficient of g in the direction
  double cD[kBS], cY[kBS];  // coefficients of pair i's D and Y
  int xcur;        // coordinate buffer (X0 or X1) that holds x
  unsigned act;    // phases this instance takes part in (kAct*; device control)
};

// Phases of an iteration under device control: the recomputed gradient after
// the switch to FP64 coordinates, the direction pass with the speculative
// trials, a pass of further line-search trials, and the step pass.
constexpr unsigned kActFresh = 1, kActDir = 2, kActSearch = 4, kActStep = 8;

// Vectors G, D and Y are stored in TV (double, or float for the mixed
// schedule); coordinates always in FP64 (holding FP32 values while an
// instance runs in FP32).
struct BBufs {
  double* X0;  // coordinates; the step pass writes x + tau d to the other buffer
  double* X1;
  void* G0;
  void* G1;
  void* D;  // m + 1 slots of N elements
  void* Y;
  size_t N;
  const BInst* inst;
  const BDyn* dyn;   // per position, or per instance when map is set
  const int* map;    // position -> instance (device control), or null
  const int* ablk;   // na + 1 prefix sums of the block counts
  int na;
  double* part;  // per-block partials, K per block
  unsigned phase;  // nonzero: only instances whose act has this bit take part
};

// A block of an instance that does not take part in this phase leaves at once
// (the whole block: the instance is uniform over it).
#define MINIM_BATCH_SKIP(B, q) \
  if ((B).phase && !((q).act & (B).phase)) return

__device__ __forceinline__ int find_pos(const int* __restrict__ ablk, int na, int b) {
  int lo = 0, hi = na - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ablk[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Block sums of the values whose bit is set in `on` (the others are 0),
// written to part[blockIdx.x * S + k] (rows of S >= K values).
template <int K, int S = K>
__device__ __forceinline__ void block_partials(const double (&v)[K], unsigned long long on, double* part) {
  __shared__ double lds[(kBlock / kWave) * K];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < K; k++) {
    double x = 0;
    if ((on >> k) & 1ull) {
      x = v[k];
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
    }
    if (lane == 0) lds[wid * K + k] = x;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    double s = lds[k];
    for (int w = 1; w < kBlock / kWave; w++) s += lds[w * K + k];
    part[(size_t)blockIdx.x * S + k] = s;
  }
}

// out[k] = sum of the block rows [b0, b0 + nb) of part (K per row, or
// rows of `stride` values), by the whole block: lane l of a wavefront adds
// rows l, l + 64, ..., then the 64 lane sums are added in a fixed tree, so the
// result depends only on those rows, wherever they lie in part.
__device__ __forceinline__ void reduce_rows(const double* __restrict__ part, int b0, int nb, int K,
                                            double* __restrict__ out, int stride = 0) {
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const size_t S = stride > 0 ? (size_t)stride : (size_t)K;
  for (int k = w; k < K; k += kBlock / kWave) {
    double s = 0;
    for (int b = lane; b < nb; b += kWave) s += part[(size_t)(b0 + b) * S + k];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
    if (lane == 0) out[k] = s;
  }
}

// res[j * K + k]: the sums of the instance at position p (j = p, or map[p]
// under device control).
__global__ __launch_bounds__(kBlock) void bfinal_kernel(const double* __restrict__ part,
                                                        const int* __restrict__ ablk, const int* __restrict__ map,
                                                        int K, double* __restrict__ res,
                                                        const BDyn* __restrict__ dyn, unsigned phase) {
  const int p = blockIdx.x;
  const int j = map ? map[p] : p;
  if (phase && !(dyn[j].act & phase)) return;
  reduce_rows(part, ablk[p], ablk[p + 1] - ablk[p], K, res + (size_t)j * K);
}

// Energy (forward bonds) and gradient of one-fluid sites, the arithmetic of the
// tiled kernel for one fluid (PhaseField.cpp:238-288, :372-378) in T (FP32
// for an instance in the first phase of the mixed schedule), vectors in TV.
// MODE bits: 1 energy, 2 gradient into G[gout] with g.g, 4 at x + tau * d,
// 8 also the dots of the gradient with every stored pair, g.D_i and g.Y_i
// (partials [2, 2 + kBS) and [2 + kBS, 2 + 2 kBS)).
template <int MODE>
constexpr int bpf_k() {
  return (MODE & 8) ? kKF : kKP;
}
// The partials a bpf pass sums (bit k: partial k).
template <int MODE>
__device__ __forceinline__ unsigned long long bpf_mask(const BDyn& q) {
  unsigned long long on = (MODE & 1 ? 1ull : 0ull) | (MODE & 2 ? 2ull : 0ull);
  if (MODE & 8) {
    const unsigned long long pm = (1ull << q.np) - 1;
    on |= (pm << kKP) | (pm << (kKP + kBS));
  }
  return on;
}

// The pass bodies add the contributions of this thread's elements of block
// lb of the instance to acc (the kernels then sum acc over the block).
template <int BR, int MODE, typename T, typename TV>
__device__ __forceinline__ void bpf_body(const BBufs& B, const BDyn& q, const BInst& I, int lb,
                                         double (&acc)[bpf_k<MODE>()]) {
  using C = BC<T>;
  const size_t N = B.N;
  const double* __restrict__ x = (q.xcur ? B.X1 : B.X0) + I.off;
  const TV* __restrict__ d = static_cast<const TV*>(B.D) + (size_t)q.dslot * B.N + I.off;
  TV* __restrict__ g = static_cast<TV*>(q.gout ? B.G1 : B.G0) + I.off;
  const T tau = (T)q.tau;
  const int Z = I.Z, Y = I.Y, X = I.X, YZ = Y * Z;
  auto val = [&](int i) -> T {
    const T xv = (T)x[i];
    if (MODE & 4) {
      const T st = tau * (T)d[i];
      return xv + st;
    }
    return xv;
  };
#pragma unroll
  for (int r = 0; r < BR; r++) {
    const int i = lb * (BR * kBlock) + r * kBlock + (int)threadIdx.x;
    if (i >= I.n) break;
    const int z = i % Z, t2 = i / Z;
    const int y = t2 % Y, xx = t2 / Y;
    const int izm = z == 0 ? i + (Z - 1) : i - 1;
    const int izp = z == Z - 1 ? i - (Z - 1) : i + 1;
    const int iym = y == 0 ? i + (Y - 1) * Z : i - Z;
    const int iyp = y == Y - 1 ? i - (Y - 1) * Z : i + Z;
    const int ixm = xx == 0 ? i + (X - 1) * YZ : i - YZ;
    const int ixp = xx == X - 1 ? i - (X - 1) * YZ : i + YZ;
    const T ck = val(i);
    const T dxp = ck - val(ixp), dyp = ck - val(iyp), dzp = ck - val(izp);
    const T a2 = ck - (T)1;
    if (MODE & 1) {
      const T a1 = ck + (T)1;
      const T be = C::fb(I) * ((a1 * a1) * (a2 * a2));
      T sq = dxp * dxp;
      sq = fma(dyp, dyp, sq);
      sq = fma(dzp, dzp, sq);
      acc[0] += (double)fma(C::fg(I), sq, be);
    }
    if (MODE & 2) {
      const T dxm = ck - val(ixm), dym = ck - val(iym), dzm = ck - val(izm);
      const T lap = ((dxm + dxp) + (dym + dyp)) + (dzm + dzp);
      const T bg = C::fbg(I) * (ck * fma(ck, ck, (T)-1));
      const TV gv = (TV)fma(C::fg2(I), lap, bg);
      g[i] = gv;
      acc[1] += (double)gv * (double)gv;
      if (MODE & 8) {
        const double a = (double)gv;
        const TV* __restrict__ Dm = static_cast<const TV*>(B.D) + I.off;
        const TV* __restrict__ Ym = static_cast<const TV*>(B.Y) + I.off;
#pragma unroll
        for (int k = 0; k < kBS; k++)
          if (k < q.np) {
            acc[kKP + k] += a * (double)Dm[(size_t)q.slot[k] * N + i];
            acc[kKP + kBS + k] += a * (double)Ym[(size_t)q.slot[k] * N + i];
          }
      }
    }
  }
}

template <int BR, int MODE, typename TV>
__global__ __launch_bounds__(kBlock) void bpf_kernel(BBufs B) {
  const int p = find_pos(B.ablk, B.na, blockIdx.x);
  const BDyn& q = B.dyn[B.map ? B.map[p] : p];
  MINIM_BATCH_SKIP(B, q);
  const BInst I = B.inst[q.j];
  const int lb = blockIdx.x - B.ablk[p];
  constexpr int K = bpf_k<MODE>();
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0;
  if (q.lp) bpf_body<BR, MODE, float, TV>(B, q, I, lb, acc);
  else bpf_body<BR, MODE, double, TV>(B, q, I, lb, acc);
  block_partials<K>(acc, bpf_mask<MODE>(q), B.part);
}

// Energies at kNT points x + tau_k d along the line-search sequence, tau_k =
// tau 2^-k (the halvings of linesearch.cpp:24-25 are exact): x and d are read
// once for all of them, and the host takes the first point that passes the
// Armijo test, exactly as trying them one after another would.
template <int BR, typename T, typename TV>
__device__ __forceinline__ void btrial_body(const BBufs& B, const BDyn& q, const BInst& I, int lb,
                                            double (&acc)[kNT]) {
  using C = BC<T>;
  const double* __restrict__ x = (q.xcur ? B.X1 : B.X0) + I.off;
  const TV* __restrict__ d = static_cast<const TV*>(B.D) + (size_t)q.dslot * B.N + I.off;
  T tk[kNT];
  tk[0] = (T)q.tau;
#pragma unroll
  for (int k = 1; k < kNT; k++) tk[k] = tk[k - 1] * (T)0.5;
  const int Z = I.Z, Y = I.Y, X = I.X, YZ = Y * Z;
  for (int r = 0; r < BR; r++) {
    const int i = lb * (BR * kBlock) + r * kBlock + (int)threadIdx.x;
    if (i >= I.n) break;
    const int z = i % Z, t2 = i / Z;
    const int y = t2 % Y, xx = t2 / Y;
    const int izp = z == Z - 1 ? i - (Z - 1) : i + 1;
    const int iyp = y == Y - 1 ? i - (Y - 1) * Z : i + Z;
    const int ixp = xx == X - 1 ? i - (X - 1) * YZ : i + YZ;
    const T x0 = (T)x[i], x1 = (T)x[ixp], x2 = (T)x[iyp], x3 = (T)x[izp];
    const T d0 = (T)d[i], d1 = (T)d[ixp], d2 = (T)d[iyp], d3 = (T)d[izp];
#pragma unroll
    for (int k = 0; k < kNT; k++) {
      const T s0 = tk[k] * d0, s1 = tk[k] * d1, s2 = tk[k] * d2, s3 = tk[k] * d3;
      const T ck = x0 + s0;
      const T dxp = ck - (x1 + s1), dyp = ck - (x2 + s2), dzp = ck - (x3 + s3);
      const T a1 = ck + (T)1, a2 = ck - (T)1;
      const T be = C::fb(I) * ((a1 * a1) * (a2 * a2));
      T sq = dxp * dxp;
      sq = fma(dyp, dyp, sq);
      sq = fma(dzp, dzp, sq);
      acc[k] += (double)fma(C::fg(I), sq, be);
    }
  }
}

template <int BR, typename TV>
__global__ __launch_bounds__(kBlock) void btrial_kernel(BBufs B) {
  const int p = find_pos(B.ablk, B.na, blockIdx.x);
  const BDyn& q = B.dyn[B.map ? B.map[p] : p];
  MINIM_BATCH_SKIP(B, q);
  const BInst I = B.inst[q.j];
  const int lb = blockIdx.x - B.ablk[p];
  double acc[kNT];
#pragma unroll
  for (int k = 0; k < kNT; k++) acc[k] = 0;
  if (q.lp) btrial_body<BR, float, TV>(B, q, I, lb, acc);
  else btrial_body<BR, double, TV>(B, q, I, lb, acc);
  block_partials<kNT>(acc, (1ull << kNT) - 1, B.part);
}

// x rounded to FP32 (the mixed schedule's first phase keeps FP32 coordinates).
template <int BR>
__global__ __launch_bounds__(kBlock) void bround_kernel(BBufs B) {
  const int p = find_pos(B.ablk, B.na, blockIdx.x);
  const BDyn& q = B.dyn[B.map ? B.map[p] : p];
  MINIM_BATCH_SKIP(B, q);
  const BInst I = B.inst[q.j];
  const int lb = blockIdx.x - B.ablk[p];
  double* __restrict__ x = (q.xcur ? B.X1 : B.X0) + I.off;
#pragma unroll
  for (int r = 0; r < BR; r++) {
    const int i = lb * (BR * kBlock) + r * kBlock + (int)threadIdx.x;
    if (i >= I.n) break;
    x[i] = (double)(float)x[i];
  }
}

// d = cg g + sum_i cD[i] D_{slot i} + cY[i] Y_{slot i} (pairs oldest first, as
// the single-State direction pass) into D[dslot]; partials d.d, d.g, d.y_i.
// P is the number of stored pairs, so every load is unconditional.
// A value that is the same in every lane of the wave (a block works on one
// instance, so its BDyn fields are), moved to a scalar register: the pointer
// arithmetic and coefficients of the passes then take no vector registers.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uni(double v) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <int BR, int P, typename TV>
__device__ __forceinline__ void bdir_body(const BBufs& B, const BDyn& q, const BInst& I, int lb,
                                          double (&acc)[kKD]) {
  const size_t N = B.N;
  const TV* __restrict__ g = static_cast<const TV*>(uni(q.gcur) ? B.G1 : B.G0) + I.off;
  TV* __restrict__ out = static_cast<TV*>(B.D) + (size_t)uni(q.dslot) * N + I.off;
  const TV* Dp[P > 0 ? P : 1];
  const TV* Yp[P > 0 ? P : 1];
  double cd[P > 0 ? P : 1], cy[P > 0 ? P : 1];
#pragma unroll
  for (int k = 0; k < P; k++) {
    const int sk = uni(q.slot[k]);
    Dp[k] = static_cast<const TV*>(B.D) + (size_t)sk * N + I.off;
    Yp[k] = static_cast<const TV*>(B.Y) + (size_t)sk * N + I.off;
    cd[k] = uni(q.cD[k]);
    cy[k] = uni(q.cY[k]);
  }
  const double cg = uni(q.cg);
#pragma unroll 2
  for (int r = 0; r < BR; r++) {
    const int i = lb * (BR * kBlock) + r * kBlock + (int)threadIdx.x;
    if (i >= I.n) break;
    const double gi = (double)g[i];
    double dv[P > 0 ? P : 1], yv[P > 0 ? P : 1];
#pragma unroll
    for (int k = 0; k < P; k++) {
      dv[k] = (double)Dp[k][i];
      yv[k] = (double)Yp[k][i];
    }
    double v = cg * gi;
#pragma unroll
    for (int k = 0; k < P; k++) v = v + cd[k] * dv[k];
#pragma unroll
    for (int k = 0; k < P; k++) v = v + cy[k] * yv[k];
    const TV vs = (TV)v;  // FP64 per element, stored as TV; the dots use the stored value
    out[i] = vs;
    v = (double)vs;
    acc[0] += v * v;
    acc[1] += v * gi;
#pragma unroll
    for (int k = 0; k < P; k++) acc[2 + k] += v * yv[k];
  }
}

// The direction pass with the number of stored pairs as a compile-time value.
template <int BR, typename TV>
__device__ __forceinline__ void bdir_any(const BBufs& B, const BDyn& q, const BInst& I, int lb,
                                         double (&acc)[kKD]) {
  switch (q.np) {
    case 0: bdir_body<BR, 0, TV>(B, q, I, lb, acc); break;
    case 1: bdir_body<BR, 1, TV>(B, q, I, lb, acc); break;
    case 2: bdir_body<BR, 2, TV>(B, q, I, lb, acc); break;
    case 3: bdir_body<BR, 3, TV>(B, q, I, lb, acc); break;
    case 4: bdir_body<BR, 4, TV>(B, q, I, lb, acc); break;
    case 5: bdir_body<BR, 5, TV>(B, q, I, lb, acc); break;
    case 6: bdir_body<BR, 6, TV>(B, q, I, lb, acc); break;
    case 7: bdir_body<BR, 7, TV>(B, q, I, lb, acc); break;
    case 8: bdir_body<BR, 8, TV>(B, q, I, lb, acc); break;
    case 9: bdir_body<BR, 9, TV>(B, q, I, lb, acc); break;
    default: bdir_body<BR, 10, TV>(B, q, I, lb, acc); break;
  }
}

template <int BR, typename TV>
__global__ __launch_bounds__(kBlock) void bdir_kernel(BBufs B) {
  const int p = find_pos(B.ablk, B.na, blockIdx.x);
  const BDyn& q = B.dyn[B.map ? B.map[p] : p];
  MINIM_BATCH_SKIP(B, q);
  const BInst I = B.inst[q.j];
  const int lb = blockIdx.x - B.ablk[p];
  double acc[kKD];
#pragma unroll
  for (int k = 0; k < kKD; k++) acc[k] = 0;
  bdir_any<BR, TV>(B, q, I, lb, acc);
  block_partials<kKD>(acc, (1ull << (2 + q.np)) - 1, B.part);
}

// The step, gradient and Gram passes of an iteration in one: x' = x + tau d
// (product rounded first, as the single-State axpy; in FP32 for an instance in
// the first phase of the mixed schedule) into the other coordinate buffer; at
// x' the energy and gradient gn = G[gout] (the stencil of bpf_body, each
// neighbour's x' formed the same way); y = gn - G[gcur] into Y[dslot]; and
// the partials, the D list being the P stored pairs then the new direction
// and the Y list the P pairs:
//   0 energy ; then from 1:
//   [0,kBS) gn.D_i ; [kBS,2kBS) gn.Y_i ; 2kBS gn.gn ; 2kBS+1 gn.y ;
//   [2kBS+2,3kBS+2) y.D_i ; [3kBS+2,4kBS+2) y.Y_i ; 4kBS+2 y.y
// Every value is the one the separate passes would compute, and each partial
// sums the same terms in the same order.
template <int BR, int P, typename T, typename TV>
__device__ __forceinline__ void bstep_body(const BBufs& B, const BDyn& q, const BInst& I, int lb,
                                           double (&acc)[kKS]) {
  using C = BC<T>;
  const size_t N = B.N;
  const int xc = uni(q.xcur), ds = uni(q.dslot);
  const double* __restrict__ x = (xc ? B.X1 : B.X0) + I.off;
  double* __restrict__ xn = (xc ? B.X0 : B.X1) + I.off;
  const TV* __restrict__ d = static_cast<const TV*>(B.D) + (size_t)ds * N + I.off;
  TV* __restrict__ gn = static_cast<TV*>(uni(q.gout) ? B.G1 : B.G0) + I.off;
  const TV* __restrict__ go = static_cast<const TV*>(uni(q.gcur) ? B.G1 : B.G0) + I.off;
  TV* __restrict__ yo = static_cast<TV*>(B.Y) + (size_t)ds * N + I.off;
  const TV* Dp[P > 0 ? P : 1];
  const TV* Yp[P > 0 ? P : 1];
#pragma unroll
  for (int k = 0; k < P; k++) {
    const int sk = uni(q.slot[k]);
    Dp[k] = static_cast<const TV*>(B.D) + (size_t)sk * N + I.off;
    Yp[k] = static_cast<const TV*>(B.Y) + (size_t)sk * N + I.off;
  }
  const T tau = (T)uni(q.tau);
  auto val = [&](int i) -> T {
    const T st = tau * (T)d[i];
    return (T)x[i] + st;
  };
  constexpr int o = 1;  // Gram block offset
  const int Z = I.Z, Y = I.Y, X = I.X, YZ = Y * Z;
#pragma unroll 2
  for (int r = 0; r < BR; r++) {
    const int i = lb * (BR * kBlock) + r * kBlock + (int)threadIdx.x;
    if (i >= I.n) break;
    const int z = i % Z, t2 = i / Z;
    const int y = t2 % Y, xx = t2 / Y;
    const int izm = z == 0 ? i + (Z - 1) : i - 1;
    const int izp = z == Z - 1 ? i - (Z - 1) : i + 1;
    const int iym = y == 0 ? i + (Y - 1) * Z : i - Z;
    const int iyp = y == Y - 1 ? i - (Y - 1) * Z : i + Z;
    const int ixm = xx == 0 ? i + (X - 1) * YZ : i - YZ;
    const int ixp = xx == X - 1 ? i - (X - 1) * YZ : i + YZ;
    const T ck = val(i);
    xn[i] = (double)ck;
    const T dxp = ck - val(ixp), dyp = ck - val(iyp), dzp = ck - val(izp);
    const T a1 = ck + (T)1, a2 = ck - (T)1;
    const T be = C::fb(I) * ((a1 * a1) * (a2 * a2));
    T sq = dxp * dxp;
    sq = fma(dyp, dyp, sq);
    sq = fma(dzp, dzp, sq);
    acc[0] += (double)fma(C::fg(I), sq, be);
    const T dxm = ck - val(ixm), dym = ck - val(iym), dzm = ck - val(izm);
    const T lap = ((dxm + dxp) + (dym + dyp)) + (dzm + dzp);
    const T bg = C::fbg(I) * (ck * fma(ck, ck, (T)-1));
    const TV gv = (TV)fma(C::fg2(I), lap, bg);
    gn[i] = gv;
    const double a = (double)gv;
    const double b = (double)go[i];
    double dv[P + 1], wv[P > 0 ? P : 1];
#pragma unroll
    for (int k = 0; k < P; k++) dv[k] = (double)Dp[k][i];
    dv[P] = (double)d[i];
#pragma unroll
    for (int k = 0; k < P; k++) wv[k] = (double)Yp[k][i];
    const TV ys = (TV)(a - b);  // stored as TV; the dots use the stored value
    yo[i] = ys;
    const double yv = (double)ys;
#pragma unroll
    for (int k = 0; k <= P; k++) {
      acc[o + k] += a * dv[k];
      acc[o + 2 * kBS + 2 + k] += yv * dv[k];
    }
#pragma unroll
    for (int k = 0; k < P; k++) {
      acc[o + kBS + k] += a * wv[k];
      acc[o + 3 * kBS + 2 + k] += yv * wv[k];
    }
    acc[o + 2 * kBS] += a * a;
    acc[o + 2 * kBS + 1] += a * yv;
    acc[o + 4 * kBS + 2] += yv * yv;
  }
}

// The partials a step pass with P stored pairs sums (bit k: partial k).
__device__ __forceinline__ unsigned long long bstep_mask(int P) {
  const unsigned long long dm = (1ull << (P + 1)) - 1, ym = (1ull << P) - 1;
  const unsigned long long gon = dm | (ym << kBS) | (3ull << (2 * kBS)) | (dm << (2 * kBS + 2)) |
                                 (ym << (3 * kBS + 2)) | (1ull << (4 * kBS + 2));
  return 1ull | (gon << 1);
}

template <int BR, int P, typename TV>
__device__ __forceinline__ void bstep_prec(const BBufs& B, const BDyn& q, const BInst& I, int lb,
                                           double (&acc)[kKS]) {
  if (q.lp) bstep_body<BR, P, float, TV>(B, q, I, lb, acc);
  else bstep_body<BR, P, double, TV>(B, q, I, lb, acc);
}

// The step pass with the number of stored pairs as a compile-time value.
template <int BR, typename TV>
__device__ __forceinline__ void bstep_any(const BBufs& B, const BDyn& q, const BInst& I, int lb,
                                          double (&acc)[kKS]) {
  switch (q.np) {
    case 0: bstep_prec<BR, 0, TV>(B, q, I, lb, acc); break;
    case 1: bstep_prec<BR, 1, TV>(B, q, I, lb, acc); break;
    case 2: bstep_prec<BR, 2, TV>(B, q, I, lb, acc); break;
    case 3: bstep_prec<BR, 3, TV>(B, q, I, lb, acc); break;
    case 4: bstep_prec<BR, 4, TV>(B, q, I, lb, acc); break;
    case 5: bstep_prec<BR, 5, TV>(B, q, I, lb, acc); break;
    case 6: bstep_prec<BR, 6, TV>(B, q, I, lb, acc); break;
    case 7: bstep_prec<BR, 7, TV>(B, q, I, lb, acc); break;
    case 8: bstep_prec<BR, 8, TV>(B, q, I, lb, acc); break;
    case 9: bstep_prec<BR, 9, TV>(B, q, I, lb, acc); break;
    default: bstep_prec<BR, 10, TV>(B, q, I, lb, acc); break;
  }
}

template <int BR, typename TV>
__global__ __launch_bounds__(kBlock) void bstep_kernel(BBufs B) {
  const int p = find_pos(B.ablk, B.na, blockIdx.x);
  const BDyn& q = B.dyn[B.map ? B.map[p] : p];
  MINIM_BATCH_SKIP(B, q);
  const BInst I = B.inst[q.j];
  const int lb = blockIdx.x - B.ablk[p];
  double acc[kKS];
#pragma unroll
  for (int k = 0; k < kKS; k++) acc[k] = 0;
  bstep_any<BR, TV>(B, q, I, lb, acc);
  block_partials<kKS>(acc, bstep_mask(q.np), B.part);
}

// Device control: one launch for the line-search pass of the instances still
// searching and the step pass of those whose search has ended (the phase is
// uniform over an instance's blocks).  Both write rows of kKS partials.
template <int BR, typename TV>
__global__ __launch_bounds__(kBlock) void bts_kernel(BBufs B) {
  const int p = find_pos(B.ablk, B.na, blockIdx.x);
  const BDyn& q = B.dyn[B.map[p]];
  if (!(q.act & (kActSearch | kActStep))) return;
  const BInst I = B.inst[q.j];
  const int lb = blockIdx.x - B.ablk[p];
  if (q.act & kActStep) {
    double acc[kKS];
#pragma unroll
    for (int k = 0; k < kKS; k++) acc[k] = 0;
    bstep_any<BR, TV>(B, q, I, lb, acc);
    block_partials<kKS>(acc, bstep_mask(q.np), B.part);
  } else {
    double acc[kNT];
#pragma unroll
    for (int k = 0; k < kNT; k++) acc[k] = 0;
    if (q.lp) btrial_body<BR, float, TV>(B, q, I, lb, acc);
    else btrial_body<BR, double, TV>(B, q, I, lb, acc);
    block_partials<kNT, kKS>(acc, (1ull << kNT) - 1, B.part);
  }
}

// ---------------------------------------------------------------- device control
// The per-instance decisions of an iteration (two-loop recursion on the Gram
// data, the Armijo tests of the backtracking search, the Gram update, the
// mixed schedule's switch and the convergence test) as kernels with one thread
// per instance, so that an iteration is a fixed sequence of launches over all
// instances, each instance taking part in the phases its act bits name.  The
// arithmetic is the host path's, operation for operation.

struct BCtl {
  int np;             // stored pairs, oldest first
  int slot[kBS];      // their history slots
  double ptau[kBS];   // their step factors (s = tau d)
  double prho[kBS];   // 1 / s.y
  int used[kBS];
  double gg, e;
  double dg[kBS], yg[kBS], dres[kBS];
  double dy[kBS][kBS], yy[kBS][kBS];
  int gcur, xcur, dslot, iters, trial, stall;
  double tau0, tau, t;
  int started, done, searching, lp, fresh, lsFailed, converged;
  long long nEnergy, nGradient;
  int iterFP32, iterFP64;
  unsigned long long tstart;  // device wall clock when the instance was admitted
  unsigned long long tdone;   // ... and when it finished
  double conv;               // the State's convergence tolerance
  int n;                     // elements
};

struct BCtlArgs {
  BCtl* ctl;
  BDyn* dyn;
  const double* res;
  int ni;
  int m;
  double maxStep, initStep, switchRms;
  int maxIter;
  int spec, backtracking;
  size_t rPF, rF;                       // offsets in res: initial and fresh gradients
  const double* part;                   // block partials of the last pass
  const double* part2;                  // ... of the speculative trials
  const int* ablk;                      // the launch grid (block prefix sums)
  const int* map;                       // position -> instance
  int* live;                            // per instance: 1 while it runs
  unsigned long long* clock0;           // device wall clock at the start
};

// Up to kNT trials of the sequence (linesearch.cpp:15-27); true when the
// search has ended, with C.tau the step factor to take
__device__ bool bc_trials(BCtl& C, const double* e) {
  for (int k = 0; k < kNT; k++) {
    C.nEnergy++;
    if (C.e - e[k] >= C.t) return true;  // accepted at x + tau d
    C.tau *= 0.5;
    C.t *= 0.5;
    if (++C.trial >= 10) {
      C.lsFailed = 1;  // x + tau0 * 2^-10, not evaluated by the reference
      return true;
    }
  }
  return false;
}

__device__ void bc_direction(const BCtlArgs& A, int j, BCtl& C);

// Instances admitted to the batch (map[0..count), one block each, its first
// thread): the gradient and energy at their initial coordinates have been
// reduced into res[rPF], the state of iteration 0, and the first direction
// (steepest descent) is set.  setClock: the first admission, which starts
// the clock.
__global__ __launch_bounds__(kBlock) void bc_admit_kernel(BCtlArgs A, const int* __restrict__ map, int count,
                                                          int setClock) {
  const int p = blockIdx.x;
  if (threadIdx.x != 0) return;
  if (setClock && p == 0) *A.clock0 = wall_clock64();
  if (p >= count) return;
  const int j = map[p];
  BCtl& C = A.ctl[j];
  C.e = A.res[A.rPF + (size_t)j * kKP];
  C.gg = A.res[A.rPF + (size_t)j * kKP + 1];
  C.gcur = 0;
  C.nGradient++;
  C.started = 1;
  C.tstart = wall_clock64();
  A.live[j] = 1;
  bc_direction(A, j, C);
}

// Instances that have just left FP32 coordinates take part in the fresh
// phase: gradient and energy at the same point in FP64, overwriting g, and its
// dots with the stored pairs.  (The update left them with no phase; the
// others keep the phase it set.)
__global__ __launch_bounds__(kBlock) void bc_begin_kernel(BCtlArgs A) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= A.ni) return;
  const BCtl& C = A.ctl[j];
  BDyn& D = A.dyn[j];
  if (!C.started || C.done || !C.fresh) return;
  D.j = j;
  D.gcur = C.gcur;
  D.gout = C.gcur;
  D.xcur = C.xcur;
  D.lp = C.lp;
  D.np = C.np;
  for (int i = 0; i < C.np; i++) D.slot[i] = C.slot[i];
  D.act = kActFresh;
}

// A block's instance's control record, copied between global memory and LDS
// by the whole block (coalesced, in parallel), so that the first thread's
// decisions read and write LDS instead of making a chain of dependent global
// accesses (the record was last written by another kernel, possibly on
// another XCD, so each of them would miss in L2).
__device__ __forceinline__ void ctl_copy(BCtl* __restrict__ dst, const BCtl* __restrict__ src) {
  static_assert(sizeof(BCtl) % 8 == 0, "BCtl is copied in 8-byte words");
  constexpr int W = sizeof(BCtl) / 8;
  const auto* s = reinterpret_cast<const unsigned long long*>(src);
  auto* d = reinterpret_cast<unsigned long long*>(dst);
  for (int i = threadIdx.x; i < W; i += kBlock) d[i] = s[i];
}

// The mixed schedule: the fresh phase's results, then the direction
// coefficients (Lbfgs.cpp:84-127), one block per position, for the instances
// that took part in the fresh phase (the update set the others' directions).
__global__ __launch_bounds__(kBlock) void bc_direction_kernel(BCtlArgs A) {
  __shared__ BCtl Cs;
  const int p = blockIdx.x, j = A.map[p];
  if (!(A.dyn[j].act & kActFresh)) return;  // uniform over the block
  ctl_copy(&Cs, A.ctl + j);
  __syncthreads();
  if (threadIdx.x == 0) bc_direction(A, j, Cs);
  __syncthreads();
  ctl_copy(A.ctl + j, &Cs);
}

__device__ void bc_direction(const BCtlArgs& A, int j, BCtl& C) {
  BDyn& D = A.dyn[j];
  if (!C.started || C.done) {
    D.act = 0;
    return;
  }
  if (C.searching) {
    D.act = kActSearch;  // D still holds the direction and D.tau the next trial
    return;
  }
  if (D.act & kActFresh) {
    const double* r = A.res + A.rF + (size_t)j * kKF;
    C.e = r[0];
    C.gg = r[1];
    for (int i = 0; i < C.np; i++) {
      C.dg[C.slot[i]] = r[kKP + i];
      C.yg[C.slot[i]] = r[kKP + kBS + i];
    }
    C.fresh = 0;
    C.nGradient++;
    A.live[j] = 1;
  }
  int spare = -1;
  for (int s = 0; s <= A.m; s++)
    if (!C.used[s]) {
      spare = s;
      break;
    }
  C.dslot = spare;
  const int P = C.np;
  D.j = j;
  D.dslot = spare;
  D.gcur = C.gcur;
  D.gout = 1 - C.gcur;
  D.xcur = C.xcur;
  D.lp = C.lp;
  D.tau = 1.0;
  D.np = P;
  for (int i = 0; i < P; i++) D.slot[i] = C.slot[i];
  if (P == 0) {
    const double gnorm = sqrt(C.gg);
    D.cg = (gnorm > 0) ? -(A.initStep / gnorm) : -1.0;
  } else {
    // in LDS (the first thread of the block runs this): indexed by the
    // run-time pair count, in registers they would take a block of them
    __shared__ double alpha[kBS], ag_s[kBS], ay[kBS];
    for (int i = 0; i < P; i++) ag_s[i] = ay[i] = 0.0;
    double ag = -1.0;
    for (int k = P - 1; k >= 0; k--) {
      double sq = ag * (C.ptau[k] * C.dg[C.slot[k]]);
      for (int r = 0; r < P; r++) sq += ay[r] * (C.ptau[k] * C.dy[C.slot[k]][C.slot[r]]);
      alpha[k] = C.prho[k] * sq;
      ay[k] -= alpha[k];
    }
    const int nw = P - 1;
    const double gamma = 1 / (C.prho[nw] * C.yy[C.slot[nw]][C.slot[nw]]);
    ag *= gamma;
    for (int r = 0; r < P; r++) ay[r] *= gamma;
    for (int k = 0; k < P; k++) {
      double yr = ag * C.yg[C.slot[k]];
      for (int r = 0; r < P; r++)
        yr += ag_s[r] * (C.ptau[r] * C.dy[C.slot[r]][C.slot[k]]) + ay[r] * C.yy[C.slot[k]][C.slot[r]];
      const double beta = C.prho[k] * yr;
      ag_s[k] += alpha[k] - beta;
    }
    D.cg = ag;
    for (int r = 0; r < P; r++) {
      D.cD[r] = ag_s[r] * C.ptau[r];
      D.cY[r] = ay[r];
    }
  }
  D.act = kActDir;
}

// The phases after the direction, step and trial passes run one block per
// position of the launch grid: the block reduces the instance's partials
// (reduce_rows, as bfinal_kernel) into LDS, then its first thread takes the
// instance's decisions.  A block whose instance does not take part leaves at
// once (the update kernel's block still counts a running instance).

// The direction's dots: maxStep cap and flip (Lbfgs.cpp:53-57, :121-124), the
// Armijo slope, and the speculative trials at x + d (linesearch.cpp:8-29).
// The instance then takes its step, or searches.
__device__ void bc_after_dir(const BCtlArgs& A, BCtl& C, BDyn& D, const double* r, const double* rt);

__global__ __launch_bounds__(kBlock) void bc_after_dir_kernel(BCtlArgs A) {
  __shared__ double r[kKD], rt[kNT];
  const int p = blockIdx.x, j = A.map[p];
  BCtl& C = A.ctl[j];
  BDyn& D = A.dyn[j];
  if (!(D.act & kActDir)) return;
  const int b0 = A.ablk[p], nb = A.ablk[p + 1] - b0;
  reduce_rows(A.part, b0, nb, kKD, r);
  if (A.spec) reduce_rows(A.part2, b0, nb, kNT, rt);
  __syncthreads();
  if (threadIdx.x == 0) bc_after_dir(A, C, D, r, rt);
}

// r: the direction's partial sums, rt: the speculative trials' energies.
__device__ void bc_after_dir(const BCtlArgs& A, BCtl& C, BDyn& D, const double* r, const double* rt) {
  const double dd = r[0], dgv = r[1];
  const int P = C.np;
  for (int i = 0; i < P; i++) C.dres[C.slot[i]] = r[2 + i];
  double tau0 = 1.0;
  if (P > 0 && A.maxStep != 0) {
    const double ss = sqrt(dd);
    if (ss > A.maxStep) tau0 = A.maxStep / ss;
  }
  double gs = tau0 * dgv;
  if (gs > 0) {
    gs = -gs;
    tau0 = -tau0;
  }
  C.tau0 = tau0;
  C.tau = tau0;
  C.trial = 0;
  C.searching = 0;
  C.lsFailed = 0;
  if (A.backtracking) {
    C.t = -0.5 * gs;
    bool ended = false;
    if (A.spec) {
      if (tau0 == 1.0) ended = bc_trials(C, rt);
      else C.nEnergy++;  // the speculative trials at x + d were not on this search's sequence
    }
    if (!ended) {
      C.searching = 1;
      D.tau = C.tau;
      D.act = kActSearch;
      return;
    }
  }
  D.tau = C.tau;
  D.act = kActStep;
}

// The next kNT trials of the instances searching; those whose search ends
// take their step in this iteration, the others search on in the next.
__device__ __forceinline__ void bc_consume(BCtl& C, BDyn& D, const double* rt) {
  const bool ended = bc_trials(C, rt);
  D.tau = C.tau;
  if (ended) {
    C.searching = 0;
    D.act = kActStep;
  }
}

// New energy and the Gram update (Lbfgs.cpp:58-80), the mixed schedule's
// switch, and Minimiser::minimise's end tests (Minimiser.cpp:27-38,
// Lbfgs.cpp:130-133).
__device__ void bc_update(const BCtlArgs& A, int j, BCtl& C, const double* r);

// After bts_kernel, one block per position.  An instance that took its step:
// the update, then the next iteration's direction, except for an instance
// that has just switched to FP64 coordinates, which recomputes its gradient
// first (no phase until bc_begin_kernel).  An instance that searched: the
// trials' Armijo tests (bc_consume); if its search has ended it takes the
// step in the next launch sequence, where it takes no new direction.
__global__ __launch_bounds__(kBlock) void bc_finish_kernel(BCtlArgs A) {
  __shared__ double rs[kKS];
  __shared__ BCtl Cs;
  const int p = blockIdx.x, j = A.map[p];
  BDyn& D = A.dyn[j];
  const int b0 = A.ablk[p], nb = A.ablk[p + 1] - b0;
  if (D.act & kActStep) {  // uniform over the block
    ctl_copy(&Cs, A.ctl + j);
    reduce_rows(A.part, b0, nb, kKS, rs);
    __syncthreads();
    if (threadIdx.x == 0) {
      bc_update(A, j, Cs, rs);
      if (Cs.fresh && !Cs.done) D.act = 0;
      else bc_direction(A, j, Cs);
    }
    __syncthreads();
    ctl_copy(A.ctl + j, &Cs);
  } else if (D.act & kActSearch) {
    reduce_rows(A.part, b0, nb, kNT, rs, kKS);
    __syncthreads();
    if (threadIdx.x == 0) bc_consume(A.ctl[j], D, rs);
  }
}

__device__ void bc_update(const BCtlArgs& A, int j, BCtl& C, const double* rsum) {
  {
    const double* r = rsum;
    C.e = r[0];
    C.nGradient++;
    r += 1;
    const double gng = r[2 * kBS], gny = r[2 * kBS + 1];
    const double yyv = r[4 * kBS + 2];
    const int ds = C.dslot;
    const int P = C.np;
    __shared__ int oldSlots[kBS];  // as alpha in bc_direction
    for (int i = 0; i < P; i++) oldSlots[i] = C.slot[i];
    const double sy = C.tau * r[2 * kBS + 2 + P];
    if (sy != 0) {
      for (int i = 0; i < P; i++) {
        const int sl = oldSlots[i];
        C.dy[ds][sl] = C.dres[sl];  // d_new.y_old from the direction pass
        C.dy[sl][ds] = r[2 * kBS + 2 + i];
        C.yy[ds][sl] = r[3 * kBS + 2 + i];
        C.yy[sl][ds] = r[3 * kBS + 2 + i];
      }
      C.dy[ds][ds] = r[2 * kBS + 2 + P];
      C.yy[ds][ds] = yyv;
      C.dg[ds] = r[P];
      C.yg[ds] = gny;
      C.used[ds] = 1;
      C.slot[C.np] = ds;
      C.ptau[C.np] = C.tau;
      C.prho[C.np] = 1 / sy;
      C.np++;
      if (C.np > A.m) {
        C.used[C.slot[0]] = 0;
        for (int i = 1; i < C.np; i++) {
          C.slot[i - 1] = C.slot[i];
          C.ptau[i - 1] = C.ptau[i];
          C.prho[i - 1] = C.prho[i];
        }
        C.np--;
      }
    }
    for (int i = 0; i < P; i++) {
      C.dg[oldSlots[i]] = r[i];
      C.yg[oldSlots[i]] = r[kBS + i];
    }
    C.gg = gng;
    C.gcur = 1 - C.gcur;
    C.xcur = 1 - C.xcur;
    C.iters++;
    const double rms = sqrt(C.gg / C.n);
    if (C.lp) {
      const double sw = A.switchRms > 0 ? A.switchRms : 1e3 * C.conv;
      C.stall = C.lsFailed ? C.stall + 1 : 0;
      C.iterFP32++;
      if (rms < sw || C.stall >= 2) {
        C.lp = 0;
        C.fresh = 1;  // no convergence test until g is recomputed
        A.live[j] = 2;  // running, waiting for the fresh phase
      }
    } else {
      C.iterFP64++;
    }
    if (!C.fresh && rms < C.conv) {
      C.converged = 1;
      C.done = 1;
    } else if (C.iters > A.maxIter) {
      C.done = 1;
    }
    if (C.done) {
      C.tdone = wall_clock64();
      A.live[j] = 0;
    }
  }
}

// ---------------------------------------------------------------- resident instances
// Resident control: one workgroup per instance runs that instance's whole
// minimisation in one launch -- the same pass bodies, taken over the
// instance's blocks in turn (each thread adds its elements' contributions in
// a fixed order, then the workgroup sums them as block_partials does), and
// the same decision functions, with workgroup barriers in place of kernel
// boundaries.  The line search loops until it ends, so an iteration costs no
// launch and no host synchronisation; small instances finish in the time
// their own iterations take, instead of waiting for launches sized for the
// whole batch.  Every instance of the batch is in flight at once (or as soon
// as a workgroup slot frees), largest first.  Sums are split differently
// from the block-per-launch passes, so results agree with the other controls
// to rounding.  One workgroup streams only a few tens of GB/s, so this pays
// for small instances (a few thousand sites) and loses for large ones.

// out[k] = the workgroup's sum of v[k] for the bits of `on` (the tree of
// block_partials), visible to every thread on return.
template <int K>
__device__ __forceinline__ void wg_sum(const double (&v)[K], unsigned long long on, double* out) {
  __shared__ double lds[(kBlock / kWave) * K];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < K; k++) {
    double x = 0;
    if ((on >> k) & 1ull) {
      x = v[k];
#pragma unroll
      for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
    }
    if (lane == 0) lds[wid * K + k] = x;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    double s = lds[k];
    for (int w = 1; w < kBlock / kWave; w++) s += lds[w * K + k];
    out[k] = s;
  }
  __syncthreads();
}

// A whole pass over the instance's blocks, the number of stored pairs fixed
// outside the loop (FP64 arithmetic).
template <int BR, int P, typename TV>
__device__ __forceinline__ void bres_dir(const BBufs& B, const BDyn& D, const BInst& I, int nblk,
                                         double (&acc)[kKD]) {
#pragma unroll 1
  for (int lb = 0; lb < nblk; lb++) bdir_body<BR, P, TV>(B, D, I, lb, acc);
}
template <int BR, int P, typename TV>
__device__ __forceinline__ void bres_step(const BBufs& B, const BDyn& D, const BInst& I, int nblk,
                                          double (&acc)[kKS]) {
#pragma unroll 1
  for (int lb = 0; lb < nblk; lb++) bstep_body<BR, P, double, TV>(B, D, I, lb, acc);
}
// D.np <= PM: the cases above PM are not compiled, so the kernel's registers
// are those of the passes with at most PM pairs.
#define MINIM_BRES_CASE(FN, K, ...) \
  case K:                           \
    if constexpr (K < PM) {         \
      FN<BR, K, TV>(__VA_ARGS__);   \
      break;                        \
    }                               \
    [[fallthrough]];
#define MINIM_BRES_SWITCH(FN, ...)               \
  switch (D.np) {                                \
    MINIM_BRES_CASE(FN, 0, __VA_ARGS__)          \
    MINIM_BRES_CASE(FN, 1, __VA_ARGS__)          \
    MINIM_BRES_CASE(FN, 2, __VA_ARGS__)          \
    MINIM_BRES_CASE(FN, 3, __VA_ARGS__)          \
    MINIM_BRES_CASE(FN, 4, __VA_ARGS__)          \
    MINIM_BRES_CASE(FN, 5, __VA_ARGS__)          \
    MINIM_BRES_CASE(FN, 6, __VA_ARGS__)          \
    MINIM_BRES_CASE(FN, 7, __VA_ARGS__)          \
    MINIM_BRES_CASE(FN, 8, __VA_ARGS__)          \
    MINIM_BRES_CASE(FN, 9, __VA_ARGS__)          \
    default: FN<BR, PM, TV>(__VA_ARGS__); break; \
  }

template <int BR, int PM, typename TV>
__global__ __launch_bounds__(kBlock) void bres_kernel(BBufs B, BCtlArgs A) {
  __shared__ BCtl Cs;
  __shared__ double rsum[kKS], rt[kNT];
  const int j = A.map[blockIdx.x];
  const BInst I = B.inst[j];
  BDyn& D = A.dyn[j];
  const int nblk = (I.n + BR * kBlock - 1) / (BR * kBlock);
  ctl_copy(&Cs, A.ctl + j);
  __syncthreads();
  {  // iteration 0: gradient and energy at the initial coordinates
    double acc[kKP] = {0.0, 0.0};
#pragma unroll 1
    for (int lb = 0; lb < nblk; lb++) bpf_body<BR, 3, double, TV>(B, D, I, lb, acc);
    wg_sum<kKP>(acc, 3ull, rsum);
    if (threadIdx.x == 0) {
      Cs.e = rsum[0];
      Cs.gg = rsum[1];
      Cs.gcur = 0;
      Cs.nGradient++;
      Cs.started = 1;
      Cs.tstart = wall_clock64();
    }
    __syncthreads();
  }
  for (;;) {
    if (threadIdx.x == 0) bc_direction(A, j, Cs);
    __syncthreads();
    if (!(D.act & kActDir)) break;  // finished
    {
      double acc[kKD];
#pragma unroll
      for (int k = 0; k < kKD; k++) acc[k] = 0;
      MINIM_BRES_SWITCH(bres_dir, B, D, I, nblk, acc)
      wg_sum<kKD>(acc, (1ull << (2 + D.np)) - 1, rsum);
      if (A.spec) {  // D.tau is 1: the speculative trials at x + 2^-k d
        double at[kNT] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
        for (int lb = 0; lb < nblk; lb++) btrial_body<BR, double, TV>(B, D, I, lb, at);
        wg_sum<kNT>(at, (1ull << kNT) - 1, rt);
      }
      if (threadIdx.x == 0) bc_after_dir(A, Cs, D, rsum, rt);
      __syncthreads();
    }
    while (D.act & kActSearch) {
      double at[kNT] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
      for (int lb = 0; lb < nblk; lb++) btrial_body<BR, double, TV>(B, D, I, lb, at);
      wg_sum<kNT>(at, (1ull << kNT) - 1, rt);
      if (threadIdx.x == 0) bc_consume(Cs, D, rt);
      __syncthreads();
    }
    {
      double acc[kKS];
#pragma unroll
      for (int k = 0; k < kKS; k++) acc[k] = 0;
      MINIM_BRES_SWITCH(bres_step, B, D, I, nblk, acc)
      wg_sum<kKS>(acc, bstep_mask(D.np), rsum);
      if (threadIdx.x == 0) bc_update(A, j, Cs, rsum);
      __syncthreads();
    }
    if (Cs.done) break;
  }
  ctl_copy(A.ctl + j, &Cs);
}

// Host side ------------------------------------------------------------------

struct Pair {
  int slot;
  double tau;
  double rho;
};

struct Inst {
  minim::State* s = nullptr;
  BInst bi{};
  int nblk = 0;
  double conv = 0;
  std::deque<Pair> pairs;
  std::vector<char> used;
  double gg = 0, e = 0;
  std::vector<double> dg, yg, dres;
  std::vector<std::vector<double>> dy, yy;
  int gcur = 0;
  int xcur = 0;  // coordinate buffer that holds x
  int iters = 0;
  bool done = false;
  minim::MinimiseStats st;
  // the current iteration
  int dslot = 0;
  double tau0 = 1, tau = 1, t = 0, gs = 0;
  int trial = 0;
  bool searching = false;
  // mixed schedule
  bool lp = false;     // FP32 coordinates and arithmetic
  bool fresh = false;  // g to be recomputed from FP64 coordinates, history kept
  bool lsFailed = false;
  int stall = 0;
};

using KFn = void (*)(BBufs);
struct Kernels {
  KFn pf_eg, pf_trial, dir, step, fresh, round, ts;
  void (*resident5)(BBufs, BCtlArgs);   // FP64 only, m <= 5
  void (*resident10)(BBufs, BCtlArgs);  // FP64 only, m <= 10
};
template <int BR, typename TV>
Kernels kernels_br() {
  Kernels k{bpf_kernel<BR, 3, TV>,  btrial_kernel<BR, TV>, bdir_kernel<BR, TV>, bstep_kernel<BR, TV>,
            bpf_kernel<BR, 11, TV>, bround_kernel<BR>,     bts_kernel<BR, TV>,  nullptr,
            nullptr};
  if constexpr (sizeof(TV) == 8) {
    k.resident5 = bres_kernel<BR, 5, TV>;
    k.resident10 = bres_kernel<BR, 10, TV>;
  }
  return k;
}
template <typename TV>
Kernels kernels_for(int br) {
  switch (br) {
    case 4: return kernels_br<4, TV>();
    case 16: return kernels_br<16, TV>();
    default: return kernels_br<8, TV>();
  }
}

// One launch group: the instances taking part, their per-phase parameters and
// block prefix sums, staged in pinned memory and copied to the device.
struct Launch {
  std::vector<int> pos;  // instance index per active position
  int na = 0, nb = 0;
};

}  // namespace
}  // namespace minim_amd

namespace minim {

std::vector<MinimiseStats> Lbfgs::minimiseBatch(const std::vector<State*>& states) {
  return batchRun(states, 0, 0, true);
}

std::vector<std::array<double, 2>> Lbfgs::autotuneBatch(const std::vector<State*>& states, std::vector<int> candidates,
                                                        int replays) {
  if (states.empty() || candidates.empty() || replays < 1)
    throw std::invalid_argument("Lbfgs::autotuneBatch: nothing to time");
  for (int c : candidates)
    if (c != 4 && c != 8 && c != 16)
      throw std::invalid_argument("Lbfgs::autotuneBatch: candidates must be 4, 8 or 16 elements per thread.");
  std::vector<std::array<double, 2>> rows;
  int best = 0;
  double bestS = 0;
  for (int c : candidates) {
    batchRun(states, c, replays, false);  // first run of each shape: code objects, buffers
    MINIM_HIP_CHECK(hipDeviceSynchronize());
    const auto t0 = std::chrono::steady_clock::now();
    batchRun(states, c, replays, false);
    MINIM_HIP_CHECK(hipDeviceSynchronize());
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    rows.push_back({(double)c, sec});
    if (best == 0 || sec < bestS) {
      best = c;
      bestS = sec;
    }
  }
  size_t N = 0;
  for (const State* s : states) N += s->ndof;
  // resident control keys its own entry: one workgroup per instance favours
  // other sizes than launches over every instance
  minim_amd::tune_store(minim_amd::tune_key(_batchControl == "resident" ? "batch-block-resident" : "batch-block",
                                            {(long)states.size(), (long)N, (long)_m}, _precMode == "mixed" ? 1 : 0),
                        {best});
  return rows;
}

std::vector<MinimiseStats> Lbfgs::batchRun(const std::vector<State*>& states, int block, long replayLimit,
                                           bool writeBack) {
  using namespace minim_amd;
  const int ni = (int)states.size();
  std::vector<MinimiseStats> out(ni);
  if (ni == 0) return out;
  if (replayLimit > 0) writeBack = false;
  if (_m > kBS - 1) throw std::invalid_argument("Lbfgs::minimiseBatch: m must be at most 10.");
  if (_precMode == "fp32")
    throw std::invalid_argument("Lbfgs::minimiseBatch runs in FP64 or the mixed schedule, not in FP32 throughout.");
  const bool resident = _batchControl == "resident";
  if (resident && _precMode == "mixed")
    throw std::invalid_argument("Lbfgs::minimiseBatch: resident control runs the FP64 schedule only.");
  // mixed: FP32 coordinates and arithmetic until sqrt(g.g/ndof) < switchRms (or
  // the line search stalls twice), then FP64 coordinates; the vectors stay in
  // FP32 throughout, as the single-State schedule does for a PhaseField
  const bool mixed = _precMode == "mixed";
  const size_t vb = mixed ? 4 : 8;
  {
    size_t nTot = 0;
    for (const State* s : states)
      if (s) nTot += s->ndof;
    if (block == 0) block = _batchBlock;
    std::vector<int> tuned;
    const char* tkey = _batchControl == "resident" ? "batch-block-resident" : "batch-block";
    if (block == 0 && tune_lookup(tune_key(tkey, {(long)ni, (long)nTot, (long)_m}, mixed ? 1 : 0), tuned) &&
        !tuned.empty())
      block = tuned[0];
    if (block != 4 && block != 8 && block != 16) block = nTot < kSmallBatch ? kBlockSmall : kBlockDefault;
  }
  const int kBE = block * kBlock;  // elements per block
  const Kernels K = mixed ? kernels_for<float>(block) : kernels_for<double>(block);

  std::vector<Inst> in(ni);
  size_t N = 0;
  long totalBlocks = 0;
  for (int j = 0; j < ni; j++) {
    State* s = states[j];
    if (!s) throw std::invalid_argument("Lbfgs::minimiseBatch: null State.");
    for (int k = 0; k < j; k++)
      if (states[k] == s) throw std::invalid_argument("Lbfgs::minimiseBatch: a State appears twice.");
    if (_precMode == "fp64") s->setPrecision(FP64);
    if (s->precision() != FP64)
      throw std::invalid_argument("Lbfgs::minimiseBatch: every State must hold FP64 coordinates.");
    const auto* pf = dynamic_cast<const PhaseField*>(s->pot.get());
    int c[3];
    double fb = 0, fg = 0;
    if (!pf || !pf->plainStencil(c, fb, fg))
      throw std::invalid_argument(
          "Lbfgs::minimiseBatch: every State must hold a one-fluid PhaseField without solid nodes, extra "
          "terms or fixed fluid.");
    const size_t n = s->ndof;
    if (n != (size_t)c[0] * c[1] * c[2] || n > (size_t)INT_MAX / 2)
      throw std::invalid_argument("Lbfgs::minimiseBatch: unsupported field size.");
    Inst& I = in[j];
    I.s = s;
    I.bi = BInst{N, (int)n, c[0], c[1], c[2], fb, fg, 4.0 * fb, 2.0 * fg,
                 (float)fb, (float)fg, (float)(4.0 * fb), (float)(2.0 * fg)};
    I.lp = mixed;
    I.nblk = (int)((n + kBE - 1) / kBE);
    I.conv = s->convergence;
    I.used.assign(_m + 1, 0);
    I.dg.assign(kBS, 0.0);
    I.yg.assign(kBS, 0.0);
    I.dres.assign(kBS, 0.0);
    I.dy.assign(kBS, std::vector<double>(kBS, 0.0));
    I.yy.assign(kBS, std::vector<double>(kBS, 0.0));
    N += n;
    totalBlocks += I.nblk;
  }
  if (totalBlocks > INT_MAX / kKS) throw std::invalid_argument("Lbfgs::minimiseBatch: too many elements.");

  hipStream_t st = static_cast<hipStream_t>(states[0]->stream());
  MINIM_HIP_CHECK(hipDeviceSynchronize());  // every State's coordinates are final
  const auto t0 = std::chrono::steady_clock::now();
  auto since = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };

  const int S = _m + 1;
  DeviceBuffer Xa(N * 8), Xb(N * 8), G0(N * vb), G1(N * vb), Dm(N * vb * S), Ym(N * vb * S);
  // per-phase staging, one copy per phase: block prefix sums, then the BDyn of
  // each instance in the launch
  const size_t ablkBytes = (sizeof(int) * (ni + 1) + 15) / 16 * 16;
  const size_t stageBytes = ablkBytes + sizeof(BDyn) * ni;
  DeviceBuffer instD(sizeof(BInst) * ni), stageD(stageBytes);
  // results of one fetch: fresh (pf + dots), direction + trials, or step
  constexpr int kRes = std::max(std::max(kKF, kKD + kNT), kKS);
  DeviceBuffer part(sizeof(double) * (size_t)totalBlocks * kKS), res(sizeof(double) * (size_t)ni * kRes);
  struct Pinned {
    void* p = nullptr;
    explicit Pinned(size_t b) { MINIM_HIP_CHECK(hipHostMalloc(&p, b)); }
    ~Pinned() { (void)hipHostFree(p); }
  };
  Pinned stageH(stageBytes), resH(sizeof(double) * (size_t)ni * kRes);
  int* ablkh = static_cast<int*>(stageH.p);
  BDyn* dynh = reinterpret_cast<BDyn*>(static_cast<char*>(stageH.p) + ablkBytes);
  double* resh = static_cast<double*>(resH.p);

  {
    std::vector<BInst> bi(ni);
    for (int j = 0; j < ni; j++) bi[j] = in[j].bi;
    MINIM_HIP_CHECK(hipMemcpyAsync(instD.get(), bi.data(), sizeof(BInst) * ni, hipMemcpyHostToDevice, st));
    for (int j = 0; j < ni; j++)
      device_copy(static_cast<double*>(Xa.get()) + in[j].bi.off, in[j].s->deviceCoords(), in[j].bi.n * 8, st);
    MINIM_HIP_CHECK(hipStreamSynchronize(st));  // bi leaves scope
  }

  BBufs B{};
  B.X0 = static_cast<double*>(Xa.get());
  B.X1 = static_cast<double*>(Xb.get());
  B.G0 = G0.get();
  B.G1 = G1.get();
  B.D = Dm.get();
  B.Y = Ym.get();
  B.N = N;
  B.inst = static_cast<const BInst*>(instD.get());
  B.ablk = static_cast<const int*>(stageD.get());
  B.dyn = reinterpret_cast<const BDyn*>(static_cast<const char*>(stageD.get()) + ablkBytes);
  B.part = static_cast<double*>(part.get());
  double* resd = static_cast<double*>(res.get());

  // Stage the instances listed in L (their BDyn already in dynh[0..na)) and
  // upload; the pinned staging is rewritten only after the next fetch.
  auto stage = [&](Launch& L) {
    L.na = (int)L.pos.size();
    ablkh[0] = 0;
    for (int p = 0; p < L.na; p++) ablkh[p + 1] = ablkh[p] + in[L.pos[p]].nblk;
    L.nb = ablkh[L.na];
    MINIM_HIP_CHECK(
        hipMemcpyAsync(stageD.get(), stageH.p, ablkBytes + sizeof(BDyn) * L.na, hipMemcpyHostToDevice, st));
    B.na = L.na;
  };
  auto finalize = [&](const Launch& L, int K, size_t resOff) {
    bfinal_kernel<<<L.na, kBlock, 0, st>>>(B.part, B.ablk, nullptr, K, resd + resOff, B.dyn, 0u);
    MINIM_HIP_CHECK(hipGetLastError());
  };
  long hostSyncs = 0, graphLaunches = 0;
  auto fetch = [&](size_t count) {
    MINIM_HIP_CHECK(hipMemcpyAsync(resh, resd, sizeof(double) * count, hipMemcpyDeviceToHost, st));
    MINIM_HIP_CHECK(hipStreamSynchronize(st));
    hostSyncs++;
  };
  auto dyn_base = [&](const Inst& I) {
    BDyn q{};
    q.j = (int)(&I - in.data());
    q.dslot = I.dslot;
    q.gcur = I.gcur;
    q.gout = 1 - I.gcur;
    q.xcur = I.xcur;
    q.lp = I.lp ? 1 : 0;
    q.tau = 0;
    return q;
  };
  auto launch = [&](KFn f, const Launch& L) {
    f<<<L.nb, kBlock, 0, st>>>(B);
    MINIM_HIP_CHECK(hipGetLastError());
  };
  const bool backtracking = linesearch == "backtracking";
  const bool spec = backtracking && _maxStep == 0;  // first trial at x + d, launched with the direction
  int maxIters = 0;

  if (_batchControl != "host") {
    // ---- device control: the control kernels (one block per instance, its
    // first thread deciding) set which instances take part in each phase, and
    // one iteration is a fixed sequence of launches -- launched as it is
    // ("device", the default) or captured as a HIP graph and replayed
    // ("graph").  An instance whose line search goes past the speculative
    // trials and one further pass carries it over to the next iteration, where
    // it takes no new direction.  The host reads the live flags every kChunk
    // iterations.  The launch grid covers the running instances only (block
    // prefix sums and a position -> instance map): when their blocks have
    // fallen below 4/5 of it, it is laid out again (and the graph captured
    // again).  Where an instance's rows lie in
    // the partials does not change its sums.
    //
    // With a budget (setBatchBudget), instances are admitted smallest first
    // while the running ones hold at most that many sites, so that the first
    // iterations are short and small instances finish early; once half of
    // the instances have finished, every remaining one is admitted, so that
    // the rest of the batch runs at the throughput of the whole.
    constexpr int kChunk = 8;
    const bool useGraph = _batchControl == "graph";
    std::vector<BDyn> dy0(ni);
    std::vector<BCtl> c0(ni);
    for (int j = 0; j < ni; j++) {
      BDyn q{};
      q.j = j;
      q.lp = mixed ? 1 : 0;
      dy0[j] = q;
      BCtl c{};
      c.conv = in[j].conv;
      c.n = in[j].bi.n;
      c.lp = mixed ? 1 : 0;
      c0[j] = c;
    }
    const size_t rPF = 0, rF = rPF + (size_t)ni * kKP, rEnd = rF + (size_t)ni * kKF;
    // the start clock at byte 8, the live flags from byte 64
    const size_t cntBytes = 64 + sizeof(int) * (size_t)ni;
    DeviceBuffer ablkA(sizeof(int) * (ni + 1)), mapA(sizeof(int) * ni), dynA(sizeof(BDyn) * ni),
        ctlA(sizeof(BCtl) * ni), resA(8 * rEnd), cntA(cntBytes), part2(sizeof(double) * (size_t)totalBlocks * kNT);
    Pinned cntH(cntBytes), cntH2(cntBytes);  // the flags of two chunks in turn
    std::memset(cntH.p, 0, cntBytes);
    MINIM_HIP_CHECK(hipMemcpyAsync(dynA.get(), dy0.data(), sizeof(BDyn) * ni, hipMemcpyHostToDevice, st));
    MINIM_HIP_CHECK(hipMemcpyAsync(ctlA.get(), c0.data(), sizeof(BCtl) * ni, hipMemcpyHostToDevice, st));
    MINIM_HIP_CHECK(hipMemcpyAsync(cntA.get(), cntH.p, cntBytes, hipMemcpyHostToDevice, st));
    MINIM_HIP_CHECK(hipStreamSynchronize(st));
    // a stream of its own: graph capture needs one other than the null stream
    hipStream_t cs = nullptr;
    MINIM_HIP_CHECK(hipStreamCreate(&cs));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } csGuard{cs};
    BBufs G = B;
    G.ablk = static_cast<const int*>(ablkA.get());
    G.map = static_cast<const int*>(mapA.get());
    G.dyn = static_cast<const BDyn*>(dynA.get());
    G.phase = 0;
    double* res = static_cast<double*>(resA.get());
    BCtlArgs A{};
    A.ctl = static_cast<BCtl*>(ctlA.get());
    A.dyn = static_cast<BDyn*>(dynA.get());
    A.res = res;
    A.ni = ni;
    A.m = _m;
    A.maxStep = _maxStep;
    A.initStep = _initStep;
    A.switchRms = _switchRms;
    A.maxIter = maxIter;
    A.spec = spec ? 1 : 0;
    A.backtracking = backtracking ? 1 : 0;
    A.rPF = rPF;
    A.rF = rF;
    A.part = G.part;
    A.part2 = static_cast<const double*>(part2.get());
    A.ablk = G.ablk;
    A.map = G.map;
    A.clock0 = reinterpret_cast<unsigned long long*>(static_cast<char*>(cntA.get()) + 8);
    A.live = reinterpret_cast<int*>(static_cast<char*>(cntA.get()) + 64);
    const int cb = (ni + kBlock - 1) / kBlock;
    // the launch grid over a list of instances
    std::vector<int> ab(ni + 1, 0), mp(ni, 0);
    int naCur = 0, nbCur = 0;
    auto layout = [&](const std::vector<int>& list) {
      naCur = 0;
      for (int j : list) {
        mp[naCur] = j;
        ab[naCur + 1] = ab[naCur] + in[j].nblk;
        naCur++;
      }
      nbCur = ab[naCur];
      G.na = naCur;
      MINIM_HIP_CHECK(hipMemcpyAsync(ablkA.get(), ab.data(), sizeof(int) * (naCur + 1), hipMemcpyHostToDevice, cs));
      MINIM_HIP_CHECK(hipMemcpyAsync(mapA.get(), mp.data(), sizeof(int) * naCur, hipMemcpyHostToDevice, cs));
      MINIM_HIP_CHECK(hipStreamSynchronize(cs));  // ab and mp are rewritten on the next layout
    };
    auto run = [&](KFn f, unsigned phase) {
      BBufs b = G;
      b.phase = phase;
      f<<<nbCur, kBlock, 0, cs>>>(b);
      MINIM_HIP_CHECK(hipGetLastError());
    };
    auto fin = [&](unsigned phase, int K, size_t off) {
      bfinal_kernel<<<naCur, kBlock, 0, cs>>>(G.part, G.ablk, G.map, K, res + off, G.dyn, phase);
      MINIM_HIP_CHECK(hipGetLastError());
    };
    auto ctl = [&](void (*k)(BCtlArgs)) {  // one thread per instance
      k<<<cb, kBlock, 0, cs>>>(A);
      MINIM_HIP_CHECK(hipGetLastError());
    };
    auto ctlp = [&](void (*k)(BCtlArgs)) {  // one block per position
      k<<<naCur, kBlock, 0, cs>>>(A);
      MINIM_HIP_CHECK(hipGetLastError());
    };
    // An iteration: 5 launches (FP64, with the speculative trials).  With
    // fresh (the mixed schedule, only when an instance has switched to FP64
    // coordinates since the host last looked), the fresh phase first; a
    // switched instance waits for it with no phase, at most kChunk iterations.
    auto iteration = [&](bool fresh) {
      if (fresh) {
        ctl(bc_begin_kernel);
        run(K.fresh, kActFresh);
        fin(kActFresh, kKF, rF);
        ctlp(bc_direction_kernel);
      }
      run(K.dir, kActDir);
      if (spec) {
        BBufs b = G;
        b.phase = kActDir;
        b.part = static_cast<double*>(part2.get());
        K.pf_trial<<<nbCur, kBlock, 0, cs>>>(b);
        MINIM_HIP_CHECK(hipGetLastError());
      }
      ctlp(bc_after_dir_kernel);
      run(K.ts, kActSearch | kActStep);
      ctlp(bc_finish_kernel);
    };
    struct GraphGuard {
      hipGraph_t g = nullptr;
      hipGraphExec_t e = nullptr;
      void reset() {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
        e = nullptr;
        g = nullptr;
      }
      ~GraphGuard() { reset(); }
    } graph;
    long captures = 0;
    auto capture = [&] {
      graph.reset();
      MINIM_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      iteration(false);
      MINIM_HIP_CHECK(hipStreamEndCapture(cs, &graph.g));
      MINIM_HIP_CHECK(hipGraphInstantiate(&graph.e, graph.g, nullptr, nullptr, 0));
      captures++;
    };

    long replays = 0;
    if (resident) {
      // every instance in one launch, one workgroup each, largest first
      std::vector<int> order(ni);
      for (int j = 0; j < ni; j++) order[j] = j;
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return in[a].bi.n > in[b].bi.n; });
      MINIM_HIP_CHECK(hipMemcpyAsync(mapA.get(), order.data(), sizeof(int) * ni, hipMemcpyHostToDevice, cs));
      bc_admit_kernel<<<1, kWave, 0, cs>>>(A, nullptr, 0, 1);  // the start clock
      MINIM_HIP_CHECK(hipGetLastError());
      // autotuneBatch's bounded runs: at most replayLimit iterations each
      if (replayLimit > 0) A.maxIter = (int)std::min<long>(maxIter, replayLimit - 1);
      auto* kres = _m <= 5 ? K.resident5 : K.resident10;
      if (!kres) throw std::logic_error("Lbfgs::minimiseBatch: no resident kernel for this storage type.");
      BBufs R = G;
      R.map = nullptr;
      R.ablk = nullptr;
      kres<<<ni, kBlock, 0, cs>>>(R, A);
      MINIM_HIP_CHECK(hipGetLastError());
      MINIM_HIP_CHECK(hipStreamSynchronize(cs));  // order leaves scope
      hostSyncs++;
    } else {
    // admission order: ascending size
    std::vector<int> order(ni);
    for (int j = 0; j < ni; j++) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return in[a].bi.n < in[b].bi.n; });
    int next = 0, finished = 0;
    long runningSites = 0;
    std::vector<char> running(ni, 0);
    auto admit = [&] {
      std::vector<int> add;
      const bool open = _batchBudget == 0 || 2 * finished >= ni;
      while (next < ni) {
        const int j = order[next];
        if (!open && runningSites > 0 && runningSites + in[j].bi.n > _batchBudget) break;
        runningSites += in[j].bi.n;
        add.push_back(j);
        next++;
      }
      return add;
    };
    // gradient and energy at the admitted instances' initial coordinates (the
    // mixed schedule rounds x to FP32 first), then the grid over every running
    // instance and the graph
    bool firstAdmission = true;
    auto start = [&](const std::vector<int>& add) {
      layout(add);
      if (mixed) {
        K.round<<<nbCur, kBlock, 0, cs>>>(G);
        MINIM_HIP_CHECK(hipGetLastError());
      }
      run(K.pf_eg, 0);
      fin(0, kKP, rPF);
      bc_admit_kernel<<<naCur, kWave, 0, cs>>>(A, G.map, naCur, firstAdmission ? 1 : 0);
      MINIM_HIP_CHECK(hipGetLastError());
      firstAdmission = false;
      for (int j : add) running[j] = 1;
    };
    auto regrid = [&] {
      std::vector<int> list;
      for (int j = 0; j < ni; j++)
        if (running[j]) list.push_back(j);
      layout(list);
      if (useGraph) capture();
    };
    start(admit());
    regrid();
    long lastAdmission = 0;
    bool freshPending = false;
    long freshChunk = -1;  // the last chunk that began with the fresh phase
    // The flags are read one chunk late: chunk c is launched, then the host
    // waits for chunk c-1's copy of the flags, so the device always has a
    // chunk queued.  Decisions on them (finished instances, admissions, the
    // fresh phase) take effect a chunk later; an instance's result does not
    // depend on when.  After an admission or a new layout (which drain the
    // stream) the flags still in flight are dropped: they predate it.
    struct Events {
      hipEvent_t e[2] = {nullptr, nullptr};
      Events() {
        for (auto& x : e) MINIM_HIP_CHECK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
      }
      ~Events() {
        for (auto& x : e) (void)hipEventDestroy(x);
      }
    } ev;
    void* flagsH[2] = {cntH.p, cntH2.p};
    long chunk = 0, unread = -1;  // the chunk whose flags are in flight, or -1
    for (;;) {
      for (int k = 0; k < kChunk; k++) {
        if (k == 0 && freshPending) {
          iteration(true);
          freshChunk = chunk;
        } else if (useGraph) {
          MINIM_HIP_CHECK(hipGraphLaunch(graph.e, cs));
          graphLaunches++;
        } else {
          iteration(false);
        }
        replays++;
      }
      const int cb = (int)(chunk & 1);
      MINIM_HIP_CHECK(hipMemcpyAsync(flagsH[cb], cntA.get(), cntBytes, hipMemcpyDeviceToHost, cs));
      MINIM_HIP_CHECK(hipEventRecord(ev.e[cb], cs));
      const long readable = unread;
      unread = chunk++;
      if (readable < 0) continue;  // nothing to read yet: queue another chunk first
      const int rb = (int)(readable & 1);
      MINIM_HIP_CHECK(hipEventSynchronize(ev.e[rb]));
      hostSyncs++;
      const int* liveH = reinterpret_cast<const int*>(static_cast<const char*>(flagsH[rb]) + 64);
      long liveBlocks = 0;
      freshPending = false;
      for (int j = 0; j < ni; j++) {
        if (running[j] && !liveH[j]) {  // finished
          running[j] = 0;
          runningSites -= in[j].bi.n;
          finished++;
        }
        if (running[j]) liveBlocks += in[j].nblk;
        // a switch seen in chunk r is served by a fresh phase in any later chunk
        if (running[j] && liveH[j] == 2 && freshChunk <= readable) freshPending = true;
      }
      if (liveBlocks == 0 && next == ni) break;
      if (replayLimit > 0 && replays >= replayLimit) break;
      // every iteration moves each running instance on by at least one
      // line-search pass, and an iteration takes at most 4 of them
      if (replays - lastAdmission > 4L * ((long)maxIter + 2) + 2 * kChunk) break;
      const std::vector<int> add = admit();
      if (!add.empty()) {
        start(add);
        regrid();
        lastAdmission = replays;
        unread = -1;
      } else if (5 * liveBlocks <= 4 * (long)nbCur) {
        regrid();
        unread = -1;
      }
    }
    }  // resident or graph / device
    batchReport.captures = captures;
    std::vector<BCtl> cf(ni);
    unsigned long long clk0 = 0;
    MINIM_HIP_CHECK(hipMemcpyAsync(cf.data(), A.ctl, sizeof(BCtl) * ni, hipMemcpyDeviceToHost, cs));
    MINIM_HIP_CHECK(hipMemcpyAsync(&clk0, A.clock0, sizeof(clk0), hipMemcpyDeviceToHost, cs));
    MINIM_HIP_CHECK(hipStreamSynchronize(cs));
    int dev = 0, rateKHz = 0;
    MINIM_HIP_CHECK(hipGetDevice(&dev));
    MINIM_HIP_CHECK(hipDeviceGetAttribute(&rateKHz, hipDeviceAttributeWallClockRate, dev));
    const double total = since();
    for (int j = 0; j < ni; j++) {
      Inst& I = in[j];
      const BCtl& c = cf[j];
      I.e = c.e;
      I.gg = c.gg;
      I.iters = c.iters;
      I.xcur = c.xcur;
      I.done = c.done != 0;
      I.st.nEnergy = c.nEnergy;
      I.st.nGradient = c.nGradient;
      I.st.iterFP32 = c.iterFP32;
      I.st.iterFP64 = c.iterFP64;
      I.st.converged = c.converged != 0;
      I.st.seconds = (c.done && rateKHz > 0) ? (double)(c.tdone - clk0) / (1e3 * rateKHz) : total;
      I.st.startSeconds = (c.started && rateKHz > 0) ? (double)(c.tstart - clk0) / (1e3 * rateKHz) : 0.0;
      if (I.iters > maxIters) maxIters = I.iters;
    }
    batchReport.iterations = maxIters;
    batchReport.replays = resident ? maxIters : replays;
  } else {

  // ---- iteration 0's gradient (and energy) at the initial coordinates
  {
    Launch L;
    for (int j = 0; j < ni; j++) {
      in[j].dslot = 0;
      BDyn q = dyn_base(in[j]);
      q.gout = 0;
      dynh[L.pos.size()] = q;
      L.pos.push_back(j);
    }
    stage(L);
    if (mixed) {  // the State's coordinates rounded to FP32 (State::setPrecision)
      K.round<<<L.nb, kBlock, 0, st>>>(B);
      MINIM_HIP_CHECK(hipGetLastError());
    }
    launch(K.pf_eg, L);
    finalize(L, kKP, 0);
    fetch((size_t)L.na * kKP);
    for (int p = 0; p < L.na; p++) {
      Inst& I = in[L.pos[p]];
      I.e = resh[p * kKP];
      I.gg = resh[p * kKP + 1];
      I.gcur = 0;
      I.st.nGradient++;
    }
  }

  for (;;) {
    Launch L;
    for (int j = 0; j < ni; j++)
      if (!in[j].done) L.pos.push_back(j);
    if (L.pos.empty()) break;
    if (replayLimit > 0 && maxIters >= replayLimit) break;

    // ---- instances that have just left FP32 coordinates: gradient and energy
    // at the same point in FP64, and its dots with the stored pairs
    {
      Launch F;
      for (int j : L.pos)
        if (in[j].fresh) {
          BDyn q = dyn_base(in[j]);
          q.gout = in[j].gcur;  // overwrite g
          q.np = (int)in[j].pairs.size();
          for (int i = 0; i < q.np; i++) q.slot[i] = in[j].pairs[i].slot;
          dynh[F.pos.size()] = q;
          F.pos.push_back(j);
        }
      if (!F.pos.empty()) {
        stage(F);
        launch(K.fresh, F);
        finalize(F, kKF, 0);
        fetch((size_t)F.na * kKF);
        for (int p = 0; p < F.na; p++) {
          Inst& I = in[F.pos[p]];
          const double* r = resh + (size_t)p * kKF;
          I.e = r[0];
          I.gg = r[1];
          for (int i = 0; i < (int)I.pairs.size(); i++) {
            I.dg[I.pairs[i].slot] = r[kKP + i];
            I.yg[I.pairs[i].slot] = r[kKP + kBS + i];
          }
          I.fresh = false;
          I.st.nGradient++;
        }
      }
    }

    // ---- direction coefficients (two-loop recursion on the Gram data,
    // Lbfgs.cpp:84-127) and the speculative first trial at x + d
    for (int p = 0; p < (int)L.pos.size(); p++) {
      Inst& I = in[L.pos[p]];
      int spare = -1;
      for (int s = 0; s <= _m; s++)
        if (!I.used[s]) {
          spare = s;
          break;
        }
      I.dslot = spare;
      BDyn q = dyn_base(I);
      q.tau = 1.0;
      const int P = (int)I.pairs.size();
      q.np = P;
      for (int i = 0; i < P; i++) q.slot[i] = I.pairs[i].slot;
      if (P == 0) {
        const double gnorm = std::sqrt(I.gg);
        q.cg = (gnorm > 0) ? -(_initStep / gnorm) : -1.0;
      } else {
        std::vector<double> alpha(P), ag_s(P, 0.0), ay(P, 0.0);
        double ag = -1.0;
        auto sdotg = [&](int i) { return I.pairs[i].tau * I.dg[I.pairs[i].slot]; };
        auto sdoty = [&](int i, int k) { return I.pairs[i].tau * I.dy[I.pairs[i].slot][I.pairs[k].slot]; };
        auto ydoty = [&](int i, int k) { return I.yy[I.pairs[i].slot][I.pairs[k].slot]; };
        for (int k = P - 1; k >= 0; k--) {
          double sq = ag * sdotg(k);
          for (int r = 0; r < P; r++) sq += ay[r] * sdoty(k, r);
          alpha[k] = I.pairs[k].rho * sq;
          ay[k] -= alpha[k];
        }
        const int nw = P - 1;
        const double gamma = 1 / (I.pairs[nw].rho * ydoty(nw, nw));
        ag *= gamma;
        for (int r = 0; r < P; r++) ay[r] *= gamma;
        for (int k = 0; k < P; k++) {
          double yr = ag * I.yg[I.pairs[k].slot];
          for (int r = 0; r < P; r++) yr += ag_s[r] * sdoty(r, k) + ay[r] * ydoty(k, r);
          const double beta = I.pairs[k].rho * yr;
          ag_s[k] += alpha[k] - beta;
        }
        q.cg = ag;
        for (int r = 0; r < P; r++) {
          q.cD[r] = ag_s[r] * I.pairs[r].tau;
          q.cY[r] = ay[r];
        }
      }
      dynh[p] = q;
    }
    stage(L);
    launch(K.dir, L);
    finalize(L, kKD, 0);
    const size_t trialOff = (size_t)L.na * kKD;
    if (spec) {
      launch(K.pf_trial, L);
      finalize(L, kNT, trialOff);
    }
    fetch(trialOff + (spec ? (size_t)L.na * kNT : 0));

    // Armijo tests on the energies of up to kNT consecutive trials of the
    // sequence (linesearch.cpp:15-27); true when the search has ended
    auto consume = [&](Inst& I, const double* e) {
      for (int k = 0; k < kNT; k++) {
        I.st.nEnergy++;
        if (I.e - e[k] >= I.t) return true;  // accepted at x + tau d
        I.tau *= 0.5;
        I.t *= 0.5;
        if (++I.trial >= 10) {
          I.lsFailed = true;  // x + tau0 * 2^-10, not evaluated by the reference
          return true;
        }
      }
      return false;
    };

    std::vector<int> search;
    for (int p = 0; p < L.na; p++) {
      Inst& I = in[L.pos[p]];
      const double* r = resh + (size_t)p * kKD;
      const double dd = r[0], dgv = r[1];
      const int P = (int)I.pairs.size();
      for (int i = 0; i < P; i++) I.dres[I.pairs[i].slot] = r[2 + i];  // d.y of each stored pair
      double tau0 = 1.0;
      if (P > 0 && _maxStep != 0) {  // Lbfgs.cpp:121-124
        const double ss = std::sqrt(dd);
        if (ss > _maxStep) tau0 = _maxStep / ss;
      }
      double gs = tau0 * dgv;  // Lbfgs.cpp:53-57
      if (gs > 0) {
        gs = -gs;
        tau0 = -tau0;
      }
      I.tau0 = tau0;
      I.tau = tau0;
      I.trial = 0;
      I.searching = false;
      I.lsFailed = false;
      if (!backtracking) continue;
      I.t = -0.5 * gs;  // linesearch.cpp:11-13
      if (spec) {
        if (tau0 == 1.0) {
          if (consume(I, resh + trialOff + (size_t)p * kNT)) continue;
        } else {
          I.st.nEnergy++;  // the speculative trials at x + d were not on this search's sequence
        }
      }
      I.searching = true;
      search.push_back(L.pos[p]);
    }

    // ---- further trials of the instances still searching (linesearch.cpp:15-27)
    while (!search.empty()) {
      Launch T;
      T.pos = search;
      for (int p = 0; p < (int)search.size(); p++) {
        BDyn q = dyn_base(in[search[p]]);
        q.tau = in[search[p]].tau;
        dynh[p] = q;
      }
      stage(T);
      launch(K.pf_trial, T);
      finalize(T, kNT, 0);
      fetch((size_t)T.na * kNT);
      std::vector<int> next;
      for (int p = 0; p < T.na; p++) {
        Inst& I = in[T.pos[p]];
        if (consume(I, resh + (size_t)p * kNT)) {
          I.searching = false;
          continue;
        }
        next.push_back(T.pos[p]);
      }
      search.swap(next);
    }

    // ---- step, new gradient and energy, Gram update (Lbfgs.cpp:58-80)
    for (int p = 0; p < L.na; p++) {
      Inst& I = in[L.pos[p]];
      BDyn q = dyn_base(I);
      q.tau = I.tau;
      q.np = (int)I.pairs.size();
      for (int i = 0; i < q.np; i++) q.slot[i] = I.pairs[i].slot;
      dynh[p] = q;
    }
    stage(L);
    launch(K.step, L);
    finalize(L, kKS, 0);
    fetch((size_t)L.na * kKS);

    for (int p = 0; p < L.na; p++) {
      Inst& I = in[L.pos[p]];
      I.e = resh[(size_t)p * kKS];
      I.st.nGradient++;
      const double* r = resh + (size_t)p * kKS + 1;
      auto gD = [&](int i) { return r[i]; };  // by position: stored pairs, then the new direction
      auto gY = [&](int i) { return r[kBS + i]; };
      const double gng = r[2 * kBS], gny = r[2 * kBS + 1];
      auto yD = [&](int i) { return r[2 * kBS + 2 + i]; };
      auto yY = [&](int i) { return r[3 * kBS + 2 + i]; };
      const double yyv = r[4 * kBS + 2];
      const int ds = I.dslot;
      const int P = (int)I.pairs.size();
      std::vector<int> oldSlots;
      for (auto& pr : I.pairs) oldSlots.push_back(pr.slot);
      const double sy = I.tau * yD(P);
      if (sy != 0) {
        for (int i = 0; i < P; i++) {
          const int sl = oldSlots[i];
          I.dy[ds][sl] = I.dres[sl];  // d_new.y_old from the direction pass
          I.dy[sl][ds] = yD(i);
          I.yy[ds][sl] = yY(i);
          I.yy[sl][ds] = yY(i);
        }
        I.dy[ds][ds] = yD(P);
        I.yy[ds][ds] = yyv;
        I.dg[ds] = gD(P);
        I.yg[ds] = gny;
        I.used[ds] = 1;
        I.pairs.push_back({ds, I.tau, 1 / sy});
        if ((int)I.pairs.size() > _m) {
          I.used[I.pairs.front().slot] = 0;
          I.pairs.pop_front();
        }
      }
      for (int i = 0; i < P; i++) {
        I.dg[oldSlots[i]] = gD(i);
        I.yg[oldSlots[i]] = gY(i);
      }
      I.gg = gng;
      I.gcur = 1 - I.gcur;
      I.xcur = 1 - I.xcur;
      I.iters++;
      const double rms = std::sqrt(I.gg / I.bi.n);
      if (I.lp) {
        // leave FP32 coordinates once they have done their job (as Lbfgs::iteration)
        const double sw = _switchRms > 0 ? _switchRms : 1e3 * I.conv;
        I.stall = I.lsFailed ? I.stall + 1 : 0;
        I.st.iterFP32++;
        if (rms < sw || I.stall >= 2) {
          I.lp = false;
          I.fresh = true;  // no convergence test until g is recomputed
        }
      } else {
        I.st.iterFP64++;
      }
      // Minimiser::minimise (Minimiser.cpp:27-38) with Lbfgs::checkConvergence (Lbfgs.cpp:130-133)
      if (!I.fresh && rms < I.conv) {
        I.st.converged = true;
        I.done = true;
      } else if (I.iters > maxIter) {
        I.done = true;
      }
      if (I.done) I.st.seconds = since();
      if (I.iters > maxIters) maxIters = I.iters;
    }
  }

  batchReport.iterations = maxIters;
  batchReport.replays = maxIters;
  batchReport.captures = 0;
  }  // host control

  batchReport.hostSyncs = hostSyncs;
  batchReport.graphLaunches = graphLaunches;
  batchReport.control = _batchControl;
  batchReport.block = block;
  if (!writeBack) {
    MINIM_HIP_CHECK(hipStreamSynchronize(st));
    return {};
  }
  for (int j = 0; j < ni; j++) {
    Inst& I = in[j];
    device_copy(I.s->deviceCoords(), static_cast<double*>(I.xcur ? Xb.get() : Xa.get()) + I.bi.off, I.bi.n * 8, st);
    I.s->invalidateHost();
    I.st.energy = I.e;
    I.st.gradRms = std::sqrt(I.gg / I.bi.n);
    out[j] = I.st;
  }
  MINIM_HIP_CHECK(hipStreamSynchronize(st));
  stats = MinimiseStats{};
  iter = maxIters;
  return out;
}

}  // namespace minim
