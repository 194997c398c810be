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
//   step        x += tau d, then energy and gradient at x (+ g.g)
//   Gram        y = g_new - g, with every dot the next recursion needs
//
// A block covers kBE consecutive elements of one instance; it finds its
// instance by binary search over the prefix sums of the block counts of the
// instances in the launch, and writes its partial sums to its own row.  A
// second small kernel sums each instance's rows in a fixed order, so every
// per-instance result is the same whatever other instances share the launch.
// The direction and Gram passes are compiled for each number of stored pairs
// (a block-uniform switch), so all their loads are unconditional.
// The host keeps each instance's (2m+1)^2 Gram data and runs the two-loop
// recursion on coefficients, as the single-State Lbfgs does; one host
// synchronisation per phase serves all instances.
#include <hip/hip_runtime.h>

#include <chrono>
#include <climits>
#include <cmath>
#include <deque>
#include <stdexcept>
#include <vector>

#include "minim/Minimiser.h"
#include "minim/PhaseField.h"
#include "minim/State.h"
#include "runtime.hpp"

namespace minim_amd {
namespace {

constexpr int kBS = 11;            // history slots (m + 1 with m <= 10)
constexpr int kBR = 8;             // elements per thread
constexpr int kBE = kBlock * kBR;  // elements per block
constexpr int kKP = 2;             // stencil partials: energy, g.g
constexpr int kNT = 4;             // line-search trials evaluated per pass: tau, tau/2, tau/4, tau/8
constexpr int kKD = 2 + kBS;       // direction: d.d, d.g, d.Y_s
constexpr int kKG = 4 * kBS + 3;   // Gram (layout in bgram_kernel)

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
};

// Vectors G, D and Y are stored in TV (double, or float for the mixed
// schedule); coordinates always in FP64 (holding FP32 values while an
// instance runs in FP32).
struct BBufs {
  double* X;
  void* G0;
  void* G1;
  void* D;  // m + 1 slots of N elements
  void* Y;
  size_t N;
  const BInst* inst;
  const BDyn* dyn;
  const int* ablk;  // na + 1 prefix sums of the block counts
  int na;
  double* part;  // per-block partials, K per block
};

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
// written to part[blockIdx.x * K + k].
template <int K>
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
    part[(size_t)blockIdx.x * K + k] = s;
  }
}

// res[p * K + k] = sum of instance p's block rows: lane l of a wavefront adds
// rows l, l + 64, ..., then the 64 lane sums are added in a fixed tree, so the
// result depends only on that instance's rows.
__global__ __launch_bounds__(kBlock) void bfinal_kernel(const double* __restrict__ part,
                                                        const int* __restrict__ ablk, int K,
                                                        double* __restrict__ res) {
  const int p = blockIdx.x, lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const int b0 = ablk[p], nb = ablk[p + 1] - b0;
  for (int k = w; k < K; k += kBlock / kWave) {
    double s = 0;
    for (int b = lane; b < nb; b += kWave) s += part[(size_t)(b0 + b) * K + k];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
    if (lane == 0) res[(size_t)p * K + k] = s;
  }
}

// Energy (forward bonds) and gradient of one-fluid sites, the arithmetic of the
// tiled kernel for one fluid (PhaseField.cpp:238-288, :372-378) in T (FP32
// for an instance in the first phase of the mixed schedule), vectors in TV.
// MODE bits: 1 energy, 2 gradient into G[gout] with g.g, 4 at x + tau * d.
template <int MODE, typename T, typename TV>
__device__ __forceinline__ void bpf_body(const BBufs& B, const BDyn& q, const BInst& I, int lb) {
  using C = BC<T>;
  const double* __restrict__ x = B.X + I.off;
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
  double acc[kKP] = {0, 0};
#pragma unroll
  for (int r = 0; r < kBR; r++) {
    const int i = lb * kBE + r * kBlock + (int)threadIdx.x;
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
    }
  }
  block_partials<kKP>(acc, (MODE & 1 ? 1ull : 0ull) | (MODE & 2 ? 2ull : 0ull), B.part);
}

template <int MODE, typename TV>
__global__ __launch_bounds__(kBlock) void bpf_kernel(BBufs B) {
  const int p = find_pos(B.ablk, B.na, blockIdx.x);
  const BDyn& q = B.dyn[p];
  const BInst I = B.inst[q.j];
  const int lb = blockIdx.x - B.ablk[p];
  if (q.lp) bpf_body<MODE, float, TV>(B, q, I, lb);
  else bpf_body<MODE, double, TV>(B, q, I, lb);
}

// Energies at kNT points x + tau_k d along the line-search sequence, tau_k =
// tau 2^-k (the halvings of linesearch.cpp:24-25 are exact): x and d are read
// once for all of them, and the host takes the first point that passes the
// Armijo test, exactly as trying them one after another would.
template <typename T, typename TV>
__device__ __forceinline__ void btrial_body(const BBufs& B, const BDyn& q, const BInst& I, int lb) {
  using C = BC<T>;
  const double* __restrict__ x = B.X + I.off;
  const TV* __restrict__ d = static_cast<const TV*>(B.D) + (size_t)q.dslot * B.N + I.off;
  T tk[kNT];
  tk[0] = (T)q.tau;
#pragma unroll
  for (int k = 1; k < kNT; k++) tk[k] = tk[k - 1] * (T)0.5;
  const int Z = I.Z, Y = I.Y, X = I.X, YZ = Y * Z;
  double acc[kNT];
#pragma unroll
  for (int k = 0; k < kNT; k++) acc[k] = 0;
  for (int r = 0; r < kBR; r++) {
    const int i = lb * kBE + r * kBlock + (int)threadIdx.x;
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
  block_partials<kNT>(acc, (1ull << kNT) - 1, B.part);
}

template <typename TV>
__global__ __launch_bounds__(kBlock) void btrial_kernel(BBufs B) {
  const int p = find_pos(B.ablk, B.na, blockIdx.x);
  const BDyn& q = B.dyn[p];
  const BInst I = B.inst[q.j];
  const int lb = blockIdx.x - B.ablk[p];
  if (q.lp) btrial_body<float, TV>(B, q, I, lb);
  else btrial_body<double, TV>(B, q, I, lb);
}

// x rounded to FP32 (the mixed schedule's first phase keeps FP32 coordinates).
__global__ __launch_bounds__(kBlock) void bround_kernel(BBufs B) {
  const int p = find_pos(B.ablk, B.na, blockIdx.x);
  const BDyn& q = B.dyn[p];
  const BInst I = B.inst[q.j];
  const int lb = blockIdx.x - B.ablk[p];
  double* __restrict__ x = B.X + I.off;
#pragma unroll
  for (int r = 0; r < kBR; r++) {
    const int i = lb * kBE + r * kBlock + (int)threadIdx.x;
    if (i >= I.n) break;
    x[i] = (double)(float)x[i];
  }
}

// d = cg g + sum_i cD[i] D_{slot i} + cY[i] Y_{slot i} (pairs oldest first, as
// the single-State direction pass) into D[dslot]; partials d.d, d.g, d.y_i.
// P is the number of stored pairs, so every load is unconditional.
template <int P, typename TV>
__device__ __forceinline__ void bdir_body(const BBufs& B, const BDyn& q, const BInst& I, int lb) {
  const size_t N = B.N;
  const TV* __restrict__ g = static_cast<const TV*>(q.gcur ? B.G1 : B.G0) + I.off;
  TV* __restrict__ out = static_cast<TV*>(B.D) + (size_t)q.dslot * N + I.off;
  const TV* Dp[P > 0 ? P : 1];
  const TV* Yp[P > 0 ? P : 1];
  double cd[P > 0 ? P : 1], cy[P > 0 ? P : 1];
#pragma unroll
  for (int k = 0; k < P; k++) {
    Dp[k] = static_cast<const TV*>(B.D) + (size_t)q.slot[k] * N + I.off;
    Yp[k] = static_cast<const TV*>(B.Y) + (size_t)q.slot[k] * N + I.off;
    cd[k] = q.cD[k];
    cy[k] = q.cY[k];
  }
  const double cg = q.cg;
  double acc[kKD];
#pragma unroll
  for (int k = 0; k < kKD; k++) acc[k] = 0;
#pragma unroll 2
  for (int r = 0; r < kBR; r++) {
    const int i = lb * kBE + r * kBlock + (int)threadIdx.x;
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
  constexpr unsigned long long on = (1ull << (2 + P)) - 1;
  block_partials<kKD>(acc, on, B.part);
}

template <typename TV>
__global__ __launch_bounds__(kBlock) void bdir_kernel(BBufs B) {
  const int p = find_pos(B.ablk, B.na, blockIdx.x);
  const BDyn& q = B.dyn[p];
  const BInst I = B.inst[q.j];
  const int lb = blockIdx.x - B.ablk[p];
  switch (q.np) {
    case 0: bdir_body<0, TV>(B, q, I, lb); break;
    case 1: bdir_body<1, TV>(B, q, I, lb); break;
    case 2: bdir_body<2, TV>(B, q, I, lb); break;
    case 3: bdir_body<3, TV>(B, q, I, lb); break;
    case 4: bdir_body<4, TV>(B, q, I, lb); break;
    case 5: bdir_body<5, TV>(B, q, I, lb); break;
    case 6: bdir_body<6, TV>(B, q, I, lb); break;
    case 7: bdir_body<7, TV>(B, q, I, lb); break;
    case 8: bdir_body<8, TV>(B, q, I, lb); break;
    case 9: bdir_body<9, TV>(B, q, I, lb); break;
    default: bdir_body<10, TV>(B, q, I, lb); break;
  }
}

// x += tau * d (product rounded first, as the single-State axpy), in FP32 for
// an instance in the first phase of the mixed schedule.
template <typename TV>
__global__ __launch_bounds__(kBlock) void baxpy_kernel(BBufs B) {
  const int p = find_pos(B.ablk, B.na, blockIdx.x);
  const BDyn& q = B.dyn[p];
  const BInst I = B.inst[q.j];
  const int lb = blockIdx.x - B.ablk[p];
  double* __restrict__ x = B.X + I.off;
  const TV* __restrict__ d = static_cast<const TV*>(B.D) + (size_t)q.dslot * B.N + I.off;
  if (q.lp) {
    const float tau = (float)q.tau;
#pragma unroll
    for (int r = 0; r < kBR; r++) {
      const int i = lb * kBE + r * kBlock + (int)threadIdx.x;
      if (i >= I.n) break;
      const float st = tau * (float)d[i];
      x[i] = (double)((float)x[i] + st);
    }
  } else {
    const double tau = q.tau;
#pragma unroll
    for (int r = 0; r < kBR; r++) {
      const int i = lb * kBE + r * kBlock + (int)threadIdx.x;
      if (i >= I.n) break;
      const double st = tau * (double)d[i];
      x[i] = x[i] + st;
    }
  }
}

// y = G[gout] - G[gcur] into Y[dslot]; partials with gn = G[gout], the D list
// being the P stored pairs then the new direction and the Y list the P pairs:
//   [0,kBS) gn.D_i ; [kBS,2kBS) gn.Y_i ; 2kBS gn.gn ; 2kBS+1 gn.y ;
//   [2kBS+2,3kBS+2) y.D_i ; [3kBS+2,4kBS+2) y.Y_i ; 4kBS+2 y.y
template <int P, typename TV>
__device__ __forceinline__ void bgram_body(const BBufs& B, const BDyn& q, const BInst& I, int lb) {
  const size_t N = B.N;
  const TV* __restrict__ gn = static_cast<const TV*>(q.gout ? B.G1 : B.G0) + I.off;
  const TV* __restrict__ go = static_cast<const TV*>(q.gcur ? B.G1 : B.G0) + I.off;
  TV* __restrict__ yo = static_cast<TV*>(B.Y) + (size_t)q.dslot * N + I.off;
  const TV* Dp[P + 1];
  const TV* Yp[P > 0 ? P : 1];
#pragma unroll
  for (int k = 0; k < P; k++) {
    Dp[k] = static_cast<const TV*>(B.D) + (size_t)q.slot[k] * N + I.off;
    Yp[k] = static_cast<const TV*>(B.Y) + (size_t)q.slot[k] * N + I.off;
  }
  Dp[P] = static_cast<const TV*>(B.D) + (size_t)q.dslot * N + I.off;
  double acc[kKG];
#pragma unroll
  for (int k = 0; k < kKG; k++) acc[k] = 0;
#pragma unroll 2
  for (int r = 0; r < kBR; r++) {
    const int i = lb * kBE + r * kBlock + (int)threadIdx.x;
    if (i >= I.n) break;
    const double a = (double)gn[i];
    const double b = (double)go[i];
    double dv[P + 1], wv[P > 0 ? P : 1];
#pragma unroll
    for (int k = 0; k <= P; k++) dv[k] = (double)Dp[k][i];
#pragma unroll
    for (int k = 0; k < P; k++) wv[k] = (double)Yp[k][i];
    const TV ys = (TV)(a - b);  // stored as TV; the dots use the stored value
    yo[i] = ys;
    const double yv = (double)ys;
#pragma unroll
    for (int k = 0; k <= P; k++) {
      acc[k] += a * dv[k];
      acc[2 * kBS + 2 + k] += yv * dv[k];
    }
#pragma unroll
    for (int k = 0; k < P; k++) {
      acc[kBS + k] += a * wv[k];
      acc[3 * kBS + 2 + k] += yv * wv[k];
    }
    acc[2 * kBS] += a * a;
    acc[2 * kBS + 1] += a * yv;
    acc[4 * kBS + 2] += yv * yv;
  }
  constexpr unsigned long long dm = (1ull << (P + 1)) - 1, ym = (1ull << P) - 1;
  constexpr unsigned long long on = dm | (ym << kBS) | (3ull << (2 * kBS)) | (dm << (2 * kBS + 2)) |
                                    (ym << (3 * kBS + 2)) | (1ull << (4 * kBS + 2));
  block_partials<kKG>(acc, on, B.part);
}

template <typename TV>
__global__ __launch_bounds__(kBlock) void bgram_kernel(BBufs B) {
  const int p = find_pos(B.ablk, B.na, blockIdx.x);
  const BDyn& q = B.dyn[p];
  const BInst I = B.inst[q.j];
  const int lb = blockIdx.x - B.ablk[p];
  switch (q.np) {
    case 0: bgram_body<0, TV>(B, q, I, lb); break;
    case 1: bgram_body<1, TV>(B, q, I, lb); break;
    case 2: bgram_body<2, TV>(B, q, I, lb); break;
    case 3: bgram_body<3, TV>(B, q, I, lb); break;
    case 4: bgram_body<4, TV>(B, q, I, lb); break;
    case 5: bgram_body<5, TV>(B, q, I, lb); break;
    case 6: bgram_body<6, TV>(B, q, I, lb); break;
    case 7: bgram_body<7, TV>(B, q, I, lb); break;
    case 8: bgram_body<8, TV>(B, q, I, lb); break;
    case 9: bgram_body<9, TV>(B, q, I, lb); break;
    default: bgram_body<10, TV>(B, q, I, lb); break;
  }
}

// Dots of a recomputed g = G[gcur] with every stored pair (after the switch
// to FP64 coordinates): [0,kBS) g.D_i ; [kBS,2kBS) g.Y_i.
template <typename TV>
__global__ __launch_bounds__(kBlock) void bref_kernel(BBufs B) {
  const int p = find_pos(B.ablk, B.na, blockIdx.x);
  const BDyn& q = B.dyn[p];
  const BInst I = B.inst[q.j];
  const int lb = blockIdx.x - B.ablk[p];
  const size_t N = B.N;
  const TV* __restrict__ g = static_cast<const TV*>(q.gcur ? B.G1 : B.G0) + I.off;
  const TV* __restrict__ D = static_cast<const TV*>(B.D) + I.off;
  const TV* __restrict__ Y = static_cast<const TV*>(B.Y) + I.off;
  double acc[2 * kBS];
#pragma unroll
  for (int k = 0; k < 2 * kBS; k++) acc[k] = 0;
  for (int r = 0; r < kBR; r++) {
    const int i = lb * kBE + r * kBlock + (int)threadIdx.x;
    if (i >= I.n) break;
    const double a = (double)g[i];
#pragma unroll
    for (int k = 0; k < kBS; k++)
      if (k < q.np) {
        acc[k] += a * (double)D[(size_t)q.slot[k] * N + i];
        acc[kBS + k] += a * (double)Y[(size_t)q.slot[k] * N + i];
      }
  }
  const unsigned long long pm = (1ull << q.np) - 1;
  block_partials<2 * kBS>(acc, pm | (pm << kBS), B.part);
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
  KFn pf_eg, pf_trial, dir, axpy, gram, ref;
};
template <typename TV>
Kernels kernels_for() {
  return Kernels{bpf_kernel<3, TV>, btrial_kernel<TV>, bdir_kernel<TV>,
                 baxpy_kernel<TV>,  bgram_kernel<TV>,  bref_kernel<TV>};
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
  using namespace minim_amd;
  const int ni = (int)states.size();
  std::vector<MinimiseStats> out(ni);
  if (ni == 0) return out;
  if (_m > kBS - 1) throw std::invalid_argument("Lbfgs::minimiseBatch: m must be at most 10.");
  if (_precMode == "fp32")
    throw std::invalid_argument("Lbfgs::minimiseBatch runs in FP64 or the mixed schedule, not in FP32 throughout.");
  // mixed: FP32 coordinates and arithmetic until sqrt(g.g/ndof) < switchRms (or
  // the line search stalls twice), then FP64 coordinates; the vectors stay in
  // FP32 throughout, as the single-State schedule does for a PhaseField
  const bool mixed = _precMode == "mixed";
  const size_t vb = mixed ? 4 : 8;
  const Kernels K = mixed ? kernels_for<float>() : kernels_for<double>();

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
  if (totalBlocks > INT_MAX / kKG) throw std::invalid_argument("Lbfgs::minimiseBatch: too many elements.");

  hipStream_t st = static_cast<hipStream_t>(states[0]->stream());
  MINIM_HIP_CHECK(hipDeviceSynchronize());  // every State's coordinates are final
  const auto t0 = std::chrono::steady_clock::now();
  auto since = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };

  const int S = _m + 1;
  DeviceBuffer X(N * 8), G0(N * vb), G1(N * vb), Dm(N * vb * S), Ym(N * vb * S);
  // per-phase staging, one copy per phase: block prefix sums, then the BDyn of
  // each instance in the launch
  const size_t ablkBytes = (sizeof(int) * (ni + 1) + 15) / 16 * 16;
  const size_t stageBytes = ablkBytes + sizeof(BDyn) * ni;
  DeviceBuffer instD(sizeof(BInst) * ni), stageD(stageBytes);
  DeviceBuffer part(sizeof(double) * (size_t)totalBlocks * kKG), res(sizeof(double) * (size_t)ni * (kKP + kKG));
  struct Pinned {
    void* p = nullptr;
    explicit Pinned(size_t b) { MINIM_HIP_CHECK(hipHostMalloc(&p, b)); }
    ~Pinned() { (void)hipHostFree(p); }
  };
  Pinned stageH(stageBytes), resH(sizeof(double) * (size_t)ni * (kKP + kKG));
  int* ablkh = static_cast<int*>(stageH.p);
  BDyn* dynh = reinterpret_cast<BDyn*>(static_cast<char*>(stageH.p) + ablkBytes);
  double* resh = static_cast<double*>(resH.p);

  {
    std::vector<BInst> bi(ni);
    for (int j = 0; j < ni; j++) bi[j] = in[j].bi;
    MINIM_HIP_CHECK(hipMemcpyAsync(instD.get(), bi.data(), sizeof(BInst) * ni, hipMemcpyHostToDevice, st));
    for (int j = 0; j < ni; j++)
      device_copy(static_cast<double*>(X.get()) + in[j].bi.off, in[j].s->deviceCoords(), in[j].bi.n * 8, st);
    MINIM_HIP_CHECK(hipStreamSynchronize(st));  // bi leaves scope
  }

  BBufs B{};
  B.X = static_cast<double*>(X.get());
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
    bfinal_kernel<<<L.na, kBlock, 0, st>>>(B.part, B.ablk, K, resd + resOff);
    MINIM_HIP_CHECK(hipGetLastError());
  };
  auto fetch = [&](size_t count) {
    MINIM_HIP_CHECK(hipMemcpyAsync(resh, resd, sizeof(double) * count, hipMemcpyDeviceToHost, st));
    MINIM_HIP_CHECK(hipStreamSynchronize(st));
  };
  auto dyn_base = [&](const Inst& I) {
    BDyn q{};
    q.j = (int)(&I - in.data());
    q.dslot = I.dslot;
    q.gcur = I.gcur;
    q.gout = 1 - I.gcur;
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
      bround_kernel<<<L.nb, kBlock, 0, st>>>(B);
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

  int maxIters = 0;
  for (;;) {
    Launch L;
    for (int j = 0; j < ni; j++)
      if (!in[j].done) L.pos.push_back(j);
    if (L.pos.empty()) break;

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
        launch(K.pf_eg, F);
        finalize(F, kKP, 0);
        launch(K.ref, F);
        finalize(F, 2 * kBS, (size_t)F.na * kKP);
        fetch((size_t)F.na * (kKP + 2 * kBS));
        for (int p = 0; p < F.na; p++) {
          Inst& I = in[F.pos[p]];
          I.e = resh[(size_t)p * kKP];
          I.gg = resh[(size_t)p * kKP + 1];
          const double* r = resh + (size_t)F.na * kKP + (size_t)p * 2 * kBS;
          for (int i = 0; i < (int)I.pairs.size(); i++) {
            I.dg[I.pairs[i].slot] = r[i];
            I.yg[I.pairs[i].slot] = r[kBS + i];
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
    launch(K.axpy, L);
    launch(K.pf_eg, L);
    finalize(L, kKP, 0);
    const size_t gramOff = (size_t)L.na * kKP;
    launch(K.gram, L);
    finalize(L, kKG, gramOff);
    fetch(gramOff + (size_t)L.na * kKG);

    for (int p = 0; p < L.na; p++) {
      Inst& I = in[L.pos[p]];
      I.e = resh[(size_t)p * kKP];
      I.st.nGradient++;
      const double* r = resh + gramOff + (size_t)p * kKG;
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

  for (int j = 0; j < ni; j++) {
    Inst& I = in[j];
    device_copy(I.s->deviceCoords(), static_cast<double*>(X.get()) + I.bi.off, I.bi.n * 8, st);
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
