// lj.hip -- Lennard-Jones potential on MI355X.
//
// Reference semantics (LjNd::init / elementEnergyGradient,
// src/potentials/LjNd.cpp:12-53): one element per pair i<j, E = 4 eps (lj12 -
// lj6), g_i += dx * (-24 eps (2 lj12 - lj6) / r^2), g_j -= the same.  The
// reference stores N(N-1)/2 heap elements and scatters into g; here one thread
// owns particle i and gathers over all j through LDS tiles of positions, so the
// gradient is written once per particle with no atomics and the energy is the
// FP64 sum of per-particle halves.
//
// Cutoff mode (no reference counterpart): truncated-and-shifted at rc in a
// periodic box, with a full Verlet list of radius rc+skin built from a cell
// list; the list is rebuilt when any particle has moved more than skin/2 from
// where the list was built.  Lists are built deterministically (cell contents
// sorted by particle id), so energies and gradients are reproducible.
#include <climits>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "common.hpp"
#include "minim/LjNd.h"
#include "runtime.hpp"
#include "vec.hpp"

namespace minim_amd {
namespace {

template <typename T, int D>
__device__ inline void load_pos(const T* __restrict__ x, const T* __restrict__ d, T t, int i, T (&p)[D]) {
#pragma unroll
  for (int k = 0; k < D; k++) {
    T c = x[(size_t)i * D + k];
    if (d) {
      T s = t * d[(size_t)i * D + k];
      c = c + s;
    }
    p[k] = c;
  }
}

// All pairs, gather form.  mode bit 1 energy, bit 2 gradient.
template <typename T, int D>
__global__ __launch_bounds__(kBlock) void lj_allpairs_kernel(const T* __restrict__ x, const T* __restrict__ d, T t,
                                                             int N, T sig6, T eps, int mode, T* __restrict__ g,
                                                             RedScratch rs) {
  __shared__ T tile[kBlock * D];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  T xi[D];
  T gi[D];
#pragma unroll
  for (int k = 0; k < D; k++) gi[k] = (T)0;
  double ei = 0.0;
  if (i < N) load_pos<T, D>(x, d, t, i, xi);
  for (int base = 0; base < N; base += kBlock) {
    const int j = base + threadIdx.x;
    if (j < N) {
      T pj[D];
      load_pos<T, D>(x, d, t, j, pj);
#pragma unroll
      for (int k = 0; k < D; k++) tile[threadIdx.x * D + k] = pj[k];
    }
    __syncthreads();
    if (i < N) {
      const int cnt = min(kBlock, N - base);
      for (int jj = 0; jj < cnt; jj++) {
        if (base + jj == i) continue;
        T dx[D], r2 = (T)0;
#pragma unroll
        for (int k = 0; k < D; k++) {
          dx[k] = xi[k] - tile[jj * D + k];
          r2 += dx[k] * dx[k];
        }
        const T r6 = r2 * r2 * r2;
        const T lj6 = sig6 / r6;
        const T lj12 = lj6 * lj6;
        if (mode & 1) ei += 0.5 * (double)((T)4 * eps * (lj12 - lj6));
        if (mode & 2) {
          const T f = (T)-24 * eps * ((T)2 * lj12 - lj6) / r2;
#pragma unroll
          for (int k = 0; k < D; k++) gi[k] += dx[k] * f;
        }
      }
    }
    __syncthreads();
  }
  if ((mode & 2) && i < N) {
#pragma unroll
    for (int k = 0; k < D; k++) g[(size_t)i * D + k] = gi[k];
  }
  double acc[1] = {ei};
  grid_sum_store<1>(acc, rs);
}

// ---------------------------------------------------------------- cutoff mode
struct BoxArgs {
  double L[3];
  double invL[3];
  double inv_cell[3];
  int nc[3];
  double rlist2;
  double rc2;
  double half_skin2;
};

// Minimum image v - L rint(v / L), with v / L taken as v * (1/L): the two
// can round to different integers only when |v| is within an ulp of L/2,
// where both images are far beyond any cutoff.
template <typename T>
__device__ inline T min_image(T v, T L, T invL) {
  return v - L * rint(v * invL);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void lj_cell_index_kernel(const T* __restrict__ x, int N, BoxArgs b,
                                                               int* __restrict__ cell, int* __restrict__ count,
                                                               double* __restrict__ xbuild) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  int c[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double v = (double)x[(size_t)i * 3 + k];
    xbuild[(size_t)i * 3 + k] = v;
    double w = v - b.L[k] * floor(v / b.L[k]);  // wrapped into [0, L)
    int ci = (int)(w * b.inv_cell[k]);
    if (ci >= b.nc[k]) ci = b.nc[k] - 1;
    if (ci < 0) ci = 0;
    c[k] = ci;
  }
  const int id = (c[0] * b.nc[1] + c[1]) * b.nc[2] + c[2];
  cell[i] = id;
  atomicAdd(&count[id], 1);
}

// Exclusive scan of cell counts by a single block (cell counts are small).
__global__ __launch_bounds__(kBlock) void lj_scan_kernel(const int* __restrict__ count, int* __restrict__ start,
                                                         int* __restrict__ cursor, int ncell) {
  __shared__ int part[kBlock];
  const int per = (ncell + kBlock - 1) / kBlock;
  const int lo = threadIdx.x * per, hi = min(lo + per, ncell);
  int s = 0;
  for (int c = lo; c < hi; c++) s += count[c];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int k = 0; k < kBlock; k++) {
      int v = part[k];
      part[k] = run;
      run += v;
    }
  }
  __syncthreads();
  int run = part[threadIdx.x];
  for (int c = lo; c < hi; c++) {
    start[c] = run;
    cursor[c] = 0;
    run += count[c];
  }
  if (threadIdx.x == kBlock - 1) start[ncell] = run;
}

__global__ __launch_bounds__(kBlock) void lj_fill_kernel(const int* __restrict__ cell, int N,
                                                         const int* __restrict__ start, int* __restrict__ cursor,
                                                         int* __restrict__ order) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  const int c = cell[i];
  const int pos = start[c] + atomicAdd(&cursor[c], 1);
  order[pos] = i;
}

// Sort each cell's members by particle id so the neighbour order is
// reproducible: one wavefront per cell; with up to 64 members each lane
// holds one and computes its rank by comparing with the others (ids are
// distinct, so the ranks are a permutation); larger cells are sorted by one
// lane with an insertion sort.
__global__ __launch_bounds__(kBlock) void lj_sort_cells_kernel(const int* __restrict__ start, int* __restrict__ order,
                                                               int ncell) {
  const int lane = threadIdx.x % kWave;
  const int c = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (c >= ncell) return;  // uniform over the wavefront
  const int lo = start[c], hi = start[c + 1], k = hi - lo;
  if (k <= kWave) {
    const int v = lane < k ? order[lo + lane] : INT_MAX;
    int rank = 0;
    for (int q = 0; q < k; q++) rank += __shfl(v, q, kWave) < v ? 1 : 0;
    if (lane < k) order[lo + rank] = v;
  } else if (lane == 0) {
    for (int a = lo + 1; a < hi; a++) {
      int v = order[a];
      int b = a - 1;
      while (b >= lo && order[b] > v) {
        order[b + 1] = order[b];
        b--;
      }
      order[b + 1] = v;
    }
  }
}

// Verlet list layout: for each group of 64 consecutive particles (one
// wavefront), maxn rows of 64 entries, so that entry k of particle i sits at
// ((i / 64) maxn + k) 64 + i % 64.  A wavefront's lanes read row k of their
// group together (one coalesced 256-byte access in the force pass), and a
// wavefront building its list writes only within its group's maxn x 64
// region, which stays cached until it is complete (with rows of all N
// particles, each lane's k-th entry fell in a different region of a list
// larger than L2).
__host__ __device__ inline size_t nbr_at(int i, int k, int maxn) {
  return ((size_t)(i >> 6) * maxn + k) * 64 + (i & 63);
}
inline size_t nbr_elems(int N, int maxn) { return (size_t)((N + 63) / 64) * 64 * maxn; }

// A coordinate wrapped into [0, L), as the cell index takes it.
__device__ __forceinline__ double wrap_box(double v, double L) { return v - L * floor(v / L); }

// Positions in cell order, wrapped into the box, each with its particle id:
// xc[p] = (x, y, z, j) for j = order[p] (FP64, the list is built in FP64 in
// either precision), so that the list build reads one contiguous 32-byte
// entry per candidate instead of an index and then three scattered
// coordinates.
template <typename T>
__global__ __launch_bounds__(kBlock) void lj_gather_cells_kernel(const T* __restrict__ x, int N, BoxArgs b,
                                                                 const int* __restrict__ order,
                                                                 double4* __restrict__ xc) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= N) return;
  const int j = order[p];
  xc[p] = make_double4(wrap_box((double)x[(size_t)j * 3], b.L[0]), wrap_box((double)x[(size_t)j * 3 + 1], b.L[1]),
                       wrap_box((double)x[(size_t)j * 3 + 2], b.L[2]), (double)j);
}

// Full Verlet list: neighbour k of particle i at nbr[nbr_at(i, k, maxn)].  The 27
// cells around i's cell are visited in a fixed order and each cell's members
// in particle-id order (lj_sort_cells_kernel), so the list -- and with it the
// order in which the force pass sums pairs -- does not depend on scheduling.
// Candidates are read U at a time, all loads issued before any is used.
// With at least 3 cells of side >= rc + skin along each axis, the image of a
// neighbour cell that lies within the list radius is the one its index wrap
// implies, so each candidate needs no minimum-image rounding: i's wrapped
// position is shifted once per neighbour cell.  Pairs are kept below the
// list radius widened by 1e-12 of itself, so no pair the minimum-image test
// keeps is lost to rounding; a pair kept in excess lies beyond the cutoff
// (or at it, for skin 0) and the force pass, which tests the cutoff itself,
// skips it.  Threads take the particles in cell order (thread p builds the
// list of particle xc[p].w), so the lanes of a wavefront lie in a few cells
// and read the same candidates instead of gathering from two dozen cells.
__global__ __launch_bounds__(kBlock) void lj_build_list_kernel(int N, BoxArgs b, const int* __restrict__ cell,
                                                               const int* __restrict__ start,
                                                               const double4* __restrict__ xc, int maxn,
                                                               int* __restrict__ nn, int* __restrict__ nbr,
                                                               int* __restrict__ overflow) {
  constexpr int U = 4;
  const int pi = blockIdx.x * kBlock + threadIdx.x;
  if (pi >= N) return;
  const double4 vi = xc[pi];
  const int i = (int)vi.w;
  const double xw0 = vi.x, xw1 = vi.y, xw2 = vi.z;
  const double rl2 = b.rlist2 * (1.0 + 1e-12);
  const int c = cell[i];
  const int cz = c % b.nc[2], cy = (c / b.nc[2]) % b.nc[1], cx = c / (b.nc[1] * b.nc[2]);
  int cnt = 0;
  for (int dx = -1; dx <= 1; dx++) {
    const int ux = cx + dx;
    const int ax = ux < 0 ? ux + b.nc[0] : (ux >= b.nc[0] ? ux - b.nc[0] : ux);
    // the neighbour cell's image sits at its wrapped position -/+ L
    const double xi0 = ux < 0 ? xw0 + b.L[0] : (ux >= b.nc[0] ? xw0 - b.L[0] : xw0);
    for (int dy = -1; dy <= 1; dy++) {
      const int uy = cy + dy;
      const int ay = uy < 0 ? uy + b.nc[1] : (uy >= b.nc[1] ? uy - b.nc[1] : uy);
      const double xi1 = uy < 0 ? xw1 + b.L[1] : (uy >= b.nc[1] ? xw1 - b.L[1] : xw1);
      for (int dz = -1; dz <= 1; dz++) {
        const int uz = cz + dz;
        const int az = uz < 0 ? uz + b.nc[2] : (uz >= b.nc[2] ? uz - b.nc[2] : uz);
        const double xi2 = uz < 0 ? xw2 + b.L[2] : (uz >= b.nc[2] ? xw2 - b.L[2] : xw2);
        const int cc = (ax * b.nc[1] + ay) * b.nc[2] + az;
        const int p1 = start[cc + 1];
        for (int p0 = start[cc]; p0 < p1; p0 += U) {
          double4 v[U];
#pragma unroll
          for (int u = 0; u < U; u++) v[u] = xc[min(p0 + u, p1 - 1)];
#pragma unroll
          for (int u = 0; u < U; u++) {
            if (p0 + u >= p1) break;
            const int j = (int)v[u].w;
            if (j == i) continue;
            const double r0 = xi0 - v[u].x, r1 = xi1 - v[u].y, r2 = xi2 - v[u].z;
            if (r0 * r0 + r1 * r1 + r2 * r2 < rl2) {
              if (cnt < maxn) nbr[nbr_at(i, cnt, maxn)] = j;
              cnt++;
            }
          }
        }
      }
    }
  }
  nn[i] = cnt < maxn ? cnt : maxn;
  if (cnt > maxn) atomicMax(overflow, cnt);
}

// Small boxes (fewer than 3 cells along a side): list by brute force.
template <typename T>
__global__ __launch_bounds__(kBlock) void lj_build_list_brute_kernel(const T* __restrict__ x, int N, BoxArgs b,
                                                                     int maxn, int* __restrict__ nn,
                                                                     int* __restrict__ nbr,
                                                                     int* __restrict__ overflow,
                                                                     double* __restrict__ xbuild) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  double xi[3];
  for (int k = 0; k < 3; k++) {
    xi[k] = (double)x[(size_t)i * 3 + k];
    xbuild[(size_t)i * 3 + k] = xi[k];
  }
  int cnt = 0;
  for (int j = 0; j < N; j++) {
    if (j == i) continue;
    double r2 = 0;
    for (int k = 0; k < 3; k++) {
      const double v = min_image(xi[k] - (double)x[(size_t)j * 3 + k], b.L[k], b.invL[k]);
      r2 += v * v;
    }
    if (r2 < b.rlist2) {
      if (cnt < maxn) nbr[nbr_at(i, cnt, maxn)] = j;
      cnt++;
    }
  }
  nn[i] = cnt < maxn ? cnt : maxn;
  if (cnt > maxn) atomicMax(overflow, cnt);
}

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double4;
};

// Positions at the evaluated point (x, or x + t*d) padded to four components,
// so that the force pass reads a neighbour with one vector access.
template <typename T>
__global__ __launch_bounds__(kBlock) void lj_pad_kernel(const T* __restrict__ x, const T* __restrict__ d, T t, int N,
                                                        typename Vec4<T>::type* __restrict__ xp) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  T p[3];
  load_pos<T, 3>(x, d, t, i, p);
  typename Vec4<T>::type v;
  v.x = p[0];
  v.y = p[1];
  v.z = p[2];
  v.w = (T)0;
  xp[i] = v;
}

// One thread per particle over its Verlet list, in chunks of U neighbours whose
// indices and positions are all requested before any is used (the list rows
// of a chunk are clamped to the last entry, and the repeats are skipped), so
// the gathers overlap instead of forming one dependent chain per neighbour.
// Positions come padded (lj_pad_kernel): one vector load per neighbour.
// Pairs are summed in list order.  The same pass counts the particles that
// have moved more than skin/2 since the list was built (reduction slot 1):
// when any has, the caller rebuilds the list and evaluates again.  Blocks are
// spread over the XCDs in contiguous ranges, so each L2 holds the positions
// of one spatially coherent range of particles.  (Two to eight lanes per
// particle, each taking every P-th entry, measured no faster at 262,144
// particles.)
template <typename T>
__global__ __launch_bounds__(kBlock) void lj_cut_kernel(const typename Vec4<T>::type* __restrict__ xp, int N,
                                                        BoxArgs b, T sig2, T eps, T eshift, int mode,
                                                        const int* __restrict__ nn, const int* __restrict__ nbr,
                                                        int maxn, const double* __restrict__ xbuild,
                                                        T* __restrict__ g, RedScratch rs) {
  constexpr int U = 8;
  using V = typename Vec4<T>::type;
  const int i = xcd_logical_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  double ei = 0.0, moved = 0.0;
  if (i < N) {
    const V vi = xp[i];
    const T xi[3] = {vi.x, vi.y, vi.z};
    double dr2 = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double v = (double)xi[k] - xbuild[(size_t)i * 3 + k];
      dr2 += v * v;
    }
    if (dr2 > b.half_skin2) moved = 1.0;
    const T L0 = (T)b.L[0], L1 = (T)b.L[1], L2 = (T)b.L[2];
    const T rc2 = (T)b.rc2;
    const T iL0 = (T)b.invL[0], iL1 = (T)b.invL[1], iL2 = (T)b.invL[2];
    T gi[3] = {(T)0, (T)0, (T)0};
    const int cnt = nn[i];
    for (int k0 = 0; k0 < cnt; k0 += U) {
      int jj[U];
#pragma unroll
      for (int u = 0; u < U; u++) jj[u] = nbr[nbr_at(i, min(k0 + u, cnt - 1), maxn)];
      V xj[U];
#pragma unroll
      for (int u = 0; u < U; u++) xj[u] = xp[jj[u]];
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (k0 + u >= cnt) break;
        const T r0 = min_image(xi[0] - xj[u].x, L0, iL0);
        const T r1 = min_image(xi[1] - xj[u].y, L1, iL1);
        const T r2v = min_image(xi[2] - xj[u].z, L2, iL2);
        const T r2 = r0 * r0 + r1 * r1 + r2v * r2v;
        if (r2 >= rc2) continue;
        const T ir2 = (T)1 / r2;  // one division per pair
        const T s2 = sig2 * ir2;
        const T lj6 = s2 * s2 * s2;
        const T lj12 = lj6 * lj6;
        if (mode & 1) ei += 0.5 * ((double)((T)4 * eps * (lj12 - lj6)) - (double)eshift);
        if (mode & 2) {
          const T f = (T)-24 * eps * ((T)2 * lj12 - lj6) * ir2;
          gi[0] += r0 * f;
          gi[1] += r1 * f;
          gi[2] += r2v * f;
        }
      }
    }
    if (mode & 2) {
      g[(size_t)i * 3] = gi[0];
      g[(size_t)i * 3 + 1] = gi[1];
      g[(size_t)i * 3 + 2] = gi[2];
    }
  }
  double acc[2] = {ei, moved};
  grid_sum_store<2>(acc, rs);
}

template <typename T>
void launch_allpairs(int D, const void* x, const void* d, double t, int N, double sigma, double eps, int mode, void* g,
                     RedScratch rs, hipStream_t s) {
  const int nb = (N + kBlock - 1) / kBlock;
  if (nb > kMaxRedBlocks) throw std::invalid_argument("LjNd: too many particles for the all-pairs kernel");
  const T* xx = (const T*)x;
  const T* dd = (const T*)d;
  T* gg = (T*)g;
  const T sig6 = (T)std::pow(sigma, 6);
  switch (D) {
    case 1: lj_allpairs_kernel<T, 1><<<nb, kBlock, 0, s>>>(xx, dd, (T)t, N, sig6, (T)eps, mode, gg, rs); break;
    case 2: lj_allpairs_kernel<T, 2><<<nb, kBlock, 0, s>>>(xx, dd, (T)t, N, sig6, (T)eps, mode, gg, rs); break;
    case 3: lj_allpairs_kernel<T, 3><<<nb, kBlock, 0, s>>>(xx, dd, (T)t, N, sig6, (T)eps, mode, gg, rs); break;
    case 4: lj_allpairs_kernel<T, 4><<<nb, kBlock, 0, s>>>(xx, dd, (T)t, N, sig6, (T)eps, mode, gg, rs); break;
    default: throw std::invalid_argument("LjNd: nDim must be 1..4 on the device path");
  }
  MINIM_HIP_CHECK(hipGetLastError());
  launch_finalize(rs, 1, nb, s);
}

}  // namespace
}  // namespace minim_amd

namespace minim {

using minim_amd::DeviceBuffer;
using minim_amd::kBlock;

struct LjNd::Impl {
  minim_amd::BoxArgs box{};
  bool cellMode = false;
  int maxn = 128;
  long builds = 0;
  bool built = false;
  DeviceBuffer cell, count, start, cursor, order, nn, nbr, flag, xbuild, xpad, xcell;
  int ncell = 0;
};

LjNd::LjNd(const LjNd& o)
    : NewPotential<LjNd>(o), nDim(o.nDim), nParticle(o.nParticle), sigma(o.sigma), epsilon(o.epsilon),
      cutoff(o.cutoff), box(o.box), skin(o.skin) {}

LjNd& LjNd::operator=(const LjNd& o) {
  NewPotential<LjNd>::operator=(o);
  nDim = o.nDim;
  nParticle = o.nParticle;
  sigma = o.sigma;
  epsilon = o.epsilon;
  cutoff = o.cutoff;
  box = o.box;
  skin = o.skin;
  impl_.reset();
  return *this;
}

LjNd::~LjNd() = default;

LjNd& LjNd::setSigma(double s) {
  sigma = s;
  return *this;
}
LjNd& LjNd::setEpsilon(double e) {
  epsilon = e;
  return *this;
}
LjNd& LjNd::setCutoff(double rc, std::array<double, 3> b, double sk) {
  if (rc <= 0 || sk < 0) throw std::invalid_argument("LjNd: cutoff must be positive and skin non-negative.");
  for (double v : b)
    if (v <= 0) throw std::invalid_argument("LjNd: box lengths must be positive.");
  cutoff = rc;
  box = b;
  skin = sk;
  return *this;
}

long LjNd::listBuilds() const { return impl_ ? impl_->builds : 0; }
// The largest list length of the last build, read back from the device (a
// statistic: it waits for the device).
int LjNd::maxNeighbours() const {
  if (!impl_ || !impl_->built) return 0;
  std::vector<int> h(nParticle);
  MINIM_HIP_CHECK(hipMemcpy(h.data(), impl_->nn.get(), sizeof(int) * h.size(), hipMemcpyDeviceToHost));
  int mx = 0;
  for (int v : h) mx = v > mx ? v : mx;
  return mx;
}

// LjNd::init (src/potentials/LjNd.cpp:12-26): nParticle from the coordinate count.
void LjNd::init(const vector<double>& coords) {
  if (nDim < 1) throw std::invalid_argument("LjNd: nDim must be positive.");
  if (coords.size() % nDim != 0)
    throw std::invalid_argument("LjNd: Length of coords must be a multiple of the number of dimensions.");
  nParticle = (int)(coords.size() / nDim);
  uploadFixed(coords.size());
  impl_ = std::make_shared<Impl>();
  if (cutoff > 0) {
    if (nDim != 3) throw std::invalid_argument("LjNd: the cutoff/periodic mode needs nDim = 3.");
    auto& b = impl_->box;
    const double rl = cutoff + skin;
    for (int k = 0; k < 3; k++) {
      if (box[k] < 2 * cutoff) throw std::invalid_argument("LjNd: box must be at least twice the cutoff.");
      b.L[k] = box[k];
      b.nc[k] = (int)std::floor(box[k] / rl);
      if (b.nc[k] < 1) b.nc[k] = 1;
      b.inv_cell[k] = b.nc[k] / box[k];
    }
    b.rlist2 = rl * rl;
    for (int k = 0; k < 3; k++) b.invL[k] = 1.0 / box[k];
    b.rc2 = cutoff * cutoff;
    b.half_skin2 = 0.25 * skin * skin;
    impl_->cellMode = b.nc[0] >= 3 && b.nc[1] >= 3 && b.nc[2] >= 3;
    impl_->ncell = b.nc[0] * b.nc[1] * b.nc[2];
    const size_t N = nParticle;
    impl_->cell.reset(sizeof(int) * N);
    impl_->count.reset(sizeof(int) * impl_->ncell);
    impl_->start.reset(sizeof(int) * (impl_->ncell + 1));
    impl_->cursor.reset(sizeof(int) * impl_->ncell);
    impl_->order.reset(sizeof(int) * N);
    impl_->nn.reset(sizeof(int) * N);
    impl_->nbr.reset(sizeof(int) * minim_amd::nbr_elems(N, impl_->maxn));
    impl_->flag.reset(sizeof(int));
    impl_->xbuild.reset(sizeof(double) * 3 * N);
    impl_->xpad.reset(sizeof(double) * 4 * N);
  }
}

namespace {
int cut_blocks(int N) { return (N + kBlock - 1) / kBlock; }

void launch_cut(const LjNd::Impl& I, int prec, const void* x, const void* d, double t, int N, double sigma,
                double epsilon, double eshift, int mode, void* g, minim_amd::RedScratch rs, hipStream_t s) {
  const int nb = cut_blocks(N);
  if (prec == FP64) {
    auto* xp = (double4*)I.xpad.get();
    minim_amd::lj_pad_kernel<double><<<nb, kBlock, 0, s>>>((const double*)x, (const double*)d, t, N, xp);
    minim_amd::lj_cut_kernel<double><<<nb, kBlock, 0, s>>>(xp, N, I.box, sigma * sigma, epsilon, eshift, mode,
                                                           (const int*)I.nn.get(), (const int*)I.nbr.get(), I.maxn,
                                                           (const double*)I.xbuild.get(), (double*)g, rs);
  } else {
    auto* xp = (float4*)I.xpad.get();
    minim_amd::lj_pad_kernel<float><<<nb, kBlock, 0, s>>>((const float*)x, (const float*)d, (float)t, N, xp);
    minim_amd::lj_cut_kernel<float><<<nb, kBlock, 0, s>>>(xp, N, I.box, (float)(sigma * sigma), (float)epsilon,
                                                          (float)eshift, mode, (const int*)I.nn.get(),
                                                          (const int*)I.nbr.get(), I.maxn,
                                                          (const double*)I.xbuild.get(), (float*)g, rs);
  }
  MINIM_HIP_CHECK(hipGetLastError());
}

// (Re)builds the Verlet list at the evaluated positions x (+ t*d).
void rebuild_list(LjNd::Impl& I, int prec, const void* x, const void* d, double t, int N, hipStream_t s) {
  const int nb = (N + kBlock - 1) / kBlock;
  // Build at the evaluated positions: materialise x + t*d when needed.
  DeviceBuffer tmp;
  const void* xb = x;
  if (d) {
    tmp.reset(3 * (size_t)N * minim_amd::prec_bytes(prec));
    minim_amd::device_copy(tmp.get(), x, 3 * (size_t)N * minim_amd::prec_bytes(prec), s);
    minim_amd::launch_axpy(prec, tmp.get(), d, t, 3 * (size_t)N, s);
    xb = tmp.get();
  }
  for (int attempt = 0; attempt < 2; attempt++) {
    MINIM_HIP_CHECK(hipMemsetAsync(I.flag.get(), 0, sizeof(int), s));
    if (I.cellMode) {
      MINIM_HIP_CHECK(hipMemsetAsync(I.count.get(), 0, sizeof(int) * I.ncell, s));
      constexpr int kCellsPerBlock = kBlock / minim_amd::kWave;  // one wavefront per cell
      const int nbc = (I.ncell + kCellsPerBlock - 1) / kCellsPerBlock;
      if (prec == FP64)
        minim_amd::lj_cell_index_kernel<double><<<nb, kBlock, 0, s>>>((const double*)xb, N, I.box, (int*)I.cell.get(),
                                                                      (int*)I.count.get(), (double*)I.xbuild.get());
      else
        minim_amd::lj_cell_index_kernel<float><<<nb, kBlock, 0, s>>>((const float*)xb, N, I.box, (int*)I.cell.get(),
                                                                     (int*)I.count.get(), (double*)I.xbuild.get());
      minim_amd::lj_scan_kernel<<<1, kBlock, 0, s>>>((const int*)I.count.get(), (int*)I.start.get(),
                                                     (int*)I.cursor.get(), I.ncell);
      minim_amd::lj_fill_kernel<<<nb, kBlock, 0, s>>>((const int*)I.cell.get(), N, (const int*)I.start.get(),
                                                      (int*)I.cursor.get(), (int*)I.order.get());
      minim_amd::lj_sort_cells_kernel<<<nbc, kBlock, 0, s>>>((const int*)I.start.get(), (int*)I.order.get(), I.ncell);
      if (I.xcell.bytes() < sizeof(double4) * (size_t)N) I.xcell.reset(sizeof(double4) * (size_t)N);
      double4* xc = static_cast<double4*>(I.xcell.get());
      if (prec == FP64) {
        minim_amd::lj_gather_cells_kernel<double><<<nb, kBlock, 0, s>>>((const double*)xb, N, I.box,
                                                                         (const int*)I.order.get(), xc);
      } else {
        minim_amd::lj_gather_cells_kernel<float><<<nb, kBlock, 0, s>>>((const float*)xb, N, I.box,
                                                                        (const int*)I.order.get(), xc);
      }
      minim_amd::lj_build_list_kernel<<<nb, kBlock, 0, s>>>(N, I.box, (const int*)I.cell.get(),
                                                             (const int*)I.start.get(), xc, I.maxn, (int*)I.nn.get(),
                                                             (int*)I.nbr.get(), (int*)I.flag.get());
    } else {
      if (prec == FP64)
        minim_amd::lj_build_list_brute_kernel<double><<<nb, kBlock, 0, s>>>(
            (const double*)xb, N, I.box, I.maxn, (int*)I.nn.get(), (int*)I.nbr.get(), (int*)I.flag.get(),
            (double*)I.xbuild.get());
      else
        minim_amd::lj_build_list_brute_kernel<float><<<nb, kBlock, 0, s>>>(
            (const float*)xb, N, I.box, I.maxn, (int*)I.nn.get(), (int*)I.nbr.get(), (int*)I.flag.get(),
            (double*)I.xbuild.get());
    }
    MINIM_HIP_CHECK(hipGetLastError());
    int over = 0;
    MINIM_HIP_CHECK(hipMemcpyAsync(&over, I.flag.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    MINIM_HIP_CHECK(hipStreamSynchronize(s));
    if (over == 0) break;
    if (attempt == 1) throw std::runtime_error("LjNd: neighbour list overflow");
    I.maxn = over + 32;
    I.nbr.reset(sizeof(int) * minim_amd::nbr_elems(N, I.maxn));
  }
  I.built = true;
  I.builds++;
}
}  // namespace

void LjNd::evaluate(int prec, const void* x, const void* d, double t, double* e, void* g, void* stream) const {
  if (!impl_) throw std::runtime_error("LjNd: not initialised (construct a State first).");
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto& ws = minim_amd::Workspace::current();
  const int mode = (e ? 1 : 0) | (g ? 2 : 0);
  if (!mode) return;
  const int N = nParticle;
  if (cutoff <= 0) {
    if (prec == FP64) minim_amd::launch_allpairs<double>(nDim, x, d, t, N, sigma, epsilon, mode, g, ws.rs(), s);
    else minim_amd::launch_allpairs<float>(nDim, x, d, t, N, sigma, epsilon, mode, g, ws.rs(), s);
  } else {
    Impl& I = *impl_;
    if ((N + kBlock - 1) / kBlock > minim_amd::kMaxRedBlocks)
      throw std::invalid_argument("LjNd: too many particles for one launch");
    const double src6 = std::pow(sigma / cutoff, 6);
    const double eshift = 4 * epsilon * (src6 * src6 - src6);
    auto launch_force = [&]() {
      launch_cut(I, prec, x, d, t, N, sigma, epsilon, eshift, mode, g, ws.rs(), s);
      minim_amd::launch_finalize(ws.rs(), 2, cut_blocks(N), s);
    };
    if (!I.built) rebuild_list(I, prec, x, d, t, N, s);
    launch_force();
    double r[2];
    ws.fetch(s, 2, r);  // energy, particles moved more than skin/2 since the list was built
    if (r[1] > 0) {
      rebuild_list(I, prec, x, d, t, N, s);
      launch_force();
      ws.fetch(s, 2, r);
    }
    if (g) constrainDevice(prec, x, g, (size_t)N * nDim, stream);
    if (e) *e = r[0];
    return;
  }
  if (g) constrainDevice(prec, x, g, (size_t)N * nDim, stream);
  if (e) {
    double r[1];
    ws.fetch(s, 1, r);
    *e = r[0];
  }
}

// Cutoff mode: the force pass without the host synchronisation of the list
// check; the moved-particle count goes to the result slot returned.
int LjNd::gradientDeferred(int prec, const void* x, void* g, void* stream) const {
  if (!impl_) throw std::runtime_error("LjNd: not initialised (construct a State first).");
  if (cutoff <= 0) return Potential::gradientDeferred(prec, x, g, stream);
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto& ws = minim_amd::Workspace::current();
  Impl& I = *impl_;
  const int N = nParticle;
  if (!I.built) rebuild_list(I, prec, x, nullptr, 0.0, N, s);
  const double src6 = std::pow(sigma / cutoff, 6);
  const double eshift = 4 * epsilon * (src6 * src6 - src6);
  launch_cut(I, prec, x, nullptr, 0.0, N, sigma, epsilon, eshift, 2, g, ws.rs(), s);
  minim_amd::launch_finalize(ws.rs(), 2, cut_blocks(N), s, 1, ws.device_result() + minim_amd::kDeferSlot);
  constrainDevice(prec, x, g, (size_t)N * nDim, stream);
  return minim_amd::kDeferSlot;
}

void LjNd::redoGradient(int prec, const void* x, void* g, void* stream) const {
  if (!impl_) throw std::runtime_error("LjNd: not initialised (construct a State first).");
  if (cutoff <= 0) return Potential::redoGradient(prec, x, g, stream);
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto& ws = minim_amd::Workspace::current();
  const int N = nParticle;
  rebuild_list(*impl_, prec, x, nullptr, 0.0, N, s);
  const double src6 = std::pow(sigma / cutoff, 6);
  const double eshift = 4 * epsilon * (src6 * src6 - src6);
  launch_cut(*impl_, prec, x, nullptr, 0.0, N, sigma, epsilon, eshift, 2, g, ws.rs(), s);
  minim_amd::launch_finalize(ws.rs(), 2, cut_blocks(N), s);
  constrainDevice(prec, x, g, (size_t)N * nDim, stream);
}

// Average duration (ms) of reps launches of the force kernel (cutoff mode:
// the list is built first when it does not exist yet; all-pairs mode: the
// kernel and its reduction).
double LjNd::componentEnergy(int component, const vector<double>& coords) const {
  return component == 0 ? energy(coords) : 0.0;
}

double LjNd::timeKernel(int prec, const void* x, bool wantE, void* g, int reps, void* stream) const {
  if (!impl_) throw std::runtime_error("LjNd: not initialised (construct a State first).");
  const int mode = (wantE ? 1 : 0) | (g ? 2 : 0);
  if (mode == 0 || reps < 1) throw std::invalid_argument("LjNd: nothing to time");
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto& ws = minim_amd::Workspace::current();
  const int N = nParticle;
  const double src6 = cutoff > 0 ? std::pow(sigma / cutoff, 6) : 0.0;
  const double eshift = 4 * epsilon * (src6 * src6 - src6);
  auto launch = [&]() {
    if (cutoff > 0) {
      launch_cut(*impl_, prec, x, nullptr, 0.0, N, sigma, epsilon, eshift, mode, g, ws.rs(), s);
    } else if (prec == FP64) {
      minim_amd::launch_allpairs<double>(nDim, x, nullptr, 0.0, N, sigma, epsilon, mode, g, ws.rs(), s);
    } else {
      minim_amd::launch_allpairs<float>(nDim, x, nullptr, 0.0, N, sigma, epsilon, mode, g, ws.rs(), s);
    }
  };
  if (cutoff > 0 && !impl_->built) rebuild_list(*impl_, prec, x, nullptr, 0.0, N, s);
  launch();
  hipEvent_t e0, e1;
  MINIM_HIP_CHECK(hipEventCreate(&e0));
  MINIM_HIP_CHECK(hipEventCreate(&e1));
  MINIM_HIP_CHECK(hipEventRecord(e0, s));
  for (int i = 0; i < reps; i++) launch();
  MINIM_HIP_CHECK(hipEventRecord(e1, s));
  MINIM_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  MINIM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  MINIM_HIP_CHECK(hipEventDestroy(e0));
  MINIM_HIP_CHECK(hipEventDestroy(e1));
  return (double)ms / reps;
}

}  // namespace minim
