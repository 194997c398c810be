// vec.hip -- fused, HBM-streaming vector kernels of the device-resident
// minimisers (L-BFGS direction and Gram passes, FIRE update, dots, axpy).
//
// The reference performs the L-BFGS two-loop recursion (src/minimisers/Lbfgs.cpp:84-127)
// as ~2m dependent dot/axpy sweeps over full vectors with allocating
// temporaries (include/utils/vec.hpp operators).  Here the recursion runs on the
// (2m+1)x(2m+1) Gram matrix of the basis {g, s_i, y_i} on the host, and the
// vectors are touched by two passes per iteration:
//   * direction pass: d = sum_j c_j b_j (+ per-site projection), fused with the
//     dots d.g, d.y_i and d.d that the next Gram update and the descent and
//     max-step checks need;
//   * Gram pass: y_new = g_new - g_old fused with every dot of g_new and y_new
//     against the stored basis.
// All accumulations are FP64; vectors are FP64 or FP32.
#include <algorithm>

#include "common.hpp"
#include "vec.hpp"

namespace minim_amd {

namespace {

// result layout: [0] = d.d, [1+j] = d.dotv[j]
template <typename T, int NF, int K>
__global__ __launch_bounds__(kBlock) void direction_kernel(DirArgs a, RedScratch rs) {
  constexpr int NFM = NF > 0 ? NF : 16;
  const int nf = NF > 0 ? NF : a.proj.nf;
  const int NK = NF > 0 ? NF : nf;
  const size_t nsite = a.n / nf;
  T* __restrict__ out = static_cast<T*>(a.out);
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0.0;

  for (size_t s = (size_t)blockIdx.x * kBlock + threadIdx.x; s < nsite; s += (size_t)gridDim.x * kBlock) {
    const size_t base = s * nf;
    double v[NFM];
#pragma unroll
    for (int k = 0; k < NK; k++) v[k] = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxBasis; j++) {
      if (j < a.nb) {
        const T* __restrict__ b = static_cast<const T*>(a.basis[j]);
        const double cj = a.coef[j];
#pragma unroll
        for (int k = 0; k < NK; k++) {
          const double p = cj * (double)b[base + k];
          v[k] = v[k] + p;
        }
      }
    }
    if (a.proj.fixmask || (a.proj.hard && nf > 1)) {
#pragma unroll
      for (int k = 0; k < NK; k++)
        if (a.proj.fixmask & (1u << k)) v[k] = 0.0;
      if (a.proj.hard && nf > 1 && a.proj.nvar > 0) {
        double comp = 0.0;
#pragma unroll
        for (int k = 0; k < NK; k++)
          if (!(a.proj.fixmask & (1u << k))) comp += v[k];
        comp *= a.proj.inv_nvar;
#pragma unroll
        for (int k = 0; k < NK; k++)
          if (!(a.proj.fixmask & (1u << k))) v[k] -= comp;
      }
    }
    T o[NFM];
#pragma unroll
    for (int k = 0; k < NK; k++) {
      o[k] = (T)v[k];
      out[base + k] = o[k];
      acc[0] += (double)o[k] * (double)o[k];
    }
#pragma unroll
    for (int j = 0; j < K - 1; j++) {
      if (j < a.ndot) {
        const T* __restrict__ w = static_cast<const T*>(a.dotv[j]);
#pragma unroll
        for (int k = 0; k < NK; k++) acc[1 + j] += (double)o[k] * (double)w[base + k];
      }
    }
  }
  grid_sum_store<K>(acc, rs);
}

// result layout with KB = basis bucket: [0,KB) g.D ; [KB,2KB) g.Y ; 2KB g.g ;
// 2KB+1 g.y ; [2KB+2,3KB+2) y.D ; [3KB+2,4KB+2) y.Y ; 4KB+2 y.y
template <typename T, int KB>
__global__ __launch_bounds__(kBlock) void gram_kernel(GramArgs a, RedScratch rs) {
  constexpr int K = 4 * KB + 3;
  const T* __restrict__ gn = static_cast<const T*>(a.gnew);
  const T* __restrict__ go = static_cast<const T*>(a.gold);
  T* __restrict__ yo = static_cast<T*>(a.y);
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0.0;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += (size_t)gridDim.x * kBlock) {
    const T gv = gn[i];
    const T yv = gv - go[i];
    yo[i] = yv;
    const double g = (double)gv, y = (double)yv;
#pragma unroll
    for (int j = 0; j < KB; j++) {
      if (j < a.nd) {
        const double dv = (double)static_cast<const T*>(a.D[j])[i];
        acc[j] += g * dv;
        acc[2 * KB + 2 + j] += y * dv;
      }
    }
#pragma unroll
    for (int j = 0; j < KB; j++) {
      if (j < a.ny) {
        const double yy = (double)static_cast<const T*>(a.Y[j])[i];
        acc[KB + j] += g * yy;
        acc[3 * KB + 2 + j] += y * yy;
      }
    }
    acc[2 * KB] += g * g;
    acc[2 * KB + 1] += g * y;
    acc[4 * KB + 2] += y * y;
  }
  grid_sum_store<K>(acc, rs);
}

// ---- compile-time-P fast paths: every load unconditional and issued together.
// Their vectors are streamed once per pass (12-14 of them), so loads and
// stores are non-temporal: measured at 256^3 x 3 with P = 5, the Gram pass
// goes from 5.2 to 5.8 TB/s (FP32) and from 5.1 to 5.6 TB/s (FP64).  A
// direction pass over FP64 sites of several components does not load a
// site's components with one instruction each (non-temporal loads of 8 of 24
// bytes per lane re-fetch lines: 4.0 TB/s) but stages 64 sites per wavefront
// through LDS (direction_staged_kernel: 5.4 TB/s, against 4.9 before).
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ __forceinline__ void st_stream(T v, T* p) {
  __builtin_nontemporal_store(v, p);
}

// L-BFGS direction with P stored pairs, sites of NF components:
// d = c0 g + sum_q c(1+q) d_q + sum_q c(1+P+q) y_q, projected; dots d.d, d.g, d.y_q.
template <typename T, int NF, int P>
__global__ __launch_bounds__(kBlock) void direction_fast_kernel(DirArgs a, RedScratch rs) {
  constexpr int NB = 1 + 2 * P;
  constexpr int K = 2 + P;
  const size_t nsite = a.n / NF;
  T* __restrict__ out = static_cast<T*>(a.out);
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0.0;
  for (size_t s = (size_t)blockIdx.x * kBlock + threadIdx.x; s < nsite; s += (size_t)gridDim.x * kBlock) {
    const size_t base = s * NF;
    T b[NB][NF];
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const T* __restrict__ src = static_cast<const T*>(a.basis[j]);
#pragma unroll
      for (int k = 0; k < NF; k++) b[j][k] = ld_stream(src + base + k);
    }
    double v[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) v[k] = 0.0;
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const double cj = a.coef[j];
#pragma unroll
      for (int k = 0; k < NF; k++) {
        const double p = cj * (double)b[j][k];
        v[k] = v[k] + p;
      }
    }
    if (NF > 1 && (a.proj.fixmask || a.proj.hard)) {
#pragma unroll
      for (int k = 0; k < NF; k++)
        if (a.proj.fixmask & (1u << k)) v[k] = 0.0;
      if (a.proj.hard && a.proj.nvar > 0) {
        double comp = 0.0;
#pragma unroll
        for (int k = 0; k < NF; k++)
          if (!(a.proj.fixmask & (1u << k))) comp += v[k];
        comp *= a.proj.inv_nvar;
#pragma unroll
        for (int k = 0; k < NF; k++)
          if (!(a.proj.fixmask & (1u << k))) v[k] -= comp;
      }
    }
#pragma unroll
    for (int k = 0; k < NF; k++) {
      const T o = (T)v[k];
      st_stream(o, out + base + k);
      const double od = (double)o;
      acc[0] += od * od;
      acc[1] += od * (double)b[0][k];
#pragma unroll
      for (int q = 0; q < P; q++) acc[2 + q] += od * (double)b[1 + P + q][k];
    }
  }
  grid_sum_store<K>(acc, rs);
}

// Exchange through LDS within one wavefront: the writes of every lane are
// visible to the reads that follow.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// direction_fast_kernel's arithmetic, site for site, with the loads and
// stores of a wavefront's 64 consecutive sites (64 NF consecutive elements of
// each vector) issued as NF coalesced non-temporal instructions and passed to
// and from one site per lane through the wavefront's slice of LDS.
template <typename T, int NF, int P>
__global__ __launch_bounds__(kBlock) void direction_staged_kernel(DirArgs a, RedScratch rs) {
  constexpr int NB = 1 + 2 * P;
  constexpr int K = 2 + P;
  constexpr int CH = kWave * NF;
  __shared__ T slice[kBlock / kWave][CH];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const size_t nchunk = (a.n + CH - 1) / CH;
  T* __restrict__ out = static_cast<T*>(a.out);
  T* __restrict__ sl = slice[w];
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0.0;
  for (size_t c = (size_t)blockIdx.x * (kBlock / kWave) + w; c < nchunk;
       c += (size_t)gridDim.x * (kBlock / kWave)) {
    const size_t e0 = c * CH;
    const int ne = (int)((a.n - e0) < (size_t)CH ? (a.n - e0) : (size_t)CH);
    T r[NB][NF];
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const T* __restrict__ src = static_cast<const T*>(a.basis[j]) + e0;
#pragma unroll
      for (int i = 0; i < NF; i++) {
        const int idx = i * kWave + lane;
        r[j][i] = idx < ne ? ld_stream(src + idx) : (T)0;
      }
    }
    T b[NB][NF];
#pragma unroll
    for (int j = 0; j < NB; j++) {
#pragma unroll
      for (int i = 0; i < NF; i++) sl[i * kWave + lane] = r[j][i];
      wave_sync();
#pragma unroll
      for (int k = 0; k < NF; k++) b[j][k] = sl[lane * NF + k];
      wave_sync();
    }
    const bool valid = lane * NF < ne;
    double v[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) v[k] = 0.0;
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const double cj = a.coef[j];
#pragma unroll
      for (int k = 0; k < NF; k++) {
        const double p = cj * (double)b[j][k];
        v[k] = v[k] + p;
      }
    }
    if (a.proj.fixmask || a.proj.hard) {
#pragma unroll
      for (int k = 0; k < NF; k++)
        if (a.proj.fixmask & (1u << k)) v[k] = 0.0;
      if (a.proj.hard && a.proj.nvar > 0) {
        double comp = 0.0;
#pragma unroll
        for (int k = 0; k < NF; k++)
          if (!(a.proj.fixmask & (1u << k))) comp += v[k];
        comp *= a.proj.inv_nvar;
#pragma unroll
        for (int k = 0; k < NF; k++)
          if (!(a.proj.fixmask & (1u << k))) v[k] -= comp;
      }
    }
#pragma unroll
    for (int k = 0; k < NF; k++) {
      const T o = (T)v[k];
      sl[lane * NF + k] = o;
      if (valid) {
        const double od = (double)o;
        acc[0] += od * od;
        acc[1] += od * (double)b[0][k];
#pragma unroll
        for (int q = 0; q < P; q++) acc[2 + q] += od * (double)b[1 + P + q][k];
      }
    }
    wave_sync();
#pragma unroll
    for (int i = 0; i < NF; i++) {
      const int idx = i * kWave + lane;
      if (idx < ne) st_stream(sl[idx], out + e0 + idx);
    }
    wave_sync();
  }
  grid_sum_store<K>(acc, rs);
}

// Gram pass with P old pairs: D = [d_0..d_{P-1}, d_new], Y = [y_0..y_{P-1}];
// result layout of GramArgs with KB = P + 1.
template <typename T, int P>
__global__ __launch_bounds__(kBlock) void gram_fast_kernel(GramArgs a, RedScratch rs) {
  constexpr int KB = P + 1;
  constexpr int K = 4 * KB + 3;
  const T* __restrict__ gn = static_cast<const T*>(a.gnew);
  const T* __restrict__ go = static_cast<const T*>(a.gold);
  T* __restrict__ yo = static_cast<T*>(a.y);
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0.0;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += (size_t)gridDim.x * kBlock) {
    T dv[P + 1], yv[P > 0 ? P : 1];
    const T gv = ld_stream(gn + i);
    const T gov = ld_stream(go + i);
#pragma unroll
    for (int j = 0; j <= P; j++) dv[j] = ld_stream(static_cast<const T*>(a.D[j]) + i);
#pragma unroll
    for (int j = 0; j < P; j++) yv[j] = ld_stream(static_cast<const T*>(a.Y[j]) + i);
    const T ynew = gv - gov;
    st_stream(ynew, yo + i);
    const double g = (double)gv, y = (double)ynew;
#pragma unroll
    for (int j = 0; j <= P; j++) {
      acc[j] += g * (double)dv[j];
      acc[2 * KB + 2 + j] += y * (double)dv[j];
    }
#pragma unroll
    for (int j = 0; j < P; j++) {
      acc[KB + j] += g * (double)yv[j];
      acc[3 * KB + 2 + j] += y * (double)yv[j];
    }
    acc[2 * KB] += g * g;
    acc[2 * KB + 1] += g * y;
    acc[4 * KB + 2] += y * y;
  }
  grid_sum_store<K>(acc, rs);
}

template <typename T, int NF>
bool dir_fast_nf(const DirArgs& a, RedScratch rs, hipStream_t s, int nb) {
  switch (a.pairs) {
#define MINIM_DIR_CASE(P)                                                   \
  case P:                                                                   \
    if constexpr (sizeof(T) == 8 && NF > 1)                                 \
      direction_staged_kernel<T, NF, P><<<nb, kBlock, 0, s>>>(a, rs);       \
    else                                                                    \
      direction_fast_kernel<T, NF, P><<<nb, kBlock, 0, s>>>(a, rs);         \
    launch_finalize(rs, 2 + P, nb, s);                                      \
    return true;
    MINIM_DIR_CASE(0) MINIM_DIR_CASE(1) MINIM_DIR_CASE(2) MINIM_DIR_CASE(3) MINIM_DIR_CASE(4)
    MINIM_DIR_CASE(5) MINIM_DIR_CASE(6) MINIM_DIR_CASE(7) MINIM_DIR_CASE(8) MINIM_DIR_CASE(9)
    MINIM_DIR_CASE(10)
#undef MINIM_DIR_CASE
  }
  return false;
}

template <typename T>
bool gram_fast(const GramArgs& a, RedScratch rs, hipStream_t s, int nb) {
  if (a.nd != a.ny + 1) return false;
  switch (a.ny) {
#define MINIM_GRAM_CASE(P)                                                  \
  case P:                                                                   \
    gram_fast_kernel<T, P><<<nb, kBlock, 0, s>>>(a, rs);                    \
    launch_finalize(rs, 4 * (P + 1) + 3, nb, s);                            \
    return true;
    MINIM_GRAM_CASE(0) MINIM_GRAM_CASE(1) MINIM_GRAM_CASE(2) MINIM_GRAM_CASE(3) MINIM_GRAM_CASE(4)
    MINIM_GRAM_CASE(5) MINIM_GRAM_CASE(6) MINIM_GRAM_CASE(7) MINIM_GRAM_CASE(8) MINIM_GRAM_CASE(9)
    MINIM_GRAM_CASE(10)
#undef MINIM_GRAM_CASE
  }
  return false;
}

struct DotPtrs {
  const void* a[8];
  const void* b[8];
};

template <typename T, int K>
__global__ __launch_bounds__(kBlock) void dots_kernel_v(DotPtrs p, int k, size_t n, RedScratch rs) {
  double acc[K];
#pragma unroll
  for (int j = 0; j < K; j++) acc[j] = 0.0;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
#pragma unroll
    for (int j = 0; j < K; j++) {
      if (j >= k) break;
      acc[j] += (double)static_cast<const T*>(p.a[j])[i] * (double)static_cast<const T*>(p.b[j])[i];
    }
  }
  grid_sum_store<K>(acc, rs);
}

template <typename T, typename TV>
__global__ __launch_bounds__(kBlock) void axpy_kernel(T* __restrict__ x, const TV* __restrict__ d, T t, size_t n) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
    const T s = t * (T)d[i];
    x[i] = x[i] + s;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void add_kernel(T* __restrict__ x, const T* __restrict__ st, size_t n) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    x[i] = x[i] + st[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void scale_kernel(T* __restrict__ v, T f, size_t n) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    v[i] = v[i] * f;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void scale_copy_kernel(T* __restrict__ o, const T* __restrict__ v, T f, size_t n) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    o[i] = f * v[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void zero_idx_kernel(T* __restrict__ v, const int* __restrict__ idx, size_t m) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < m; i += (size_t)gridDim.x * kBlock)
    v[idx[i]] = (T)0;
}

template <typename TD, typename TS>
__global__ __launch_bounds__(kBlock) void convert_kernel(TD* __restrict__ dst, const TS* __restrict__ src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    dst[i] = (TD)src[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void fire_kernel(int mode, T* __restrict__ v, const T* __restrict__ g,
                                                      T* __restrict__ step, T f1, T f2, T dt, size_t n) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
    T nv;
    if (mode == 1) {
      const T t1 = f1 * v[i];
      const T t2 = f2 * g[i];
      nv = t1 - t2;
    } else {
      nv = f2 * g[i];
    }
    v[i] = nv;
    step[i] = dt * nv;
  }
}

// NaN-propagating maximum
__device__ inline double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }

template <typename T>
__global__ __launch_bounds__(kBlock) void group_norm_max_kernel(const T* __restrict__ v, size_t ngroup, int dim,
                                                                 double* __restrict__ partial) {
  double m = 0.0;
  for (size_t q = (size_t)blockIdx.x * kBlock + threadIdx.x; q < ngroup; q += (size_t)gridDim.x * kBlock) {
    double s = 0.0;
    for (int k = 0; k < dim; k++) {
      const double c = (double)v[q * dim + k];
      s += c * c;
    }
    m = nan_max(m, s);
  }
  for (int off = kWave / 2; off > 0; off >>= 1) m = nan_max(m, __shfl_xor(m, off, kWave));
  __shared__ double lds[kBlock / kWave];
  if (threadIdx.x % kWave == 0) lds[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / kWave; w++) m = nan_max(m, lds[w]);
    partial[blockIdx.x] = m;
  }
}

__global__ __launch_bounds__(kBlock) void max_sqrt_finalize_kernel(const double* __restrict__ partial, int nb,
                                                                   double* __restrict__ result) {
  double m = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) m = nan_max(m, partial[b]);
  for (int off = kWave / 2; off > 0; off >>= 1) m = nan_max(m, __shfl_xor(m, off, kWave));
  __shared__ double lds[kBlock / kWave];
  if (threadIdx.x % kWave == 0) lds[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / kWave; w++) m = nan_max(m, lds[w]);
    result[0] = sqrt(m);
  }
}

int bucket(int k) {
  if (k <= 8) return 8;
  if (k <= 16) return 16;
  if (k <= 32) return 32;
  if (k <= 64) return 64;
  return 96;
}

template <typename T, int NF>
void dir_nf(const DirArgs& a, RedScratch rs, hipStream_t s, int nb) {
  const int K = bucket(a.ndot + 1);
  switch (K) {
    case 8: direction_kernel<T, NF, 8><<<nb, kBlock, 0, s>>>(a, rs); break;
    case 16: direction_kernel<T, NF, 16><<<nb, kBlock, 0, s>>>(a, rs); break;
    default: direction_kernel<T, NF, 32><<<nb, kBlock, 0, s>>>(a, rs); break;
  }
  launch_finalize(rs, K > 32 ? 32 : K, nb, s);
}

template <typename T>
void dir_t(const DirArgs& a, RedScratch rs, hipStream_t s) {
  const int nf = a.proj.nf < 1 ? 1 : a.proj.nf;
  // the fast kernels: 512 blocks (2 per CU); with 12 concurrent streams fewer,
  // longer-running blocks reach the higher rate (256^3 x 3, P = 5: FP32 418 us
  // at 512 blocks against 463 us at 2048).  The general kernel, which loops
  // over its basis at run time, keeps 2048 (1.3 ms against 2.4 ms at 512 for
  // 16 basis vectors over 50 M elements).
  const bool fast = a.pairs >= 0 && a.pairs <= kFastPairs && (nf == 1 || nf == 3);
  const int nb = a.blocks > 0 ? std::min(a.blocks, kMaxRedBlocks) : grid_blocks(a.n / nf, 2, fast ? 512 : 2048);
  if (fast) {
    if (nf == 1 && dir_fast_nf<T, 1>(a, rs, s, nb)) return;
    if (nf == 3 && dir_fast_nf<T, 3>(a, rs, s, nb)) return;
  }
  switch (nf) {
    case 1: dir_nf<T, 1>(a, rs, s, nb); break;
    case 2: dir_nf<T, 2>(a, rs, s, nb); break;
    case 3: dir_nf<T, 3>(a, rs, s, nb); break;
    case 4: dir_nf<T, 4>(a, rs, s, nb); break;
    default: dir_nf<T, 0>(a, rs, s, nb); break;
  }
}

template <typename T>
void gram_t(const GramArgs& a, RedScratch rs, hipStream_t s) {
  // the fast kernels: 512 blocks (2 per CU); the fused y/Gram pass streams
  // 2P+3 vectors at once and reaches a higher rate with fewer, longer-running
  // blocks (256^3 x 3, P = 5: FP32 489 us at 512 blocks, 511 us at 1024, 539
  // us at 4096).  The general kernel keeps 1024.
  const bool fast = a.ny <= kFastPairs && a.nd == a.ny + 1;
  const int nb = a.blocks > 0 ? std::min(a.blocks, kMaxRedBlocks) : grid_blocks(a.n, 4, fast ? 512 : 1024);
  if (fast && gram_fast<T>(a, rs, s, nb)) return;
  const int KB = gram_bucket(a.nd > a.ny ? a.nd : a.ny);
  switch (KB) {
    case 6: gram_kernel<T, 6><<<nb, kBlock, 0, s>>>(a, rs); break;
    case 11: gram_kernel<T, 11><<<nb, kBlock, 0, s>>>(a, rs); break;
    default: gram_kernel<T, 16><<<nb, kBlock, 0, s>>>(a, rs); break;
  }
  launch_finalize(rs, 4 * KB + 3, nb, s);
}

}  // namespace

void launch_direction(const DirArgs& a, int prec, RedScratch rs, hipStream_t s) {
  if (a.nb > kMaxBasis || a.ndot > kMaxBasis) throw std::invalid_argument("direction: basis too large");
  if (prec == F64) dir_t<double>(a, rs, s);
  else dir_t<float>(a, rs, s);
  MINIM_HIP_CHECK(hipGetLastError());
}

int gram_bucket(int k) {
  if (k <= 6) return 6;
  if (k <= 11) return 11;
  return 16;
}

int gram_kb(int nd, int ny) {
  if (nd == ny + 1 && ny <= kFastPairs) return ny + 1;
  return gram_bucket(nd > ny ? nd : ny);
}

void launch_gram(const GramArgs& a, int prec, RedScratch rs, hipStream_t s) {
  if (a.nd > 16 || a.ny > 16) throw std::invalid_argument("gram: history too long");
  if (prec == F64) gram_t<double>(a, rs, s);
  else gram_t<float>(a, rs, s);
  MINIM_HIP_CHECK(hipGetLastError());
}

void launch_dots(int prec, const void* const* a, const void* const* b, int k, size_t n, RedScratch rs,
                 hipStream_t s) {
  if (k < 1 || k > 8) throw std::invalid_argument("dots: 1..8 pairs");
  DotPtrs p{};
  for (int j = 0; j < k; j++) {
    p.a[j] = a[j];
    p.b[j] = b[j];
  }
  const int nb = grid_blocks(n, 4, 2048);
  if (prec == F64) dots_kernel_v<double, 8><<<nb, kBlock, 0, s>>>(p, k, n, rs);
  else dots_kernel_v<float, 8><<<nb, kBlock, 0, s>>>(p, k, n, rs);
  MINIM_HIP_CHECK(hipGetLastError());
  launch_finalize(rs, k, nb, s);
}

void launch_axpy(int prec, void* x, const void* d, double t, size_t n, hipStream_t s) {
  launch_axpy_mixed(prec, x, prec, d, t, n, s);
}

void launch_axpy_mixed(int px, void* x, int pv, const void* d, double t, size_t n, hipStream_t s) {
  const int nb = grid_blocks(n, 4, 4096);
  if (px == F64 && pv == F64) axpy_kernel<double, double><<<nb, kBlock, 0, s>>>((double*)x, (const double*)d, t, n);
  else if (px == F32 && pv == F32) axpy_kernel<float, float><<<nb, kBlock, 0, s>>>((float*)x, (const float*)d, (float)t, n);
  else if (px == F64 && pv == F32) axpy_kernel<double, float><<<nb, kBlock, 0, s>>>((double*)x, (const float*)d, t, n);
  else throw std::invalid_argument("axpy: FP32 coordinates with an FP64 direction");
  MINIM_HIP_CHECK(hipGetLastError());
}

void launch_add(int prec, void* x, const void* st, size_t n, hipStream_t s) {
  const int nb = grid_blocks(n, 4, 4096);
  if (prec == F64) add_kernel<double><<<nb, kBlock, 0, s>>>((double*)x, (const double*)st, n);
  else add_kernel<float><<<nb, kBlock, 0, s>>>((float*)x, (const float*)st, n);
  MINIM_HIP_CHECK(hipGetLastError());
}

void launch_scale(int prec, void* v, double f, size_t n, hipStream_t s) {
  const int nb = grid_blocks(n, 4, 4096);
  if (prec == F64) scale_kernel<double><<<nb, kBlock, 0, s>>>((double*)v, f, n);
  else scale_kernel<float><<<nb, kBlock, 0, s>>>((float*)v, (float)f, n);
  MINIM_HIP_CHECK(hipGetLastError());
}

void launch_scale_copy(int prec, void* o, const void* v, double f, size_t n, hipStream_t s) {
  const int nb = grid_blocks(n, 4, 4096);
  if (prec == F64) scale_copy_kernel<double><<<nb, kBlock, 0, s>>>((double*)o, (const double*)v, f, n);
  else scale_copy_kernel<float><<<nb, kBlock, 0, s>>>((float*)o, (const float*)v, (float)f, n);
  MINIM_HIP_CHECK(hipGetLastError());
}

void launch_zero_idx(int prec, void* v, const int* idx, size_t m, hipStream_t s) {
  if (m == 0) return;
  const int nb = grid_blocks(m, 1, 1024);
  if (prec == F64) zero_idx_kernel<double><<<nb, kBlock, 0, s>>>((double*)v, idx, m);
  else zero_idx_kernel<float><<<nb, kBlock, 0, s>>>((float*)v, idx, m);
  MINIM_HIP_CHECK(hipGetLastError());
}

void launch_convert(int pd, void* dst, int ps, const void* src, size_t n, hipStream_t s) {
  const int nb = grid_blocks(n, 4, 4096);
  if (pd == F64 && ps == F32) convert_kernel<double, float><<<nb, kBlock, 0, s>>>((double*)dst, (const float*)src, n);
  else if (pd == F32 && ps == F64) convert_kernel<float, double><<<nb, kBlock, 0, s>>>((float*)dst, (const double*)src, n);
  else if (pd == F64) convert_kernel<double, double><<<nb, kBlock, 0, s>>>((double*)dst, (const double*)src, n);
  else convert_kernel<float, float><<<nb, kBlock, 0, s>>>((float*)dst, (const float*)src, n);
  MINIM_HIP_CHECK(hipGetLastError());
}

void launch_group_norm_max(int prec, const void* v, size_t n, int dim, RedScratch rs, hipStream_t s) {
  if (dim < 1) throw std::invalid_argument("group norm: dim must be positive");
  const size_t ngroup = n / (size_t)dim;
  const int nb = grid_blocks(ngroup, 4, 2048);
  if (prec == F64) group_norm_max_kernel<double><<<nb, kBlock, 0, s>>>((const double*)v, ngroup, dim, rs.partials);
  else group_norm_max_kernel<float><<<nb, kBlock, 0, s>>>((const float*)v, ngroup, dim, rs.partials);
  MINIM_HIP_CHECK(hipGetLastError());
  max_sqrt_finalize_kernel<<<1, kBlock, 0, s>>>(rs.partials, nb, rs.result);
  MINIM_HIP_CHECK(hipGetLastError());
}

void launch_fire_update(int prec, int mode, void* v, const void* g, void* step, double f1, double f2,
                        double dt, size_t n, hipStream_t s) {
  const int nb = grid_blocks(n, 4, 4096);
  if (prec == F64)
    fire_kernel<double><<<nb, kBlock, 0, s>>>(mode, (double*)v, (const double*)g, (double*)step, f1, f2, dt, n);
  else
    fire_kernel<float><<<nb, kBlock, 0, s>>>(mode, (float*)v, (const float*)g, (float*)step, (float)f1, (float)f2,
                                             (float)dt, n);
  MINIM_HIP_CHECK(hipGetLastError());
}

}  // namespace minim_amd
