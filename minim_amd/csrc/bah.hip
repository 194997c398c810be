// bah.hip -- bar-and-hinge elastic sheet on MI355X.
//
// Element kernels follow BarAndHinge::stretching (src/potentials/BarAndHinge.cpp:269-290,
// including its gradient factor 2*k*dl/l, twice the derivative of its
// 0.5*k*dl^2 energy), bending (:293-356) and forceEnergy (:359-369); element
// order and the all-DOFs-fixed skip follow BarAndHinge::init (:16-67).
//
// The reference loops over heap elements and scatters into g.  Here the
// energy is one thread per element with an FP64 reduction, and the gradient is
// gathered: one thread per node walks its incident bars and hinges (CSR built
// once on the host, in element order) and adds each element's contribution to
// that node, so every gradient entry is written once, without atomics, in a
// fixed order.  A bar's contribution is recomputed at each of its two nodes.
// For a hinge, the energy pass (or, for a gradient without energy, the hinge
// pass) stores per hinge what the gather needs beyond the node positions:
// in FP32 the terms its four nodes share -- k dtheta times each scaled
// normal, and the two skews (8 values) -- so that the gather recomputes no
// geometry; in FP64 only the angle deviation, the gather recomputing the
// geometry at each node.  Measured at 1024^2 nodes: FP32 gather 114 -> 77 us
// with the terms (energy pass 54 -> 70 us); FP64 gather 158 -> 149 us but
// energy pass 78 -> 127 us (64 B instead of 8 B written per hinge), so FP64
// keeps the angle alone.
//
// Rigidities, rest lengths and rest angles that are the same for every
// element are passed as scalars instead of per-element streams.
//
// Bending angle: the FP64 path uses the reference's acos form (bit-level
// parity); the FP32 path uses atan2(|b2| n1.b3, n1.n2), the same angle without
// acos's loss of precision near 0 and pi.
#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <unordered_map>

#include "common.hpp"
#include "minim/BarAndHinge.h"
#include "runtime.hpp"
#include "vec.hpp"

namespace minim_amd {
namespace {

struct BaHArgs {
  int nnode, nbar, nhinge;
  const int* bar;     // 2 per bar
  const double* bk;   // per bar, or null: every bar has bkU
  const double* bl0;  // ... bl0U
  const int* hinge;   // 4 per hinge
  const double* hk;   // per hinge, or null: hkU
  const double* ht0;  // ... ht0U
  double bkU, bl0U, hkU, ht0U;
  const int* bptr;    // node -> incident bars CSR (nnode+1)
  const int* bidx;    // (bar << 1) | local
  const int* hptr;    // node -> incident hinges CSR
  const int* hidx;    // (hinge << 2) | local
  const double* fz;   // per-node z force or nullptr
  const unsigned char* nodeActive;  // force element exists for the node
};

template <typename T>
__device__ inline void ld3(const T* __restrict__ x, const T* __restrict__ d, T t, int n, T (&p)[3]) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    T c = x[(size_t)n * 3 + k];
    if (d) {
      T s = t * d[(size_t)n * 3 + k];
      c = c + s;
    }
    p[k] = c;
  }
}

template <typename T>
__device__ inline void cross(const T (&a)[3], const T (&b)[3], T (&c)[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename T>
__device__ inline T dot3(const T (&a)[3], const T (&b)[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Hinge geometry shared by energy and gradient.
template <typename T>
struct HingeGeo {
  T b1[3], b2[3], b3[3], n1[3], n2[3];
  T b2m, n1sq, n2sq, dtheta;
};

// Hinge geometry from the four node positions, without the angle.
template <typename T>
__device__ inline void hinge_geo_vec(const T (&x1)[3], const T (&x2)[3], const T (&x3)[3], const T (&x4)[3],
                                     HingeGeo<T>& G) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    G.b1[k] = x2[k] - x1[k];
    G.b2[k] = x3[k] - x2[k];
    G.b3[k] = x4[k] - x3[k];
  }
  G.b2m = sqrt(dot3(G.b2, G.b2));
  cross(G.b1, G.b2, G.n1);
  cross(G.b2, G.b3, G.n2);
  G.n1sq = dot3(G.n1, G.n1);
  G.n2sq = dot3(G.n2, G.n2);
}

// The bending angle's deviation from the rest angle t0.
template <typename T>
__device__ inline void hinge_dtheta(HingeGeo<T>& G, T t0) {
  const T n1b3 = dot3(G.n1, G.b3);
  const T pi = (T)3.14159265358979323846;
  T theta;
  if (sizeof(T) == 8) {
    T c = dot3(G.n1, G.n2) / sqrt(G.n1sq * G.n2sq);
    c = c < (T)-1 ? (T)-1 : c;
    c = c > (T)1 ? (T)1 : c;
    theta = (n1b3 >= 0) ? acos(c) : (T)2 * pi - acos(c);
  } else {
    const T yv = G.b2m * n1b3;
    const T at = atan2(yv, dot3(G.n1, G.n2));
    theta = (yv >= 0) ? at : (T)2 * pi + at;
  }
  // dtheta = fmod(theta - t0 + pi, 2 pi) - pi (BarAndHinge.cpp:324).  With
  // |a| < 2b, fmod(a, b) is a, a - b or a + b, each exact (Sterbenz), so the
  // library call is needed only for rest angles far outside [0, 2 pi].
  const T a = (theta - t0) + pi, b2 = (T)2 * pi;
  T r;
  if (a < (T)2 * b2 && a > -(T)2 * b2) r = a >= b2 ? a - b2 : (a <= -b2 ? a + b2 : a);
  else r = fmod(a, b2);
  G.dtheta = r - pi;
}

// Hinge geometry from the four node positions and the rest angle.
template <typename T>
__device__ inline void hinge_geo_at(const T (&x1)[3], const T (&x2)[3], const T (&x3)[3], const T (&x4)[3], T t0,
                                    HingeGeo<T>& G) {
  hinge_geo_vec(x1, x2, x3, x4, G);
  hinge_dtheta(G, t0);
}

// Node positions and rest angle of hinge h (one int4 of node ids).
template <typename T>
struct HingeIn {
  T x1[3], x2[3], x3[3], x4[3];
  T t0, k;
};

template <typename T>
__device__ inline void hinge_load(const T* __restrict__ x, const T* __restrict__ d, T t, const BaHArgs& a, int h,
                                  HingeIn<T>& H) {
  const int4 nd = reinterpret_cast<const int4*>(a.hinge)[h];
  ld3(x, d, t, nd.x, H.x1);
  ld3(x, d, t, nd.y, H.x2);
  ld3(x, d, t, nd.z, H.x3);
  ld3(x, d, t, nd.w, H.x4);
  H.t0 = (T)(a.ht0 ? a.ht0[h] : a.ht0U);
  H.k = (T)(a.hk ? a.hk[h] : a.hkU);
}

// A hinge's gradient (BarAndHinge.cpp:336-354) through the terms its four
// nodes share: A = k dtheta |b2| n1 / |n1|^2, B = k dtheta |b2| n2 / |n2|^2,
// skew1 = -b1.b2 / |b2|^2, skew2 = -b3.b2 / |b2|^2 (8 values, HingeTerms);
// node 0 gets -A, node 1 (1 - skew1) A - skew2 B, node 2 skew1 A - (1 - skew2) B,
// node 3 B.
template <typename T>
struct alignas(sizeof(T) * 4) HingeTerms {
  T A[3], s1, B[3], s2;
};

template <typename T>
__device__ inline void hinge_terms(const HingeGeo<T>& G, T k, HingeTerms<T>& H) {
  const T gf = k * (G.dtheta);
  const T f1 = G.b2m / G.n1sq, f2 = G.b2m / G.n2sq;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    H.A[c] = gf * (f1 * G.n1[c]);
    H.B[c] = gf * (f2 * G.n2[c]);
  }
  const T bb = G.b2m * G.b2m;
  H.s1 = -dot3(G.b1, G.b2) / bb;
  H.s2 = -dot3(G.b3, G.b2) / bb;
}

// What the energy or hinge pass stores per hinge for the gradient gather.
template <typename T>
struct HingeStore {
  using type = T;  // FP64: the angle deviation
};
template <>
struct HingeStore<float> {
  using type = HingeTerms<float>;
};
template <typename T>
using HStore = typename HingeStore<T>::type;
template <typename T>
constexpr bool kHingeTerms = sizeof(HStore<T>) != sizeof(T);

template <typename T>
__device__ inline void hinge_store(const HingeGeo<T>& G, T k, HStore<T>* __restrict__ out, size_t h) {
  if constexpr (kHingeTerms<T>) {
    HingeTerms<T> tm;
    hinge_terms(G, k, tm);
    out[h] = tm;
  } else {
    out[h] = G.dtheta;
  }
}

template <typename T>
__device__ inline void hinge_grad_at(const HingeTerms<T>& H, int loc, T (&gn)[3]) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    T v;
    if (loc == 0) v = -H.A[c];
    else if (loc == 1) v = ((T)1 - H.s1) * H.A[c] - H.s2 * H.B[c];
    else if (loc == 2) v = H.s1 * H.A[c] - ((T)1 - H.s2) * H.B[c];
    else v = H.B[c];
    gn[c] += v;
  }
}

// Energy: bars, then hinges, then the external force, each a grid-stride
// loop that takes U elements per step and issues all their loads first.
// With SD, each hinge's gradient terms also go to ht (for the gradient
// kernel, which then recomputes no hinge geometry).
template <typename T, bool SD>
__global__ __launch_bounds__(kBlock) void bah_energy_kernel(BaHArgs a, const T* __restrict__ x,
                                                            const T* __restrict__ d, T t, RedScratch rs,
                                                            HStore<T>* __restrict__ ht) {
  constexpr int U = 4;
  double acc[1] = {0.0};
  const size_t stride = (size_t)gridDim.x * kBlock;
  // contiguous ranges of blocks per XCD: neighbouring elements share nodes in one L2
  const size_t tid = (size_t)xcd_logical_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  for (size_t w0 = tid; w0 < (size_t)a.nbar; w0 += U * stride) {
    T p1[U][3], p2[U][3], k[U], l0[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const size_t w = w0 + u * stride;
      const int b = (int)(w < (size_t)a.nbar ? w : (size_t)a.nbar - 1);
      const int2 nd = reinterpret_cast<const int2*>(a.bar)[b];
      ld3(x, d, t, nd.x, p1[u]);
      ld3(x, d, t, nd.y, p2[u]);
      k[u] = (T)(a.bk ? a.bk[b] : a.bkU);
      l0[u] = (T)(a.bl0 ? a.bl0[b] : a.bl0U);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (w0 + u * stride >= (size_t)a.nbar) break;
      T dx[3] = {p1[u][0] - p2[u][0], p1[u][1] - p2[u][1], p1[u][2] - p2[u][2]};
      const T l = sqrt(dot3(dx, dx));
      const T dl = l - l0[u];
      acc[0] += (double)((T)0.5 * k[u] * (dl * dl));
    }
  }
  for (size_t w0 = tid; w0 < (size_t)a.nhinge; w0 += U * stride) {
    HingeIn<T> H[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const size_t w = w0 + u * stride;
      hinge_load(x, d, t, a, (int)(w < (size_t)a.nhinge ? w : (size_t)a.nhinge - 1), H[u]);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (w0 + u * stride >= (size_t)a.nhinge) break;
      HingeGeo<T> G;
      hinge_geo_at(H[u].x1, H[u].x2, H[u].x3, H[u].x4, H[u].t0, G);
      acc[0] += (double)((T)0.5 * H[u].k * (G.dtheta * G.dtheta));
      if (SD) hinge_store(G, H[u].k, ht, w0 + u * stride);
    }
  }
  if (a.fz) {
    for (size_t n = tid; n < (size_t)a.nnode; n += stride) {
      if (a.nodeActive[n]) {
        T p[3];
        ld3(x, d, t, (int)n, p);
        acc[0] -= a.fz[n] * (double)p[2];
      }
    }
  }
  grid_sum_store<1>(acc, rs);
}

// What the gradient gather needs of every hinge (gradient without energy).
template <typename T>
__global__ __launch_bounds__(kBlock) void bah_hinge_kernel(BaHArgs a, const T* __restrict__ x,
                                                           HStore<T>* __restrict__ ht) {
  for (size_t h = (size_t)blockIdx.x * kBlock + threadIdx.x; h < (size_t)a.nhinge; h += (size_t)gridDim.x * kBlock) {
    HingeIn<T> H;
    hinge_load<T>(x, nullptr, (T)0, a, (int)h, H);
    HingeGeo<T> G;
    hinge_geo_at(H.x1, H.x2, H.x3, H.x4, H.t0, G);
    hinge_store(G, H.k, ht, h);
  }
}

// Gradient, gathered per node over its incident bars and hinges (CSR in
// element order), U elements at a time with all their loads issued first;
// contributions are added in list order.  Hinge terms (FP32) or angle
// deviations (FP64) come from ht, stored by the energy or the hinge kernel.
// TG: the gradient's storage (FP32 for FP64 arithmetic in the mixed schedule).
template <typename T, typename TG = T>
__global__ __launch_bounds__(kBlock) void bah_gradient_kernel(BaHArgs a, const T* __restrict__ x,
                                                              const HStore<T>* __restrict__ ht,
                                                              TG* __restrict__ g) {
  constexpr int U = 4;
  for (int n = xcd_logical_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; n < a.nnode;
       n += gridDim.x * kBlock) {
    T gn[3] = {(T)0, (T)0, (T)0};
    const int qb0 = a.bptr[n], qb1 = a.bptr[n + 1];
    for (int q0 = qb0; q0 < qb1; q0 += U) {
      T p1[U][3], p2[U][3], k[U], l0[U];
      int loc[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int code = a.bidx[min(q0 + u, qb1 - 1)];
        const int b = code >> 1;
        loc[u] = code & 1;
        const int2 nd = reinterpret_cast<const int2*>(a.bar)[b];
        ld3<T>(x, nullptr, (T)0, nd.x, p1[u]);
        ld3<T>(x, nullptr, (T)0, nd.y, p2[u]);
        k[u] = (T)(a.bk ? a.bk[b] : a.bkU);
        l0[u] = (T)(a.bl0 ? a.bl0[b] : a.bl0U);
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (q0 + u >= qb1) break;
        T dx[3] = {p1[u][0] - p2[u][0], p1[u][1] - p2[u][1], p1[u][2] - p2[u][2]};
        const T l = sqrt(dot3(dx, dx));
        const T dl = l - l0[u];
        // stretching gradient 2 k (l - l0) / l * dx (BarAndHinge.cpp:282)
        const T gf = (T)2 * k[u] * dl / l;
        const T sg = loc[u] == 0 ? (T)1 : (T)-1;
#pragma unroll
        for (int c = 0; c < 3; c++) gn[c] += sg * (gf * dx[c]);
      }
    }
    const int qh0 = a.hptr[n], qh1 = a.hptr[n + 1];
    for (int q0 = qh0; q0 < qh1; q0 += U) {
      int loc[U];
      if constexpr (kHingeTerms<T>) {
        HingeTerms<T> H[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int code = a.hidx[min(q0 + u, qh1 - 1)];
          loc[u] = code & 3;
          H[u] = ht[code >> 2];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          if (q0 + u >= qh1) break;
          hinge_grad_at(H[u], loc[u], gn);
        }
      } else {
        HingeIn<T> H[U];
        T dt[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int code = a.hidx[min(q0 + u, qh1 - 1)];
          loc[u] = code & 3;
          hinge_load<T>(x, nullptr, (T)0, a, code >> 2, H[u]);
          dt[u] = ht[code >> 2];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          if (q0 + u >= qh1) break;
          HingeGeo<T> G;
          hinge_geo_vec(H[u].x1, H[u].x2, H[u].x3, H[u].x4, G);
          G.dtheta = dt[u];
          HingeTerms<T> tm;
          hinge_terms(G, H[u].k, tm);
          hinge_grad_at(tm, loc[u], gn);
        }
      }
    }
    if (a.fz && a.nodeActive[n]) gn[2] -= (T)a.fz[n];
#pragma unroll
    for (int k = 0; k < 3; k++) g[(size_t)n * 3 + k] = (TG)gn[k];
  }
}

}  // namespace
}  // namespace minim_amd

namespace minim {

using minim_amd::DeviceBuffer;
using minim_amd::kBlock;

struct BarAndHinge::Dev {
  minim_amd::BaHArgs args{};
  DeviceBuffer mem;
  DeviceBuffer ht;      // per-hinge store of the last gradient evaluation (HStore, either precision)
  DeviceBuffer xtrial;  // node coordinates at a line-search trial point (FP64-sized)
  std::vector<int> activeBars, activeHinges;  // element index -> original bond/hinge index
  // The coordinates (and their precision) for which a fused step's energy
  // pass kept the hinge store with no gradient yet; any other evaluation
  // clears it, and the next gradient-only evaluation at exactly those
  // coordinates skips the hinge pass.
  mutable const void* keptAt = nullptr;
  mutable int keptPrec = -1;
};

BarAndHinge& BarAndHinge::setTriangulation(const vector2d<int>& triList) {
  tri_ = triList;
  return *this;
}
BarAndHinge& BarAndHinge::setBondList(const vector2d<int>& b) {
  bonds_ = b;
  return *this;
}
BarAndHinge& BarAndHinge::setHingeList(const vector2d<int>& h) {
  hinges_ = h;
  return *this;
}
BarAndHinge& BarAndHinge::setRigidity(double kb, double kh) {
  kBond_ = {kb};
  kHinge_ = {kh};
  return *this;
}
BarAndHinge& BarAndHinge::setRigidity(const vector<double>& kb, const vector<double>& kh) {
  if (!kb.empty()) kBond_ = kb;
  if (!kh.empty()) kHinge_ = kh;
  return *this;
}
BarAndHinge& BarAndHinge::setLength0(double l) {
  length0_ = {l};
  return *this;
}
BarAndHinge& BarAndHinge::setLength0(const vector<double>& l) {
  length0_ = l;
  return *this;
}
BarAndHinge& BarAndHinge::setTheta0(double t) {
  theta0_ = {t};
  return *this;
}
BarAndHinge& BarAndHinge::setTheta0(const vector<double>& t) {
  theta0_ = t;
  return *this;
}
BarAndHinge& BarAndHinge::setForce(const vector<double>& f) {
  force_ = f;
  return *this;
}

namespace {

// Bond list of BarAndHinge::computeBondList (BarAndHinge.cpp:70-93): sorted
// node pairs of every triangle, lexicographically sorted and unique.
vector2d<int> bonds_from_tris(const vector2d<int>& tris) {
  std::vector<std::pair<int, int>> e;
  e.reserve(tris.size() * 3);
  for (const auto& t0 : tris) {
    if (t0.size() != 3) throw std::invalid_argument("BarAndHinge: triangles need three nodes.");
    int t[3] = {t0[0], t0[1], t0[2]};
    std::sort(t, t + 3);
    e.push_back({t[0], t[1]});
    e.push_back({t[0], t[2]});
    e.push_back({t[1], t[2]});
  }
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  vector2d<int> out;
  out.reserve(e.size());
  for (auto& p : e) out.push_back({p.first, p.second});
  return out;
}

// Hinge list of BarAndHinge::computeHingeList (BarAndHinge.cpp:96-147) in
// the same order, built in O(T) from an edge -> triangles map instead of the
// reference's all-pairs triangle loop: for each triangle it1 and each later
// triangle it2 sharing exactly two nodes, nodes = [unshared of it1, the next
// two of it1 in cyclic order, unshared of it2].
vector2d<int> hinges_from_tris(const vector2d<int>& tris) {
  std::unordered_map<long long, std::vector<int>> edge;
  edge.reserve(tris.size() * 3);
  auto key = [](int a, int b) {
    if (a > b) std::swap(a, b);
    return ((long long)a << 32) | (unsigned)b;
  };
  for (int it = 0; it < (int)tris.size(); it++)
    for (int k = 0; k < 3; k++) edge[key(tris[it][k], tris[it][(k + 1) % 3])].push_back(it);
  vector2d<int> out;
  for (int it1 = 0; it1 < (int)tris.size(); it1++) {
    std::vector<int> cand;
    for (int k = 0; k < 3; k++)
      for (int it2 : edge[key(tris[it1][k], tris[it1][(k + 1) % 3])])
        if (it2 > it1) cand.push_back(it2);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    for (int it2 : cand) {
      bool s1[3] = {false, false, false}, s2[3] = {false, false, false};
      int nShared = 0;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
          if (tris[it1][a] == tris[it2][b]) {
            nShared++;
            s1[a] = true;
            s2[b] = true;
          }
      if (nShared != 2) continue;
      std::vector<int> nodes(4);
      for (int a = 0; a < 3; a++)
        if (!s1[a]) {
          nodes[0] = tris[it1][a];
          nodes[1] = tris[it1][(a + 1) % 3];
          nodes[2] = tris[it1][(a + 2) % 3];
        }
      for (int b = 0; b < 3; b++)
        if (!s2[b]) nodes[3] = tris[it2][b];
      out.push_back(nodes);
    }
  }
  return out;
}

double pi() { return 4 * std::atan(1.0); }

}  // namespace

// BarAndHinge::init (BarAndHinge.cpp:16-67)
void BarAndHinge::init(const vector<double>& coords) {
  if (coords.size() % 3 != 0) throw std::invalid_argument("BarAndHinge: coordinates must be 3 per node.");
  const int nnode = (int)(coords.size() / 3);
  if (bonds_.empty()) {
    if (tri_.empty()) throw std::invalid_argument("BarAndHinge: The triangulation has not been set.");
    bonds_ = bonds_from_tris(tri_);
  }
  if (hinges_.empty()) {
    if (tri_.empty()) throw std::invalid_argument("BarAndHinge: The triangulation has not been set.");
    hinges_ = hinges_from_tris(tri_);
  }
  for (auto& b : bonds_)
    if (b.size() < 2 || b[0] < 0 || b[1] < 0 || b[0] >= nnode || b[1] >= nnode)
      throw std::invalid_argument("BarAndHinge: bond references a node outside the coordinates.");
  for (auto& h : hinges_) {
    if (h.size() < 4) throw std::invalid_argument("BarAndHinge: hinges need four nodes.");
    for (int k = 0; k < 4; k++)
      if (h[k] < 0 || h[k] >= nnode) throw std::invalid_argument("BarAndHinge: hinge references a node outside the coordinates.");
  }
  // computeRigidities (:150-160), computeLength0 (:213-226), computeTheta0 (:229-248)
  if (kBond_.empty() || kHinge_.empty())
    throw std::invalid_argument("BarAndHinge requires the rigidities to be set (thickness-derived rigidities are not supported)");
  if (kBond_.size() == 1) kBond_ = vector<double>(bonds_.size(), kBond_[0]);
  if (kHinge_.size() == 1) kHinge_ = vector<double>(hinges_.size(), kHinge_[0]);
  if (kBond_.size() != bonds_.size() || kHinge_.size() != hinges_.size())
    throw std::invalid_argument("BarAndHinge: rigidity arrays do not match the element lists.");
  auto X = [&](int n, int k) { return coords[3 * (size_t)n + k]; };
  if (length0_.empty()) {
    length0_.resize(bonds_.size());
    for (size_t i = 0; i < bonds_.size(); i++) {
      double s = 0;
      for (int k = 0; k < 3; k++) {
        double v = X(bonds_[i][0], k) - X(bonds_[i][1], k);
        s += v * v;
      }
      length0_[i] = std::sqrt(s);
    }
  } else if (length0_.size() == 1) {
    length0_ = vector<double>(bonds_.size(), length0_[0]);
  }
  if (theta0_.empty()) {
    theta0_.resize(hinges_.size());
    for (size_t i = 0; i < hinges_.size(); i++) {
      const auto& h = hinges_[i];
      double b1[3], b2[3], b3[3];
      for (int k = 0; k < 3; k++) {
        b1[k] = X(h[1], k) - X(h[0], k);
        b2[k] = X(h[2], k) - X(h[1], k);
        b3[k] = X(h[3], k) - X(h[2], k);
      }
      double n1[3] = {b1[1] * b2[2] - b1[2] * b2[1], b1[2] * b2[0] - b1[0] * b2[2], b1[0] * b2[1] - b1[1] * b2[0]};
      double n2[3] = {b2[1] * b3[2] - b2[2] * b3[1], b2[2] * b3[0] - b2[0] * b3[2], b2[0] * b3[1] - b2[1] * b3[0]};
      double nn1 = std::sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
      double nn2 = std::sqrt(n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2]);
      double c = (n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2]) / (nn1 * nn2);
      double n1b3 = n1[0] * b3[0] + n1[1] * b3[1] + n1[2] * b3[2];
      theta0_[i] = (n1b3 >= 0) ? std::acos(c) : 2 * pi() - std::acos(c);
    }
  } else if (theta0_.size() == 1) {
    theta0_ = vector<double>(hinges_.size(), theta0_[0]);
  }
  if (length0_.size() != bonds_.size() || theta0_.size() != hinges_.size())
    throw std::invalid_argument("BarAndHinge: rest length/angle arrays do not match the element lists.");
  if (!force_.empty() && force_.size() != 3 && force_.size() != coords.size())
    throw std::invalid_argument("BarAndHinge: force must have 3 or 3*nNode entries.");
  uploadFixed(coords.size());
  nNodeHint = nnode;
  upload();
}

void BarAndHinge::setHingeTheta0(int iHinge, double t0) {
  theta0_.at(iHinge) = t0;
  if (dev_) upload();
}

void BarAndHinge::upload() {
  const int nnode = nNodeHint;
  std::vector<char> fixed(3 * (size_t)nnode, 0);
  for (int i : fixedDofs())
    if (i >= 0 && (size_t)i < fixed.size()) fixed[i] = 1;
  auto allFixed = [&](const int* nd, int k) {
    for (int q = 0; q < k; q++)
      for (int c = 0; c < 3; c++)
        if (!fixed[3 * (size_t)nd[q] + c]) return false;
    return true;
  };
  auto dev = std::make_shared<Dev>();
  std::vector<int> bar, hin;
  std::vector<double> bk, bl, hk, ht;
  for (size_t i = 0; i < bonds_.size(); i++) {
    int nd[2] = {bonds_[i][0], bonds_[i][1]};
    if (allFixed(nd, 2)) continue;
    bar.push_back(nd[0]);
    bar.push_back(nd[1]);
    bk.push_back(kBond_[i]);
    bl.push_back(length0_[i]);
    dev->activeBars.push_back((int)i);
  }
  for (size_t i = 0; i < hinges_.size(); i++) {
    int nd[4] = {hinges_[i][0], hinges_[i][1], hinges_[i][2], hinges_[i][3]};
    if (allFixed(nd, 4)) continue;
    for (int k = 0; k < 4; k++) hin.push_back(nd[k]);
    hk.push_back(kHinge_[i]);
    ht.push_back(theta0_[i]);
    dev->activeHinges.push_back((int)i);
  }
  const int nb = (int)bk.size(), nh = (int)hk.size();
  // node -> incident element CSR, in element order
  std::vector<int> bptr(nnode + 1, 0), hptr(nnode + 1, 0);
  for (int b = 0; b < nb; b++)
    for (int l = 0; l < 2; l++) bptr[bar[2 * b + l] + 1]++;
  for (int h = 0; h < nh; h++)
    for (int l = 0; l < 4; l++) hptr[hin[4 * h + l] + 1]++;
  for (int n = 0; n < nnode; n++) {
    bptr[n + 1] += bptr[n];
    hptr[n + 1] += hptr[n];
  }
  std::vector<int> bidx(bptr[nnode]), hidx(hptr[nnode]);
  {
    std::vector<int> cb(bptr.begin(), bptr.end() - 1), ch(hptr.begin(), hptr.end() - 1);
    for (int b = 0; b < nb; b++)
      for (int l = 0; l < 2; l++) bidx[cb[bar[2 * b + l]]++] = (b << 1) | l;
    for (int h = 0; h < nh; h++)
      for (int l = 0; l < 4; l++) hidx[ch[hin[4 * h + l]]++] = (h << 2) | l;
  }
  std::vector<double> fz;
  std::vector<unsigned char> active(nnode, 0);
  if (!force_.empty()) {
    fz.resize(nnode);
    for (int n = 0; n < nnode; n++) {
      fz[n] = force_.size() == 3 ? force_[2] : force_[3 * (size_t)n + 2];
      int nd[1] = {n};
      active[n] = allFixed(nd, 1) ? 0 : 1;
    }
  }
  // one allocation, 256-byte aligned sub-buffers
  size_t off = 0;
  auto reserve = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t o_bar = reserve(sizeof(int) * bar.size()), o_bk = reserve(sizeof(double) * bk.size()),
               o_bl = reserve(sizeof(double) * bl.size()), o_hin = reserve(sizeof(int) * hin.size()),
               o_hk = reserve(sizeof(double) * hk.size()), o_ht = reserve(sizeof(double) * ht.size()),
               o_bptr = reserve(sizeof(int) * bptr.size()), o_bidx = reserve(sizeof(int) * bidx.size()),
               o_hptr = reserve(sizeof(int) * hptr.size()), o_hidx = reserve(sizeof(int) * hidx.size()),
               o_fz = reserve(sizeof(double) * fz.size()), o_act = reserve(active.size());
  dev->mem.reset(off + 256);
  char* base = static_cast<char*>(dev->mem.get());
  auto put = [&](size_t o, const void* src, size_t bytes) {
    if (bytes) MINIM_HIP_CHECK(hipMemcpy(base + o, src, bytes, hipMemcpyHostToDevice));
  };
  put(o_bar, bar.data(), sizeof(int) * bar.size());
  put(o_bk, bk.data(), sizeof(double) * bk.size());
  put(o_bl, bl.data(), sizeof(double) * bl.size());
  put(o_hin, hin.data(), sizeof(int) * hin.size());
  put(o_hk, hk.data(), sizeof(double) * hk.size());
  put(o_ht, ht.data(), sizeof(double) * ht.size());
  put(o_bptr, bptr.data(), sizeof(int) * bptr.size());
  put(o_bidx, bidx.data(), sizeof(int) * bidx.size());
  put(o_hptr, hptr.data(), sizeof(int) * hptr.size());
  put(o_hidx, hidx.data(), sizeof(int) * hidx.size());
  put(o_fz, fz.data(), sizeof(double) * fz.size());
  put(o_act, active.data(), active.size());
  auto& a = dev->args;
  a.nnode = nnode;
  a.nbar = nb;
  a.nhinge = nh;
  a.bar = (const int*)(base + o_bar);
  // parameters shared by every element go as scalars
  auto uniform = [](const std::vector<double>& v) {
    for (double x : v)
      if (!(x == v[0]) || std::signbit(x) != std::signbit(v[0])) return false;
    return true;
  };
  a.bk = uniform(bk) ? nullptr : (const double*)(base + o_bk);
  a.bl0 = uniform(bl) ? nullptr : (const double*)(base + o_bl);
  a.hinge = (const int*)(base + o_hin);
  a.hk = uniform(hk) ? nullptr : (const double*)(base + o_hk);
  a.ht0 = uniform(ht) ? nullptr : (const double*)(base + o_ht);
  a.bkU = bk.empty() ? 0.0 : bk[0];
  a.bl0U = bl.empty() ? 0.0 : bl[0];
  a.hkU = hk.empty() ? 0.0 : hk[0];
  a.ht0U = ht.empty() ? 0.0 : ht[0];
  a.bptr = (const int*)(base + o_bptr);
  a.bidx = (const int*)(base + o_bidx);
  a.hptr = (const int*)(base + o_hptr);
  a.hidx = (const int*)(base + o_hidx);
  a.fz = fz.empty() ? nullptr : (const double*)(base + o_fz);
  a.nodeActive = (const unsigned char*)(base + o_act);
  dev->ht.reset(std::max(sizeof(minim_amd::HStore<double>), sizeof(minim_amd::HStore<float>)) *
                (size_t)(nh > 0 ? nh : 1));
  dev->xtrial.reset(3 * sizeof(double) * (size_t)(nnode > 0 ? nnode : 1));
  dev_ = dev;
}

namespace {
// Energy kernel (into rs.result[0] when wantE, and to eDev when given) and
// gradient kernel (when g; the hinge terms come from the energy kernel, or
// from the hinge kernel without an energy).  prec: the coordinates and the
// arithmetic; pv: the direction d and the gradient g (FP32 against FP64
// coordinates in the mixed schedule).
void launch_bah(const BarAndHinge::Dev& dv, int prec, int pv, const void* x, const void* d, double t, bool wantE,
                void* g, minim_amd::RedScratch rs, hipStream_t s, double* eDev = nullptr, bool keep = false) {
  const auto& a = dv.args;
  if (pv != prec && !(prec == FP64 && pv == FP32))
    throw std::invalid_argument("BarAndHinge: FP32 coordinates with FP64 vectors are not supported.");
  const bool reuse = !d && !wantE && g && dv.keptAt == x && dv.keptPrec == prec;
  dv.keptAt = nullptr;
  auto go = [&](auto tag, auto gtag) {
    using T = decltype(tag);
    using TG = decltype(gtag);
    auto* ht = static_cast<minim_amd::HStore<T>*>(dv.ht.get());
    const T* xx = (const T*)x;
    if (d) {
      // a trial point: materialised once (x + t*d, as the element loads would
      // form it), so that the energy pass reads one array per node
      T* xt = static_cast<T*>(dv.xtrial.get());
      const size_t n3 = 3 * (size_t)a.nnode;
      minim_amd::device_copy(xt, x, n3 * sizeof(T), s);
      minim_amd::launch_axpy_mixed(prec, xt, pv, d, t, n3, s);
      xx = xt;
    }
    if (wantE) {
      const size_t tot = (size_t)a.nbar + a.nhinge + a.nnode;
      const int nb = minim_amd::grid_blocks(tot, 4, 2048);
      if (g || keep)
        minim_amd::bah_energy_kernel<T, true><<<nb, kBlock, 0, s>>>(a, xx, (const T*)nullptr, (T)0, rs, ht);
      else
        minim_amd::bah_energy_kernel<T, false><<<nb, kBlock, 0, s>>>(a, xx, (const T*)nullptr, (T)0, rs, ht);
      MINIM_HIP_CHECK(hipGetLastError());
      minim_amd::launch_finalize(rs, 1, nb, s, 0, eDev);
    } else if (g && a.nhinge > 0 && !reuse) {
      minim_amd::bah_hinge_kernel<T><<<minim_amd::grid_blocks(a.nhinge, 1, 1 << 20), kBlock, 0, s>>>(a, xx, ht);
      MINIM_HIP_CHECK(hipGetLastError());
    }
    if (g) {
      const int nb = minim_amd::grid_blocks(a.nnode, 1, 1 << 20);
      minim_amd::bah_gradient_kernel<T, TG><<<nb, kBlock, 0, s>>>(a, xx, ht, (TG*)g);
      MINIM_HIP_CHECK(hipGetLastError());
    }
  };
  if (prec == FP64 && pv == FP64) go(0.0, 0.0);
  else if (prec == FP64) go(0.0, 0.0f);
  else go(0.0f, 0.0f);
  if (keep && wantE && !d) {
    dv.keptAt = x;
    dv.keptPrec = prec;
  }
}
}  // namespace

void BarAndHinge::evaluate(int prec, const void* x, const void* d, double t, double* e, void* g,
                           void* stream) const {
  if (!dev_) throw std::runtime_error("BarAndHinge: not initialised (construct a State first).");
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto& ws = minim_amd::Workspace::current();
  const auto& a = dev_->args;
  if (g && d) throw std::invalid_argument("BarAndHinge: gradients at a trial point are not supported.");
  if (e || g) {
    launch_bah(*dev_, prec, prec, x, d, t, e != nullptr, g, ws.rs(), s);
    if (g) constrainDevice(prec, x, g, 3 * (size_t)a.nnode, stream);
  }
  if (e) {
    double r[1];
    ws.fetch(s, 1, r);
    *e = r[0];
  }
}

// evaluate without the host synchronisation: the energy stays on the device
// (e_dev, or the calling thread's reduction result slot 0).
void BarAndHinge::evaluateAsync(int prec, const void* x, const void* d, double t, bool wantE, double* e_dev, void* g,
                                void* stream) const {
  if (!dev_) throw std::runtime_error("BarAndHinge: not initialised (construct a State first).");
  if (g && d) throw std::invalid_argument("BarAndHinge: gradients at a trial point are not supported.");
  if (!wantE && !g) return;
  hipStream_t s = static_cast<hipStream_t>(stream);
  launch_bah(*dev_, prec, prec, x, d, t, wantE, g, minim_amd::Workspace::current().rs(), s, wantE ? e_dev : nullptr);
  if (g) constrainDevice(prec, x, g, 3 * (size_t)dev_->args.nnode, stream);
}

// The mixed schedule's second phase: FP64 coordinates and arithmetic, the
// direction and the gradient in FP32.  Only single-DOF constraints (their
// projection does not read the coordinates).
bool BarAndHinge::supportsMixed() const { return !hasNormalConstraints(); }

void BarAndHinge::evaluateMixedAsync(const void* x, const void* d, double t, bool wantE, double* e_dev, void* g,
                                     void* stream) const {
  if (!dev_) throw std::runtime_error("BarAndHinge: not initialised (construct a State first).");
  if (g && d) throw std::invalid_argument("BarAndHinge: gradients at a trial point are not supported.");
  if (!supportsMixed()) throw std::invalid_argument("BarAndHinge: no mixed evaluation with normal constraints.");
  if (!wantE && !g) return;
  hipStream_t s = static_cast<hipStream_t>(stream);
  launch_bah(*dev_, FP64, FP32, x, d, t, wantE, g, minim_amd::Workspace::current().rs(), s, wantE ? e_dev : nullptr);
  if (g) constrainDevice(FP32, nullptr, g, 3 * (size_t)dev_->args.nnode, stream);
}

// An accepted first trial then needs no second energy pass and no separate
// hinge pass for its gradient; xout holds exactly what the in-place update
// x += t*d would (the same copy-and-axpy as the trial point).  With g null
// the gradient is deferred: the energy pass keeps the hinge store for xout,
// and the gradient evaluation at xout that follows runs only the gather.
bool BarAndHinge::supportsStep() const { return dev_ != nullptr; }
bool BarAndHinge::stepDefersGradient() const { return true; }

void BarAndHinge::evaluateStepAsync(int px, int pv, const void* x, const void* d, double t, void* xout, void* g,
                                    void* stream) const {
  if (!dev_) throw std::runtime_error("BarAndHinge: not initialised (construct a State first).");
  if (pv != px && !supportsMixed())
    throw std::invalid_argument("BarAndHinge: no mixed evaluation with normal constraints.");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t n3 = 3 * (size_t)dev_->args.nnode;
  minim_amd::device_copy(xout, x, n3 * (px == FP64 ? sizeof(double) : sizeof(float)), s);
  minim_amd::launch_axpy_mixed(px, xout, pv, d, t, n3, s);
  launch_bah(*dev_, px, pv, xout, nullptr, 0.0, true, g, minim_amd::Workspace::current().rs(), s, nullptr, !g);
  if (g) constrainDevice(pv, pv == px ? xout : nullptr, g, n3, stream);
}

void BarAndHinge::evaluateMixed(const void* x, const void* d, double t, double* e, void* g, void* stream) const {
  evaluateMixedAsync(x, d, t, e != nullptr, nullptr, g, stream);
  if (e) {
    double r[1];
    minim_amd::Workspace::current().fetch(static_cast<hipStream_t>(stream), 1, r);
    *e = r[0];
  }
}

// The energy of one element type (State::componentEnergy, src/State.cpp:307-312):
// the energy pass over bars (type 0), hinges (1) or the external force on the
// nodes (2) alone, in FP64 at the host coordinates.
double BarAndHinge::componentEnergy(int component, const vector<double>& coords) const {
  if (!dev_) throw std::runtime_error("BarAndHinge: not initialised (construct a State first).");
  auto a = dev_->args;
  if ((size_t)coords.size() != 3 * (size_t)a.nnode)
    throw std::invalid_argument("BarAndHinge: coordinates have the wrong size.");
  if (component != 0) a.nbar = 0;
  if (component != 1) a.nhinge = 0;
  if (component != 2) a.fz = nullptr;
  if (component < 0 || component > 2 || (component == 2 && !dev_->args.fz)) return 0.0;
  hipStream_t s = nullptr;
  auto& ws = minim_amd::Workspace::current();
  void* xd = minim_amd::Staging::current().device(0, coords.size() * sizeof(double));
  minim_amd::upload_raw(coords.data(), coords.size(), FP64, xd, s);
  const size_t tot = (size_t)a.nbar + a.nhinge + a.nnode;
  const int nb = minim_amd::grid_blocks(tot, 4, 2048);
  minim_amd::bah_energy_kernel<double, false>
      <<<nb, kBlock, 0, s>>>(a, (const double*)xd, (const double*)nullptr, 0.0, ws.rs(),
                             (minim_amd::HStore<double>*)nullptr);
  MINIM_HIP_CHECK(hipGetLastError());
  minim_amd::launch_finalize(ws.rs(), 1, nb, s);
  double r[1];
  ws.fetch(s, 1, r);
  return r[0];
}

// Average duration (ms) of reps evaluations on the device (energy kernel with
// its reduction, gradient kernel; no constraints, no host copy).
double BarAndHinge::timeKernel(int prec, const void* x, bool wantE, void* g, int reps, void* stream) const {
  if (!dev_) throw std::runtime_error("BarAndHinge: not initialised (construct a State first).");
  if ((!wantE && !g) || reps < 1) throw std::invalid_argument("BarAndHinge: nothing to time");
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto& ws = minim_amd::Workspace::current();
  launch_bah(*dev_, prec, prec, x, nullptr, 0.0, wantE, g, ws.rs(), s);
  hipEvent_t e0, e1;
  MINIM_HIP_CHECK(hipEventCreate(&e0));
  MINIM_HIP_CHECK(hipEventCreate(&e1));
  MINIM_HIP_CHECK(hipEventRecord(e0, s));
  for (int i = 0; i < reps; i++) launch_bah(*dev_, prec, prec, x, nullptr, 0.0, wantE, g, ws.rs(), s);
  MINIM_HIP_CHECK(hipEventRecord(e1, s));
  MINIM_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  MINIM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  MINIM_HIP_CHECK(hipEventDestroy(e0));
  MINIM_HIP_CHECK(hipEventDestroy(e1));
  return (double)ms / reps;
}

}  // namespace minim
