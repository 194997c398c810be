// vec.hpp -- launch interface of the fused vector kernels (vec.hip) used by the
// device-resident minimisers.
#pragma once

#include "common.hpp"

namespace minim_amd {

constexpr int kMaxBasis = 40;  // 2m+1 basis vectors for m <= 15 (+ spare)
constexpr int kMaxM = 15;

// Per-site projection applied to a freshly built direction: fixed components
// zeroed, then the mean over the variable components removed (PhaseField hard
// density constraint).  nf == 1 and fixmask == 0 means no projection.
struct SiteProj {
  int nf;
  unsigned fixmask;
  int hard;
  int nvar;
  double inv_nvar;
};

// out = sum_j coef[j] * basis[j]  (FP64 per element, stored as T), projected,
// then rs.result[0] = out.out and rs.result[1+j] = out.dotv[j] (j < ndot <= 31).
struct DirArgs {
  const void* basis[kMaxBasis];
  double coef[kMaxBasis];
  int nb;
  const void* dotv[kMaxBasis];
  int ndot;
  void* out;
  size_t n;
  SiteProj proj;
  // >= 0 when the arguments follow the L-BFGS layout with P = pairs stored:
  // basis = [g, d_0..d_{P-1}, y_0..y_{P-1}], dotv = [g, y_0..y_{P-1}]; selects
  // kernels whose loads are all unconditional (compile-time P).
  int pairs = -1;
  int blocks = 0;  // launch size (an autotuned shape); 0 = the heuristic
};
constexpr int kFastPairs = 10;  // largest P with a compile-time kernel
void launch_direction(const DirArgs& a, int prec, RedScratch rs, hipStream_t s);

// y = gnew - gold, then dots written to rs.result with KB = gram_bucket(max(nd, ny)):
//   [0,KB) gnew.D[j] ; [KB,2KB) gnew.Y[j] ; 2KB gnew.gnew ; 2KB+1 gnew.y ;
//   [2KB+2,3KB+2) y.D[j] ; [3KB+2,4KB+2) y.Y[j] ; 4KB+2 y.y
struct GramArgs {
  const void* gnew;
  const void* gold;
  void* y;
  const void* D[kMaxBasis];
  int nd;
  const void* Y[kMaxBasis];
  int ny;
  size_t n;
  int blocks = 0;  // launch size (an autotuned shape); 0 = the heuristic
};
int gram_bucket(int k);
// KB of the result layout above for a launch with these counts
int gram_kb(int nd, int ny);
void launch_gram(const GramArgs& a, int prec, RedScratch rs, hipStream_t s);

// dots a[j].b[j] for j < k -> rs.result[0..k)
void launch_dots(int prec, const void* const* a, const void* const* b, int k, size_t n, RedScratch rs,
                 hipStream_t s);

// x = x + t*d (element-wise, product rounded first)
void launch_axpy(int prec, void* x, const void* d, double t, size_t n, hipStream_t s);
// as launch_axpy with x in precision px and d in pv (pv == px, or FP32 d
// widened exactly into FP64 x)
void launch_axpy_mixed(int px, void* x, int pv, const void* d, double t, size_t n, hipStream_t s);

// dst = src converted between precisions (prec_dst, prec_src)
void launch_convert(int prec_dst, void* dst, int prec_src, const void* src, size_t n, hipStream_t s);

// FIRE velocity update (Fire.cpp:50-57) fused with step = dt*v (:70):
//  mode 1: v = (f1*v) - (f2*g) ; mode 0: v = f2*g ;  then step = dt*v
void launch_fire_update(int prec, int mode, void* v, const void* g, void* step, double f1, double f2,
                        double dt, size_t n, hipStream_t s);

// x = x + step (element-wise); used where the step vector is explicit
void launch_add(int prec, void* x, const void* step, size_t n, hipStream_t s);

// out = f * in (element-wise)
void launch_scale_copy(int prec, void* out, const void* in, double f, size_t n, hipStream_t s);

// v *= f (element-wise)
void launch_scale(int prec, void* v, double f, size_t n, hipStream_t s);

// rs.result[0] = the largest Euclidean norm over groups of dim consecutive
// entries of v (NaN if any entry is NaN)
void launch_group_norm_max(int prec, const void* v, size_t n, int dim, RedScratch rs, hipStream_t s);

// zero the entries listed in idx (single-DOF constraints, Potential.cpp:104)
void launch_zero_idx(int prec, void* v, const int* idx, size_t nidx, hipStream_t s);

}  // namespace minim_amd
