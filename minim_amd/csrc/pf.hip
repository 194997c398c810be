// pf.hip -- PhaseField energy/gradient stencil for MI355X (gfx950).
//
// Replaces PhaseField::energyGradient (src/potentials/PhaseField.cpp:662-703),
// fluidEnergy (:365-393), phaseGradient (:238-288, "no solid" branch),
// densityConstraintEnergy (:457-473) and the gradient projection of
// specialisedConstraints (:580-612) that State::basicEG applies after every
// gradient (src/State.cpp:48-56).
//
// The reference scatters +-factor*(c_i - c_j) into three gradient entries per
// site, fluid and direction with OpenMP atomics.  With a uniform node volume the
// scatter sums to the atomic-free gather form
//     g_i = bulk'(c_i) + 2 f * sum_d (2 c_i - c_{i-e_d} - c_{i+e_d})
//     E   = sum_i [ bulk(c_i) + f * sum_d 0.5((c_i-c_{i-e_d})^2 + (c_i-c_{i+e_d})^2) ]
// so every site is produced by exactly one thread and written once.
//
// Work decomposition.  The grid is canonicalised (size-1 dimensions moved to
// the front, which keeps the linear index) to X x Y x Z with Z contiguous.  A
// block owns a TZ x TY tile of the (y,z) plane (64 x 4 when Z >= 64, so that
// y neighbours are mostly loaded by the same block) and marches along a chunk
// of x planes; each thread owns one (y,z) column.  The march advances U planes
// per step: all U new column values and the 4U y/z neighbours are issued
// before anything is waited on, so every thread keeps U sites of HBM traffic
// in flight.  x neighbours stay in registers.  A dimension of size 1 needs no
// branch: its two "neighbours" are the site itself and contribute exact zeros,
// as the reference's skip of such dimensions does (PhaseField.cpp:245).
// Blocks are ordered x-chunk major and spread over the 8 XCDs in contiguous
// ranges, so all tiles of an x slab (which share halo rows) and consecutive
// slabs (which share one halo plane) meet in the same L2.
#include "common.hpp"
#include "pf.hpp"

namespace minim_amd {

namespace {

template <int NF>
struct NFM {
  static constexpr int value = NF > 0 ? NF : kPFMaxFluid;
};

// T: precision of the coordinates and of the arithmetic; TV: precision of the
// direction and gradient vectors (T = double, TV = float in the mixed path).
template <typename T, typename TV, int NF, bool DIR>
__device__ inline void load_site(const T* __restrict__ x, const TV* __restrict__ d, T t, size_t site, int nf,
                                 T (&v)[NFM<NF>::value]) {
  const size_t base = site * (size_t)nf;
  const int NK = NF > 0 ? NF : nf;
#pragma unroll
  for (int k = 0; k < NK; k++) {
    T c = x[base + k];
    if (DIR) {
      T s = t * (T)d[base + k];  // trial point x + t*d, rounded as x + (t*d)
      c = c + s;
    }
    v[k] = c;
  }
}

// Energy and gradient of one site from its own value c, x neighbours (cm, cp)
// and y/z neighbours.  Returns the site energy in FP64; writes gv.
template <typename T, int NF, bool WE, bool WG>
__device__ inline double site_eg(const PFArgs& a, int nf, const T (&fb)[NFM<NF>::value],
                                 const T (&fg)[NFM<NF>::value], T coef, T inv_nvar, const T* c, const T* cm,
                                 const T* cp, const T* ym, const T* yp, const T* zm, const T* zp,
                                 T (&gv)[NFM<NF>::value]) {
  const int NK = NF > 0 ? NF : nf;
  T rho = (T)-1;
  double esite = 0.0;
#pragma unroll
  for (int k = 0; k < NK; k++) {
    const T ck = c[k];
    T bulk_e, bulk_g;
    if (NF == 1 || (NF == 0 && nf == 1)) {
      // (kappa/16) v (c+1)^2 (c-1)^2 and its derivative (PhaseField.cpp:373-378)
      const T a1 = ck + (T)1, a2 = ck - (T)1;
      bulk_e = fb[k] * (a1 * a1) * (a2 * a2);
      bulk_g = fb[k] * (T)4 * ck * (ck * ck - (T)1);
    } else {
      // 0.5 kappa v c^2 (c-1)^2 and its derivative (PhaseField.cpp:380-385)
      const T a2 = ck - (T)1;
      bulk_e = fb[k] * (ck * ck) * (a2 * a2);
      bulk_g = fb[k] * (T)2 * ck * a2 * ((T)2 * ck - (T)1);
    }
    const T gxm = ck - cm[k], gxp = ck - cp[k];
    const T gym = ck - ym[k], gyp = ck - yp[k];
    const T gzm = ck - zm[k], gzp = ck - zp[k];
    if (WE) {
      T grad2 = (T)0.5 * (gxm * gxm + gxp * gxp);
      grad2 += (T)0.5 * (gym * gym + gyp * gyp);
      grad2 += (T)0.5 * (gzm * gzm + gzp * gzp);
      esite += (double)bulk_e + (double)(fg[k] * grad2);
    }
    if (WG) {
      T lap = gxm + gxp;
      lap += gym + gyp;
      lap += gzm + gzp;
      gv[k] = bulk_g + (T)2 * fg[k] * lap;
    }
    rho += ck;
  }
  if (a.density == kDensitySoft && nf > 1) {
    // densityConstraintEnergy (PhaseField.cpp:457-473)
    if (WE) esite += (double)(coef * (rho * rho));
    if (WG) {
      const T gr = (T)2 * coef * rho;
#pragma unroll
      for (int k = 0; k < NK; k++) gv[k] += gr;
    }
  }
  if (WG) {
    // specialisedConstraints: fixed fluids, then the hard density projection
    if (a.fixmask) {
#pragma unroll
      for (int k = 0; k < NK; k++)
        if (a.fixmask & (1u << k)) gv[k] = (T)0;
    }
    if (a.density == kDensityHard && nf > 1 && a.nvar > 0) {
      T comp = (T)0;
#pragma unroll
      for (int k = 0; k < NK; k++)
        if (!(a.fixmask & (1u << k))) comp += gv[k];
      comp *= inv_nvar;
#pragma unroll
      for (int k = 0; k < NK; k++)
        if (!(a.fixmask & (1u << k))) gv[k] -= comp;
    }
  }
  return esite;
}

// MODE bits: 1 = energy, 2 = gradient (+ g.g), 4 = trial direction.
// U = planes per march step.
template <typename T, typename TV, int NF, int MODE, int U>
__global__ __launch_bounds__(kBlock) void pf_kernel(PFArgs a, const T* __restrict__ x,
                                                    const TV* __restrict__ d, T t, TV* __restrict__ g,
                                                    RedScratch rs) {
  constexpr bool WE = MODE & 1, WG = MODE & 2, DIR = MODE & 4;
  constexpr int M = NFM<NF>::value;
  const int nf = NF > 0 ? NF : a.nf;
  const int NK = NF > 0 ? NF : nf;

  const int L = xcd_logical_block(blockIdx.x, gridDim.x);
  const int xc = L / a.ntile, tile = L - xc * a.ntile;
  const int ty_tile = tile / a.ntz, tz_tile = tile - ty_tile * a.ntz;
  const int ly = threadIdx.x / a.TZ, lz = threadIdx.x - ly * a.TZ;
  const int y = ty_tile * a.TY + ly, z = tz_tile * a.TZ + lz;

  double acc[2] = {0.0, 0.0};  // energy, g.g

  if (ly < a.TY && y < a.Y && z < a.Z) {
    T fb[M], fg[M];
#pragma unroll
    for (int k = 0; k < NK; k++) {
      fb[k] = (T)a.fb[k];
      fg[k] = (T)a.fg[k];
    }
    const T coef = (T)a.coef;
    const T inv_nvar = (T)a.inv_nvar;
    const int p = y * a.Z + z;
    const int pzm = (z == 0) ? p + a.Z - 1 : p - 1;
    const int pzp = (z == a.Z - 1) ? p - (a.Z - 1) : p + 1;
    const int pym = (y == 0) ? p + (a.Y - 1) * a.Z : p - a.Z;
    const int pyp = (y == a.Y - 1) ? p - (a.Y - 1) * a.Z : p + a.Z;
    const size_t YZ = (size_t)a.YZ;
    const int X = a.X;
    const int x0 = xc * a.xchunk;
    const int x1 = min(x0 + a.xchunk, X);
    auto wrapx = [X](int v) { return v >= X ? v - X : (v < 0 ? v + X : v); };

    // col[0] = plane xi-1, col[1] = plane xi, col[2..U+1] = planes xi+1..xi+U
    T col[U + 2][M];
    load_site<T, TV, NF, DIR>(x, d, t, (size_t)wrapx(x0 - 1) * YZ + p, nf, col[0]);
    load_site<T, TV, NF, DIR>(x, d, t, (size_t)x0 * YZ + p, nf, col[1]);

    int xi = x0;
    for (; xi + U <= x1; xi += U) {
      T ym[U][M], yp[U][M], zm[U][M], zp[U][M];
#pragma unroll
      for (int u = 0; u < U; u++) load_site<T, TV, NF, DIR>(x, d, t, (size_t)wrapx(xi + 1 + u) * YZ + p, nf, col[2 + u]);
#pragma unroll
      for (int u = 0; u < U; u++) {
        const size_t plane = (size_t)(xi + u) * YZ;
        load_site<T, TV, NF, DIR>(x, d, t, plane + pym, nf, ym[u]);
        load_site<T, TV, NF, DIR>(x, d, t, plane + pyp, nf, yp[u]);
        load_site<T, TV, NF, DIR>(x, d, t, plane + pzm, nf, zm[u]);
        load_site<T, TV, NF, DIR>(x, d, t, plane + pzp, nf, zp[u]);
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        T gv[M];
        const double es = site_eg<T, NF, WE, WG>(a, nf, fb, fg, coef, inv_nvar, col[1 + u], col[u], col[2 + u],
                                                 ym[u], yp[u], zm[u], zp[u], gv);
        if (WE) acc[0] += es;
        if (WG) {
          const size_t base = ((size_t)(xi + u) * YZ + p) * (size_t)nf;
#pragma unroll
          for (int k = 0; k < NK; k++) {
            const TV gs = (TV)gv[k];
            g[base + k] = gs;
            acc[1] += (double)gs * (double)gs;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < NK; k++) {
        col[0][k] = col[U][k];
        col[1][k] = col[U + 1][k];
      }
    }
    // remaining planes one at a time
    for (; xi < x1; xi++) {
      T ym[M], yp[M], zm[M], zp[M];
      load_site<T, TV, NF, DIR>(x, d, t, (size_t)wrapx(xi + 1) * YZ + p, nf, col[2]);
      const size_t plane = (size_t)xi * YZ;
      load_site<T, TV, NF, DIR>(x, d, t, plane + pym, nf, ym);
      load_site<T, TV, NF, DIR>(x, d, t, plane + pyp, nf, yp);
      load_site<T, TV, NF, DIR>(x, d, t, plane + pzm, nf, zm);
      load_site<T, TV, NF, DIR>(x, d, t, plane + pzp, nf, zp);
      T gv[M];
      const double es =
          site_eg<T, NF, WE, WG>(a, nf, fb, fg, coef, inv_nvar, col[1], col[0], col[2], ym, yp, zm, zp, gv);
      if (WE) acc[0] += es;
      if (WG) {
        const size_t base = (plane + p) * (size_t)nf;
#pragma unroll
        for (int k = 0; k < NK; k++) {
          const TV gs = (TV)gv[k];
          g[base + k] = gs;
          acc[1] += (double)gs * (double)gs;
        }
      }
#pragma unroll
      for (int k = 0; k < NK; k++) {
        col[0][k] = col[1][k];
        col[1][k] = col[2][k];
      }
    }
  }
  grid_sum_store<2>(acc, rs);
}

// Hard density projection / fixed-fluid zeroing of an arbitrary vector
// (Potential::applyConstraints -> PhaseField::specialisedConstraints).
template <typename T, int NF>
__global__ __launch_bounds__(kBlock) void pf_constrain_kernel(PFArgs a, T* __restrict__ v, size_t nsite) {
  constexpr int M = NFM<NF>::value;
  const int nf = NF > 0 ? NF : a.nf;
  const int NK = NF > 0 ? NF : nf;
  const T inv_nvar = (T)a.inv_nvar;
  for (size_t s = (size_t)blockIdx.x * kBlock + threadIdx.x; s < nsite; s += (size_t)gridDim.x * kBlock) {
    T gv[M];
    const size_t base = s * nf;
#pragma unroll
    for (int k = 0; k < NK; k++) {
      gv[k] = v[base + k];
      if (a.fixmask & (1u << k)) gv[k] = (T)0;
    }
    if (a.density == kDensityHard && nf > 1 && a.nvar > 0) {
      T comp = (T)0;
#pragma unroll
      for (int k = 0; k < NK; k++)
        if (!(a.fixmask & (1u << k))) comp += gv[k];
      comp *= inv_nvar;
#pragma unroll
      for (int k = 0; k < NK; k++)
        if (!(a.fixmask & (1u << k))) gv[k] -= comp;
    }
#pragma unroll
    for (int k = 0; k < NK; k++) v[base + k] = gv[k];
  }
}

// Planes per march step: enough HBM loads in flight per thread while keeping
// the register file within reach of 4+ waves per SIMD.
template <typename T, int NF>
constexpr int march_unroll() {
  return NF == 0 ? 1 : (sizeof(T) * NF <= 12 ? 4 : (sizeof(T) * NF <= 24 ? 2 : 1));
}

template <typename T, typename TV, int NF>
void launch_eval_nf(const PFArgs& a, int mode, const void* x, const void* d, double t, void* g,
                    RedScratch rs, hipStream_t s, int nblocks) {
  constexpr int U = march_unroll<T, NF>();
  const T* xx = static_cast<const T*>(x);
  const TV* dd = static_cast<const TV*>(d);
  TV* gg = static_cast<TV*>(g);
  const T tt = (T)t;
  switch (mode) {
    case 1: pf_kernel<T, TV, NF, 1, U><<<nblocks, kBlock, 0, s>>>(a, xx, dd, tt, gg, rs); break;
    case 5: pf_kernel<T, TV, NF, 5, U><<<nblocks, kBlock, 0, s>>>(a, xx, dd, tt, gg, rs); break;
    case 2: pf_kernel<T, TV, NF, 2, U><<<nblocks, kBlock, 0, s>>>(a, xx, dd, tt, gg, rs); break;
    case 3: pf_kernel<T, TV, NF, 3, U><<<nblocks, kBlock, 0, s>>>(a, xx, dd, tt, gg, rs); break;
    default: throw std::invalid_argument("pf: unsupported evaluation mode");
  }
}

template <typename T, typename TV>
void launch_eval_t(const PFArgs& a, int mode, const void* x, const void* d, double t, void* g,
                   RedScratch rs, hipStream_t s, int nblocks) {
  switch (a.nf) {
    case 1: launch_eval_nf<T, TV, 1>(a, mode, x, d, t, g, rs, s, nblocks); break;
    case 2: launch_eval_nf<T, TV, 2>(a, mode, x, d, t, g, rs, s, nblocks); break;
    case 3: launch_eval_nf<T, TV, 3>(a, mode, x, d, t, g, rs, s, nblocks); break;
    case 4: launch_eval_nf<T, TV, 4>(a, mode, x, d, t, g, rs, s, nblocks); break;
    default:
      if (sizeof(T) != sizeof(TV)) throw std::invalid_argument("PhaseField: mixed precision needs nFluid <= 4");
      launch_eval_nf<T, T, 0>(a, mode, x, d, t, g, rs, s, nblocks);
      break;
  }
}

template <typename T>
void launch_constrain_t(const PFArgs& a, void* v, size_t nsite, hipStream_t s) {
  const int nb = grid_blocks(nsite, 4, 4096);
  T* vv = static_cast<T*>(v);
  switch (a.nf) {
    case 1: pf_constrain_kernel<T, 1><<<nb, kBlock, 0, s>>>(a, vv, nsite); break;
    case 2: pf_constrain_kernel<T, 2><<<nb, kBlock, 0, s>>>(a, vv, nsite); break;
    case 3: pf_constrain_kernel<T, 3><<<nb, kBlock, 0, s>>>(a, vv, nsite); break;
    case 4: pf_constrain_kernel<T, 4><<<nb, kBlock, 0, s>>>(a, vv, nsite); break;
    default: pf_constrain_kernel<T, 0><<<nb, kBlock, 0, s>>>(a, vv, nsite); break;
  }
}

}  // namespace

void pf_args_finalize(PFArgs& a, int target_blocks) {
  a.YZ = a.Y * a.Z;
  a.actX = a.X > 1;
  a.actY = a.Y > 1;
  a.actZ = a.Z > 1;
  // (y,z) tile of one block: 64 z x 4 y when rows are long enough, else whole
  // rows stacked up to the block size
  a.TZ = a.Z >= 64 ? 64 : a.Z;
  a.TY = kBlock / a.TZ;
  if (a.TY > a.Y) a.TY = a.Y;
  a.ntz = (a.Z + a.TZ - 1) / a.TZ;
  a.nty = (a.Y + a.TY - 1) / a.TY;
  a.ntile = a.ntz * a.nty;
  int xchunk = a.xchunk_override > 0 ? a.xchunk_override : 1;
  if (a.xchunk_override <= 0) {
    // Longest x-march that still gives ~target_blocks blocks (a chunk re-reads
    // two halo planes, so longer chunks cost less traffic).
    const long want = target_blocks > 0 ? target_blocks : 2048;
    long nxc = (want + a.ntile - 1) / a.ntile;
    if (nxc < 1) nxc = 1;
    if (nxc > a.X) nxc = a.X;
    xchunk = (int)((a.X + nxc - 1) / nxc);
  }
  if (xchunk < 1) xchunk = 1;
  if (xchunk > a.X) xchunk = a.X;
  a.xchunk = xchunk;
  a.nxc = (a.X + xchunk - 1) / xchunk;
  while ((long)a.ntile * a.nxc > kMaxRedBlocks) {
    a.xchunk *= 2;
    a.nxc = (a.X + a.xchunk - 1) / a.xchunk;
  }
  if ((long)a.ntile * a.nxc > kMaxRedBlocks) throw std::invalid_argument("PhaseField: grid too large for one launch");
  a.nblocks = a.ntile * a.nxc;
}

void pf_launch_eval(const PFArgs& a, int prec, int pv, int mode, const void* x, const void* d, double t,
                    void* g, RedScratch rs, hipStream_t s) {
  if (prec == F64 && pv == F64)
    launch_eval_t<double, double>(a, mode, x, d, t, g, rs, s, a.nblocks);
  else if (prec == F32 && pv == F32)
    launch_eval_t<float, float>(a, mode, x, d, t, g, rs, s, a.nblocks);
  else if (prec == F64 && pv == F32)
    launch_eval_t<double, float>(a, mode, x, d, t, g, rs, s, a.nblocks);
  else
    throw std::invalid_argument("PhaseField: FP32 coordinates with FP64 vectors are not supported");
  MINIM_HIP_CHECK(hipGetLastError());
  launch_finalize(rs, 2, a.nblocks, s);
}

void pf_launch_constrain(const PFArgs& a, int prec, void* v, size_t nsite, hipStream_t s) {
  if (prec == F64)
    launch_constrain_t<double>(a, v, nsite, s);
  else
    launch_constrain_t<float>(a, v, nsite, s);
  MINIM_HIP_CHECK(hipGetLastError());
}

}  // namespace minim_amd
