// pf.hip -- PhaseField energy/gradient stencil for MI355X (gfx950).
//
// Replaces PhaseField::energyGradient (src/potentials/PhaseField.cpp:662-703),
// fluidEnergy (:365-393), phaseGradient (:238-288, "no solid" branch),
// densityConstraintEnergy (:457-473) and the gradient projection of
// specialisedConstraints (:580-612) that State::basicEG applies after every
// gradient (src/State.cpp:48-56).
//
// The reference scatters +-factor*(c_i - c_j) into three gradient entries per
// site, fluid and direction with OpenMP atomics.  With a uniform node volume
// the scatter sums to the atomic-free gather form
//     g_i = bulk'(c_i) + 2 f * sum_d (2 c_i - c_{i-e_d} - c_{i+e_d})
// and the gradient energy, 0.5 f sum_i sum_d ((c_i-c_{i-e_d})^2 + (c_i-c_{i+e_d})^2),
// counts every bond twice, so it equals f * sum_i sum_d (c_i - c_{i+e_d})^2:
// each site adds only its three forward bonds.  Every site is produced by
// exactly one thread and written once.  Dimensions of size 1 need no branch:
// both "neighbours" are the site itself and contribute exact zeros, as the
// reference's skip of such dimensions does (PhaseField.cpp:245).
//
// Tiled kernel (nFluid <= 4), see pf_tile.hpp for the decomposition.  Each
// block streams the x planes of a 64 x 4R (z, y) tile through a two-plane LDS
// ring: while plane i is computed from LDS (y/z neighbours) and registers
// (x neighbours, own column), the loads of plane i+2 are in flight; one
// LDS-only barrier per plane (no wait on outstanding global loads or stores)
// separates the reads of a buffer from its refill.  Loads and stores are one
// global_load/store_dwordxN per site at a 32-bit offset from the plane base.
// FP32 gradients are written with streaming (nt) stores.  The block order is
// x-chunk major and the blocks of a chunk are kept on one XCD, so the halo
// rows shared by neighbouring tiles meet in one L2.
//
// Generic kernel (nFluid > 4): one thread per (y, z) column, marching x with
// the y/z neighbours read through the L1.
#include "common.hpp"
#include "pf.hpp"
#include "pf_tile.hpp"

namespace minim_amd {

namespace {

// Barrier that orders LDS traffic only: waits for this wave's LDS operations,
// not for its outstanding global loads and stores.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

template <typename T>
__device__ __forceinline__ T fmad(T a, T b, T c) {
  return fma(a, b, c);
}

enum { DM_NONE = 0, DM_HARD = 1, DM_SOFT = 2 };

// ---------------------------------------------------------------- tiled kernel
template <typename T, int NF>
struct SiteV {
  T v[NF];
};

// One site at a 32-bit byte offset from a uniform base (one dwordxN access).
template <typename T, int NF>
__device__ __forceinline__ void ld_site(const T* __restrict__ base, unsigned boff, T (&v)[NF]) {
  const SiteV<T, NF> s = *reinterpret_cast<const SiteV<T, NF>*>(reinterpret_cast<const char*>(base) + boff);
#pragma unroll
  for (int k = 0; k < NF; k++) v[k] = s.v[k];
}

// Streaming (nt) stores: the gradient is not re-read by this kernel (FP32
// 4.2 -> 6.9 TB/s in a copy test; FP64 E+G 174 -> 169 us at 256^3 x 3).
template <typename T, int NF>
__device__ __forceinline__ void st_site(T* __restrict__ base, unsigned boff, const T (&v)[NF]) {
  T* p = reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff);
#pragma unroll
  for (int k = 0; k < NF; k++) __builtin_nontemporal_store(v[k], p + k);
}

// Per-precision coefficient tables of PFArgs.
template <typename T>
struct PFC;
template <>
struct PFC<double> {
  __device__ static double fb(const PFArgs& a, int k) { return a.fb[k]; }
  __device__ static double fbg(const PFArgs& a, int k) { return a.fbg[k]; }
  __device__ static double fg(const PFArgs& a, int k) { return a.fg[k]; }
  __device__ static double fg2(const PFArgs& a, int k) { return a.fg2[k]; }
  __device__ static double coef(const PFArgs& a) { return a.coef; }
  __device__ static double coef2(const PFArgs& a) { return a.coef2; }
  __device__ static double inv_nvar(const PFArgs& a) { return a.inv_nvar; }
};
template <>
struct PFC<float> {
  __device__ static float fb(const PFArgs& a, int k) { return a.fb_f[k]; }
  __device__ static float fbg(const PFArgs& a, int k) { return a.fbg_f[k]; }
  __device__ static float fg(const PFArgs& a, int k) { return a.fg_f[k]; }
  __device__ static float fg2(const PFArgs& a, int k) { return a.fg2_f[k]; }
  __device__ static float coef(const PFArgs& a) { return a.coef_f; }
  __device__ static float coef2(const PFArgs& a) { return a.coef2_f; }
  __device__ static float inv_nvar(const PFArgs& a) { return a.inv_nvar_f; }
};

// Energy (forward bonds) and constrained gradient of one site.
template <typename T, int NF, int DM, bool WE, bool WG>
__device__ __forceinline__ void site_tile(const PFArgs& a, const T (&c)[NF], const T (&xm)[NF], const T (&xp)[NF],
                                          const T (&ym)[NF], const T (&yp)[NF], const T (&zm)[NF],
                                          const T (&zp)[NF], T& e, T (&g)[NF]) {
  using C = PFC<T>;
  T rho = (T)-1;
  e = (T)0;
#pragma unroll
  for (int k = 0; k < NF; k++) {
    const T ck = c[k];
    const T dxm = ck - xm[k], dxp = ck - xp[k];
    const T dym = ck - ym[k], dyp = ck - yp[k];
    const T dzm = ck - zm[k], dzp = ck - zp[k];
    const T a2 = ck - (T)1;
    if (WE) {
      // bulk: (kappa/16) v (c+1)^2 (c-1)^2 or 0.5 kappa v c^2 (c-1)^2 (PhaseField.cpp:373-385)
      T be;
      if (NF == 1) {
        const T a1 = ck + (T)1;
        be = C::fb(a, k) * ((a1 * a1) * (a2 * a2));
      } else {
        be = C::fb(a, k) * ((ck * ck) * (a2 * a2));
      }
      T sq = dxp * dxp;
      sq = fmad(dyp, dyp, sq);
      sq = fmad(dzp, dzp, sq);
      e = e + be;
      e = fmad(C::fg(a, k), sq, e);
    }
    if (WG) {
      const T lap = ((dxm + dxp) + (dym + dyp)) + (dzm + dzp);
      T bg;
      if (NF == 1) bg = C::fbg(a, k) * (ck * fmad(ck, ck, (T)-1));
      else bg = C::fbg(a, k) * ((ck * a2) * fmad((T)2, ck, (T)-1));
      g[k] = fmad(C::fg2(a, k), lap, bg);
    }
    if (DM == DM_SOFT) rho = rho + ck;
  }
  if (DM == DM_SOFT) {
    // densityConstraintEnergy (PhaseField.cpp:457-473)
    if (WE) e = fmad(C::coef(a), rho * rho, e);
    if (WG) {
      const T gr = C::coef2(a) * rho;
#pragma unroll
      for (int k = 0; k < NF; k++) g[k] = g[k] + gr;
    }
  }
  if (WG) {
    // specialisedConstraints: fixed fluids zeroed, then the hard density projection
    if (a.fixmask) {
#pragma unroll
      for (int k = 0; k < NF; k++)
        if (a.fixmask & (1u << k)) g[k] = (T)0;
    }
    if (DM == DM_HARD) {
      T comp = g[0];
#pragma unroll
      for (int k = 1; k < NF; k++) comp = comp + g[k];  // fixed entries are zero
      comp = comp * C::inv_nvar(a);
#pragma unroll
      for (int k = 0; k < NF; k++)
        if (!(a.fixmask & (1u << k))) g[k] = g[k] - comp;
    }
  }
}

// Plane data of the R own rows and the halo site, as loaded (x, and d for a
// trial point).
template <typename T, typename TV, int NF, int R, bool DIR>
struct PlaneIn {
  T x[R][NF];
  T hx[NF];
  TV d[DIR ? R : 1][NF];
  TV hd[NF];
};

template <typename T, typename TV, int NF, int R, bool DIR>
__device__ __forceinline__ void load_plane(const T* __restrict__ xp, const TV* __restrict__ dp,
                                           const unsigned (&ox)[R], const unsigned (&od)[R], unsigned hox,
                                           unsigned hod, PlaneIn<T, TV, NF, R, DIR>& p) {
#pragma unroll
  for (int r = 0; r < R; r++) ld_site<T, NF>(xp, ox[r], p.x[r]);
  ld_site<T, NF>(xp, hox, p.hx);
  if constexpr (DIR) {
#pragma unroll
    for (int r = 0; r < R; r++) ld_site<TV, NF>(dp, od[r], p.d[r]);
    ld_site<TV, NF>(dp, hod, p.hd);
  }
}

// Coordinates of a loaded plane: x, or the trial point x + (t*d).
template <typename T, typename TV, int NF, int R, bool DIR>
__device__ __forceinline__ void resolve_plane(PlaneIn<T, TV, NF, R, DIR>& p, T t) {
  if constexpr (DIR) {
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int k = 0; k < NF; k++) {
        const T s = t * (T)p.d[r][k];
        p.x[r][k] = p.x[r][k] + s;
      }
#pragma unroll
    for (int k = 0; k < NF; k++) {
      const T s = t * (T)p.hd[k];
      p.hx[k] = p.hx[k] + s;
    }
  }
}

// MODE bits: 1 = energy, 2 = gradient (+ g.g), 4 = trial point x + t*d,
// 8 = the trial point also written to xo (an accepted line-search step and
// the gradient there in one pass).
template <typename T, typename TV, int NF, int DM, int MODE, int R>
__global__ __launch_bounds__(kBlock) void pf_tile_kernel(PFArgs a, TileGeom G, const T* __restrict__ x,
                                                         const TV* __restrict__ d, T t, TV* __restrict__ g,
                                                         RedScratch rs, T* __restrict__ xo) {
  constexpr bool WE = MODE & 1, WG = MODE & 2, DIR = MODE & 4, XW = (MODE & 8) && DIR;
  constexpr int LH = kTileWaves * R + 2, LW = kTileZ + 2;
  __shared__ T lds[2][LH][LW][NF];

  TileThread th;
  tile_thread(G, xcd_logical_block(blockIdx.x, gridDim.x), threadIdx.x, th);
  const size_t plane = (size_t)G.Y * G.Z * NF;  // elements per x plane
  const int X = G.X;
  unsigned ox[R], od[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    ox[r] = th.site[r] * (unsigned)(NF * sizeof(T));
    od[r] = th.site[r] * (unsigned)(NF * sizeof(TV));
  }
  const unsigned hox = th.hsite * (unsigned)(NF * sizeof(T));
  const unsigned hod = th.hsite * (unsigned)(NF * sizeof(TV));
  const bool halo = th.hrow >= 0;
  const int x0 = th.x0, x1 = th.x1;

  T A[R][NF], B[R][NF], Cn[R][NF];  // planes i-1, i, i+1 of the own rows
  {
    PlaneIn<T, TV, NF, R, DIR> pm, p0, p1;
    const size_t sm = (size_t)tile_plane(x0 - 1, X) * plane, s0 = (size_t)tile_plane(x0, X) * plane,
                 s1 = (size_t)tile_plane(x0 + 1, X) * plane;
    load_plane<T, TV, NF, R, DIR>(x + sm, d + sm, ox, od, hox, hod, pm);
    load_plane<T, TV, NF, R, DIR>(x + s0, d + s0, ox, od, hox, hod, p0);
    load_plane<T, TV, NF, R, DIR>(x + s1, d + s1, ox, od, hox, hod, p1);
    resolve_plane<T, TV, NF, R, DIR>(pm, t);
    resolve_plane<T, TV, NF, R, DIR>(p0, t);
    resolve_plane<T, TV, NF, R, DIR>(p1, t);
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int k = 0; k < NF; k++) {
        A[r][k] = pm.x[r][k];
        B[r][k] = p0.x[r][k];
        Cn[r][k] = p1.x[r][k];
        lds[0][th.lrow0 + r][th.lcol][k] = p0.x[r][k];
        lds[1][th.lrow0 + r][th.lcol][k] = p1.x[r][k];
      }
    if (halo)
#pragma unroll
      for (int k = 0; k < NF; k++) {
        lds[0][th.hrow][th.hcol][k] = p0.hx[k];
        lds[1][th.hrow][th.hcol][k] = p1.hx[k];
      }
  }
  __syncthreads();

  double acc[2] = {0.0, 0.0};  // energy, g.g
#pragma nounroll
  for (int i = x0; i < x1; i++) {
    const int b = (i - x0) & 1;
    // plane i+2 in flight while plane i is computed (wrapped: always a valid plane)
    PlaneIn<T, TV, NF, R, DIR> q;
    const size_t sq = (size_t)tile_plane(i + 2, X) * plane;
    load_plane<T, TV, NF, R, DIR>(x + sq, d + sq, ox, od, hox, hod, q);

    TV* gp = g + (size_t)i * plane;
    T* xop = xo + (size_t)i * plane;
    T eacc = (T)0, gacc = (T)0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      T ym[NF], yp[NF], zm[NF], zp[NF], gv[NF], es;
#pragma unroll
      for (int k = 0; k < NF; k++) {
        ym[k] = (r > 0) ? B[r > 0 ? r - 1 : 0][k] : lds[b][th.lrow0 + r - 1][th.lcol][k];
        yp[k] = (r < R - 1) ? B[r < R - 1 ? r + 1 : 0][k] : lds[b][th.lrow0 + r + 1][th.lcol][k];
        zm[k] = lds[b][th.lrow0 + r][th.lcol - 1][k];
        zp[k] = lds[b][th.lrow0 + r][th.lcol + 1][k];
      }
      site_tile<T, NF, DM, WE, WG>(a, B[r], A[r], Cn[r], ym, yp, zm, zp, es, gv);
      if (th.own & (1u << r)) {
        if (XW) st_site<T, NF>(xop, ox[r], B[r]);
        if (WE) eacc = eacc + es;
        if (WG) {
          TV gs[NF];
#pragma unroll
          for (int k = 0; k < NF; k++) {
            gs[k] = (TV)gv[k];
            gacc = fmad((T)gs[k], (T)gs[k], gacc);
          }
          st_site<TV, NF>(gp, od[r], gs);
        }
      }
    }
    if (WE) acc[0] += (double)eacc;
    if (WG) acc[1] += (double)gacc;

    lds_barrier();  // every wave is done reading buffer b
    resolve_plane<T, TV, NF, R, DIR>(q, t);
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int k = 0; k < NF; k++) {
        lds[b][th.lrow0 + r][th.lcol][k] = q.x[r][k];
        A[r][k] = B[r][k];
        B[r][k] = Cn[r][k];
        Cn[r][k] = q.x[r][k];
      }
    if (halo)
#pragma unroll
      for (int k = 0; k < NF; k++) lds[b][th.hrow][th.hcol][k] = q.hx[k];
  }
  grid_sum_store<2>(acc, rs);
}

// Rows per thread: the three plane copies of the own rows plus the plane in
// flight must stay in registers.
template <typename T, int NF>
constexpr int tile_rows() {
  return sizeof(T) * NF <= 24 ? 4 : 2;
}

// ---------------------------------------------------------------- generic kernel (nFluid > 4)
template <typename T>
__device__ inline void load_site_g(const T* __restrict__ x, const T* __restrict__ d, T t, size_t site, int nf,
                                   bool dir, T (&v)[kPFMaxFluid]) {
  const size_t base = site * (size_t)nf;
  for (int k = 0; k < nf; k++) {
    T c = x[base + k];
    if (dir) {
      const T s = t * d[base + k];
      c = c + s;
    }
    v[k] = c;
  }
}

template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void pf_generic_kernel(PFArgs a, const T* __restrict__ x,
                                                            const T* __restrict__ d, T t, T* __restrict__ g,
                                                            RedScratch rs) {
  constexpr bool WE = MODE & 1, WG = MODE & 2, DIR = MODE & 4;
  const int nf = a.nf;
  const int L = xcd_logical_block(blockIdx.x, gridDim.x);
  const int xc = L / a.ntile, tile = L - xc * a.ntile;
  const int ty_tile = tile / a.ntz, tz_tile = tile - ty_tile * a.ntz;
  const int ly = threadIdx.x / a.TZ, lz = threadIdx.x - ly * a.TZ;
  const int y = ty_tile * a.TY + ly, z = tz_tile * a.TZ + lz;
  double acc[2] = {0.0, 0.0};
  if (ly < a.TY && y < a.Y && z < a.Z) {
    const int p = y * a.Z + z;
    const int pzm = (z == 0) ? p + a.Z - 1 : p - 1;
    const int pzp = (z == a.Z - 1) ? p - (a.Z - 1) : p + 1;
    const int pym = (y == 0) ? p + (a.Y - 1) * a.Z : p - a.Z;
    const int pyp = (y == a.Y - 1) ? p - (a.Y - 1) * a.Z : p + a.Z;
    const size_t YZ = (size_t)a.YZ;
    const int X = a.X;
    auto wrapx = [X](int v) { return v >= X ? v - X : (v < 0 ? v + X : v); };
    const int x0 = xc * a.xchunk, x1 = min(x0 + a.xchunk, X);
    T cm[kPFMaxFluid], c[kPFMaxFluid], cp[kPFMaxFluid];
    load_site_g<T>(x, d, t, (size_t)wrapx(x0 - 1) * YZ + p, nf, DIR, cm);
    load_site_g<T>(x, d, t, (size_t)x0 * YZ + p, nf, DIR, c);
    for (int xi = x0; xi < x1; xi++) {
      T ym[kPFMaxFluid], yp[kPFMaxFluid], zm[kPFMaxFluid], zp[kPFMaxFluid];
      const size_t pl = (size_t)xi * YZ;
      load_site_g<T>(x, d, t, (size_t)wrapx(xi + 1) * YZ + p, nf, DIR, cp);
      load_site_g<T>(x, d, t, pl + pym, nf, DIR, ym);
      load_site_g<T>(x, d, t, pl + pyp, nf, DIR, yp);
      load_site_g<T>(x, d, t, pl + pzm, nf, DIR, zm);
      load_site_g<T>(x, d, t, pl + pzp, nf, DIR, zp);
      T gv[kPFMaxFluid];
      T rho = (T)-1;
      T es = (T)0;
      for (int k = 0; k < nf; k++) {
        const T ck = c[k];
        const T a2 = ck - (T)1;
        const T dxm = ck - cm[k], dxp = ck - cp[k], dym = ck - ym[k], dyp = ck - yp[k];
        const T dzm = ck - zm[k], dzp = ck - zp[k];
        if (WE) {
          const T be = PFC<T>::fb(a, k) * ((ck * ck) * (a2 * a2));
          T sq = dxp * dxp;
          sq = fmad(dyp, dyp, sq);
          sq = fmad(dzp, dzp, sq);
          es = es + be;
          es = fmad(PFC<T>::fg(a, k), sq, es);
        }
        if (WG) {
          const T lap = ((dxm + dxp) + (dym + dyp)) + (dzm + dzp);
          const T bg = PFC<T>::fbg(a, k) * ((ck * a2) * fmad((T)2, ck, (T)-1));
          gv[k] = fmad(PFC<T>::fg2(a, k), lap, bg);
        }
        rho = rho + ck;
      }
      if (a.density == kDensitySoft) {
        if (WE) es = fmad(PFC<T>::coef(a), rho * rho, es);
        if (WG)
          for (int k = 0; k < nf; k++) gv[k] = gv[k] + PFC<T>::coef2(a) * rho;
      }
      if (WE) acc[0] += (double)es;
      if (WG) {
        for (int k = 0; k < nf; k++)
          if (a.fixmask & (1u << k)) gv[k] = (T)0;
        if (a.density == kDensityHard && a.nvar > 0) {
          T comp = (T)0;
          for (int k = 0; k < nf; k++) comp = comp + gv[k];
          comp = comp * PFC<T>::inv_nvar(a);
          for (int k = 0; k < nf; k++)
            if (!(a.fixmask & (1u << k))) gv[k] = gv[k] - comp;
        }
        const size_t base = (pl + p) * (size_t)nf;
        T gacc = (T)0;
        for (int k = 0; k < nf; k++) {
          g[base + k] = gv[k];
          gacc = fmad(gv[k], gv[k], gacc);
        }
        acc[1] += (double)gacc;
      }
      for (int k = 0; k < nf; k++) {
        cm[k] = c[k];
        c[k] = cp[k];
      }
    }
  }
  grid_sum_store<2>(acc, rs);
}

// ---------------------------------------------------------------- constraints
// Hard density projection / fixed-fluid zeroing of an arbitrary vector
// (Potential::applyConstraints -> PhaseField::specialisedConstraints).
template <typename T, int NF>
__global__ __launch_bounds__(kBlock) void pf_constrain_kernel(PFArgs a, T* __restrict__ v, size_t nsite) {
  constexpr int M = NF > 0 ? NF : kPFMaxFluid;
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

// ---------------------------------------------------------------- launches
constexpr int kTileTargetBlocks = 512;  // two blocks of 4 waves per CU

int density_mode(const PFArgs& a) {
  if (a.nf > 1 && a.density == kDensityHard && a.nvar > 0) return DM_HARD;
  if (a.nf > 1 && a.density == kDensitySoft) return DM_SOFT;
  return DM_NONE;
}

template <typename T, typename TV, int NF>
TileGeom tile_geom_for(const PFArgs& a) {
  TileGeom G;
  const int pi = sizeof(T) == 8 ? 0 : 1;
  const bool user = a.target_blocks > 0 || a.xchunk_override > 0;
  const int target = a.target_blocks > 0 ? a.target_blocks
                     : (!user && a.tuned_target[pi] > 0) ? a.tuned_target[pi]
                                                         : kTileTargetBlocks;
  const int xo = user ? a.xchunk_override : a.tuned_xchunk[pi];
  if (!tile_geom(a.X, a.Y, a.Z, tile_rows<T, NF>(), xo, target, kMaxRedBlocks, G))
    throw std::invalid_argument("PhaseField: grid too large for one launch");
  if ((double)a.Y * a.Z * NF * sizeof(T) >= 4294967296.0)
    throw std::invalid_argument("PhaseField: one x plane of the grid exceeds 4 GiB");
  return G;
}

template <typename T, typename TV, int NF, int DM, int MODE>
void launch_tile(const PFArgs& a, const TileGeom& G, const void* x, const void* d, double t, void* g, RedScratch rs,
                 hipStream_t s, void* xo) {
  constexpr int R = tile_rows<T, NF>();
  // without a direction d is never read; without a gradient g is never written;
  // without mode 8 xo is never written
  const TV* dd = static_cast<const TV*>((MODE & 4) ? d : x);
  TV* gg = static_cast<TV*>((MODE & 2) ? g : const_cast<void*>(x));
  T* xx = static_cast<T*>((MODE & 8) ? xo : const_cast<void*>(x));
  pf_tile_kernel<T, TV, NF, DM, MODE, R><<<G.nblocks, kBlock, 0, s>>>(a, G, static_cast<const T*>(x), dd, (T)t, gg,
                                                                      rs, xx);
}

template <typename T, typename TV, int NF, int MODE>
void launch_tile_dm(const PFArgs& a, const TileGeom& G, const void* x, const void* d, double t, void* g,
                    RedScratch rs, hipStream_t s, void* xo) {
  const int dm = density_mode(a);
  if constexpr (NF == 1) {
    launch_tile<T, TV, NF, DM_NONE, MODE>(a, G, x, d, t, g, rs, s, xo);
  } else {
    if (dm == DM_SOFT) {
      launch_tile<T, TV, NF, DM_SOFT, MODE>(a, G, x, d, t, g, rs, s, xo);
    } else if constexpr ((MODE & 2) != 0) {
      if (dm == DM_HARD) launch_tile<T, TV, NF, DM_HARD, MODE>(a, G, x, d, t, g, rs, s, xo);
      else launch_tile<T, TV, NF, DM_NONE, MODE>(a, G, x, d, t, g, rs, s, xo);
    } else {
      launch_tile<T, TV, NF, DM_NONE, MODE>(a, G, x, d, t, g, rs, s, xo);  // the projection acts on g only
    }
  }
}

template <typename T, typename TV, int NF>
int launch_tile_nf(const PFArgs& a, int mode, const void* x, const void* d, double t, void* g, RedScratch rs,
                   hipStream_t s, void* xo) {
  const TileGeom G = tile_geom_for<T, TV, NF>(a);
  switch (mode) {
    case 1: launch_tile_dm<T, T, NF, 1>(a, G, x, d, t, g, rs, s, xo); break;  // no vector is touched
    case 2: launch_tile_dm<T, TV, NF, 2>(a, G, x, d, t, g, rs, s, xo); break;
    case 3: launch_tile_dm<T, TV, NF, 3>(a, G, x, d, t, g, rs, s, xo); break;
    case 5: launch_tile_dm<T, TV, NF, 5>(a, G, x, d, t, g, rs, s, xo); break;
    case 15: launch_tile_dm<T, TV, NF, 15>(a, G, x, d, t, g, rs, s, xo); break;
    default: throw std::invalid_argument("pf: unsupported evaluation mode");
  }
  return G.nblocks;
}

template <typename T, typename TV>
int launch_eval_t(const PFArgs& a, int mode, const void* x, const void* d, double t, void* g, RedScratch rs,
                  hipStream_t s, void* xo) {
  switch (a.nf) {
    case 1: return launch_tile_nf<T, TV, 1>(a, mode, x, d, t, g, rs, s, xo);
    case 2: return launch_tile_nf<T, TV, 2>(a, mode, x, d, t, g, rs, s, xo);
    case 3: return launch_tile_nf<T, TV, 3>(a, mode, x, d, t, g, rs, s, xo);
    case 4: return launch_tile_nf<T, TV, 4>(a, mode, x, d, t, g, rs, s, xo);
    default: break;
  }
  if (mode & 8) throw std::invalid_argument("PhaseField: the fused step needs nFluid <= 4");
  if constexpr (sizeof(T) != sizeof(TV)) {
    throw std::invalid_argument("PhaseField: mixed precision needs nFluid <= 4");
  } else {
    const T* xx = static_cast<const T*>(x);
    const T* dd = static_cast<const T*>(d);
    T* gg = static_cast<T*>(g);
    const T tt = (T)t;
    switch (mode) {
      case 1: pf_generic_kernel<T, 1><<<a.nblocks, kBlock, 0, s>>>(a, xx, dd, tt, gg, rs); break;
      case 2: pf_generic_kernel<T, 2><<<a.nblocks, kBlock, 0, s>>>(a, xx, dd, tt, gg, rs); break;
      case 3: pf_generic_kernel<T, 3><<<a.nblocks, kBlock, 0, s>>>(a, xx, dd, tt, gg, rs); break;
      case 5: pf_generic_kernel<T, 5><<<a.nblocks, kBlock, 0, s>>>(a, xx, dd, tt, gg, rs); break;
      default: throw std::invalid_argument("pf: unsupported evaluation mode");
    }
    return a.nblocks;
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
  a.target_blocks = target_blocks;
  for (int k = 0; k < kPFMaxFluid; k++) {
    a.fbg[k] = (a.nf == 1 ? 4.0 : 2.0) * a.fb[k];
    a.fg2[k] = 2.0 * a.fg[k];
    a.fb_f[k] = (float)a.fb[k];
    a.fbg_f[k] = (float)a.fbg[k];
    a.fg_f[k] = (float)a.fg[k];
    a.fg2_f[k] = (float)a.fg2[k];
  }
  a.coef2 = 2.0 * a.coef;
  a.coef_f = (float)a.coef;
  a.coef2_f = (float)a.coef2;
  a.inv_nvar_f = (float)a.inv_nvar;
  // launch shape of the generic kernel: (y,z) tiles of 64 z x 4 y (or whole
  // rows stacked up to the block size), each marching over an x chunk
  a.TZ = a.Z >= 64 ? 64 : a.Z;
  a.TY = kBlock / a.TZ;
  if (a.TY > a.Y) a.TY = a.Y;
  a.ntz = (a.Z + a.TZ - 1) / a.TZ;
  a.nty = (a.Y + a.TY - 1) / a.TY;
  a.ntile = a.ntz * a.nty;
  int xchunk = a.xchunk_override > 0 ? a.xchunk_override : 1;
  if (a.xchunk_override <= 0) {
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
  while ((long)a.ntile * a.nxc > kMaxRedBlocks && a.xchunk < a.X) {
    a.xchunk *= 2;
    if (a.xchunk > a.X) a.xchunk = a.X;
    a.nxc = (a.X + a.xchunk - 1) / a.xchunk;
  }
  if ((long)a.ntile * a.nxc > kMaxRedBlocks) throw std::invalid_argument("PhaseField: grid too large for one launch");
  a.nblocks = a.ntile * a.nxc;
}

void pf_launch_eval(const PFArgs& a, int prec, int pv, int mode, const void* x, const void* d, double t,
                    void* g, RedScratch rs, hipStream_t s, bool finalize, void* xout, hipEvent_t evBefore,
                    hipEvent_t evAfter, double* eOut) {
  int nb;
  if (evBefore) MINIM_HIP_CHECK(hipEventRecord(evBefore, s));
  if (xout) {
    if (mode != 7 || a.use_ext || a.volfix)
      throw std::invalid_argument("PhaseField: the fused step needs energy, gradient and a direction, and the "
                                  "tiled kernel");
    mode |= 8;
  }
  if (a.use_ext) {
    PFCall c{};
    c.terms = kTermsEnergy;
    nb = pf_launch_ext(a, c, prec, pv, mode, x, d, t, g, rs, s);
  } else if (prec == F64 && pv == F64)
    nb = launch_eval_t<double, double>(a, mode, x, d, t, g, rs, s, xout);
  else if (prec == F32 && pv == F32)
    nb = launch_eval_t<float, float>(a, mode, x, d, t, g, rs, s, xout);
  else if (prec == F64 && pv == F32)
    nb = launch_eval_t<double, float>(a, mode, x, d, t, g, rs, s, xout);
  else
    throw std::invalid_argument("PhaseField: FP32 coordinates with FP64 vectors are not supported");
  MINIM_HIP_CHECK(hipGetLastError());
  if (evAfter) MINIM_HIP_CHECK(hipEventRecord(evAfter, s));
  if (finalize) {
    launch_finalize(rs, 2, nb, s, 0, (mode & 1) ? eOut : nullptr);
    // fixed volumes: the global part of the gradient projection; g.g anew
    if ((mode & 2) && a.volfix) pf_launch_volume_projection(a, pv, g, (size_t)a.X * a.YZ, true, rs, s);
  }
}

void pf_launch_constrain(const PFArgs& a, int prec, void* v, size_t nsite, RedScratch rs, hipStream_t s) {
  if (prec == F64)
    launch_constrain_t<double>(a, v, nsite, s);
  else
    launch_constrain_t<float>(a, v, nsite, s);
  MINIM_HIP_CHECK(hipGetLastError());
  if (a.volfix) pf_launch_volume_projection(a, prec, v, nsite, false, rs, s);
}

}  // namespace minim_amd
