// pf_ext.hip -- the general PhaseField kernel for MI355X (gfx950): solid
// nodes and every term of PhaseField::energyGradient beyond the bulk, gradient
// and density terms of the tiled kernel (pf.hip).
//
// Replaces, on one GPU: fluidEnergy with per-site node volumes and per-site
// surface tensions (src/potentials/PhaseField.cpp:365-393), phaseGradient with
// its one-sided branches next to solid nodes (:238-288), surfaceEnergy
// (:476-487), pressureEnergy (:396-421), densityConstraintEnergy (:424-440),
// forceEnergy (:490-518), ffConfinementEnergy (:521-537), the gradient of
// volumeConstraintEnergy (:540-577, energyComponents only), and the fixed-volume
// projection of specialisedConstraints (:614-658).
//
// Gather form with solid nodes.  The reference scatters, for each fluid site i
// and dimension d, f_i * h * (c_i - c_j) into i and its neighbours j, with
// h = 1/2 when both neighbours along d are fluid, h = 1 towards the fluid side
// when the other is solid, and nothing towards a solid site (f_i includes the
// node volume, so solid sites scatter nothing).  Collecting the scatter at
// each site gives, per bond (i, j):
//     g_i += 2 (f_i h_i->j + f_j h_j->i) (c_i - c_j),
// and the energy of site i is f_i sum_j h_i->j (c_i - c_j)^2.  The bond weights
// h are precomputed per site (the code word, pf.hpp) from the solid layout,
// so each thread reads its site's and its six neighbours' codes and factors
// and writes its own gradient entries once: no atomics.
//
// Two kernels.  The tiled one (pf_ext_tile_kernel) follows pf.hip's tiled
// stencil: x planes of a 64 x 4R tile stream through a two-plane LDS ring with
// the site codes beside the concentrations, one vector load per site.  With
// spatially uniform tensions the node volume enters every factor as the
// ratio v/res^3 of the site kind (0, k/8 or 1, exact), so a bond's weight
//     w = (v_i/res^3) h_i->j + (v_j/res^3) h_j->i
// is the same for every fluid and g_i,k = bulk'_k + 2 f_k sum_j w (c_i - c_j)
// with the bulk-site factor f_k: six weights per site serve all fluids.  The
// generic kernel (pf_ext_kernel: one thread per (y, z) column, neighbours
// through the L1/L2, factor tables by kind in LDS) covers per-site tensions,
// more than four fluids and the single-term launches of energyComponents.
#include "common.hpp"
#include "pf.hpp"
#include "pf_tile.hpp"

namespace minim_amd {

namespace {

template <typename T>
__device__ __forceinline__ T fmad(T a, T b, T c) {
  return fma(a, b, c);
}

// Bond weight (0, 1/2 or 1) of dimension j, side 0 (minus) or 1 (plus).
template <typename T>
__device__ __forceinline__ T bond_h(unsigned q, int j, int side) {
  return (T)((q >> (kPFCodeShift + 4 * j + 2 * side)) & 3u) * (T)0.5;
}

__device__ __forceinline__ int site_kind(unsigned q) { return (int)(q & 15u); }

// Coordinates, code and (with per-site tensions, PS) gradient factors of a site.
template <typename T, int M, bool PS>
struct Site {
  T c[M];
  T F[PS ? M : 1];
  unsigned q;
};

// The nf values of one site as one vector access (dwordx3 for three FP32
// fluids) when the fluid count is a compile-time constant.
template <typename T, int N>
struct SiteVec {
  T v[N];
};

template <typename T, typename TV, int NF, int M, bool PS, bool DIR>
__device__ __forceinline__ void load_site(const PFArgs& a, const T* __restrict__ x, const TV* __restrict__ d, T t,
                                          size_t site, int nf, Site<T, M, PS>& s) {
  const size_t base = site * (size_t)nf;
  if constexpr (NF > 0) {
    const SiteVec<T, NF> xv = *reinterpret_cast<const SiteVec<T, NF>*>(x + base);
#pragma unroll
    for (int k = 0; k < NF; k++) s.c[k] = xv.v[k];
    if constexpr (DIR) {
      const SiteVec<TV, NF> dv = *reinterpret_cast<const SiteVec<TV, NF>*>(d + base);
#pragma unroll
      for (int k = 0; k < NF; k++) {
        const T st = t * (T)dv.v[k];
        s.c[k] = s.c[k] + st;
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < M; k++) {
      if (k < nf) {
        T c = x[base + k];
        if constexpr (DIR) {
          const T st = t * (T)d[base + k];
          c = c + st;
        }
        s.c[k] = c;
      }
    }
  }
  if constexpr (PS) {
#pragma unroll
    for (int k = 0; k < M; k++)
      if (k < nf) s.F[k] = (T)a.Fs[base + k];
  }
  s.q = a.code ? (unsigned)a.code[site] : a.ucode;
}

// MODE bits: 1 = energy, 2 = gradient (+ g.g), 4 = trial point x + t*d.
// PS: per-site surface tensions (factors read per DOF instead of by kind).
template <typename T, typename TV, int NF, int MODE, bool PS>
__global__ __launch_bounds__(kBlock) void pf_ext_kernel(PFArgs a, PFCall cl, const T* __restrict__ x,
                                                        const TV* __restrict__ d, T t, TV* __restrict__ g,
                                                        RedScratch rs) {
  constexpr bool WE = MODE & 1, WG = MODE & 2, DIR = MODE & 4;
  constexpr int M = NF > 0 ? NF : kPFMaxFluid;
  const int nf = NF > 0 ? NF : a.nf;
  __shared__ T sB[kPFMaxFluid * kPFKinds], sF[kPFMaxFluid * kPFKinds], sV[kPFKinds], sA[kPFKinds];
  for (int i = threadIdx.x; i < nf * kPFKinds; i += kBlock) {
    sB[i] = (T)a.ext->B[i / kPFKinds][i % kPFKinds];
    sF[i] = (T)a.ext->F[i / kPFKinds][i % kPFKinds];
  }
  if (threadIdx.x < kPFKinds) {
    sV[threadIdx.x] = (T)a.ext->vol[threadIdx.x];
    sA[threadIdx.x] = (T)a.ext->area[threadIdx.x];
  }
  __syncthreads();

  const unsigned terms = cl.terms;
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
    const PFExt* E = a.ext;
    const T res = (T)E->res;
    Site<T, M, PS> sm, sc, sp;
    load_site<T, TV, NF, M, PS, DIR>(a, x, d, t, (size_t)wrapx(x0 - 1) * YZ + p, nf, sm);
    load_site<T, TV, NF, M, PS, DIR>(a, x, d, t, (size_t)x0 * YZ + p, nf, sc);
    for (int xi = x0; xi < x1; xi++) {
      const size_t pl = (size_t)xi * YZ;
      const size_t site = pl + p;
      load_site<T, TV, NF, M, PS, DIR>(a, x, d, t, (size_t)wrapx(xi + 1) * YZ + p, nf, sp);
      Site<T, M, PS> nb[4];  // y-, y+, z-, z+
      load_site<T, TV, NF, M, PS, DIR>(a, x, d, t, pl + pym, nf, nb[0]);
      load_site<T, TV, NF, M, PS, DIR>(a, x, d, t, pl + pyp, nf, nb[1]);
      load_site<T, TV, NF, M, PS, DIR>(a, x, d, t, pl + pzm, nf, nb[2]);
      load_site<T, TV, NF, M, PS, DIR>(a, x, d, t, pl + pzp, nf, nb[3]);
      const Site<T, M, PS>* mside[3] = {&sm, &nb[0], &nb[2]};
      const Site<T, M, PS>* pside[3] = {&sp, &nb[1], &nb[3]};
      const int kind = site_kind(sc.q);
      const T v = sV[kind];
      T gv[M];
      T es = (T)0;
#pragma unroll
      for (int k = 0; k < M; k++) {
        if (k >= nf) continue;
        const T ck = sc.c[k];
        T gk = (T)0;
        if (terms & kTermFluid) {
          T Bk, Fk;
          if constexpr (PS) {
            Bk = (T)a.Bs[site * nf + k];
            Fk = sc.F[k];
          } else {
            Bk = sB[k * kPFKinds + kind];
            Fk = sF[k * kPFKinds + kind];
          }
          const T a2 = ck - (T)1;
          if (nf == 1) {
            const T a1 = ck + (T)1;
            if (WE) es = es + Bk * ((a1 * a1) * (a2 * a2));
            if (WG) gk = (Bk * (T)4) * (ck * fmad(ck, ck, (T)-1));
          } else {
            if (WE) es = es + Bk * ((ck * ck) * (a2 * a2));
            if (WG) gk = (Bk * (T)2) * ((ck * a2) * fmad((T)2, ck, (T)-1));
          }
          T sq = (T)0, lap = (T)0;
#pragma unroll
          for (int j = 0; j < 3; j++) {
            const Site<T, M, PS>& sa = *mside[j];
            const Site<T, M, PS>& sb = *pside[j];
            const T hm = bond_h<T>(sc.q, j, 0), hp = bond_h<T>(sc.q, j, 1);
            const T dm = ck - sa.c[k], dp = ck - sb.c[k];
            if (WE) {
              sq = fmad(hm * dm, dm, sq);
              sq = fmad(hp * dp, dp, sq);
            }
            if (WG) {
              T Fa, Fb;
              if constexpr (PS) {
                Fa = sa.F[k];
                Fb = sb.F[k];
              } else {
                Fa = sF[k * kPFKinds + site_kind(sa.q)];
                Fb = sF[k * kPFKinds + site_kind(sb.q)];
              }
              const T wm = fmad(Fa, bond_h<T>(sa.q, j, 1), Fk * hm);
              const T wp = fmad(Fb, bond_h<T>(sb.q, j, 0), Fk * hp);
              lap = fmad(wm, dm, lap);
              lap = fmad(wp, dp, lap);
            }
          }
          if (WE) es = fmad(Fk, sq, es);
          if (WG) gk = fmad((T)2, lap, gk);
        }
        if ((terms & kTermPressure) && a.any_pres) {
          const T pk = (T)E->pres[k];
          if (pk != (T)0) {
            if (nf == 1) {
              if (WE) es = es - pk * (((T)0.5 * (ck + (T)1)) * v);
              if (WG) gk = gk - ((T)0.5 * pk) * v;
            } else {
              if (WE) es = es - pk * (ck * v);
              if (WG) gk = gk - pk * v;
            }
          }
        }
        if ((terms & kTermConfinement) && a.csign) {
          const T coef = (T)E->conf[k];
          if (coef != (T)0) {
            const int s0 = a.csign[site * nf + k];
            const T u = ck - (T)0.5;
            if ((s0 > 0 && u < (T)0) || (s0 < 0 && u > (T)0)) {
              if (WE) es = es + coef * (u * u);
              if (WG) gk = gk + (coef * (T)2) * u;
            }
          }
        }
        if (terms & kTermVolume) {
          const T vg = (T)cl.volg[k];
          if (vg != (T)0) {
            T w = ((T)4 * ck) * ((T)1 - ck);
            if (w < (T)0) w = (T)0;
            gk = gk + (vg * v) * w;
          }
        }
        gv[k] = gk;
      }
      // surfaceEnergy: binary fluids only
      if ((terms & kTermSurface) && a.wet && nf == 1) {
        const T w = (T)a.wet[site];
        if (w != (T)0) {
          const T phi = sc.c[0], ar = sA[kind];
          if (WE) es = es + (w * ((((phi * phi) * phi) / (T)3 - phi) - (T)2 / (T)3)) * ar;
          if (WG) gv[0] = gv[0] + (w * (phi * phi - (T)1)) * ar;
        }
      }
      // densityConstraintEnergy: soft constraint
      if ((terms & kTermDensity) && nf > 1 && a.density == kDensitySoft) {
        T rho = (T)-1;
#pragma unroll
        for (int k = 0; k < M; k++)
          if (k < nf) rho = rho + sc.c[k];
        const T coef = (T)a.coef;
        if (WE) es = es + coef * (rho * rho);
        if (WG) {
          const T gr = ((T)2 * coef) * rho;
#pragma unroll
          for (int k = 0; k < M; k++)
            if (k < nf) gv[k] = gv[k] + gr;
        }
      }
      // forceEnergy: not on solid sites; without a gradient only the first
      // contributing fluid adds its energy (PhaseField.cpp:511)
      if ((terms & kTermForce) && a.any_force && kind != 0) {
        const T co[3] = {(T)xi - (T)E->fc[0], (T)y - (T)E->fc[1], (T)z - (T)E->fc[2]};
#pragma unroll
        for (int k = 0; k < M; k++) {
          if (k >= nf) continue;
          const T fm = (T)E->fmag[k];
          if (fm == (T)0) continue;
          const T cv = (nf == 1) ? (T)0.5 * ((T)1 + sc.c[k]) : sc.c[k];
          if (cv < (T)0.01) continue;
          T dot = (T)0;
          dot = dot + co[0] * (T)E->fn[k][0];
          dot = dot + co[1] * (T)E->fn[k][1];
          dot = dot + co[2] * (T)E->fn[k][2];
          const T h = -dot * res;
          const T epc = (v * fm) * h;
          if (WE) es = es + cv * epc;
          if (!WG) break;
          gv[k] = gv[k] + ((nf == 1) ? (T)0.5 * epc : epc);
        }
      }
      if (WE) acc[0] += (double)es;
      if (WG) {
        if (!cl.raw) {
          // specialisedConstraints: fixed fluids zeroed, hard density projection
#pragma unroll
          for (int k = 0; k < M; k++)
            if (k < nf && (a.fixmask & (1u << k))) gv[k] = (T)0;
          if (a.density == kDensityHard && nf > 1 && a.nvar > 0) {
            T comp = (T)0;
#pragma unroll
            for (int k = 0; k < M; k++)
              if (k < nf) comp = comp + gv[k];
            comp = comp * (T)a.inv_nvar;
#pragma unroll
            for (int k = 0; k < M; k++)
              if (k < nf && !(a.fixmask & (1u << k))) gv[k] = gv[k] - comp;
          }
        }
        const size_t base = site * (size_t)nf;
        T gacc = (T)0;
        if constexpr (NF > 0) {
          SiteVec<TV, NF> o;
#pragma unroll
          for (int k = 0; k < NF; k++) {
            o.v[k] = (TV)gv[k];
            gacc = fmad((T)o.v[k], (T)o.v[k], gacc);
          }
          if constexpr (sizeof(TV) == 4) {
            // streaming: the gradient is not re-read by this kernel
#pragma unroll
            for (int k = 0; k < NF; k++) __builtin_nontemporal_store(o.v[k], g + base + k);
          } else {
            *reinterpret_cast<SiteVec<TV, NF>*>(g + base) = o;
          }
        } else {
#pragma unroll
          for (int k = 0; k < M; k++) {
            if (k >= nf) continue;
            const TV gs = (TV)gv[k];
            g[base + k] = gs;
            gacc = fmad((T)gs, (T)gs, gacc);
          }
        }
        acc[1] += (double)gacc;
      }
      sm = sc;
      sc = sp;
    }
  }
  grid_sum_store<2>(acc, rs);
}

// ---------------------------------------------------------------- tiled kernel
// Barrier that orders LDS traffic only (no wait on outstanding global accesses).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Node volume of a site kind relative to a bulk site: 0 (solid), 1 (bulk) or
// type/8 (surface types 1..7), all exact.
template <typename T>
__device__ __forceinline__ T vrel_of(unsigned q) {
  const int k = (int)(q & 15u);
  return k >= 2 ? (T)(k - 1) * (T)0.125 : (T)k;
}

template <typename T>
struct PFF;  // bulk-site factors of the tiled stencil by precision
template <>
struct PFF<double> {
  __device__ static double fb(const PFArgs& a, int k) { return a.fb[k]; }
  __device__ static double fbg(const PFArgs& a, int k) { return a.fbg[k]; }
  __device__ static double fg(const PFArgs& a, int k) { return a.fg[k]; }
  __device__ static double fg2(const PFArgs& a, int k) { return a.fg2[k]; }
};
template <>
struct PFF<float> {
  __device__ static float fb(const PFArgs& a, int k) { return a.fb_f[k]; }
  __device__ static float fbg(const PFArgs& a, int k) { return a.fbg_f[k]; }
  __device__ static float fg(const PFArgs& a, int k) { return a.fg_f[k]; }
  __device__ static float fg2(const PFArgs& a, int k) { return a.fg2_f[k]; }
};

// The constants of the extra terms, passed by value so that they sit in the
// kernel-argument segment (scalar loads, uniform registers) in the arithmetic
// precision.
template <typename T>
struct ExtK {
  T pres[kPFMaxFluid];
  T conf[kPFMaxFluid];
  T fmag[kPFMaxFluid];
  T fn[kPFMaxFluid][3];
  T fc[3];
  T area[kPFKinds];
  T v1, res;
};

template <typename T>
ExtK<T> ext_constants(const PFExt& E) {
  ExtK<T> k{};
  for (int i = 0; i < kPFMaxFluid; i++) {
    k.pres[i] = (T)E.pres[i];
    k.conf[i] = (T)E.conf[i];
    k.fmag[i] = (T)E.fmag[i];
    for (int j = 0; j < 3; j++) k.fn[i][j] = (T)E.fn[i][j];
  }
  for (int j = 0; j < 3; j++) k.fc[j] = (T)E.fc[j];
  for (int i = 0; i < kPFKinds; i++) k.area[i] = (T)E.area[i];
  k.v1 = (T)E.vol[1];
  k.res = (T)E.res;
  return k;
}

// Concentrations (x, or x + t*d) and forward-bond codes of the R own rows and
// the halo site of one x plane.
template <typename T, typename TV, int NF, int R, bool DIR>
struct ExtPlane {
  T x[R][NF];
  T hx[NF];
  TV d[DIR ? R : 1][NF];
  TV hd[NF];
  unsigned q[R];
  unsigned hq;
};

template <typename T, int NF>
__device__ __forceinline__ void ld_vec(const T* __restrict__ base, unsigned boff, T (&v)[NF]) {
  const SiteVec<T, NF> s = *reinterpret_cast<const SiteVec<T, NF>*>(reinterpret_cast<const char*>(base) + boff);
#pragma unroll
  for (int k = 0; k < NF; k++) v[k] = s.v[k];
}

template <typename T, typename TV, int NF, int R, bool DIR>
__device__ __forceinline__ void load_ext_plane(const PFArgs& a, size_t pl, const T* __restrict__ x,
                                               const TV* __restrict__ d, const unsigned (&site)[R], unsigned hsite,
                                               ExtPlane<T, TV, NF, R, DIR>& p) {
  const T* xp = x + pl * NF;
#pragma unroll
  for (int r = 0; r < R; r++) ld_vec<T, NF>(xp, site[r] * (unsigned)(NF * sizeof(T)), p.x[r]);
  ld_vec<T, NF>(xp, hsite * (unsigned)(NF * sizeof(T)), p.hx);
  if constexpr (DIR) {
    const TV* dp = d + pl * NF;
#pragma unroll
    for (int r = 0; r < R; r++) ld_vec<TV, NF>(dp, site[r] * (unsigned)(NF * sizeof(TV)), p.d[r]);
    ld_vec<TV, NF>(dp, hsite * (unsigned)(NF * sizeof(TV)), p.hd);
  }
  if (a.wcode) {
    const uint32_t* qp = a.wcode + pl;
#pragma unroll
    for (int r = 0; r < R; r++) p.q[r] = qp[site[r]];
    p.hq = qp[hsite];
  } else {
#pragma unroll
    for (int r = 0; r < R; r++) p.q[r] = a.wucode;
    p.hq = a.wucode;
  }
}

template <typename T, typename TV, int NF, int R, bool DIR>
__device__ __forceinline__ void resolve_ext_plane(ExtPlane<T, TV, NF, R, DIR>& p, T t) {
  if constexpr (DIR) {
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int k = 0; k < NF; k++) {
        const T st = t * (T)p.d[r][k];
        p.x[r][k] = p.x[r][k] + st;
      }
#pragma unroll
    for (int k = 0; k < NF; k++) {
      const T st = t * (T)p.hd[k];
      p.hx[k] = p.hx[k] + st;
    }
  }
}

// Weight of the forward bond along canonical dimension j from a forward-bond code.
template <typename T>
__device__ __forceinline__ T fwd_w(unsigned q, int j) {
  return (T)((q >> (kPFWeightShift + kPFWeightBits * j)) & ((1u << kPFWeightBits) - 1u)) * (T)0.0625;
}

// Energy (the site's bulk term and its three forward bonds) and constrained
// gradient of one site from its six neighbours, its forward-bond code qc and
// the codes of its minus neighbours (whose forward bonds end here).  FULL adds
// the force, confinement and wetting terms.
template <typename T, int NF, bool WE, bool WG, bool FULL>
__device__ __forceinline__ void site_ext(const PFArgs& a, const ExtK<T>& E, const T (&c)[NF], const T (&xm)[NF],
                                         const T (&xp)[NF], const T (&ym)[NF], const T (&yp)[NF], const T (&zm)[NF],
                                         const T (&zp)[NF], unsigned qc, unsigned qxm, unsigned qym, unsigned qzm,
                                         int xi, int y, int z, size_t site, T& e, T (&g)[NF]) {
  using C = PFF<T>;
  const T vc = vrel_of<T>(qc);
  const T wfx = fwd_w<T>(qc, 0), wfy = fwd_w<T>(qc, 1), wfz = fwd_w<T>(qc, 2);
  const T wbx = fwd_w<T>(qxm, 0), wby = fwd_w<T>(qym, 1), wbz = fwd_w<T>(qzm, 2);
  e = (T)0;
#pragma unroll
  for (int k = 0; k < NF; k++) {
    const T ck = c[k];
    const T a2 = ck - (T)1;
    const T dxm = ck - xm[k], dxp = ck - xp[k];
    const T dym = ck - ym[k], dyp = ck - yp[k];
    const T dzm = ck - zm[k], dzp = ck - zp[k];
    if (WE) {
      T be;
      if (NF == 1) {
        const T a1 = ck + (T)1;
        be = (C::fb(a, k) * vc) * ((a1 * a1) * (a2 * a2));
      } else {
        be = (C::fb(a, k) * vc) * ((ck * ck) * (a2 * a2));
      }
      T sq = (wfx * dxp) * dxp;
      sq = fmad(wfy * dyp, dyp, sq);
      sq = fmad(wfz * dzp, dzp, sq);
      e = e + be;
      e = fmad(C::fg(a, k), sq, e);
    }
    if (WG) {
      T lap = wfx * dxp;
      lap = fmad(wbx, dxm, lap);
      lap = fmad(wfy, dyp, lap);
      lap = fmad(wby, dym, lap);
      lap = fmad(wfz, dzp, lap);
      lap = fmad(wbz, dzm, lap);
      T bg;
      if (NF == 1) bg = (C::fbg(a, k) * vc) * (ck * fmad(ck, ck, (T)-1));
      else bg = (C::fbg(a, k) * vc) * ((ck * a2) * fmad((T)2, ck, (T)-1));
      g[k] = fmad(C::fg2(a, k), lap, bg);
    }
  }
  const T v = vc * E.v1;
  if (a.any_pres) {
#pragma unroll
    for (int k = 0; k < NF; k++) {
      const T pk = E.pres[k];
      if (pk == (T)0) continue;
      if (NF == 1) {
        if (WE) e = e - pk * (((T)0.5 * (c[k] + (T)1)) * v);
        if (WG) g[k] = g[k] - ((T)0.5 * pk) * v;
      } else {
        if (WE) e = e - pk * (c[k] * v);
        if (WG) g[k] = g[k] - pk * v;
      }
    }
  }
  if (NF > 1 && a.density == kDensitySoft) {
    T rho = (T)-1;
#pragma unroll
    for (int k = 0; k < NF; k++) rho = rho + c[k];
    const T coef = (T)a.coef;
    if (WE) e = e + coef * (rho * rho);
    if (WG) {
      const T gr = ((T)2 * coef) * rho;
#pragma unroll
      for (int k = 0; k < NF; k++) g[k] = g[k] + gr;
    }
  }
  if constexpr (FULL) {
    if (a.csign) {
#pragma unroll
      for (int k = 0; k < NF; k++) {
        const T coef = E.conf[k];
        if (coef == (T)0) continue;
        const int s0 = a.csign[site * NF + k];
        const T u = c[k] - (T)0.5;
        if ((s0 > 0 && u < (T)0) || (s0 < 0 && u > (T)0)) {
          if (WE) e = e + coef * (u * u);
          if (WG) g[k] = g[k] + (coef * (T)2) * u;
        }
      }
    }
    if (NF == 1 && a.wet) {
      const T w = (T)a.wet[site];
      if (w != (T)0) {
        const T phi = c[0], ar = E.area[qc & 15u];
        if (WE) e = e + (w * ((((phi * phi) * phi) / (T)3 - phi) - (T)2 / (T)3)) * ar;
        if (WG) g[0] = g[0] + (w * (phi * phi - (T)1)) * ar;
      }
    }
    if (a.any_force && (qc & 15u) != 0u) {
      const T co[3] = {(T)xi - E.fc[0], (T)y - E.fc[1], (T)z - E.fc[2]};
#pragma unroll
      for (int k = 0; k < NF; k++) {
        const T fm = E.fmag[k];
        if (fm == (T)0) continue;
        const T cv = (NF == 1) ? (T)0.5 * ((T)1 + c[k]) : c[k];
        if (cv < (T)0.01) continue;
        T dot = (T)0;
        dot = dot + co[0] * E.fn[k][0];
        dot = dot + co[1] * E.fn[k][1];
        dot = dot + co[2] * E.fn[k][2];
        const T h = -dot * E.res;
        const T epc = (v * fm) * h;
        if (WE) e = e + cv * epc;
        if (!WG) break;
        g[k] = g[k] + ((NF == 1) ? (T)0.5 * epc : epc);
      }
    }
  }
  if (WG) {
    if (a.fixmask) {
#pragma unroll
      for (int k = 0; k < NF; k++)
        if (a.fixmask & (1u << k)) g[k] = (T)0;
    }
    if (NF > 1 && a.density == kDensityHard && a.nvar > 0) {
      T comp = g[0];
#pragma unroll
      for (int k = 1; k < NF; k++) comp = comp + g[k];
      comp = comp * (T)a.inv_nvar;
#pragma unroll
      for (int k = 0; k < NF; k++)
        if (!(a.fixmask & (1u << k))) g[k] = g[k] - comp;
    }
  }
}

// MODE bits: 1 = energy, 2 = gradient (+ g.g), 4 = trial point x + t*d.
// FULL: the force, confinement and wetting terms are compiled in.
template <typename T, typename TV, int NF, int MODE, int R, bool FULL>
__global__ __launch_bounds__(kBlock) void pf_ext_tile_kernel(PFArgs a, ExtK<T> K, TileGeom G, const T* __restrict__ x,
                                                             const TV* __restrict__ d, T t, TV* __restrict__ g,
                                                             RedScratch rs) {
  constexpr bool WE = MODE & 1, WG = MODE & 2, DIR = MODE & 4;
  constexpr int LH = kTileWaves * R + 2, LW = kTileZ + 2;
  __shared__ T lds[2][LH][LW][NF];
  __shared__ unsigned lq[2][LH][LW];

  TileThread th;
  tile_thread(G, xcd_logical_block(blockIdx.x, gridDim.x), threadIdx.x, th);
  const size_t YZ = (size_t)G.Y * G.Z;
  const int X = G.X;
  unsigned site[R];
  int yy[R], zz[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    site[r] = th.site[r];
    yy[r] = (int)(th.site[r] / (unsigned)G.Z);
    zz[r] = (int)(th.site[r] - (unsigned)yy[r] * (unsigned)G.Z);
  }
  const bool halo = th.hrow >= 0;
  const int x0 = th.x0, x1 = th.x1;

  T A[R][NF], B[R][NF], Cn[R][NF];
  unsigned qA[R], qB[R];  // forward-bond codes of planes i-1 and i
  {
    ExtPlane<T, TV, NF, R, DIR> pm, p0, p1;
    load_ext_plane<T, TV, NF, R, DIR>(a, (size_t)tile_plane(x0 - 1, X) * YZ, x, d, site, th.hsite, pm);
    load_ext_plane<T, TV, NF, R, DIR>(a, (size_t)tile_plane(x0, X) * YZ, x, d, site, th.hsite, p0);
    load_ext_plane<T, TV, NF, R, DIR>(a, (size_t)tile_plane(x0 + 1, X) * YZ, x, d, site, th.hsite, p1);
    resolve_ext_plane<T, TV, NF, R, DIR>(pm, t);
    resolve_ext_plane<T, TV, NF, R, DIR>(p0, t);
    resolve_ext_plane<T, TV, NF, R, DIR>(p1, t);
#pragma unroll
    for (int r = 0; r < R; r++) {
#pragma unroll
      for (int k = 0; k < NF; k++) {
        A[r][k] = pm.x[r][k];
        B[r][k] = p0.x[r][k];
        Cn[r][k] = p1.x[r][k];
        lds[0][th.lrow0 + r][th.lcol][k] = p0.x[r][k];
        lds[1][th.lrow0 + r][th.lcol][k] = p1.x[r][k];
      }
      qA[r] = pm.q[r];
      qB[r] = p0.q[r];
      lq[0][th.lrow0 + r][th.lcol] = p0.q[r];
      lq[1][th.lrow0 + r][th.lcol] = p1.q[r];
    }
    if (halo) {
#pragma unroll
      for (int k = 0; k < NF; k++) {
        lds[0][th.hrow][th.hcol][k] = p0.hx[k];
        lds[1][th.hrow][th.hcol][k] = p1.hx[k];
      }
      lq[0][th.hrow][th.hcol] = p0.hq;
      lq[1][th.hrow][th.hcol] = p1.hq;
    }
  }
  __syncthreads();

  double acc[2] = {0.0, 0.0};
#pragma nounroll
  for (int i = x0; i < x1; i++) {
    const int b = (i - x0) & 1;
    ExtPlane<T, TV, NF, R, DIR> q;
    load_ext_plane<T, TV, NF, R, DIR>(a, (size_t)tile_plane(i + 2, X) * YZ, x, d, site, th.hsite, q);

    const size_t pl = (size_t)i * YZ;
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
      const unsigned qym = (r > 0) ? qB[r > 0 ? r - 1 : 0] : lq[b][th.lrow0 + r - 1][th.lcol];
      const unsigned qzm = lq[b][th.lrow0 + r][th.lcol - 1];
      site_ext<T, NF, WE, WG, FULL>(a, K, B[r], A[r], Cn[r], ym, yp, zm, zp, qB[r], qA[r], qym, qzm, i, yy[r], zz[r],
                                    pl + site[r], es, gv);
      if (th.own & (1u << r)) {
        if (WE) eacc = eacc + es;
        if (WG) {
          SiteVec<TV, NF> o;
#pragma unroll
          for (int k = 0; k < NF; k++) {
            o.v[k] = (TV)gv[k];
            gacc = fmad((T)o.v[k], (T)o.v[k], gacc);
          }
          TV* gp = g + (pl + site[r]) * NF;  // streaming stores, as the plain stencil's
#pragma unroll
          for (int k = 0; k < NF; k++) __builtin_nontemporal_store(o.v[k], gp + k);
        }
      }
    }
    if (WE) acc[0] += (double)eacc;
    if (WG) acc[1] += (double)gacc;

    lds_barrier();  // every wave is done reading buffer b
    resolve_ext_plane<T, TV, NF, R, DIR>(q, t);
#pragma unroll
    for (int r = 0; r < R; r++) {
      // plane i+1's codes were stored in buffer b^1 last iteration; keep them in registers too
      const unsigned q1 = lq[b ^ 1][th.lrow0 + r][th.lcol];
#pragma unroll
      for (int k = 0; k < NF; k++) {
        lds[b][th.lrow0 + r][th.lcol][k] = q.x[r][k];
        A[r][k] = B[r][k];
        B[r][k] = Cn[r][k];
        Cn[r][k] = q.x[r][k];
      }
      lq[b][th.lrow0 + r][th.lcol] = q.q[r];
      qA[r] = qB[r];
      qB[r] = q1;
    }
    if (halo) {
#pragma unroll
      for (int k = 0; k < NF; k++) lds[b][th.hrow][th.hcol][k] = q.hx[k];
      lq[b][th.hrow][th.hcol] = q.hq;
    }
  }
  grid_sum_store<2>(acc, rs);
}

// ---------------------------------------------------------------- volumes
__device__ __forceinline__ double site_vol(const PFArgs& a, size_t s) {
  const unsigned q = a.code ? (unsigned)a.code[s] : a.ucode;
  return a.ext->vol[q & 15u];
}

// Per variable fluid k: sum_i v_i (u_ik - [hard] sum_var u_ik' / nvar), the
// numerator of the reference's component (PhaseField.cpp:625-641).
template <typename T, int NF>
__global__ __launch_bounds__(kBlock) void pf_volproj_reduce_kernel(PFArgs a, const T* __restrict__ u,
                                                                   size_t nsite, RedScratch rs) {
  constexpr int M = NF > 0 ? NF : kPFMaxFluid;
  const int nf = NF > 0 ? NF : a.nf;
  const bool hard = a.density == kDensityHard;
  const double nv = (double)a.nvar;
  double acc[M];
#pragma unroll
  for (int k = 0; k < M; k++) acc[k] = 0.0;
  for (size_t s = (size_t)blockIdx.x * kBlock + threadIdx.x; s < nsite; s += (size_t)gridDim.x * kBlock) {
    const double v = site_vol(a, s);
    double w[M];
#pragma unroll
    for (int k = 0; k < M; k++)
      if (k < nf) w[k] = (double)u[s * nf + k];
#pragma unroll
    for (int k = 0; k < M; k++) {
      if (k >= nf || (a.fixmask & (1u << k))) continue;
      acc[k] += w[k] * v;
      if (hard) {
#pragma unroll
        for (int k2 = 0; k2 < M; k2++)
          if (k2 < nf && !(a.fixmask & (1u << k2))) acc[k] -= (w[k2] * v) / nv;
      }
    }
  }
  grid_sum_store<M>(acc, rs);
}

// u -= comp_k v (and the hard correction), comp_k = sum_k / tv2
// (PhaseField.cpp:643-657); optionally the new u.u.
template <typename T, int NF, bool NORM>
__global__ __launch_bounds__(kBlock) void pf_volproj_apply_kernel(PFArgs a, T* __restrict__ u, size_t nsite,
                                                                  const double* __restrict__ sums, RedScratch rs) {
  constexpr int M = NF > 0 ? NF : kPFMaxFluid;
  const int nf = NF > 0 ? NF : a.nf;
  const bool hard = a.density == kDensityHard;
  const double nv = (double)a.nvar;
  double comp[M];
#pragma unroll
  for (int k = 0; k < M; k++) comp[k] = (k < nf) ? sums[k] / a.tv2 : 0.0;
  double acc[1] = {0.0};
  for (size_t s = (size_t)blockIdx.x * kBlock + threadIdx.x; s < nsite; s += (size_t)gridDim.x * kBlock) {
    const double v = site_vol(a, s);
    double w[M];
#pragma unroll
    for (int k = 0; k < M; k++)
      if (k < nf) w[k] = (double)u[s * nf + k];
#pragma unroll
    for (int k = 0; k < M; k++) {
      if (k >= nf || (a.fixmask & (1u << k))) continue;
      w[k] -= comp[k] * v;
      if (hard) {
        const double c = (comp[k] * v) / nv;
#pragma unroll
        for (int k2 = 0; k2 < M; k2++)
          if (k2 < nf && !(a.fixmask & (1u << k2))) w[k2] += c;
      }
    }
#pragma unroll
    for (int k = 0; k < M; k++) {
      if (k >= nf) continue;
      const T o = (T)w[k];
      u[s * nf + k] = o;
      if (NORM) acc[0] += (double)o * (double)o;
    }
  }
  if (NORM) grid_sum_store<1>(acc, rs);
}

// Fluid volumes: sum_i c_ik v_i, or (c+1)/2 v for one fluid.
template <typename T, int NF>
__global__ __launch_bounds__(kBlock) void pf_volumes_kernel(PFArgs a, const T* __restrict__ x, size_t nsite,
                                                            RedScratch rs) {
  constexpr int M = NF > 0 ? NF : kPFMaxFluid;
  const int nf = NF > 0 ? NF : a.nf;
  double acc[M];
#pragma unroll
  for (int k = 0; k < M; k++) acc[k] = 0.0;
  for (size_t s = (size_t)blockIdx.x * kBlock + threadIdx.x; s < nsite; s += (size_t)gridDim.x * kBlock) {
    const double v = site_vol(a, s);
#pragma unroll
    for (int k = 0; k < M; k++) {
      if (k >= nf) continue;
      const double c = (double)x[s * nf + k];
      acc[k] += (nf == 1) ? (0.5 * (c + 1)) * v : c * v;
    }
  }
  grid_sum_store<M>(acc, rs);
}

// ---------------------------------------------------------------- launches
template <typename T, typename TV, int NF, bool PS>
void ext_nf(const PFArgs& a, const PFCall& c, int mode, const void* x, const void* d, double t, void* g,
            RedScratch rs, hipStream_t s) {
  const T* xx = static_cast<const T*>(x);
  const TV* dd = static_cast<const TV*>((mode & 4) ? d : nullptr);
  TV* gg = static_cast<TV*>((mode & 2) ? g : nullptr);
  const T tt = (T)t;
  switch (mode) {
    case 1: pf_ext_kernel<T, TV, NF, 1, PS><<<a.nblocks, kBlock, 0, s>>>(a, c, xx, dd, tt, gg, rs); break;
    case 2: pf_ext_kernel<T, TV, NF, 2, PS><<<a.nblocks, kBlock, 0, s>>>(a, c, xx, dd, tt, gg, rs); break;
    case 3: pf_ext_kernel<T, TV, NF, 3, PS><<<a.nblocks, kBlock, 0, s>>>(a, c, xx, dd, tt, gg, rs); break;
    case 5: pf_ext_kernel<T, TV, NF, 5, PS><<<a.nblocks, kBlock, 0, s>>>(a, c, xx, dd, tt, gg, rs); break;
    default: throw std::invalid_argument("pf: unsupported evaluation mode");
  }
}

constexpr int kExtTileTarget = 512;

// Rows per thread of the tiled kernel: three plane copies of the own rows and
// the plane in flight stay in registers.
template <typename T, int NF>
constexpr int ext_rows() {
  return sizeof(T) * NF <= 12 ? 4 : 2;
}

template <typename T, typename TV, int NF, bool FULL>
int ext_tile_full(const PFArgs& a, int mode, const void* x, const void* d, double t, void* g, RedScratch rs,
                  hipStream_t s) {
  constexpr int R = ext_rows<T, NF>();
  TileGeom G;
  const int pi = sizeof(T) == 8 ? 0 : 1;
  const bool user = a.target_blocks > 0 || a.xchunk_override > 0;
  const int target = a.target_blocks > 0 ? a.target_blocks
                     : (!user && a.tuned_target[pi] > 0) ? a.tuned_target[pi]
                                                         : kExtTileTarget;
  const int xo = user ? a.xchunk_override : a.tuned_xchunk[pi];
  if (!tile_geom(a.X, a.Y, a.Z, R, xo, target, kMaxRedBlocks, G))
    throw std::invalid_argument("PhaseField: grid too large for one launch");
  const T* xx = static_cast<const T*>(x);
  const TV* dd = static_cast<const TV*>((mode & 4) ? d : x);
  TV* gg = static_cast<TV*>((mode & 2) ? g : const_cast<void*>(x));
  const T tt = (T)t;
  const ExtK<T> K = ext_constants<T>(*a.ext_host);
  switch (mode) {
    case 1:
      pf_ext_tile_kernel<T, T, NF, 1, R, FULL><<<G.nblocks, kBlock, 0, s>>>(a, K, G, xx, (const T*)xx, tt, (T*)nullptr, rs);
      break;
    case 2: pf_ext_tile_kernel<T, TV, NF, 2, R, FULL><<<G.nblocks, kBlock, 0, s>>>(a, K, G, xx, dd, tt, gg, rs); break;
    case 3: pf_ext_tile_kernel<T, TV, NF, 3, R, FULL><<<G.nblocks, kBlock, 0, s>>>(a, K, G, xx, dd, tt, gg, rs); break;
    case 5: pf_ext_tile_kernel<T, TV, NF, 5, R, FULL><<<G.nblocks, kBlock, 0, s>>>(a, K, G, xx, dd, tt, gg, rs); break;
    default: throw std::invalid_argument("pf: unsupported evaluation mode");
  }
  return G.nblocks;
}

template <typename T, typename TV, int NF>
int ext_tile_nf(const PFArgs& a, int mode, const void* x, const void* d, double t, void* g, RedScratch rs,
                hipStream_t s) {
  if (a.any_force || a.csign || a.wet) return ext_tile_full<T, TV, NF, true>(a, mode, x, d, t, g, rs, s);
  return ext_tile_full<T, TV, NF, false>(a, mode, x, d, t, g, rs, s);
}

template <typename T, typename TV>
int ext_t(const PFArgs& a, const PFCall& c, int mode, const void* x, const void* d, double t, void* g,
          RedScratch rs, hipStream_t s) {
  // the tiled kernel: every term of energyGradient, uniform tensions, up to
  // four fluids, planes addressable with 32-bit byte offsets
  const bool tiled = !a.Fs && c.terms == kTermsEnergy && !c.raw && a.nf <= 4 &&
                     (double)a.Y * a.Z * a.nf * sizeof(T) < 4294967296.0;
  if (tiled) {
    switch (a.nf) {
      case 1: return ext_tile_nf<T, TV, 1>(a, mode, x, d, t, g, rs, s);
      case 2: return ext_tile_nf<T, TV, 2>(a, mode, x, d, t, g, rs, s);
      case 3: return ext_tile_nf<T, TV, 3>(a, mode, x, d, t, g, rs, s);
      default: return ext_tile_nf<T, TV, 4>(a, mode, x, d, t, g, rs, s);
    }
  }
  if (a.Fs) {  // per-site tensions: runtime fluid count
    ext_nf<T, TV, 0, true>(a, c, mode, x, d, t, g, rs, s);
    return a.nblocks;
  }
  switch (a.nf) {
    case 1: ext_nf<T, TV, 1, false>(a, c, mode, x, d, t, g, rs, s); break;
    case 2: ext_nf<T, TV, 2, false>(a, c, mode, x, d, t, g, rs, s); break;
    case 3: ext_nf<T, TV, 3, false>(a, c, mode, x, d, t, g, rs, s); break;
    case 4: ext_nf<T, TV, 4, false>(a, c, mode, x, d, t, g, rs, s); break;
    default: ext_nf<T, TV, 0, false>(a, c, mode, x, d, t, g, rs, s); break;
  }
  return a.nblocks;
}

template <typename T>
void volproj_t(const PFArgs& a, void* v, size_t nsite, bool norm, RedScratch rs, hipStream_t s) {
  T* u = static_cast<T*>(v);
  const int nb = grid_blocks(nsite, 4, 2048);
  RedScratch slot{rs.partials, rs.result + kPFVolSlot};
  switch (a.nf) {
    case 2: pf_volproj_reduce_kernel<T, 2><<<nb, kBlock, 0, s>>>(a, u, nsite, slot); break;
    case 3: pf_volproj_reduce_kernel<T, 3><<<nb, kBlock, 0, s>>>(a, u, nsite, slot); break;
    case 4: pf_volproj_reduce_kernel<T, 4><<<nb, kBlock, 0, s>>>(a, u, nsite, slot); break;
    default: pf_volproj_reduce_kernel<T, 0><<<nb, kBlock, 0, s>>>(a, u, nsite, slot); break;
  }
  const int K = (a.nf >= 2 && a.nf <= 4) ? a.nf : kPFMaxFluid;
  launch_finalize(slot, K, nb, s);
  const double* sums = rs.result + kPFVolSlot;
  if (norm) {
    switch (a.nf) {
      case 2: pf_volproj_apply_kernel<T, 2, true><<<nb, kBlock, 0, s>>>(a, u, nsite, sums, rs); break;
      case 3: pf_volproj_apply_kernel<T, 3, true><<<nb, kBlock, 0, s>>>(a, u, nsite, sums, rs); break;
      case 4: pf_volproj_apply_kernel<T, 4, true><<<nb, kBlock, 0, s>>>(a, u, nsite, sums, rs); break;
      default: pf_volproj_apply_kernel<T, 0, true><<<nb, kBlock, 0, s>>>(a, u, nsite, sums, rs); break;
    }
    launch_finalize(RedScratch{rs.partials, rs.result + 1}, 1, nb, s);
  } else {
    switch (a.nf) {
      case 2: pf_volproj_apply_kernel<T, 2, false><<<nb, kBlock, 0, s>>>(a, u, nsite, sums, rs); break;
      case 3: pf_volproj_apply_kernel<T, 3, false><<<nb, kBlock, 0, s>>>(a, u, nsite, sums, rs); break;
      case 4: pf_volproj_apply_kernel<T, 4, false><<<nb, kBlock, 0, s>>>(a, u, nsite, sums, rs); break;
      default: pf_volproj_apply_kernel<T, 0, false><<<nb, kBlock, 0, s>>>(a, u, nsite, sums, rs); break;
    }
  }
}

template <typename T>
void volumes_t(const PFArgs& a, const void* x, size_t nsite, RedScratch rs, hipStream_t s) {
  const T* xx = static_cast<const T*>(x);
  const int nb = grid_blocks(nsite, 4, 2048);
  switch (a.nf) {
    case 1: pf_volumes_kernel<T, 1><<<nb, kBlock, 0, s>>>(a, xx, nsite, rs); break;
    case 2: pf_volumes_kernel<T, 2><<<nb, kBlock, 0, s>>>(a, xx, nsite, rs); break;
    case 3: pf_volumes_kernel<T, 3><<<nb, kBlock, 0, s>>>(a, xx, nsite, rs); break;
    case 4: pf_volumes_kernel<T, 4><<<nb, kBlock, 0, s>>>(a, xx, nsite, rs); break;
    default: pf_volumes_kernel<T, 0><<<nb, kBlock, 0, s>>>(a, xx, nsite, rs); break;
  }
  const int K = (a.nf <= 4) ? a.nf : kPFMaxFluid;
  launch_finalize(rs, K, nb, s);
}

}  // namespace

int pf_launch_ext(const PFArgs& a, const PFCall& c, int prec, int pv, int mode, const void* x, const void* d,
                  double t, void* g, RedScratch rs, hipStream_t s) {
  if (!a.ext) throw std::runtime_error("PhaseField: general kernel constants missing");
  int nb;
  if (prec == F64 && pv == F64)
    nb = ext_t<double, double>(a, c, mode, x, d, t, g, rs, s);
  else if (prec == F32 && pv == F32)
    nb = ext_t<float, float>(a, c, mode, x, d, t, g, rs, s);
  else if (prec == F64 && pv == F32)
    nb = ext_t<double, float>(a, c, mode, x, d, t, g, rs, s);
  else
    throw std::invalid_argument("PhaseField: FP32 coordinates with FP64 vectors are not supported");
  MINIM_HIP_CHECK(hipGetLastError());
  return nb;
}

void pf_launch_volume_projection(const PFArgs& a, int prec, void* v, size_t nsite, bool norm, RedScratch rs,
                                 hipStream_t s) {
  if (!a.ext) throw std::runtime_error("PhaseField: general kernel constants missing");
  if (prec == F64) volproj_t<double>(a, v, nsite, norm, rs, s);
  else volproj_t<float>(a, v, nsite, norm, rs, s);
  MINIM_HIP_CHECK(hipGetLastError());
}

void pf_launch_volumes(const PFArgs& a, int prec, const void* x, size_t nsite, RedScratch rs, hipStream_t s) {
  if (!a.ext) throw std::runtime_error("PhaseField: general kernel constants missing");
  if (prec == F64) volumes_t<double>(a, x, nsite, rs, s);
  else volumes_t<float>(a, x, nsite, rs, s);
  MINIM_HIP_CHECK(hipGetLastError());
}

}  // namespace minim_amd
