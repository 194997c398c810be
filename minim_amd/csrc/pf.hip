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
// Work decomposition: the grid is canonicalised (size-1 dimensions moved to the
// front, which keeps the linear index) to X x Y x Z with Z contiguous.  A thread
// owns one (y,z) column and marches along x over a chunk of planes keeping the
// x-1, x, x+1 values of all fluids in registers; y and z neighbours are
// re-reads of values that adjacent threads of the same block (or of the block
// on the neighbouring row, placed on the same XCD) load as their own centres,
// so they are served by L1/L2.  HBM traffic per evaluation is one read of the
// coordinates (two with a trial direction) and one write of the gradient.
#include "common.hpp"
#include "pf.hpp"

namespace minim_amd {

namespace {

template <int NF>
struct NFM {
  static constexpr int value = NF > 0 ? NF : kPFMaxFluid;
};

template <typename T, int NF>
__device__ inline void load_site(const T* __restrict__ x, const T* __restrict__ d, T t, size_t site,
                                 int nf, T (&v)[NFM<NF>::value], bool dir) {
  const size_t base = site * (size_t)nf;
  const int NK = NF > 0 ? NF : nf;
#pragma unroll
  for (int k = 0; k < NK; k++) {
    T c = x[base + k];
    if (dir) {
      T s = t * d[base + k];  // trial point x + t*d, rounded as x + (t*d)
      c = c + s;
    }
    v[k] = c;
  }
}

// MODE bits: 1 = energy, 2 = gradient (+ g.g), 4 = trial direction
template <typename T, int NF, int MODE>
__global__ __launch_bounds__(kBlock) void pf_kernel(PFArgs a, const T* __restrict__ x,
                                                    const T* __restrict__ d, T t, T* __restrict__ g,
                                                    RedScratch rs) {
  constexpr bool WE = MODE & 1, WG = MODE & 2, DIR = MODE & 4;
  constexpr int M = NFM<NF>::value;
  const int nf = NF > 0 ? NF : a.nf;
  const int NK = NF > 0 ? NF : nf;

  const int L = xcd_logical_block(blockIdx.x, gridDim.x);
  const int rc = L / a.nxc, xc = L - rc * a.nxc;
  const int p = rc * kBlock + threadIdx.x;

  double acc[2] = {0.0, 0.0};  // energy, g.g

  if (p < a.YZ) {
    T fb[M], fg[M];
#pragma unroll
    for (int k = 0; k < NK; k++) {
      fb[k] = (T)a.fb[k];
      fg[k] = (T)a.fg[k];
    }
    const T coef = (T)a.coef;
    const T inv_nvar = (T)a.inv_nvar;
    const int y = p / a.Z, z = p - y * a.Z;
    const int pzm = (z == 0) ? p + a.Z - 1 : p - 1;
    const int pzp = (z == a.Z - 1) ? p - (a.Z - 1) : p + 1;
    const int pym = (y == 0) ? p + (a.Y - 1) * a.Z : p - a.Z;
    const int pyp = (y == a.Y - 1) ? p - (a.Y - 1) * a.Z : p + a.Z;
    const size_t YZ = (size_t)a.YZ;
    const int x0 = xc * a.xchunk;
    const int x1 = min(x0 + a.xchunk, a.X);

    T cm[M], c[M], cp[M];
    load_site<T, NF>(x, d, t, (size_t)x0 * YZ + p, nf, c, DIR);
    if (a.actX) load_site<T, NF>(x, d, t, (size_t)(x0 == 0 ? a.X - 1 : x0 - 1) * YZ + p, nf, cm, DIR);

    for (int xi = x0; xi < x1; xi++) {
      const size_t plane = (size_t)xi * YZ;
      if (a.actX) load_site<T, NF>(x, d, t, (size_t)(xi == a.X - 1 ? 0 : xi + 1) * YZ + p, nf, cp, DIR);
      T ym[M], yp[M], zm[M], zp[M];
      if (a.actY) {
        load_site<T, NF>(x, d, t, plane + pym, nf, ym, DIR);
        load_site<T, NF>(x, d, t, plane + pyp, nf, yp, DIR);
      }
      if (a.actZ) {
        load_site<T, NF>(x, d, t, plane + pzm, nf, zm, DIR);
        load_site<T, NF>(x, d, t, plane + pzp, nf, zp, DIR);
      }

      T gv[M];
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
        T grad2 = (T)0, lap = (T)0;
        if (a.actX) {
          const T gm = ck - cm[k], gp = ck - cp[k];
          if (WE) grad2 += (T)0.5 * (gm * gm + gp * gp);
          if (WG) lap += gm + gp;
        }
        if (a.actY) {
          const T gm = ck - ym[k], gp = ck - yp[k];
          if (WE) grad2 += (T)0.5 * (gm * gm + gp * gp);
          if (WG) lap += gm + gp;
        }
        if (a.actZ) {
          const T gm = ck - zm[k], gp = ck - zp[k];
          if (WE) grad2 += (T)0.5 * (gm * gm + gp * gp);
          if (WG) lap += gm + gp;
        }
        if (WE) esite += (double)bulk_e + (double)(fg[k] * grad2);
        if (WG) gv[k] = bulk_g + (T)2 * fg[k] * lap;
        rho += ck;
      }
      if (a.density == kDensitySoft && nf > 1) {
        // densityConstraintEnergy (PhaseField.cpp:457-473)
        if (WE) esite += (double)(coef * (rho * rho));
        if (WG) {
          const T gr = (T)2 * coef * rho;
#pragma unroll
          for (int k = 0; k < NK; k++) {
            gv[k] += gr;
          }
        }
      }
      if (WE) acc[0] += esite;
      if (WG) {
        // specialisedConstraints: fixed fluids, then the hard density projection
        if (a.fixmask) {
#pragma unroll
          for (int k = 0; k < NK; k++) {
            if (a.fixmask & (1u << k)) gv[k] = (T)0;
          }
        }
        if (a.density == kDensityHard && nf > 1 && a.nvar > 0) {
          T comp = (T)0;
#pragma unroll
          for (int k = 0; k < NK; k++) {
            if (!(a.fixmask & (1u << k))) comp += gv[k];
          }
          comp *= inv_nvar;
#pragma unroll
          for (int k = 0; k < NK; k++) {
            if (!(a.fixmask & (1u << k))) gv[k] -= comp;
          }
        }
        const size_t base = (plane + p) * (size_t)nf;
#pragma unroll
        for (int k = 0; k < NK; k++) {
          g[base + k] = gv[k];
          acc[1] += (double)gv[k] * (double)gv[k];
        }
      }
      if (a.actX) {
#pragma unroll
        for (int k = 0; k < NK; k++) {
          cm[k] = c[k];
          c[k] = cp[k];
        }
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
      for (int k = 0; k < NK; k++) {
        if (!(a.fixmask & (1u << k))) comp += gv[k];
      }
      comp *= inv_nvar;
#pragma unroll
      for (int k = 0; k < NK; k++) {
        if (!(a.fixmask & (1u << k))) gv[k] -= comp;
      }
    }
#pragma unroll
    for (int k = 0; k < NK; k++) {
      v[base + k] = gv[k];
    }
  }
}

template <typename T, int NF>
void launch_eval_nf(const PFArgs& a, int mode, const void* x, const void* d, double t, void* g,
                    RedScratch rs, hipStream_t s, int nblocks) {
  const T* xx = static_cast<const T*>(x);
  const T* dd = static_cast<const T*>(d);
  T* gg = static_cast<T*>(g);
  const T tt = (T)t;
  switch (mode) {
    case 1: pf_kernel<T, NF, 1><<<nblocks, kBlock, 0, s>>>(a, xx, dd, tt, gg, rs); break;
    case 5: pf_kernel<T, NF, 5><<<nblocks, kBlock, 0, s>>>(a, xx, dd, tt, gg, rs); break;
    case 2: pf_kernel<T, NF, 2><<<nblocks, kBlock, 0, s>>>(a, xx, dd, tt, gg, rs); break;
    case 3: pf_kernel<T, NF, 3><<<nblocks, kBlock, 0, s>>>(a, xx, dd, tt, gg, rs); break;
    default: throw std::invalid_argument("pf: unsupported evaluation mode");
  }
}

template <typename T>
void launch_eval_t(const PFArgs& a, int mode, const void* x, const void* d, double t, void* g,
                   RedScratch rs, hipStream_t s, int nblocks) {
  switch (a.nf) {
    case 1: launch_eval_nf<T, 1>(a, mode, x, d, t, g, rs, s, nblocks); break;
    case 2: launch_eval_nf<T, 2>(a, mode, x, d, t, g, rs, s, nblocks); break;
    case 3: launch_eval_nf<T, 3>(a, mode, x, d, t, g, rs, s, nblocks); break;
    case 4: launch_eval_nf<T, 4>(a, mode, x, d, t, g, rs, s, nblocks); break;
    default: launch_eval_nf<T, 0>(a, mode, x, d, t, g, rs, s, nblocks); break;
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
  const int nrc = (a.YZ + kBlock - 1) / kBlock;
  int xchunk = a.xchunk_override > 0 ? a.xchunk_override : 1;
  if (a.xchunk_override <= 0) {
    // Longest x-march that still gives ~target_blocks blocks (halo planes are
    // re-read once per chunk, so longer chunks cost less traffic).
    const long want = target_blocks > 0 ? target_blocks : 2048;
    long nxc = (want + nrc - 1) / nrc;
    if (nxc < 1) nxc = 1;
    if (nxc > a.X) nxc = a.X;
    xchunk = (int)((a.X + nxc - 1) / nxc);
  }
  if (xchunk < 1) xchunk = 1;
  if (xchunk > a.X) xchunk = a.X;
  a.xchunk = xchunk;
  a.nxc = (a.X + xchunk - 1) / xchunk;
  while ((long)nrc * a.nxc > kMaxRedBlocks) {
    a.xchunk *= 2;
    a.nxc = (a.X + a.xchunk - 1) / a.xchunk;
  }
  a.nblocks = nrc * a.nxc;
}

void pf_launch_eval(const PFArgs& a, int prec, int mode, const void* x, const void* d, double t,
                    void* g, RedScratch rs, hipStream_t s) {
  if (prec == F64)
    launch_eval_t<double>(a, mode, x, d, t, g, rs, s, a.nblocks);
  else
    launch_eval_t<float>(a, mode, x, d, t, g, rs, s, a.nblocks);
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
