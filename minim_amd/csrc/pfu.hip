// pfu.hip -- the element form of the phase field (PhaseFieldUnstructured,
// src/potentials/PhaseFieldUnstructured.cpp) on MI355X (gfx950).
//
// The reference evaluates a list of elements one after another
// (State::elementEG, src/State.cpp:58-79): a seven-site gradient element per
// site and fluid (or per site for the "fixed" density mode, whose last fluid
// is implicit), one-DOF pressure, surface, force and confinement elements and
// a soft-density element per site, then the volume term over the whole field
// (blockEnergyGradient, :280-308).  Here the host groups the elements into
// chunks of nearby sites; a block takes one chunk at a time, one thread per
// element, accumulates the energy in FP64 for a deterministic two-stage
// reduction and the gradient contributions in an FP64 LDS accumulator
// (ds_add_f64), then stores the gradient of every site only its chunk touches
// and writes the partial sums of the sites other chunks also touch to a
// scratch slot; a second pass sums those partials per site in chunk order.
// No global atomics; the LDS adds of one chunk land in the order its threads
// reach them, so the gradient's last bits may vary from launch to launch.  Each element keeps the reference's formula,
// including its choices, so that the element form serves as the independent
// cross-check of the gather-form kernels it is in the reference.
#include "common.hpp"
#include "pfu.hpp"

namespace minim_amd {

namespace {

template <typename T, typename TV, bool DIR>
struct Coords {
  const T* __restrict__ x;
  const TV* __restrict__ d;
  T t;
  __device__ __forceinline__ T operator()(int i) const {
    T c = x[i];
    if constexpr (DIR) {
      const T s = t * (T)d[i];
      c = c + s;
    }
    return c;
  }
};

// The gradient term of one seven-site element (phaseGradient,
// PhaseFieldUnstructured.cpp:311-515): node 0 is the site, nodes (1,6), (2,5),
// (3,4) its x, y, z neighbours; self[i] marks a neighbour that is the site
// itself (the reference's stand-in for a solid neighbour).  Contributions to
// the nodes' gradients go to gl.
template <typename T, bool WE, bool WG>
__device__ __forceinline__ void phase7(const T (&c)[kPFUNodes], const bool (&self)[kPFUNodes], T factor, T& e,
                                       T (&gl)[kPFUNodes]) {
  const int mp[3][2] = {{1, 6}, {2, 5}, {3, 4}};
  T grad2 = (T)0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int m = mp[d][0], p = mp[d][1];
    if (!self[m] && !self[p]) {
      const T gm = c[0] - c[m], gp = c[0] - c[p];
      if (WE) grad2 = grad2 + (T)0.5 * (gm * gm + gp * gp);
      if (WG) {
        gl[0] = gl[0] + factor * (gm + gp);
        gl[m] = gl[m] - factor * gm;
        gl[p] = gl[p] - factor * gp;
      }
    } else if (!self[p]) {
      const T gp = c[0] - c[p];
      if (WE) grad2 = grad2 + gp * gp;
      if (WG) {
        gl[0] = gl[0] + (factor * (T)2) * gp;
        gl[p] = gl[p] - (factor * (T)2) * gp;
      }
    } else if (!self[m]) {
      const T gm = c[0] - c[m];
      if (WE) grad2 = grad2 + gm * gm;
      if (WG) {
        gl[0] = gl[0] + (factor * (T)2) * gm;
        gl[m] = gl[m] - (factor * (T)2) * gm;
      }
    }
  }
  if (WE) e = e + factor * grad2;
}

// MODE bits: 1 = energy, 2 = gradient, 4 = at x + t*d.  NF: the fluid count
// when it is 1..4 (index arithmetic with constants), 0 for u.nf.
template <typename T, typename TV, int MODE, int NF>
__global__ __launch_bounds__(kBlock) void pfu_chunk_kernel(PFUArgs u, const T* __restrict__ x,
                                                           const TV* __restrict__ d, T t, TV* __restrict__ g,
                                                           RedScratch rs) {
  constexpr bool WE = MODE & 1, WG = MODE & 2, DIR = MODE & 4;
  __shared__ uint32_t ssite[kPFUChunkDofs];
  __shared__ double sacc[WG ? kPFUChunkDofs : 1];
  const Coords<T, TV, DIR> cx{x, d, t};
  const int nf = NF > 0 ? NF : u.nf;
  // gradient contribution v to fluid k of local site l
  auto add = [&](int l, int k, T v) {
    if (WG && v != (T)0) unsafeAtomicAdd(&sacc[l * nf + k], (double)v);
  };
  auto gsite = [&](int l) { return (int)(ssite[l] & ~kPFUShared); };
  double acc[2] = {0.0, 0.0};
  for (int c = blockIdx.x; c < u.nchunk; c += gridDim.x) {
    const uint32_t s0 = u.chunk_site[c];
    const int ns = (int)(u.chunk_site[c + 1] - s0);
    for (int i = threadIdx.x; i < ns; i += kBlock) ssite[i] = u.sites[s0 + i];
    if (WG)
      for (int i = threadIdx.x; i < ns * nf; i += kBlock) sacc[i] = 0.0;
    __syncthreads();
    const uint32_t e1 = u.chunk_el[c + 1];
    for (uint32_t ie = u.chunk_el[c] + threadIdx.x; ie < e1; ie += kBlock) {
      const uint4 r = u.rec[ie];
      const int loc[kPFUNodes] = {(int)(r.x & 0xffffu), (int)(r.x >> 16), (int)(r.y & 0xffffu), (int)(r.y >> 16),
                                  (int)(r.z & 0xffffu), (int)(r.z >> 16), (int)(r.w & 0xffffu)};
      const int type = (int)((r.w >> 16) & 0xffu), f = (int)(r.w >> 24);
      const int site0 = gsite(loc[0]);
      T e = (T)0;
      switch (type) {
        case kPFUFluid: {  // fluidEnergy (:518-538)
          const T vol = (T)u.node_vol[site0];
          T c[kPFUNodes], gl[kPFUNodes];
          bool self[kPFUNodes];
#pragma unroll
          for (int i = 0; i < kPFUNodes; i++) {
            c[i] = cx(gsite(loc[i]) * nf + f);
            gl[i] = (T)0;
            self[i] = (i > 0) && loc[i] == loc[0];
          }
          const T c0 = c[0];
          if (nf == 1) {
            const T factor = ((T)u.kappa[0] / (T)16) * vol;
            if (WE) e = e + (factor * ((c0 + (T)1) * (c0 + (T)1))) * ((c0 - (T)1) * (c0 - (T)1));
            if (WG) gl[0] = gl[0] + ((factor * (T)4) * c0) * (c0 * c0 - (T)1);
          } else {
            const T factor = ((T)0.5 * (T)u.kappa[f]) * vol;
            if (WE) e = e + (factor * (c0 * c0)) * ((c0 - (T)1) * (c0 - (T)1));
            if (WG) gl[0] = gl[0] + (((factor * (T)2) * c0) * (c0 - (T)1)) * ((T)2 * c0 - (T)1);
          }
          T gfac = (nf == 1) ? ((T)0.25 * (T)u.kappaP[0]) * vol : ((T)0.5 * (T)u.kappaP[f]) * vol;
          gfac = gfac / (T)u.res2;
          phase7<T, WE, WG>(c, self, gfac, e, gl);
#pragma unroll
          for (int i = 0; i < kPFUNodes; i++) add(loc[i], f, gl[i]);
          break;
        }
        case kPFUFluidAll: {  // fluidEnergyAll (:541-586): the last fluid is 1 - the others
          const T vol = (T)u.node_vol[site0];
          int sn[kPFUNodes];
          T cN[kPFUNodes], gN[kPFUNodes];
          bool none[kPFUNodes];
#pragma unroll
          for (int i = 0; i < kPFUNodes; i++) {
            sn[i] = gsite(loc[i]) * nf;
            cN[i] = (T)1;
            gN[i] = (T)0;
            none[i] = false;  // the reference passes distinct indices here: no node counts as solid
          }
          for (int k = 0; k < nf - 1; k++)
#pragma unroll
            for (int i = 0; i < kPFUNodes; i++) cN[i] = cN[i] - cx(sn[i] + k);
          const T bfN = ((T)0.5 * (T)u.kappa[nf - 1]) * vol;
          const T gfN = ((T)0.5 * (T)u.kappaP[nf - 1]) * vol;  // not divided by res^2 (:555)
          if (WE) e = e + (bfN * (cN[0] * cN[0])) * ((cN[0] - (T)1) * (cN[0] - (T)1));
          if (WG) gN[0] = gN[0] + (((bfN * (T)2) * cN[0]) * (cN[0] - (T)1)) * ((T)2 * cN[0] - (T)1);
          phase7<T, WE, WG>(cN, none, gfN, e, gN);
          for (int k = 0; k < nf - 1; k++) {
            T c[kPFUNodes], gl[kPFUNodes];
            bool self[kPFUNodes];
#pragma unroll
            for (int i = 0; i < kPFUNodes; i++) {
              c[i] = cx(sn[i] + k);
              gl[i] = (T)0;
              self[i] = (i > 0) && loc[i] == loc[0];
            }
            const T c0 = c[0];
            const T bf = ((T)0.5 * (T)u.kappa[k]) * vol;
            if (WE) e = e + (bf * (c0 * c0)) * ((c0 - (T)1) * (c0 - (T)1));
            if (WG) gl[0] = gl[0] + (((bf * (T)2) * c0) * (c0 - (T)1)) * ((T)2 * c0 - (T)1);
            T gf = ((T)0.5 * (T)u.kappaP[k]) * vol;
            gf = gf / (T)u.res2;
            phase7<T, WE, WG>(c, self, gf, e, gl);
#pragma unroll
            for (int i = 0; i < kPFUNodes; i++) add(loc[i], k, gl[i] - gN[i]);
          }
          break;
        }
        case kPFUPressure: {  // pressureEnergy (:613-628)
          const T pr = (T)u.pressure[f];
          const T v = (T)u.node_vol[site0];
          const T c = cx(site0 * nf + f);
          if (nf == 1) {
            if (WE) e = e - pr * (((T)0.5 * (c + (T)1)) * v);
            add(loc[0], f, -(((T)0.5 * pr) * v));
          } else {
            if (WE) e = e - pr * (c * v);
            add(loc[0], f, -(pr * v));
          }
          break;
        }
        case kPFUDensity: {  // densityConstraintEnergy (:631-642): three DOFs of the site
          if (nf == 1) break;
          const T coef = (T)u.coef_density;
          const T rho = ((cx(site0 * nf) + cx(site0 * nf + 1)) + cx(site0 * nf + 2)) - (T)1;
          if (WE) e = e + coef * (rho * rho);
          const T gr = ((T)2 * coef) * rho;
          add(loc[0], 0, gr);
          add(loc[0], 1, gr);
          add(loc[0], 2, gr);
          break;
        }
        case kPFUSurface: {  // surfaceEnergy (:645-654): binary fluids only
          if (nf != 1) break;
          const double* p = u.par + 3 * (size_t)u.pidx[ie];
          const T area = (T)p[0], w = (T)p[1];
          const T phi = cx(site0 * nf + f);
          if (WE) e = e + (w * ((((phi * phi) * phi) / (T)3 - phi) - (T)2 / (T)3)) * area;
          add(loc[0], f, (w * (phi * phi - (T)1)) * area);
          break;
        }
        case kPFUForce: {  // forceEnergy (:657-676)
          const T c = (nf == 1) ? (T)0.5 * ((T)1 + cx(site0 * nf + f)) : cx(site0 * nf + f);
          if (c < (T)0.01) break;
          const double* p = u.par + 3 * (size_t)u.pidx[ie];
          const T vol = (T)p[0], fm = (T)p[1], h = (T)p[2];
          if (WE) e = e + ((c * vol) * fm) * h;
          add(loc[0], f, (nf == 1) ? ((((T)0.5 * vol) * fm) * h) : ((vol * fm) * h));
          break;
        }
        case kPFUConfinement: {  // ffConfinementEnergy (:679-692)
          const T c = cx(site0 * nf + f);
          const double* p = u.par + 3 * (size_t)u.pidx[ie];
          const T coef = ((T)p[0] * (T)u.st_mean) * (T)u.res2;
          const T c0 = (T)p[1];
          if ((c0 - (T)0.5) * (c - (T)0.5) < (T)0) {
            if (WE) e = e + coef * ((c - (T)0.5) * (c - (T)0.5));
            add(loc[0], f, (coef * (T)2) * (c - (T)0.5));
          }
          break;
        }
        default:
          break;
      }
      if (WE) acc[0] += (double)e;
    }
    __syncthreads();
    if (WG) {
      // sites only this chunk touches: the gradient; the others: partials
      TV* __restrict__ part = static_cast<TV*>(u.scratch) + (size_t)s0 * nf;
      for (int i = threadIdx.x; i < ns * nf; i += kBlock) {
        const int l = i / nf;
        const uint32_t en = ssite[l];
        const TV v = (TV)sacc[i];
        if (en & kPFUShared)
          part[i] = v;
        else
          g[(size_t)(en & ~kPFUShared) * nf + (i - l * nf)] = v;
      }
      __syncthreads();
    }
  }
  grid_sum_store<2>(acc, rs);
}

// The gradient of the sites that are not exclusive to one chunk: the sum of
// their partials in chunk order (zero for a site no element touches).
template <typename T>
__global__ __launch_bounds__(kBlock) void pfu_gather_kernel(PFUArgs u, T* __restrict__ g) {
  const int nf = u.nf;
  const T* __restrict__ part = static_cast<const T*>(u.scratch);
  const size_t n = u.ngather * nf;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
    const size_t j = i / nf;
    const int k = (int)(i - j * nf);
    const uint32_t p0 = u.gpos_off[j], p1 = u.gpos_off[j + 1];
    double sum = 0.0;
    for (uint32_t p = p0; p < p1; p++) sum += (double)part[(size_t)u.gpos[p] * nf + k];
    g[(size_t)u.gsite[j] * nf + k] = (T)sum;
  }
}

// Fluid volumes for the volume term (:284-293).
template <typename T, typename TV, bool DIR>
__global__ __launch_bounds__(kBlock) void pfu_volumes_kernel(PFUArgs u, const T* __restrict__ x,
                                                             const TV* __restrict__ d, T t, RedScratch rs) {
  const Coords<T, TV, DIR> cx{x, d, t};
  const int nf = u.nf;
  double acc[kPFMaxFluid];
#pragma unroll
  for (int k = 0; k < kPFMaxFluid; k++) acc[k] = 0.0;
  for (size_t s = (size_t)blockIdx.x * kBlock + threadIdx.x; s < u.ngrid; s += (size_t)gridDim.x * kBlock) {
    const double v = u.node_vol[s];
#pragma unroll
    for (int k = 0; k < kPFMaxFluid; k++) {
      if (k >= nf) continue;
      const double c = (double)cx((int)(s * nf + k));
      acc[k] += (nf == 1) ? (0.5 * (c + 1)) * v : c * v;
    }
  }
  grid_sum_store<kPFMaxFluid>(acc, rs);
}

// volCoef * volDiff per fluid, and (with the energy) the term's energy added to
// the elements' (:295-300).  One thread.
__global__ void pfu_volfix_kernel(PFUArgs u, const double* __restrict__ vol, double* __restrict__ energy,
                                  int want_energy) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double e = want_energy ? energy[0] : 0.0;
  for (int k = 0; k < u.nf; k++) {
    const double diff = vol[k] - u.volume[k];
    e += u.vol_coef * (diff * diff) / 1;
    u.volg[k] = u.vol_coef * diff;
  }
  if (want_energy) energy[0] = e;
}

// The volume term's gradient, weighted towards interfaces (:301-306).
template <typename T, typename TV, bool DIR>
__global__ __launch_bounds__(kBlock) void pfu_volgrad_kernel(PFUArgs u, const T* __restrict__ x,
                                                             const TV* __restrict__ d, T t, TV* __restrict__ g) {
  const Coords<T, TV, DIR> cx{x, d, t};
  const size_t nd = u.ngrid * u.nf;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < nd; i += (size_t)gridDim.x * kBlock) {
    const int k = (int)(i % u.nf);
    const T c = cx((int)i);
    T w = ((T)4 * c) * ((T)1 - c);
    if (w < (T)0) w = (T)0;
    const T add = (((T)u.volg[k] * (T)u.node_vol[i / u.nf]) * w);
    g[i] = (TV)((T)g[i] + add);
  }
}

template <typename T, int MODE>
void launch_chunk(const PFUArgs& u, int nb, const T* x, const T* d, T t, T* g, RedScratch rs, hipStream_t s) {
  switch (u.nf) {
    case 1: pfu_chunk_kernel<T, T, MODE, 1><<<nb, kBlock, 0, s>>>(u, x, d, t, g, rs); break;
    case 2: pfu_chunk_kernel<T, T, MODE, 2><<<nb, kBlock, 0, s>>>(u, x, d, t, g, rs); break;
    case 3: pfu_chunk_kernel<T, T, MODE, 3><<<nb, kBlock, 0, s>>>(u, x, d, t, g, rs); break;
    case 4: pfu_chunk_kernel<T, T, MODE, 4><<<nb, kBlock, 0, s>>>(u, x, d, t, g, rs); break;
    default: pfu_chunk_kernel<T, T, MODE, 0><<<nb, kBlock, 0, s>>>(u, x, d, t, g, rs); break;
  }
}

// Blocks of a chunk-kernel launch: each loops over chunks (at most
// kMaxRedBlocks blocks write reduction partials).
int chunk_blocks(const PFUArgs& u) { return u.nchunk < 1 ? 1 : (u.nchunk < 2048 ? u.nchunk : 2048); }

template <typename T>
void eval_t(const PFUArgs& u, int mode, const void* x, const void* d, double t, void* g, RedScratch rs,
            hipStream_t s) {
  const T* xx = static_cast<const T*>(x);
  const T* dd = static_cast<const T*>((mode & 4) ? d : x);
  T* gg = static_cast<T*>(g);
  const T tt = (T)t;
  const size_t nd = u.ngrid * u.nf;
  const int nb = chunk_blocks(u);
  switch (mode) {
    case 1: launch_chunk<T, 1>(u, nb, xx, dd, tt, gg, rs, s); break;
    case 2: launch_chunk<T, 2>(u, nb, xx, dd, tt, gg, rs, s); break;
    case 3: launch_chunk<T, 3>(u, nb, xx, dd, tt, gg, rs, s); break;
    case 5: launch_chunk<T, 5>(u, nb, xx, dd, tt, gg, rs, s); break;
    default: throw std::invalid_argument("pfu: unsupported evaluation mode");
  }
  if ((mode & 2) && u.ngather)
    pfu_gather_kernel<T><<<grid_blocks(u.ngather * u.nf, 1, 8192), kBlock, 0, s>>>(u, gg);
  launch_finalize(rs, 2, nb, s);
  if (u.volfix) {
    RedScratch slot{rs.partials, rs.result + kPFVolSlot};
    const int nbv = grid_blocks(u.ngrid, 4, 2048);
    if (mode & 4)
      pfu_volumes_kernel<T, T, true><<<nbv, kBlock, 0, s>>>(u, xx, dd, tt, slot);
    else
      pfu_volumes_kernel<T, T, false><<<nbv, kBlock, 0, s>>>(u, xx, dd, tt, slot);
    launch_finalize(slot, kPFMaxFluid, nbv, s);
    pfu_volfix_kernel<<<1, 64, 0, s>>>(u, rs.result + kPFVolSlot, rs.result, (mode & 1) ? 1 : 0);
    if (mode & 2) pfu_volgrad_kernel<T, T, false><<<grid_blocks(nd, 4, 4096), kBlock, 0, s>>>(u, xx, dd, tt, gg);
  }
}

}  // namespace

void pfu_launch_eval(const PFUArgs& u, int prec, int mode, const void* x, const void* d, double t, void* g,
                     RedScratch rs, hipStream_t s) {
  if ((mode & 4) && (mode & 2)) throw std::invalid_argument("pfu: gradients at a trial point are not supported");
  if (u.ngrid * (size_t)u.nf >= (size_t)1 << 31) throw std::invalid_argument("pfu: more than 2^31 DOFs");
  if (prec == F64) eval_t<double>(u, mode, x, d, t, g, rs, s);
  else eval_t<float>(u, mode, x, d, t, g, rs, s);
  MINIM_HIP_CHECK(hipGetLastError());
}

double pfu_time_elements(const PFUArgs& u, int prec, int mode, const void* x, void* g, int reps, RedScratch rs,
                         hipStream_t s) {
  const int nb = chunk_blocks(u);
  auto launch = [&]() {
    if (prec == F64) {
      const double* xx = static_cast<const double*>(x);
      if (mode == 1) {
        launch_chunk<double, 1>(u, nb, xx, xx, 0.0, (double*)g, rs, s);
      } else {
        launch_chunk<double, 3>(u, nb, xx, xx, 0.0, (double*)g, rs, s);
        if (u.ngather)
          pfu_gather_kernel<double><<<grid_blocks(u.ngather * u.nf, 1, 8192), kBlock, 0, s>>>(u, (double*)g);
      }
    } else {
      const float* xx = static_cast<const float*>(x);
      if (mode == 1) {
        launch_chunk<float, 1>(u, nb, xx, xx, 0.0f, (float*)g, rs, s);
      } else {
        launch_chunk<float, 3>(u, nb, xx, xx, 0.0f, (float*)g, rs, s);
        if (u.ngather)
          pfu_gather_kernel<float><<<grid_blocks(u.ngather * u.nf, 1, 8192), kBlock, 0, s>>>(u, (float*)g);
      }
    }
  };
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

}  // namespace minim_amd
