// pfu.hpp -- launch interface of the element-form PhaseField kernels (pfu.hip).
#pragma once

#include <cstdint>

#include "common.hpp"
#include "pf.hpp"

namespace minim_amd {

// Element types of PhaseFieldUnstructured (PhaseFieldUnstructured.cpp:22-31).
enum { kPFUFluid = 0, kPFUFluidAll = 1, kPFUPressure = 3, kPFUDensity = 4, kPFUSurface = 5, kPFUForce = 6,
       kPFUConfinement = 7 };
constexpr int kPFUNodes = 7;          // a site and its six face neighbours
constexpr int kPFUChunkDofs = 4096;   // DOFs (sites x fluids) one chunk may touch: its LDS accumulator
constexpr uint32_t kPFUShared = 0x80000000u;  // site-table flag: other chunks touch this site too

// Device element list.  Elements are grouped into chunks of nearby sites
// (one chunk per block at a time); a chunk lists the sites its elements touch,
// and an element names its nodes by their position in that list.  Record of
// one element (16 bytes): the seven local site indices (16 bits each), the
// type (8 bits) and the fluid (8 bits).  Surface, force and confinement
// elements carry their parameters in par, through pidx.
struct PFUArgs {
  size_t nel;
  int nchunk;
  const uint4* rec;           // nel
  const uint32_t* pidx;       // nel (read for surface, force and confinement elements)
  const double* par;          // 3 per parameterised element
  const uint32_t* chunk_el;   // nchunk + 1: element range of each chunk
  const uint32_t* chunk_site; // nchunk + 1: site-table range of each chunk
  const uint32_t* sites;      // site tables: global site index | kPFUShared
  // gradient partials of shared sites: one slot per site-table entry and
  // fluid (storage precision), summed per site in a fixed order by a second
  // pass over the sites that are not exclusive to one chunk (untouched ones
  // included: they are written as zero)
  void* scratch;
  size_t ngather;
  const uint32_t* gsite;      // ngather sites
  const uint32_t* gpos_off;   // ngather + 1
  const uint32_t* gpos;       // site-table positions of each gathered site, in chunk order
  const double* node_vol;     // ngrid
  int nf;
  size_t ngrid;
  double kappa[kPFMaxFluid], kappaP[kPFMaxFluid];
  double res2;             // resolution^2
  double coef_density;     // densityConst * mean tension * res^2
  double st_mean;          // mean surface tension (confinement)
  double pressure[kPFMaxFluid];
  // volume constraint as an energy term
  int volfix;
  double vol_coef;         // volConst * mean tension / res^4
  const double* volume;    // nf targets (device)
  double* volg;            // nf factors volCoef * volDiff (device scratch)
};

// Energy (rs.result[0]) and/or gradient (g, unconstrained) of every element
// and the volume term, at x or x + t*d.  prec is the storage precision of x,
// d and g.
void pfu_launch_eval(const PFUArgs& u, int prec, int mode, const void* x, const void* d, double t, void* g,
                     RedScratch rs, hipStream_t s);
// Average duration (ms) of reps launches of the element kernel (with the
// gather pass of the shared sites when mode has the gradient).
double pfu_time_elements(const PFUArgs& u, int prec, int mode, const void* x, void* g, int reps, RedScratch rs,
                         hipStream_t s);

}  // namespace minim_amd
