// pfu.hpp -- launch interface of the element-form PhaseField kernels (pfu.hip).
#pragma once

#include <cstdint>

#include "common.hpp"
#include "pf.hpp"

namespace minim_amd {

// Element types of PhaseFieldUnstructured (PhaseFieldUnstructured.cpp:22-31).
enum { kPFUFluid = 0, kPFUFluidAll = 1, kPFUPressure = 3, kPFUDensity = 4, kPFUSurface = 5, kPFUForce = 6,
       kPFUConfinement = 7 };
constexpr int kPFUNodes = 7;  // a site and its six face neighbours

// Device element list (structure of arrays) and the potential's constants.
struct PFUArgs {
  size_t nel;
  const int8_t* type;      // nel
  const int32_t* idof;     // nel * kPFUNodes (unused slots -1; the density element uses 3)
  const double* par;       // nel * 3
  const double* node_vol;  // ngrid
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
// d and g.  The gradient is accumulated with atomics.
void pfu_launch_eval(const PFUArgs& u, int prec, int mode, const void* x, const void* d, double t, void* g,
                     RedScratch rs, hipStream_t s);
// Average duration (ms) of reps launches of the element kernel alone.
double pfu_time_elements(const PFUArgs& u, int prec, int mode, const void* x, void* g, int reps, RedScratch rs,
                         hipStream_t s);

}  // namespace minim_amd
