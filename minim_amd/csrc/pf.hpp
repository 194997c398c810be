// pf.hpp -- launch interface of the PhaseField stencil kernels (pf.hip).
#pragma once

#include "common.hpp"

namespace minim_amd {

constexpr int kPFMaxFluid = 16;
enum { kDensityHard = 0, kDensitySoft = 1, kDensityNone = 2, kDensityFixed = 3 };

// Everything a stencil launch needs, passed by value as a kernel argument.
struct PFArgs {
  // canonical grid: size-1 dimensions first, Z contiguous
  int X, Y, Z, YZ;
  int actX, actY, actZ;
  int nf;
  double fb[kPFMaxFluid];  // bulk factor per fluid (kappa/16 v, or 0.5 kappa v)
  double fg[kPFMaxFluid];  // gradient factor per fluid ((0.25|0.5) kappa' v / res^2)
  // derived constants of the tiled kernel, FP64 and rounded to FP32:
  // fbg = 4 fb (one fluid) or 2 fb, fg2 = 2 fg, coef2 = 2 coef
  double fbg[kPFMaxFluid], fg2[kPFMaxFluid], coef2;
  float fb_f[kPFMaxFluid], fbg_f[kPFMaxFluid], fg_f[kPFMaxFluid], fg2_f[kPFMaxFluid];
  float coef_f, coef2_f, inv_nvar_f;
  unsigned fixmask;        // bit k set: fluid k fixed (gradient zeroed)
  int density;             // kDensity*
  double coef;             // soft-constraint coefficient
  double inv_nvar;         // 1 / number of variable fluids
  int nvar;
  // launch shape: TZ x TY tiles of the (y,z) plane, ntile of them, each
  // marching over nxc chunks of xchunk planes
  int TZ, TY, ntz, nty, ntile;
  int xchunk, nxc, nblocks;
  int xchunk_override;     // >0 forces the x-march length (autotuning)
  int target_blocks;       // >0 overrides the launch-size heuristic
};

// Fills YZ, act*, the derived constants and the launch shape of the generic
// kernel from X, Y, Z, nf, fb, fg, coef, inv_nvar.
void pf_args_finalize(PFArgs& a, int target_blocks);

// mode bits: 1 energy, 2 gradient (plus g.g), 4 evaluate at x + t*d.
// prec: coordinates (and arithmetic); pv: the d and g vectors (prec, or FP32
// with FP64 coordinates).  Results: rs.result[0] = energy, [1] = g.g (device).
void pf_launch_eval(const PFArgs& a, int prec, int pv, int mode, const void* x, const void* d, double t,
                    void* g, RedScratch rs, hipStream_t s, bool finalize = true);
void pf_launch_constrain(const PFArgs& a, int prec, void* v, size_t nsite, hipStream_t s);

}  // namespace minim_amd
