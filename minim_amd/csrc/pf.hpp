// pf.hpp -- launch interface of the PhaseField stencil kernels (pf.hip).
#pragma once

#include "common.hpp"

namespace minim_amd {

constexpr int kPFMaxFluid = 16;
enum { kDensityHard = 0, kDensitySoft = 1, kDensityNone = 2, kDensityFixed = 3 };

// Site kinds of the general kernel: 0 = solid, 1 + the reference's surface
// type 0..7 (PhaseField::getType, PhaseField.cpp:899-958).
constexpr int kPFKinds = 9;

// Per-site code (uint16): bits 0-3 the kind; for canonical dimension j, bits
// 4+4j..5+4j and 6+4j..7+4j weight the bond to the minus and the plus
// neighbour: 0, 1 (one half) or 2 (one).  A bond to a solid site weighs 0;
// the opposite bond of a site whose other neighbour is solid weighs 1
// (phaseGradient's one-sided branches, PhaseField.cpp:265-286).
constexpr int kPFCodeShift = 4;
// Forward-bond code (uint32) of the tiled general kernel: bits 0-3 the kind;
// for canonical dimension j, 6 bits at 4+6j hold 16 w(i, i+e_j), the weight
// (v_i h_i->j + v_j h_j->i) / res^3 of the bond to the plus neighbour (0..32).
constexpr int kPFWeightShift = 4;
constexpr int kPFWeightBits = 6;

// Constants of the general kernel, resident in device memory.
struct PFExt {
  double B[kPFMaxFluid][kPFKinds];  // bulk factor (kappa/16 | 0.5 kappa) * node volume
  double F[kPFMaxFluid][kPFKinds];  // gradient factor (0.25 | 0.5) kappa' * node volume / res^2
  double vol[kPFKinds];             // node volume
  double area[kPFKinds];            // solid surface area
  double pres[kPFMaxFluid];         // pressure; 0 when unset or the fluid is fixed
  double fmag[kPFMaxFluid];         // force magnitude
  double fn[kPFMaxFluid][3];        // force direction in canonical dimensions
  double fc[3];                     // grid centre (N - 1) / 2 in canonical dimensions
  double res;
  double conf[kPFMaxFluid];         // confinement strength * mean tension * res^2
};

// Term selection of one launch of the general kernel.
enum {
  kTermFluid = 1,
  kTermSurface = 2,
  kTermPressure = 4,
  kTermDensity = 8,
  kTermForce = 16,
  kTermConfinement = 32,
  kTermVolume = 64,
  kTermsEnergy = 63  // everything PhaseField::energyGradient adds
};
struct PFCall {
  unsigned terms;
  int raw;                   // 1: no constraints applied to the gradient
  double volg[kPFMaxFluid];  // volume-constraint factor volCoef * volDiff per fluid
};

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
  int xchunk_override;     // >0 forces the x-march length (the caller's setting)
  int target_blocks;       // >0 overrides the launch-size heuristic (the caller's setting)
  // launch shape chosen by the autotuner per coordinate precision
  // ([0] FP64, [1] FP32), used where the caller set none; 0 = heuristic
  int tuned_target[2];
  int tuned_xchunk[2];
  // general kernel: used when any solid node or extra term is set (or forced)
  int use_ext;
  const PFExt* ext;        // device
  const PFExt* ext_host;   // the same constants in host memory (launch set-up)
  const uint16_t* code;    // per-site code, or NULL: ucode at every site
  unsigned ucode;
  const uint32_t* wcode;   // per-site forward-bond code, or NULL: wucode at every site
  unsigned wucode;
  const double* wet;       // per-site wetting parameter (binary fluids), or NULL
  const int8_t* csign;     // per-DOF sign of (initial c - 0.5), or NULL (no confinement)
  const double* Bs;        // per-DOF bulk factors with per-site tensions, or NULL
  const double* Fs;        // per-DOF gradient factors with per-site tensions, or NULL
  int any_pres, any_force;
  // fixed-volume projection after the per-site one
  int volfix;
  double tv2;              // sum of squared node volumes
};

// Fills YZ, act*, the derived constants and the launch shape of the generic
// kernel from X, Y, Z, nf, fb, fg, coef, inv_nvar.
void pf_args_finalize(PFArgs& a, int target_blocks);

// mode bits: 1 energy, 2 gradient (plus g.g), 4 evaluate at x + t*d.
// prec: coordinates (and arithmetic); pv: the d and g vectors (prec, or FP32
// with FP64 coordinates).  Results: rs.result[0] = energy, [1] = g.g (device).
// Runs the general kernel when a.use_ext; with fixed volumes and a gradient
// the volume projection follows (not when finalize is false: timing only).
// With xout (mode 7, tiled kernel only), the point x + t*d is also written to
// xout (precision prec): an accepted step and its gradient in one pass.
// evBefore / evAfter, when given, are recorded on s right before and right
// after the evaluation kernel (benchmark timing of the kernel inside a step).
// eOut (with finalize): the energy is also written to this device slot.
void pf_launch_eval(const PFArgs& a, int prec, int pv, int mode, const void* x, const void* d, double t,
                    void* g, RedScratch rs, hipStream_t s, bool finalize = true, void* xout = nullptr,
                    hipEvent_t evBefore = nullptr, hipEvent_t evAfter = nullptr, double* eOut = nullptr);
// Potential::applyConstraints: fixed fluids, hard density projection, then
// (fixed volumes) the global projection, which uses rs as
// pf_launch_volume_projection does.
void pf_launch_constrain(const PFArgs& a, int prec, void* v, size_t nsite, RedScratch rs, hipStream_t s);

// General kernel (pf_ext.hip): the terms selected by c, gradient projected
// per site unless c.raw.  Results as pf_launch_eval; returns the block count.
int pf_launch_ext(const PFArgs& a, const PFCall& c, int prec, int pv, int mode, const void* x, const void* d,
                  double t, void* g, RedScratch rs, hipStream_t s);
// Fixed-volume projection of v (specialisedConstraints, PhaseField.cpp:614-658),
// after the per-site one.  Uses rs.partials and rs.result[80, 96); with norm,
// rs.result[1] = v.v afterwards.
constexpr int kPFVolSlot = 80;
void pf_launch_volume_projection(const PFArgs& a, int prec, void* v, size_t nsite, bool norm, RedScratch rs,
                                 hipStream_t s);
// Fluid volumes (volumeConstraintEnergy, PhaseField.cpp:540-553):
// rs.result[k] = sum_i c_ik v_i (binary fluids: (c+1)/2 v).
void pf_launch_volumes(const PFArgs& a, int prec, const void* x, size_t nsite, RedScratch rs, hipStream_t s);

}  // namespace minim_amd
