// PhaseField.h -- device PhaseField potential, mirroring the reference's
// PhaseField (include/potentials/PhaseField.h:13-123): bulk, gradient and
// density-constraint terms, solid nodes with wetting, pressure, fixed volumes,
// external forces, confinement and per-site surface tensions on a periodic grid.
#ifndef MINIM_AMD_PHASEFIELD_H
#define MINIM_AMD_PHASEFIELD_H

#include <array>
#include <functional>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "Potential.h"

namespace minim {

class PhaseField : public NewPotential<PhaseField> {
 public:
  int potentialType() const override { return Potential::GRID; }
  bool deviceNative() const override { return true; }

  // System size
  int nFluid = 1;
  double resolution = 1;
  PhaseField& setNFluid(int nFluid);
  PhaseField& setGridSize(vector<int> gridSize);
  PhaseField& setResolution(double resolution);

  // Fluid interfaces.  Order: 1-2, 1-3, ..., 2-3 (include/potentials/PhaseField.h:25)
  vector<double> interfaceSize;
  vector<double> surfaceTension = {1};
  PhaseField& setInterfaceSize(double interfaceSize);
  PhaseField& setInterfaceSize(vector<double> interfaceSize);
  PhaseField& setSurfaceTension(double surfaceTension);
  PhaseField& setSurfaceTension(vector<double> surfaceTension);
  // per-site tensions: fn(x, y, z) gives the nParams values of a site
  PhaseField& setSurfaceTension(std::function<vector<double>(int, int, int)> surfaceTensionFn);

  // Density constraint: "hard"/"gradient" (default), "soft"/"energy penalty", "none", "fixed"
  int densityConstraint = 0;
  double densityConst = 1;
  PhaseField& setDensityConstraint(std::string method);

  // Total volume / pressure constraints (PhaseField.h:37-44)
  bool volumeFixed = false;
  double volConst = 0.01;
  vector<double> pressure;
  vector<double> volume;
  PhaseField& setPressure(vector<double> pressure);
  PhaseField& setVolume(vector<double> volume, double volConst = 0.01);
  PhaseField& setVolumeFixed(bool volumeFixed, double volConst = 0.01);

  // Solid nodes (PhaseField.h:47-54)
  vector<char> solid;
  vector<double> contactAngle;
  PhaseField& setSolid(vector<char> solid);
  PhaseField& setSolid(std::function<bool(int, int, int)> solidFn);
  PhaseField& setContactAngle(double contactAngle);
  PhaseField& setContactAngle(vector<double> contactAngle);
  PhaseField& setContactAngle(std::function<double(int, int, int)> contactAngleFn);

  // External force (PhaseField.h:57-58)
  vector<vector<double>> force;
  PhaseField& setForce(vector<double> force, vector<int> iFluid = {});

  // Frozen fluids (gradient zeroed) and the diffuse-solid method (PhaseField.h:61-68)
  vector<char> fixFluid;
  vector<double> confinementStrength;
  PhaseField& setFixFluid(int iFluid, bool fix = true);
  PhaseField& setConfinement(vector<double> strength);
  vector<double> diffuseSolid(vector<char> solid, int iFluid = 0, bool twoStep = false);
  static vector<double> diffuseSolid(vector<char> solid, PhaseField potential, int iFluid = 0, bool twoStep = false);
  static vector<double> diffuseSolid(vector<char> solid, vector<int> gridSize, int nFluid = 2, int iFluid = 0,
                                     bool twoStep = false);

  // Overrides
  void init(const vector<double>& coords) override;
  void evaluate(int prec, const void* x, const void* d, double t, double* e, void* g,
                void* stream) const override;
  void evaluateAsync(int prec, const void* x, const void* d, double t, bool wantE, double* e_dev, void* g,
                     void* stream) const override;
  bool supportsMixed() const override;
  bool supportsStep() const override;
  void evaluateStepAsync(int px, int pv, const void* x, const void* d, double t, void* xout, void* g,
                         void* stream) const override;
  void evaluateMixed(const void* x, const void* d, double t, double* e, void* g, void* stream) const override;
  void evaluateMixedAsync(const void* x, const void* d, double t, bool wantE, double* e_dev, void* g,
                          void* stream) const override;
  double timeKernel(int prec, const void* x, bool wantE, void* g, int reps, void* stream) const override;
  void timeSteps(int prec, const void* x, void* g, double* e_dev, int steps, double warmMs, void* stream,
                 double* out) const override;
  // evaluateAsync with the cross-block reduction in device scratch the caller
  // provides (partial sums in scratch[0, nPartials), 96 results after them);
  // no device memory is allocated.  Throws std::invalid_argument when the
  // potential's tables are not resident yet or the scratch is too small.
  void evaluateAsyncScratch(int prec, const void* x, const void* d, double t, bool wantE, double* e_dev, void* g,
                            double* scratch, size_t nPartials, void* stream) const;

  // Launch-shape autotuner of the evaluation kernel this field uses (after
  // init): times one energy+gradient launch (HIP events, kernel only, the
  // average of reps) at coords for every pair of target block count and
  // x-march length (0 = automatic), keeps the fastest for precision prec (FP64
  // or FP32) and stores it in the process-wide cache, where every later
  // PhaseField of the same shape and kernel finds it at init.  Returns rows
  // (target, march, ms).  stencilTargetBlocks / stencilXChunk, when set, still
  // take precedence.  The shape changes the tiling only: gradients are the
  // same bit for bit.
  std::vector<std::array<double, 3>> autotune(const vector<double>& coords, int prec,
                                              vector<int> targets = {256, 512, 1024, 2048},
                                              vector<int> marches = {0, 4, 8, 16, 32}, int reps = 10);
  // The tuned (target blocks, x-march) for a precision; (0, 0) = the heuristic.
  std::pair<int, int> tunedShape(int prec) const;
  void constrainDevice(int prec, const void* x, void* v, size_t n, void* stream) const override;
  SiteProjection projection() const override;
  bool hasGlobalConstraints() const override { return volumeFixed; }
  // One fluid on a periodic grid with no solid nodes, extra terms or fixed
  // fluid, after init: the canonical grid (size-1 dimensions first, the last
  // one contiguous) and the bulk and gradient factors of the plain stencil
  // (kappa/16 v and 0.25 kappa' v / res^2, PhaseField.cpp:372-390), as used by
  // Lbfgs::minimiseBatch.  False for any other field.
  bool plainStencil(int canon[3], double& fb, double& fg) const;

  // PhaseField::energyComponents (PhaseField.cpp:706-744): energy and gradient
  // norm (unconstrained) of each term at the given coordinates, computed on
  // the GPU in FP64.
  std::map<std::string, vector<double>> energyComponents(const vector<double>& coords) const;

  // Read only (after init)
  int nParams = 0;
  double surfaceTensionMean = 0;
  vector<double> kappa;   // nParams, or nGrid * nParams with per-site tensions
  vector<double> kappaP;
  int nGrid = 0;
  vector<double> nodeVol;      // per site, filled when solid nodes are present
  vector<double> surfaceArea;  // (uniform otherwise: resolution^3 and 0)
  double totalVolume2 = 0;     // sum of squared node volumes (fixed volume)
  vector<double> fMag;
  vector2d<double> fNorm;

  // Launch tuning: x-march length of the stencil and target block count
  // (0 = automatic for both); stencilKernel 1 forces the general kernel
  // (solid-aware, every term) where the tiled one would run.
  int stencilXChunk = 0;
  int stencilTargetBlocks = 0;
  int stencilKernel = 0;

 private:
  void setDefaults();
  void checkArraySizes();
  void assignFluidCoefficients();
  std::shared_ptr<void> args_;  // PFArgs
};

// PhaseField::getType (PhaseField.cpp:899-958) of every site of a periodic
// grid: -1 solid, 0 bulk fluid, 1..7 surface types.  Throws
// std::runtime_error where the reference finds an undefined surface type.
std::vector<int> pf_site_types(const int grid[3], const char* solid);

}  // namespace minim

#endif
