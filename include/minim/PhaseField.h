// PhaseField.h -- device PhaseField potential, mirroring the reference's
// PhaseField (include/potentials/PhaseField.h:13-123) for bulk, gradient and
// density-constraint terms on a periodic grid without solid nodes.
#ifndef MINIM_AMD_PHASEFIELD_H
#define MINIM_AMD_PHASEFIELD_H

#include <string>
#include <vector>

#include "Potential.h"

namespace minim {

class PhaseField : public NewPotential<PhaseField> {
 public:
  int potentialType() const override { return Potential::GRID; }

  // System size
  int nFluid = 1;
  double resolution = 1;
  vector<int> gridSize;
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

  // Density constraint: "hard"/"gradient" (default), "soft"/"energy penalty", "none", "fixed"
  int densityConstraint = 0;
  double densityConst = 1;
  PhaseField& setDensityConstraint(std::string method);

  // Frozen fluids (gradient zeroed)
  vector<char> fixFluid;
  PhaseField& setFixFluid(int iFluid, bool fix = true);

  // Overrides
  void init(const vector<double>& coords) override;
  void evaluate(int prec, const void* x, const void* d, double t, double* e, void* g,
                void* stream) const override;
  void evaluateAsync(int prec, const void* x, const void* d, double t, bool wantE, double* e_dev, void* g,
                     void* stream) const override;
  bool supportsMixed() const override { return nFluid <= 4; }
  void evaluateMixed(const void* x, const void* d, double t, double* e, void* g, void* stream) const override;
  void evaluateMixedAsync(const void* x, const void* d, double t, bool wantE, double* e_dev, void* g,
                          void* stream) const override;
  double timeKernel(int prec, const void* x, bool wantE, void* g, int reps, void* stream) const override;
  void applyConstraints(int prec, void* v, size_t n, void* stream) const override;
  SiteProjection projection() const override;

  // Read only (after init)
  int nParams = 0;
  double surfaceTensionMean = 0;
  vector<double> kappa;
  vector<double> kappaP;
  int nGrid = 0;

  // Launch tuning: x-march length of the stencil and target block count
  // (0 = automatic for both).
  int stencilXChunk = 0;
  int stencilTargetBlocks = 0;

 private:
  void setDefaults();
  void checkArraySizes();
  void assignFluidCoefficients();
  std::shared_ptr<void> args_;  // PFArgs
};

}  // namespace minim

#endif
