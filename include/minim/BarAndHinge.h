// BarAndHinge.h -- device bar-and-hinge elastic sheet, mirroring the
// reference's BarAndHinge (include/potentials/BarAndHinge.h:11-70) for bars,
// hinges and the external z force (element types 0, 1, 2).  The substrate wall
// (type 3) and thickness-derived rigidities are not provided.
#ifndef MINIM_AMD_BARANDHINGE_H
#define MINIM_AMD_BARANDHINGE_H

#include <vector>

#include "Potential.h"

namespace minim {

class BarAndHinge : public NewPotential<BarAndHinge> {
 public:
  int potentialType() const override { return Potential::UNSTRUCTURED; }
  bool deviceNative() const override { return true; }
  // element types 0 bars, 1 hinges, 2 external force (BarAndHinge.cpp:35-66)
  double componentEnergy(int component, const vector<double>& coords) const override;

  BarAndHinge& setTriangulation(const vector2d<int>& triList);
  BarAndHinge& setBondList(const vector2d<int>& bondList);
  BarAndHinge& setHingeList(const vector2d<int>& hingeList);
  BarAndHinge& setRigidity(double kBond, double kHinge);
  BarAndHinge& setRigidity(const vector<double>& kBond, const vector<double>& kHinge);
  BarAndHinge& setLength0(double length0);
  BarAndHinge& setLength0(const vector<double>& length0);
  BarAndHinge& setTheta0(double theta0);
  BarAndHinge& setTheta0(const vector<double>& theta0);
  BarAndHinge& setForce(const vector<double>& force);

  void init(const vector<double>& coords) override;
  void evaluate(int prec, const void* x, const void* d, double t, double* e, void* g,
                void* stream) const override;
  void evaluateAsync(int prec, const void* x, const void* d, double t, bool wantE, double* e_dev, void* g,
                     void* stream) const override;
  // FP64 coordinates and arithmetic with FP32 direction and gradient (the
  // mixed schedule's second phase); not with normal constraints.
  bool supportsMixed() const override;
  void evaluateMixed(const void* x, const void* d, double t, double* e, void* g, void* stream) const override;
  void evaluateMixedAsync(const void* x, const void* d, double t, bool wantE, double* e_dev, void* g,
                          void* stream) const override;
  // Fused accepted step: x + t*d into xout (formed as the line-search trials
  // form it), then energy and gradient there in one energy pass that keeps
  // what the gradient gather needs; d and g in precision pv (px, or FP32
  // against FP64 coordinates without normal constraints).
  bool supportsStep() const override;
  bool stepDefersGradient() const override;
  void evaluateStepAsync(int px, int pv, const void* x, const void* d, double t, void* xout, void* g,
                         void* stream) const override;
  double timeKernel(int prec, const void* x, bool wantE, void* g, int reps, void* stream) const override;

  // Element lists after init, in the reference's element order.
  vector2d<int> bondList() const { return bonds_; }
  vector2d<int> hingeList() const { return hinges_; }
  vector<double> kBondList() const { return kBond_; }
  vector<double> kHingeList() const { return kHinge_; }
  vector<double> length0List() const { return length0_; }
  vector<double> theta0List() const { return theta0_; }
  // Overwrite one hinge's rest angle (the reference's tests edit element parameters).
  void setHingeTheta0(int iHinge, double theta0);

  int nNodeHint = 0;

  struct Dev;

 private:
  vector2d<int> tri_, bonds_, hinges_;
  vector<double> kBond_, kHinge_, length0_, theta0_, force_;
  std::shared_ptr<Dev> dev_;
  void upload();
};

}  // namespace minim

#endif
