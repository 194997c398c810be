// PhaseFieldUnstructured.h -- the element form of the phase field on the GPU,
// mirroring the reference's PhaseFieldUnstructured
// (include/potentials/PhaseFieldUnstructured.h:13-104): the same setters,
// elements built as the reference builds them, evaluated one element per GPU
// thread.  It exists beside PhaseField as the reference keeps it: an
// independent formulation of the same physics.
#ifndef MINIM_AMD_PHASEFIELDUNSTRUCTURED_H
#define MINIM_AMD_PHASEFIELDUNSTRUCTURED_H

#include <functional>
#include <string>
#include <vector>

#include "Potential.h"

namespace minim {

class PhaseFieldUnstructured : public NewPotential<PhaseFieldUnstructured> {
 public:
  int potentialType() const override { return Potential::UNSTRUCTURED; }
  bool deviceNative() const override { return true; }

  // System size
  int nFluid = 1;
  double resolution = 1;
  PhaseFieldUnstructured& setNFluid(int nFluid);
  PhaseFieldUnstructured& setGridSize(vector<int> gridSize);
  PhaseFieldUnstructured& setResolution(double resolution);

  // Fluid interfaces
  vector<double> interfaceSize;
  vector<double> surfaceTension = {1};
  PhaseFieldUnstructured& setInterfaceSize(double interfaceSize);
  PhaseFieldUnstructured& setInterfaceSize(vector<double> interfaceSize);
  PhaseFieldUnstructured& setSurfaceTension(double surfaceTension);
  PhaseFieldUnstructured& setSurfaceTension(vector<double> surfaceTension);

  // Density constraint: "hard"/"gradient" (default), "soft"/"energy penalty", "none", "fixed"
  int densityConstraint = 0;
  double densityConst = 1;
  PhaseFieldUnstructured& setDensityConstraint(std::string method);

  // Total volume (an energy term here) / pressure
  bool volumeFixed = false;
  double volConst = 0.01;
  vector<double> pressure;
  vector<double> volume;
  PhaseFieldUnstructured& setPressure(vector<double> pressure);
  PhaseFieldUnstructured& setVolume(vector<double> volume, double volConst = 0.01);
  PhaseFieldUnstructured& setVolumeFixed(bool volumeFixed, double volConst = 0.01);

  // Solid nodes
  vector<char> solid;
  vector<double> contactAngle;
  PhaseFieldUnstructured& setSolid(vector<char> solid);
  PhaseFieldUnstructured& setSolid(std::function<bool(int, int, int)> solidFn);
  PhaseFieldUnstructured& setContactAngle(vector<double> contactAngle);
  PhaseFieldUnstructured& setContactAngle(std::function<double(int, int, int)> contactAngleFn);

  // External force
  vector<vector<double>> force;
  PhaseFieldUnstructured& setForce(vector<double> force, vector<int> iFluid = {});

  // Diffuse solid method
  vector<char> fixFluid;
  vector<double> confinementStrength;
  PhaseFieldUnstructured& setFixFluid(int iFluid, bool fix = true);
  PhaseFieldUnstructured& setConfinement(vector<double> strength);
  vector<double> diffuseSolid(vector<char> solid, int iFluid = 0, bool twoStep = false);
  static vector<double> diffuseSolid(vector<char> solid, PhaseFieldUnstructured potential, int iFluid = 0,
                                     bool twoStep = false);
  static vector<double> diffuseSolid(vector<char> solid, vector<int> gridSize, int nFluid = 2, int iFluid = 0,
                                     bool twoStep = false);

  // Overrides
  void init(const vector<double>& coords) override;
  void evaluate(int prec, const void* x, const void* d, double t, double* e, void* g,
                void* stream) const override;
  void evaluateAsync(int prec, const void* x, const void* d, double t, bool wantE, double* e_dev, void* g,
                     void* stream) const override;
  double timeKernel(int prec, const void* x, bool wantE, void* g, int reps, void* stream) const override;
  void constrainDevice(int prec, const void* x, void* v, size_t n, void* stream) const override;
  double componentEnergy(int component, const vector<double>& coords) const override;
  SiteProjection projection() const override;

  // Read only (after init)
  int nGrid = 0;
  double surfaceTensionMean = 0;
  vector<double> kappa;
  vector<double> kappaP;
  vector<double> nodeVol;  // per DOF, as the reference's initLocal stores it
  vector<int> fluidType;   // per DOF
  size_t nElements = 0;    // energy elements built by init
  size_t listBytes = 0;    // device bytes of the element list (records, parameters, chunk and site tables)

 private:
  void setDefaults();
  void checkArraySizes();
  void assignFluidCoefficients();
  std::shared_ptr<void> dev_;
};

}  // namespace minim

#endif
