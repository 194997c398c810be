// State.h -- a system state whose coordinates are resident in GPU memory,
// mirroring the reference's State (include/State.h:12-76).  On one GPU every
// block/proc/all variant is the whole domain, as the reference's single-rank
// Communicator makes them (src/Communicator.cpp:40-173).
#ifndef MINIM_AMD_STATE_H
#define MINIM_AMD_STATE_H

#include <memory>
#include <vector>

#include "Potential.h"

namespace minim {

class State {
 public:
  size_t ndof;
  double convergence;
  std::unique_ptr<Potential> pot;
  bool usesThisProc = true;

  // precision: FP64 or FP32 storage of the device coordinates
  State(const Potential& pot, const vector<double>& coords, int precision = FP64);
  State(const State& state);
  State& operator=(const State& state);
  ~State();

  // Energy / gradient (src/State.cpp:84-192)
  double energy() const;
  double energy(const vector<double>& coords) const;
  vector<double> gradient() const;
  vector<double> gradient(const vector<double>& coords) const;
  void energyGradient(double* e, vector<double>* g) const;
  void energyGradient(const vector<double>& coords, double* e, vector<double>* g) const;

  // Coordinates (src/State.cpp:277-304)
  double operator[](int i);
  vector<double> coords() const;
  void coords(const vector<double>& in);
  const vector<double>& blockCoords() const;
  void blockCoords(const vector<double>& in);

  // Single-GPU equivalents of the parallel variants
  double blockEnergy() const { return energy(); }
  double blockEnergy(const vector<double>& c) const { return energy(c); }
  vector<double> blockGradient() const { return gradient(); }
  vector<double> blockGradient(const vector<double>& c) const { return gradient(c); }
  void blockEnergyGradient(double* e, vector<double>* g) const { energyGradient(e, g); }
  void blockEnergyGradient(const vector<double>& c, double* e, vector<double>* g) const { energyGradient(c, e, g); }
  double procEnergy() const { return energy(); }
  double procEnergy(const vector<double>& c) const { return energy(c); }
  vector<double> procGradient() const { return gradient(); }
  vector<double> procGradient(const vector<double>& c) const { return gradient(c); }
  void procEnergyGradient(double* e, vector<double>* g) const { energyGradient(e, g); }
  void procEnergyGradient(const vector<double>& c, double* e, vector<double>* g) const { energyGradient(c, e, g); }
  double allEnergy() const { return energy(); }
  vector<double> allGradient() const { return gradient(); }
  void allEnergyGradient(double* e, vector<double>* g) const { energyGradient(e, g); }
  vector<double> allCoords() const { return coords(); }
  void communicate() {}

  // Constraints (src/State.cpp:321-323) on a host vector
  void applyConstraints(vector<double>& data) const;

  // Failure early completion (src/State.cpp:326-328)
  bool isFailed = false;
  void failed();

  // ---- device access ----
  int precision() const { return prec_; }
  void* deviceCoords() const;            // ndof entries of precision()
  void setPrecision(int precision);      // converts the device coordinates in place
  void* stream() const { return stream_; }
  void setStream(void* s) { stream_ = s; }
  void invalidateHost() { hostValid_ = false; }
  // incremented whenever coordinates are replaced from outside a minimiser
  unsigned long coordsVersion() const { return version_; }
  // A device buffer of the coordinates' size and precision, and the exchange
  // of the coordinates with such a buffer (a minimiser's accepted step written
  // by a fused evaluation; deviceCoords() changes).
  std::shared_ptr<void> newCoordsBuffer() const;
  void swapDeviceCoords(std::shared_ptr<void>& buffer);

 private:
  int prec_ = FP64;
  std::shared_ptr<void> x_;  // device coordinates
  void* stream_ = nullptr;
  mutable vector<double> host_;
  mutable bool hostValid_ = false;
  unsigned long version_ = 0;
};

}  // namespace minim

#endif
