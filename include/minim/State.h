// State.h -- a system state whose coordinates are resident in GPU memory,
// mirroring the reference's State (include/State.h:12-76).  On one GPU every
// block/proc/all variant is the whole domain, as the reference's single-rank
// Communicator makes them (src/Communicator.cpp:40-173).
#ifndef MINIM_AMD_STATE_H
#define MINIM_AMD_STATE_H

#include <memory>
#include <vector>

#include "Communicator.h"
#include "Potential.h"

namespace minim {

class State;

// The reference's public host copy of the coordinates (State::_coords,
// include/State.h:76), kept here as a view of the device coordinates: a read
// copies them to the host once after each change, an assignment uploads.
class HostCoords {
 public:
  explicit HostCoords(State* s) : s_(s) {}
  HostCoords(const HostCoords&) = delete;  // bound to its State
  HostCoords& operator=(const HostCoords& o);
  HostCoords& operator=(const vector<double>& v);
  operator const vector<double>&() const { return get(); }
  const vector<double>& get() const;
  double operator[](size_t i) const { return get()[i]; }
  size_t size() const { return get().size(); }
  bool empty() const { return get().empty(); }
  const double* data() const { return get().data(); }
  vector<double>::const_iterator begin() const { return get().begin(); }
  vector<double>::const_iterator end() const { return get().end(); }

 private:
  State* s_;
};

class State {
 public:
  size_t ndof;
  double convergence;
  std::unique_ptr<Potential> pot;
  std::unique_ptr<Communicator> comm;
  bool usesThisProc = true;

  // State(pot, coords, ranks) (src/State.cpp:12-26): clones and initialises
  // the potential, sets up the communicator, uploads coords in FP64.  This
  // process is rank 0 of one: a non-empty rank list without 0 leaves the
  // State unused here (usesThisProc false), as in the reference.
  State(const Potential& pot, const vector<double>& coords, const vector<int>& ranks = {});
  // The same with the device coordinates stored in the given precision.
  State(const Potential& pot, const vector<double>& coords, Precision precision, const vector<int>& ranks = {});
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

  // Energy of the potential's elements of one type (src/State.cpp:307-312).
  double componentEnergy(int component) const;

  void communicate() {}

  // Constraints (src/State.cpp:321-323) on a host vector
  void applyConstraints(vector<double>& data) const;

  // Failure early completion (src/State.cpp:326-328)
  bool isFailed = false;
  void failed();

  HostCoords _coords{this};

  // Energy and/or gradient on raw host arrays of ndof doubles (coords null:
  // the State's own coordinates; e or g null: not computed), the form the
  // vector overloads above and the C ABI share.
  void energyGradientRaw(const double* coords, double* e, double* g) const;

  // ---- device access ----
  int precision() const { return prec_; }
  void* deviceCoords() const;            // ndof entries of precision()
  void setPrecision(int precision);      // converts the device coordinates in place
  void setPrecision(Precision precision) { setPrecision(static_cast<int>(precision)); }
  void* stream() const { return stream_; }
  void setStream(void* s) { stream_ = s; }
  // Declares that the device coordinates changed outside this class (a write
  // through deviceCoords(), a minimiser's step): the host copy is refreshed on
  // its next read and cached energies of the old coordinates are dropped.
  void markCoordsChanged() {
    hostValid_ = false;
    version_++;
  }
  void invalidateHost() { markCoordsChanged(); }
  // Changes whenever the coordinates change.
  unsigned long coordsVersion() const { return version_; }
  // A device buffer of the coordinates' size and precision, and the exchange
  // of the coordinates with such a buffer (a minimiser's accepted step written
  // by a fused evaluation; deviceCoords() changes).
  std::shared_ptr<void> newCoordsBuffer() const;
  void swapDeviceCoords(std::shared_ptr<void>& buffer);

 private:
  void construct(const Potential& p, const vector<double>& coords, int precision, const vector<int>& ranks);
  void bindPotential();
  int prec_ = FP64;
  std::shared_ptr<void> x_;  // device coordinates
  void* stream_ = nullptr;
  mutable vector<double> host_;
  mutable bool hostValid_ = false;
  unsigned long version_ = 0;
  friend class HostCoords;
};

}  // namespace minim

#endif
