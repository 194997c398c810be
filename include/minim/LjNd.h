// LjNd.h -- device Lennard-Jones potential, mirroring the reference's LjNd /
// Lj2d / Lj3d (include/potentials/LjNd.h:12-44).
//
// Default semantics are the reference's: every pair interacts, no cutoff, no
// periodic box (src/potentials/LjNd.cpp:29-53).  setCutoff() switches to a
// truncated-and-shifted potential in a periodic orthorhombic box evaluated
// with a cell-built Verlet neighbour list (no reference counterpart).
#ifndef MINIM_AMD_LJND_H
#define MINIM_AMD_LJND_H

#include <array>
#include <vector>

#include "Potential.h"

namespace minim {

class LjNd : public NewPotential<LjNd> {
 public:
  int potentialType() const override { return Potential::UNSTRUCTURED; }
  bool deviceNative() const override { return true; }
  // the reference's elements are all pairs, of type 0 (LjNd.cpp:12-26)
  double componentEnergy(int component, const vector<double>& coords) const override;

  int nDim;
  int nParticle = 0;
  double sigma = 1;
  double epsilon = 1;

  LjNd(int nDim) : nDim(nDim) {}
  LjNd(int nDim, double sigma, double epsilon) : nDim(nDim), sigma(sigma), epsilon(epsilon) {}
  LjNd(const LjNd& o);
  LjNd& operator=(const LjNd& o);
  ~LjNd() override;

  LjNd& setSigma(double sigma);
  LjNd& setEpsilon(double epsilon);
  // Truncated-and-shifted at rc, periodic box, Verlet list with the given skin.
  LjNd& setCutoff(double rc, std::array<double, 3> box, double skin = 0.3);

  double cutoff = 0;  // 0 = all pairs (reference semantics)
  std::array<double, 3> box{{0, 0, 0}};
  double skin = 0.3;

  void init(const vector<double>& coords) override;
  void evaluate(int prec, const void* x, const void* d, double t, double* e, void* g,
                void* stream) const override;
  double timeKernel(int prec, const void* x, bool wantE, void* g, int reps, void* stream) const override;
  int gradientDeferred(int prec, const void* x, void* g, void* stream) const override;
  void redoGradient(int prec, const void* x, void* g, void* stream) const override;

  // Statistics of the neighbour list (cutoff mode) of this object's device
  // state: a State evaluates its own copy of the potential (state.pot).
  // listBuilds counts builds; maxNeighbours is the longest list of the last
  // build (it reads the list lengths back from the device).
  long listBuilds() const;
  int maxNeighbours() const;

  struct Impl;

 private:
  std::shared_ptr<Impl> impl_;  // device neighbour-list state (rebuilt lazily)
};

class Lj2d : public LjNd {
 public:
  Lj2d() : LjNd(2) {}
  Lj2d(double sigma, double epsilon) : LjNd(2, sigma, epsilon) {}
  std::unique_ptr<Potential> clone() const override { return std::make_unique<Lj2d>(*this); }
};

class Lj3d : public LjNd {
 public:
  Lj3d() : LjNd(3) {}
  Lj3d(double sigma, double epsilon) : LjNd(3, sigma, epsilon) {}
  std::unique_ptr<Potential> clone() const override { return std::make_unique<Lj3d>(*this); }
};

}  // namespace minim

#endif
