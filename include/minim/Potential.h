// Potential.h -- device plugin interface of minim_amd, mirroring the
// reference's Potential (include/Potential.h:16-108) for the energy/gradient
// hot path.  Coordinates and gradients live in GPU memory; host vectors are
// only used at set-up and by the State convenience accessors.
#ifndef MINIM_AMD_POTENTIAL_H
#define MINIM_AMD_POTENTIAL_H

#include <cstddef>
#include <memory>
#include <vector>

namespace minim {
using std::vector;
template <typename T>
using vector2d = vector<vector<T>>;

// Storage/arithmetic precision of a State's device vectors.
constexpr int FP64 = 0;
constexpr int FP32 = 1;

// Per-site projection that a potential's constraints apply to any vector in
// its DOF space (used by the fused L-BFGS direction kernel): components with a
// bit set in fixmask are zeroed, then with hard != 0 the mean over the other
// components of each group of nf consecutive DOFs is removed.
struct SiteProjection {
  int nf = 1;
  unsigned fixmask = 0;
  int hard = 0;
  int nvar = 1;
  double inv_nvar = 1.0;
};

class Potential {
 public:
  double convergence = 1e-6;  // include/Potential.h:25

  virtual ~Potential() = default;
  virtual std::unique_ptr<Potential> clone() const = 0;

  enum { SERIAL = 0, UNSTRUCTURED = 1, GRID = 2 };  // include/Potential.h:74-78
  virtual int potentialType() const = 0;

  // Single-DOF constraints (Potential::setConstraints(vector<int>), src/Potential.cpp:80-85).
  vector<int> fixedDofs;
  Potential& setConstraints(vector<int> iFix);
  bool isFixed(int index) const;

  // ---- device plugin interface ----
  // Host set-up from the initial coordinates (PhaseField::init + initLocal,
  // LjNd::init, BarAndHinge::init).  Throws std::invalid_argument on bad input.
  virtual void init(const vector<double>& coords) = 0;
  // Evaluate at x (or at x + t*d when d != nullptr).  prec is FP64 or FP32 for
  // all device vectors.  e (host) and g (device, constraints applied, as
  // State::basicEG does in src/State.cpp:48-56) are optional.
  virtual void evaluate(int prec, const void* x, const void* d, double t, double* e, void* g,
                        void* stream) const = 0;
  // As evaluate, but the energy (when wantE) is written to device memory
  // e_dev without synchronising the stream.
  virtual void evaluateAsync(int prec, const void* x, const void* d, double t, bool wantE, double* e_dev,
                             void* g, void* stream) const;
  // Mixed storage: FP64 coordinates x and FP64 arithmetic, with the direction
  // d and the gradient g held in FP32.  Potentials that implement it return
  // true from supportsMixed(); the default evaluateMixed throws
  // std::invalid_argument.
  virtual bool supportsMixed() const { return false; }
  // As evaluateMixed, without a host synchronisation (energy, when wantE, to
  // the device double e_dev; with e_dev NULL it is left where evaluateAsync
  // leaves it).
  virtual void evaluateMixedAsync(const void* x, const void* d, double t, bool wantE, double* e_dev, void* g,
                                  void* stream) const;
  // Benchmark support: average duration (ms, HIP events on the stream) of
  // reps back-to-back launches of the evaluation kernel alone, without the
  // cross-block reduction that follows it.  The default throws
  // std::invalid_argument.
  virtual double timeKernel(int prec, const void* x, bool wantE, void* g, int reps, void* stream) const;
  virtual void evaluateMixed(const void* x, const void* d, double t, double* e, void* g, void* stream) const;
  // Potential::applyConstraints (src/Potential.cpp:102-115) on a device vector.
  virtual void applyConstraints(int prec, void* v, size_t n, void* stream) const;
  // The projection applyConstraints performs, for fused kernels.  Potentials
  // whose constraints are only fixed DOFs report nf = 1, fixmask = 0: their
  // basis vectors already hold exact zeros there.
  virtual SiteProjection projection() const { return SiteProjection{}; }
  // Fused accepted step: the energy (left on the device where evaluateAsync
  // leaves it) and the constrained gradient g at x + t*d, with x + t*d also
  // written to xout (precision px, like x); d and g have precision pv (px, or
  // FP32 against FP64 coordinates).  Potentials that implement it return true
  // from supportsStep(); the default evaluateStepAsync throws
  // std::invalid_argument.
  virtual bool supportsStep() const { return false; }
  virtual void evaluateStepAsync(int px, int pv, const void* x, const void* d, double t, void* xout, void* g,
                                 void* stream) const;
  // The gradient launched without a host synchronisation, for potentials
  // whose evaluation carries a validity check (LjNd's neighbour list).
  // Returns the reduction result slot that will hold a nonzero value when the
  // gradient must be recomputed with redoGradient (read it with the caller's
  // next fetch), or -1 when the gradient is complete (the default: evaluate).
  virtual int gradientDeferred(int prec, const void* x, void* g, void* stream) const;
  virtual void redoGradient(int prec, const void* x, void* g, void* stream) const;
  // True when applyConstraints also couples all DOFs (PhaseField's fixed
  // total volumes, PhaseField.cpp:614-658): minimisers then apply it to each
  // step in addition to the per-site projection.
  virtual bool hasGlobalConstraints() const { return false; }

 protected:
  std::shared_ptr<void> fixedIdxDev_;  // device int[fixedDofs.size()]
  size_t nFixedDev_ = 0;
  void uploadFixed(size_t ndof);
};

template <typename Derived>
class NewPotential : public Potential {
 public:
  std::unique_ptr<Potential> clone() const override {
    return std::make_unique<Derived>(static_cast<const Derived&>(*this));
  }
  Derived& setConstraints(vector<int> iFix) { return static_cast<Derived&>(Potential::setConstraints(iFix)); }
};

}  // namespace minim

#endif
