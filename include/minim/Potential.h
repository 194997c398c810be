// Potential.h -- the reference's Potential plugin interface (include/Potential.h:16-104)
// over GPU-resident coordinates.
//
// Two faces:
//  * the reference's host interface, unchanged: energy / gradient /
//    energyGradient on host vectors, the function-pointer constructors,
//    energy elements (setElements, elementEnergyGradient, blockEnergyGradient),
//    constraints (single DOFs and normal projections, applyConstraints,
//    correctConstraints, specialisedConstraints), init / initLocal, clone,
//    potentialType, newState.  A user-defined potential written against it
//    compiles and runs unchanged: the State keeps its coordinates on the GPU
//    and hands them to these functions on the host for each evaluation.
//  * the device interface that the library's own potentials (PhaseField,
//    PhaseFieldUnstructured, LjNd, BarAndHinge) implement with HIP kernels:
//    evaluate(...) on device vectors.  For them the host functions above run
//    on the device too (the coordinates are uploaded, the result downloaded).
#ifndef MINIM_AMD_POTENTIAL_H
#define MINIM_AMD_POTENTIAL_H

#include <cstddef>
#include <functional>
#include <memory>
#include <vector>

namespace minim {
using std::vector;
template <typename T>
using vector2d = vector<vector<T>>;
class State;
class Communicator;

// Storage/arithmetic precision of device vectors (codes of the C ABI).
constexpr int FP64 = 0;
constexpr int FP32 = 1;
// Precision of a State's device coordinates (State's constructor; a scoped
// enum, so that a brace-enclosed rank list can never convert to it).
enum class Precision { FP64 = 0, FP32 = 1 };

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
  typedef double (*EFunc)(const vector<double>&);
  typedef vector<double> (*GFunc)(const vector<double>&);
  typedef void (*EGFunc)(const vector<double>&, double*, vector<double>*);
  EFunc _energy;
  GFunc _gradient;
  EGFunc _energyGradient;

 public:
  double convergence = 1e-6;  // include/Potential.h:25

  // ---- basic user functions (include/Potential.h:27-37, src/Potential.cpp:15-41)
  Potential(EFunc energy, GFunc gradient);
  Potential(EGFunc energyGradient);
  // A derived class overrides energyGradient, or energy and gradient (or
  // elementEnergyGradient for UNSTRUCTURED potentials), as in the reference.
  virtual double energy(const vector<double>& coords) const;
  virtual vector<double> gradient(const vector<double>& coords) const;
  virtual void energyGradient(const vector<double>& coords, const Communicator& comm, double* e,
                              vector<double>* g) const;

  State newState(int ndof, const vector<int>& ranks = {});
  State newState(const vector<double>& coords, const vector<int>& ranks = {});

  // ---- constraints (include/Potential.h:40-60, src/Potential.cpp:80-146)
  struct Constraint {
    using NormalFn = std::function<vector<double>(const vector<int>&, const vector<double>&)>;
    using CorrectionFn = std::function<void(const vector<int>&, vector<double>&)>;
    vector<int> idof;
    NormalFn normalFn;
    CorrectionFn correction = nullptr;
    vector<double> normal(const vector<double>& normalVec) const { return normalFn(idof, normalVec); }
  };
  vector<Constraint> constraints;

  Potential& setConstraints(vector<int> iFix);
  Potential& setConstraints(vector2d<int> idofs, vector<double> normal);
  Potential& setConstraints(vector2d<int> idofs, Constraint::NormalFn normal,
                            Constraint::CorrectionFn correction = nullptr);

  void applyConstraints(const vector<double>& coords, const Communicator& comm, vector<double>& step) const;
  void correctConstraints(vector<double>& coords) const;
  virtual void specialisedConstraints(const vector<double>& coords, const Communicator& comm,
                                      vector<double>& data) const {}

  bool isFixed(int index) const;
  vector<char> isFixed(const vector<int>& indicies) const;

  // ---- internal (include/Potential.h:62-71)
  virtual void init(const vector<double>& coords) {}
  virtual void initLocal(const vector<double>& coords, const Communicator& comm) {}
  bool isSerial() const;

  Potential(const Potential& o);
  Potential& operator=(const Potential& o);
  virtual ~Potential();
  virtual std::unique_ptr<Potential> clone() const { return std::make_unique<Potential>(*this); }

  enum { SERIAL = 0, UNSTRUCTURED = 1, GRID = 2 };  // include/Potential.h:74-78
  virtual int potentialType() const { return SERIAL; }
  std::unique_ptr<Communicator> newComm() const;

  // ---- UNSTRUCTURED: energy elements (include/Potential.h:83-97)
  bool distributed = false;
  struct Element {
    int type;
    vector<int> idof;
    vector<double> parameters;
  };
  vector<Element> elements;
  vector<Element> elements_halo;
  Potential& setElements(vector<Element> elements);
  Potential& setElements(vector2d<int> idofs);
  Potential& setElements(vector2d<int> idofs, vector<int> types, vector2d<double> parameters);
  virtual void elementEnergyGradient(const vector<double>& coords, const Element& el, double* e,
                                     vector<double>* g) const {}
  virtual void blockEnergyGradient(const vector<double>& coords, const Communicator& comm, double* e,
                                   vector<double>* g) const {}

  // ---- GRID (include/Potential.h:99-104)
  int dofPerNode = 1;
  int haloWidth = 1;
  vector<int> gridSize;
  vector<int> commArray;
  Potential& setCommArray(vector<int> commArray);

  // ---- the reference's State evaluation on host vectors (basicEG / elementEG,
  // src/State.cpp:48-79): e zeroed, g sized and zeroed, the elements (or
  // energyGradient), then the constraints on g.
  void stateEnergyGradient(const vector<double>& coords, const Communicator& comm, double* e,
                           vector<double>* g) const;
  // Energy of the elements of one type (State::componentEnergy,
  // src/State.cpp:307-312).  The library's potentials evaluate it on the
  // device without building a host element list.
  virtual double componentEnergy(int component, const vector<double>& coords) const;

  // ---- device interface
  // True for potentials whose evaluation runs as HIP kernels; false for
  // potentials defined through the host functions above.
  virtual bool deviceNative() const { return false; }
  // Evaluate at x (or at x + t*d when d != nullptr).  prec is FP64 or FP32 for
  // all device vectors.  e (host) and g (device, constraints applied, as
  // State::basicEG does in src/State.cpp:48-56) are optional.  The default
  // copies x (and d) to the host, runs stateEnergyGradient there and copies g
  // back: a user-defined potential's code runs where the user wrote it.
  virtual void evaluate(int prec, const void* x, const void* d, double t, double* e, void* g,
                        void* stream) const;
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
  // Benchmark support: `steps` whole energy+gradient evaluations at x on the
  // stream (as evaluateAsync, energy copied to e_dev), after untimed
  // evaluations for warmMs of host time (steady clocks).  HIP events bracket
  // the loop and, inside it, each launch of the evaluation kernel, so that the
  // kernel's times come from the very loop whose step time is reported.
  // out[0] = ms per step over the loop; out[1..4] = median, min, max and mean
  // of the per-launch kernel durations (ms).  The default throws
  // std::invalid_argument.
  virtual void timeSteps(int prec, const void* x, void* g, double* e_dev, int steps, double warmMs, void* stream,
                         double* out) const;
  virtual void evaluateMixed(const void* x, const void* d, double t, double* e, void* g, void* stream) const;
  // Potential::applyConstraints (src/Potential.cpp:102-115) on a device vector
  // v of n entries, with x the coordinates the constraints are taken at (their
  // normals depend on them; x may be null when no constraint needs it).
  // Single-DOF constraints run as a kernel; normal constraints, and all
  // constraints of a potential that is not deviceNative, through a host copy.
  virtual void constrainDevice(int prec, const void* x, void* v, size_t n, void* stream) const;
  // The projection constrainDevice performs per site, for fused kernels.
  // Potentials whose constraints are only fixed DOFs report nf = 1,
  // fixmask = 0: their basis vectors already hold exact zeros there.
  virtual SiteProjection projection() const { return SiteProjection{}; }
  // Fused accepted step: the energy (left on the device where evaluateAsync
  // leaves it) and the constrained gradient g at x + t*d, with x + t*d also
  // written to xout (precision px, like x); d and g have precision pv (px, or
  // FP32 against FP64 coordinates).  Potentials that implement it return true
  // from supportsStep(); the default evaluateStepAsync throws
  // std::invalid_argument.
  virtual bool supportsStep() const { return false; }
  // True when evaluateStepAsync also takes g = nullptr: the energy at
  // x + t*d and x + t*d in xout, keeping what the gradient there needs, so
  // that the next gradient evaluation at xout (with no other evaluation in
  // between) costs less.  The line search then pays for that gradient only
  // when it accepts the trial.
  virtual bool stepDefersGradient() const { return false; }
  virtual void evaluateStepAsync(int px, int pv, const void* x, const void* d, double t, void* xout, void* g,
                                 void* stream) const;
  // The gradient launched without a host synchronisation, for potentials
  // whose evaluation carries a validity check (LjNd's neighbour list).
  // Returns the reduction result slot that will hold a nonzero value when the
  // gradient must be recomputed with redoGradient (read it with the caller's
  // next fetch), or -1 when the gradient is complete (the default: evaluate).
  virtual int gradientDeferred(int prec, const void* x, void* g, void* stream) const;
  virtual void redoGradient(int prec, const void* x, void* g, void* stream) const;
  // True when constrainDevice does more than the per-site projection:
  // constraints that couple all DOFs (PhaseField's fixed total volumes,
  // PhaseField.cpp:614-658), normal constraints, or the constraints of a
  // potential evaluated on the host.  Minimisers then apply it to each step
  // (Lbfgs.cpp:51, Fire.cpp:71).
  virtual bool hasGlobalConstraints() const;
  // True when that step projection can change a direction that the per-site
  // projection left in place (normal constraints, host potentials): L-BFGS
  // then recomputes the direction's dots after it.
  bool constraintsMoveSteps() const;

  // The single-DOF constraints, in the order they were set.
  vector<int> fixedDofs() const;
  // Number of DOFs of the coordinates the potential was initialised with, and
  // the communicator of the State that owns it (set by State; a copy of the
  // potential keeps the first and drops the second).
  size_t ndofHint = 0;
  const Communicator* stateComm = nullptr;

 protected:
  Potential() : _energy(nullptr), _gradient(nullptr), _energyGradient(nullptr) {}
  bool hasNormalConstraints() const;
  // Fixed DOFs (kernel) and normal constraints (host copy) on a device vector.
  void constrainCommon(int prec, const void* x, void* v, size_t n, void* stream) const;
  // The host functions energy / gradient / energyGradient of deviceNative
  // potentials: upload, evaluate, download.
  void deviceEnergyGradient(const vector<double>& coords, double* e, vector<double>* g) const;
  std::shared_ptr<void> fixedIdxDev_;  // device int[number of fixed DOFs]
  size_t nFixedDev_ = 0;
  void uploadFixed(size_t ndof);
};

template <typename Derived>
class NewPotential : public Potential {
 public:
  std::unique_ptr<Potential> clone() const override {
    return std::make_unique<Derived>(static_cast<const Derived&>(*this));
  }

  Derived& setElements(vector<Element> elements) {
    return static_cast<Derived&>(Potential::setElements(elements));
  }
  Derived& setElements(vector2d<int> idofs) { return static_cast<Derived&>(Potential::setElements(idofs)); }
  Derived& setElements(vector2d<int> idofs, vector<int> types, vector2d<double> parameters) {
    return static_cast<Derived&>(Potential::setElements(idofs, types, parameters));
  }

  Derived& setConstraints(vector<int> iFix) { return static_cast<Derived&>(Potential::setConstraints(iFix)); }
  Derived& setConstraints(vector2d<int> idofs, vector<double> normal) {
    return static_cast<Derived&>(Potential::setConstraints(idofs, normal));
  }
  Derived& setConstraints(vector2d<int> idofs, Constraint::NormalFn normal,
                          Constraint::CorrectionFn correction = nullptr) {
    return static_cast<Derived&>(Potential::setConstraints(idofs, normal, correction));
  }
};

}  // namespace minim

#endif
