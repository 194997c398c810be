// potential.cpp -- Potential base class: the reference's host plugin interface
// (src/Potential.cpp:15-146) and its bridge to device vectors.
//
// A potential that is not deviceNative (one a user defines through energy /
// gradient / energyGradient / elementEnergyGradient, or the function-pointer
// constructors) is evaluated by copying the device coordinates to the host,
// running the reference's State evaluation there (basicEG / elementEG,
// src/State.cpp:48-79) and copying the gradient back.  That is the user's own
// code running where the user wrote it; the library's potentials never take
// this path (they override evaluate with HIP kernels).
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "minim/Communicator.h"
#include "minim/Potential.h"
#include "minim/State.h"
#include "runtime.hpp"
#include "vec.hpp"

namespace minim {

using minim_amd::DeviceBuffer;
using minim_amd::DVec;

Potential::Potential(EFunc energy, GFunc gradient)
    : _energy(energy), _gradient(gradient), _energyGradient(nullptr) {}

Potential::Potential(EGFunc energyGradient) : _energy(nullptr), _gradient(nullptr), _energyGradient(energyGradient) {}

Potential::Potential(const Potential& o)
    : _energy(o._energy),
      _gradient(o._gradient),
      _energyGradient(o._energyGradient),
      convergence(o.convergence),
      constraints(o.constraints),
      distributed(o.distributed),
      elements(o.elements),
      elements_halo(o.elements_halo),
      dofPerNode(o.dofPerNode),
      haloWidth(o.haloWidth),
      gridSize(o.gridSize),
      commArray(o.commArray),
      ndofHint(o.ndofHint),
      stateComm(nullptr),
      fixedIdxDev_(o.fixedIdxDev_),
      nFixedDev_(o.nFixedDev_) {}

Potential& Potential::operator=(const Potential& o) {
  if (this == &o) return *this;
  _energy = o._energy;
  _gradient = o._gradient;
  _energyGradient = o._energyGradient;
  convergence = o.convergence;
  constraints = o.constraints;
  distributed = o.distributed;
  elements = o.elements;
  elements_halo = o.elements_halo;
  dofPerNode = o.dofPerNode;
  haloWidth = o.haloWidth;
  gridSize = o.gridSize;
  commArray = o.commArray;
  ndofHint = o.ndofHint;
  stateComm = nullptr;
  fixedIdxDev_ = o.fixedIdxDev_;
  nFixedDev_ = o.nFixedDev_;
  return *this;
}

Potential::~Potential() = default;

// ---- basic user functions (src/Potential.cpp:24-41)
double Potential::energy(const vector<double>& coords) const {
  if (_energy) return _energy(coords);
  if (deviceNative()) {
    double e = 0;
    deviceEnergyGradient(coords, &e, nullptr);
    return e;
  }
  throw std::runtime_error(
      "Potential: No energy function defined. If using parallel you must call energyGradient instead.");
}

vector<double> Potential::gradient(const vector<double>& coords) const {
  if (_gradient) return _gradient(coords);
  if (deviceNative()) {
    vector<double> g;
    deviceEnergyGradient(coords, nullptr, &g);
    return g;
  }
  throw std::runtime_error(
      "Potential: No gradient function defined. If using parallel you must call energyGradient instead.");
}

void Potential::energyGradient(const vector<double>& coords, const Communicator& comm, double* e,
                               vector<double>* g) const {
  (void)comm;
  if (_energyGradient) return _energyGradient(coords, e, g);
  if (deviceNative()) {
    // the reference's potentials accumulate into *e (PhaseField.cpp:697-700)
    // and size and zero *g themselves (:663)
    double ev = 0;
    deviceEnergyGradient(coords, e ? &ev : nullptr, g);
    if (e) *e += ev;
    return;
  }
  if (e != nullptr) *e = energy(coords);
  if (g != nullptr) *g = gradient(coords);
}

State Potential::newState(const vector<double>& coords, const vector<int>& ranks) {
  return State(*this, coords, ranks);
}

State Potential::newState(int ndof, const vector<int>& ranks) {
  vector<double> coords(ndof);
  return newState(coords, ranks);
}

// ---- elements (src/Potential.cpp:53-77)
Potential& Potential::setElements(vector<Element> els) {
  elements = els;
  return *this;
}

Potential& Potential::setElements(vector2d<int> idofs) {
  elements = {};
  for (const auto& idof : idofs) elements.push_back({0, idof, {}});
  return *this;
}

Potential& Potential::setElements(vector2d<int> idofs, vector<int> types, vector2d<double> parameters) {
  elements = {};
  const size_t ne = idofs.size();
  for (size_t i = 0; i < ne; i++) elements.push_back({types[i], idofs[i], parameters[i]});
  return *this;
}

// ---- constraints (src/Potential.cpp:80-146)
Potential& Potential::setConstraints(vector<int> iFix) {
  for (int i : iFix) constraints.push_back({{i}, nullptr, nullptr});
  fixedIdxDev_.reset();
  nFixedDev_ = 0;
  return *this;
}

Potential& Potential::setConstraints(vector2d<int> idofs, vector<double> normal) {
  for (const vector<int>& idof : idofs)
    constraints.push_back({idof, [normal](const vector<int>&, const vector<double>&) { return normal; }, nullptr});
  return *this;
}

Potential& Potential::setConstraints(vector2d<int> idofs, Constraint::NormalFn normal,
                                     Constraint::CorrectionFn correction) {
  for (const vector<int>& idof : idofs) constraints.push_back({idof, normal, correction});
  return *this;
}

// Remove from step the component along each constraint's normal, zero the
// single-DOF constraints, then the potential's own constraints.
void Potential::applyConstraints(const vector<double>& coords, const Communicator& comm,
                                 vector<double>& step) const {
  for (const auto& c : constraints) {
    if (c.idof.size() == 1) {
      step[c.idof[0]] = 0;
      continue;
    }
    const vector<double> normal = c.normal(coords);
    const size_t k = normal.size();
    double normalSq = 0, gradNormal = 0;
    for (size_t i = 0; i < k; i++) normalSq += normal[i] * normal[i];
    for (size_t i = 0; i < k; i++) gradNormal += step[c.idof[i]] * normal[i];
    for (size_t i = 0; i < k; i++) step[c.idof[i]] -= gradNormal * normal[i] / normalSq;
  }
  specialisedConstraints(coords, comm, step);
}

void Potential::correctConstraints(vector<double>& coords) const {
  for (const auto& c : constraints)
    if (c.correction) c.correction(c.idof, coords);
}

bool Potential::isFixed(int index) const {
  for (const auto& c : constraints)
    if (c.idof.size() == 1 && c.idof[0] == index) return true;
  return false;
}

vector<char> Potential::isFixed(const vector<int>& indicies) const {
  vector<int> iFixed = fixedDofs();
  std::sort(iFixed.begin(), iFixed.end());
  vector<char> fixed(indicies.size(), false);
  for (size_t i = 0; i < indicies.size(); i++)
    fixed[i] = std::binary_search(iFixed.begin(), iFixed.end(), indicies[i]);
  return fixed;
}

vector<int> Potential::fixedDofs() const {
  vector<int> out;
  for (const auto& c : constraints)
    if (c.idof.size() == 1) out.push_back(c.idof[0]);
  return out;
}

bool Potential::hasNormalConstraints() const {
  for (const auto& c : constraints)
    if (c.idof.size() != 1) return true;
  return false;
}

bool Potential::isSerial() const { return potentialType() == SERIAL; }

// Potential::newComm (src/Potential.cpp:149-158)
std::unique_ptr<Communicator> Potential::newComm() const {
  if (potentialType() == UNSTRUCTURED && distributed)
    throw std::invalid_argument(
        "Potential: You cannot create a State with a Potential that has already been distributed.");
  if (potentialType() == GRID) return std::make_unique<CommGrid>(haloWidth);
  return std::make_unique<CommUnstructured>();
}

Potential& Potential::setCommArray(vector<int> ca) {
  commArray = ca;
  return *this;
}

// ---- the reference's State evaluation (basicEG, elementEG: src/State.cpp:48-79)
void Potential::stateEnergyGradient(const vector<double>& coords, const Communicator& comm, double* e,
                                    vector<double>* g) const {
  if (e) *e = 0;
  if (g) *g = vector<double>(coords.size());
  if (potentialType() == UNSTRUCTURED) {
    for (const auto& el : elements) elementEnergyGradient(coords, el, e, g);
    blockEnergyGradient(coords, comm, e, g);
  } else {
    energyGradient(coords, comm, e, g);
  }
  if (g) applyConstraints(coords, comm, *g);
}

double Potential::componentEnergy(int component, const vector<double>& coords) const {
  double e = 0;
  for (const auto& el : elements)
    if (el.type == component) elementEnergyGradient(coords, el, &e, nullptr);
  return e;
}

// ---- device bridge
namespace {
// A single-rank communicator for evaluations of a potential that no State owns.
const Communicator& comm_for(const Potential& p, size_t n, std::unique_ptr<Communicator>& holder) {
  if (p.stateComm) return *p.stateComm;
  holder = p.newComm();
  holder->setup(const_cast<Potential&>(p), n, {});
  return *holder;
}
}  // namespace

void Potential::evaluate(int prec, const void* x, const void* d, double t, double* e, void* g,
                         void* stream) const {
  if (deviceNative()) throw std::logic_error("Potential: evaluate not implemented by a device potential");
  const size_t n = ndofHint;
  if (n == 0) throw std::runtime_error("Potential: evaluated before init.");
  hipStream_t s = static_cast<hipStream_t>(stream);
  vector<double> xh(n);
  minim_amd::download_raw(x, prec, n, xh.data(), s);
  if (d) {
    // the reference's trial point x + step, with step = t*d an exact scaling
    vector<double> dh(n);
    minim_amd::download_raw(d, prec, n, dh.data(), s);
    for (size_t i = 0; i < n; i++) xh[i] = xh[i] + t * dh[i];
  }
  std::unique_ptr<Communicator> holder;
  const Communicator& comm = comm_for(*this, n, holder);
  double ev = 0;
  vector<double> gh;
  stateEnergyGradient(xh, comm, e ? &ev : nullptr, g ? &gh : nullptr);
  if (e) *e = ev;
  if (g) {
    if (gh.size() != n) throw std::invalid_argument("Potential: gradient has the wrong size.");
    minim_amd::upload_raw(gh.data(), n, prec, g, s);
  }
}

void Potential::deviceEnergyGradient(const vector<double>& coords, double* e, vector<double>* g) const {
  hipStream_t s = nullptr;
  const size_t n = coords.size();
  auto& st = minim_amd::Staging::current();
  void* xd = st.device(0, n * sizeof(double));
  minim_amd::upload_raw(coords.data(), n, FP64, xd, s);
  void* gd = g ? st.device(1, n * sizeof(double)) : nullptr;
  double ev = 0;
  evaluate(FP64, xd, nullptr, 0.0, e ? &ev : nullptr, gd, s);
  if (e) *e = ev;
  if (g) {
    g->resize(n);
    minim_amd::download_raw(gd, FP64, n, g->data(), s);
  }
}

// The constraints of the base class on a device vector: single DOFs alone by
// a kernel; with normal constraints, all of them in the order they were set
// (as applyConstraints interleaves them) on a host copy of x and v.
void Potential::constrainCommon(int prec, const void* x, void* v, size_t n, void* stream) const {
  if (constraints.empty()) return;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!hasNormalConstraints()) {
    if (nFixedDev_ != constraints.size()) const_cast<Potential*>(this)->uploadFixed(n);
    minim_amd::launch_zero_idx(prec, v, static_cast<const int*>(fixedIdxDev_.get()), nFixedDev_, s);
    return;
  }
  if (!x) throw std::invalid_argument("Potential: normal constraints need the coordinates.");
  vector<double> xh(n), vh(n);
  minim_amd::download_raw(x, prec, n, xh.data(), s);
  minim_amd::download_raw(v, prec, n, vh.data(), s);
  for (const auto& c : constraints) {
    if (c.idof.size() == 1) {
      vh[c.idof[0]] = 0;
      continue;
    }
    const vector<double> normal = c.normal(xh);
    const size_t k = normal.size();
    double normalSq = 0, gradNormal = 0;
    for (size_t i = 0; i < k; i++) normalSq += normal[i] * normal[i];
    for (size_t i = 0; i < k; i++) gradNormal += vh[c.idof[i]] * normal[i];
    for (size_t i = 0; i < k; i++) vh[c.idof[i]] -= gradNormal * normal[i] / normalSq;
  }
  minim_amd::upload_raw(vh.data(), n, prec, v, s);
}

void Potential::constrainDevice(int prec, const void* x, void* v, size_t n, void* stream) const {
  if (deviceNative()) {
    constrainCommon(prec, x, v, n, stream);
    return;
  }
  // a host-evaluated potential: the whole of applyConstraints on the host
  if (!x) throw std::invalid_argument("Potential: constraints of a host potential need the coordinates.");
  hipStream_t s = static_cast<hipStream_t>(stream);
  vector<double> xh(n), vh(n);
  minim_amd::download_raw(x, prec, n, xh.data(), s);
  minim_amd::download_raw(v, prec, n, vh.data(), s);
  std::unique_ptr<Communicator> holder;
  applyConstraints(xh, comm_for(*this, n, holder), vh);
  minim_amd::upload_raw(vh.data(), n, prec, v, s);
}

bool Potential::hasGlobalConstraints() const { return !deviceNative() || hasNormalConstraints(); }

bool Potential::constraintsMoveSteps() const { return !deviceNative() || hasNormalConstraints(); }

void Potential::uploadFixed(size_t ndof) {
  const vector<int> fixed = fixedDofs();
  if (fixed.empty()) {
    fixedIdxDev_.reset();
    nFixedDev_ = 0;
    return;
  }
  for (int i : fixed)
    if (i < 0 || (size_t)i >= ndof) throw std::invalid_argument("Potential: constrained DOF index out of range.");
  auto buf = std::make_shared<DeviceBuffer>(sizeof(int) * fixed.size());
  MINIM_HIP_CHECK(hipMemcpy(buf->get(), fixed.data(), sizeof(int) * fixed.size(), hipMemcpyHostToDevice));
  nFixedDev_ = fixed.size();
  fixedIdxDev_ = std::shared_ptr<void>(buf, buf->get());
}

int Potential::gradientDeferred(int prec, const void* x, void* g, void* stream) const {
  evaluate(prec, x, nullptr, 0.0, nullptr, g, stream);
  return -1;
}

void Potential::redoGradient(int prec, const void* x, void* g, void* stream) const {
  evaluate(prec, x, nullptr, 0.0, nullptr, g, stream);
}

void Potential::evaluateStepAsync(int, int, const void*, const void*, double, void*, void*, void*) const {
  throw std::invalid_argument("this potential has no fused step evaluation");
}

double Potential::timeKernel(int, const void*, bool, void*, int, void*) const {
  throw std::invalid_argument("Potential: kernel timing is not available for this potential.");
}

void Potential::timeSteps(int, const void*, void*, double*, int, double, void*, double*) const {
  throw std::invalid_argument("Potential: step timing is not available for this potential.");
}

void Potential::evaluateMixedAsync(const void*, const void*, double, bool, double*, void*, void*) const {
  throw std::invalid_argument("Potential: mixed FP64/FP32 storage is not supported by this potential.");
}

void Potential::evaluateMixed(const void*, const void*, double, double*, void*, void*) const {
  throw std::invalid_argument("Potential: mixed FP64/FP32 storage is not supported by this potential.");
}

void Potential::evaluateAsync(int prec, const void* x, const void* d, double t, bool wantE, double* e_dev, void* g,
                              void* stream) const {
  double e = 0;
  evaluate(prec, x, d, t, wantE ? &e : nullptr, g, stream);
  if (wantE) {
    // the value goes where the library's own potentials leave it: e_dev, or
    // the calling thread's reduction result slot 0
    double* dst = e_dev ? e_dev : minim_amd::Workspace::current().device_result();
    MINIM_HIP_CHECK(hipMemcpyAsync(dst, &e, sizeof(double), hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)));
    MINIM_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  }
}

}  // namespace minim
