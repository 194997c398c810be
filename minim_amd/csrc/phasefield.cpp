// phasefield.cpp -- host side of the device PhaseField potential.
//
// Set-up mirrors PhaseField::init (src/potentials/PhaseField.cpp:137-158),
// setDefaults (:961-967), checkArraySizes (:970-995) and
// assignFluidCoefficients (:78-134); evaluation launches the gather-form
// stencil of pf.hip.
#include <cmath>
#include <stdexcept>
#include <vector>

#include "minim/PhaseField.h"
#include "pf.hpp"
#include "runtime.hpp"

namespace minim {

using minim_amd::PFArgs;

namespace {
const PFArgs& args_of(const std::shared_ptr<void>& p) {
  if (!p) throw std::runtime_error("PhaseField: not initialised (construct a State first).");
  return *static_cast<const PFArgs*>(p.get());
}
}  // namespace

PhaseField& PhaseField::setNFluid(int n) {
  nFluid = n;
  return *this;
}

PhaseField& PhaseField::setGridSize(vector<int> g) {
  gridSize = g;
  nGrid = 1;
  for (int s : g) nGrid *= s;
  return *this;
}

PhaseField& PhaseField::setResolution(double r) {
  resolution = r;
  return *this;
}

// The scalar setters size their vectors by the current nFluid, as the
// reference's do (PhaseField.cpp:764-777): call setNFluid first.
PhaseField& PhaseField::setInterfaceSize(double w) {
  interfaceSize = vector<double>(nFluid, w);
  return *this;
}

PhaseField& PhaseField::setInterfaceSize(vector<double> w) {
  interfaceSize = w;
  return *this;
}

PhaseField& PhaseField::setSurfaceTension(double s) {
  surfaceTension = vector<double>(nFluid, s);
  return *this;
}

PhaseField& PhaseField::setSurfaceTension(vector<double> s) {
  surfaceTension = s;
  return *this;
}

// PhaseField::setDensityConstraint (PhaseField.cpp:802-815)
PhaseField& PhaseField::setDensityConstraint(std::string method) {
  if (method == "gradient" || method == "hard") {
    densityConstraint = minim_amd::kDensityHard;
  } else if (method == "energy penalty" || method == "soft") {
    densityConstraint = minim_amd::kDensitySoft;
  } else if (method == "none") {
    densityConstraint = minim_amd::kDensityNone;
  } else if (method == "fixed") {
    densityConstraint = minim_amd::kDensityFixed;
  } else {
    throw std::invalid_argument(
        "PhaseField: Invalid density constraint. Allowed methods are: fixed, none, gradient / hard, or energy "
        "penalty / soft");
  }
  return *this;
}

// PhaseField::setFixFluid (PhaseField.cpp:891-895)
PhaseField& PhaseField::setFixFluid(int iFluid, bool fix) {
  if ((int)fixFluid.size() != nFluid) fixFluid = vector<char>(nFluid, false);
  fixFluid.at(iFluid) = fix;
  return *this;
}

void PhaseField::setDefaults() {
  if (interfaceSize.empty()) interfaceSize = vector<double>(nFluid, resolution);
  if (fixFluid.empty()) fixFluid = vector<char>(nFluid, false);
}

void PhaseField::checkArraySizes() {
  if ((int)interfaceSize.size() != nFluid) throw std::invalid_argument("PhaseField: Invalid size interfaceSize array.");
  if ((int)fixFluid.size() != nFluid) throw std::invalid_argument("PhaseField: Invalid size of fixed fluid array.");
}

void PhaseField::assignFluidCoefficients() {
  vector<double> st = surfaceTension;
  if (nGrid > 1 && (int)st.size() == nGrid * nParams)
    throw std::invalid_argument("PhaseField: spatially varying surface tension is not supported by the device path.");
  if ((int)st.size() == 1) {
    st = vector<double>(nParams, st[0]);
  } else if ((int)st.size() != nParams) {
    throw std::invalid_argument("PhaseField: Invalid size of surfaceTension array.");
  }
  double sum = 0;
  for (double v : st) sum += v;
  surfaceTensionMean = sum / st.size();

  vector<double> iw = interfaceSize;
  if ((int)iw.size() == 1) {
    iw = vector<double>(nParams, iw[0]);
  } else if ((int)iw.size() != nParams) {
    throw std::invalid_argument("PhaseField: Invalid size of interfaceSize array.");
  }

  kappa.assign(nParams, 0.0);
  kappaP.assign(nParams, 0.0);
  if (nFluid <= 2) {
    for (int i = 0; i < nParams; i++) {
      kappa[i] = (3 * st[i]) / iw[i];
      kappaP[i] = (3 * st[i]) * iw[i];
    }
  } else {
    // pairwise tensions (1-2, 1-3, 2-3, ...) to per-fluid coefficients
    vector<double> ks(nParams), kps(nParams);
    for (int i = 0; i < nParams; i++) {
      ks[i] = (6 * st[i]) / iw[i];
      kps[i] = (6 * st[i]) * iw[i];
    }
    const int l = nFluid - 1;
    kappa[0] = 0.5 * (ks[0] + ks[1] - ks[l]);
    kappa[1] = 0.5 * (ks[0] - ks[1] + ks[l]);
    kappa[2] = 0.5 * (-ks[0] + ks[1] + ks[l]);
    kappaP[0] = 0.5 * (kps[0] + kps[1] - kps[l]);
    kappaP[1] = 0.5 * (kps[0] - kps[1] + kps[l]);
    kappaP[2] = 0.5 * (-kps[0] + kps[1] + kps[l]);
    for (int i = 3; i < nFluid; i++) {
      kappa[i] = ks[i - 1] - kappa[0];
      kappaP[i] = kps[i - 1] - kappaP[0];
    }
  }
}

void PhaseField::init(const vector<double>& coords) {
  if (gridSize.size() != 3) throw std::invalid_argument("PhaseField: gridSize must have three entries.");
  for (int s : gridSize)
    if (s < 1) throw std::invalid_argument("PhaseField: grid sizes must be positive.");
  nGrid = gridSize[0] * gridSize[1] * gridSize[2];
  if (nFluid < 1 || nFluid > minim_amd::kPFMaxFluid)
    throw std::invalid_argument("PhaseField: nFluid must be between 1 and 16.");
  if ((long)coords.size() != (long)nGrid * nFluid)
    throw std::invalid_argument("PhaseField: Size of coordinates array does not match the grid size.");
  if (!fixedDofs.empty())
    throw std::invalid_argument("PhaseField: single-DOF constraints are not supported; use setFixFluid.");
  nParams = nFluid;
  setDefaults();
  checkArraySizes();
  assignFluidCoefficients();
  convergence = 1e-8 * surfaceTensionMean * std::pow(resolution, 2);
  if (densityConstraint == minim_amd::kDensityFixed) fixFluid[nFluid - 1] = true;

  auto a = std::make_shared<PFArgs>();
  *a = PFArgs{};
  // canonical grid: drop size-1 dimensions to the front (keeps the linear index)
  int dims[3], nd = 0;
  for (int i = 0; i < 3; i++)
    if (gridSize[i] > 1) dims[nd++] = gridSize[i];
  int canon[3] = {1, 1, 1};
  for (int i = 0; i < nd; i++) canon[3 - nd + i] = dims[i];
  a->X = canon[0];
  a->Y = canon[1];
  a->Z = canon[2];
  a->nf = nFluid;
  const double nodeVol = 1 * std::pow(resolution, 3);
  for (int k = 0; k < nFluid; k++) {
    a->fb[k] = (nFluid == 1) ? kappa[k] / 16 * nodeVol : 0.5 * kappa[k] * nodeVol;
    double f = (nFluid == 1) ? 0.25 * kappaP[k] * nodeVol : 0.5 * kappaP[k] * nodeVol;
    a->fg[k] = f / std::pow(resolution, 2);
  }
  a->fixmask = 0;
  int nvar = 0;
  for (int k = 0; k < nFluid; k++) {
    if (fixFluid[k]) a->fixmask |= (1u << k);
    else nvar++;
  }
  a->nvar = nvar;
  a->inv_nvar = nvar > 0 ? 1.0 / nvar : 0.0;
  a->density = densityConstraint;
  a->coef = densityConst * surfaceTensionMean * std::pow(resolution, 2);
  a->xchunk_override = stencilXChunk;
  minim_amd::pf_args_finalize(*a, stencilTargetBlocks);
  args_ = a;
}

namespace {
void pf_eval_sync(const PFArgs& a, int px, int pv, const void* x, const void* d, double t, double* e, void* g,
                  void* stream) {
  int mode = (e ? 1 : 0) | (g ? 2 : 0) | (d ? 4 : 0);
  if (mode == 0) return;
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto& ws = minim_amd::Workspace::current();
  if ((mode & 4) && (mode & 2)) {
    throw std::invalid_argument("PhaseField: gradients at a trial point are not supported.");
  }
  minim_amd::pf_launch_eval(a, px, pv, mode, x, d, t, g, ws.rs(), s);
  if (e) {
    double r[1];
    ws.fetch(s, 1, r);
    *e = r[0];
  }
}
}  // namespace

void PhaseField::evaluate(int prec, const void* x, const void* d, double t, double* e, void* g,
                          void* stream) const {
  pf_eval_sync(args_of(args_), prec, prec, x, d, t, e, g, stream);
}

double PhaseField::timeKernel(int prec, const void* x, bool wantE, void* g, int reps, void* stream) const {
  const PFArgs& a = args_of(args_);
  const int mode = (wantE ? 1 : 0) | (g ? 2 : 0);
  if (mode == 0 || reps < 1) throw std::invalid_argument("PhaseField: nothing to time");
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto& ws = minim_amd::Workspace::current();
  minim_amd::pf_launch_eval(a, prec, prec, mode, x, nullptr, 0.0, g, ws.rs(), s, false);  // warm-up
  hipEvent_t e0, e1;
  MINIM_HIP_CHECK(hipEventCreate(&e0));
  MINIM_HIP_CHECK(hipEventCreate(&e1));
  MINIM_HIP_CHECK(hipEventRecord(e0, s));
  for (int i = 0; i < reps; i++) minim_amd::pf_launch_eval(a, prec, prec, mode, x, nullptr, 0.0, g, ws.rs(), s, false);
  MINIM_HIP_CHECK(hipEventRecord(e1, s));
  MINIM_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  MINIM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  MINIM_HIP_CHECK(hipEventDestroy(e0));
  MINIM_HIP_CHECK(hipEventDestroy(e1));
  return (double)ms / reps;
}

void PhaseField::evaluateMixed(const void* x, const void* d, double t, double* e, void* g, void* stream) const {
  pf_eval_sync(args_of(args_), FP64, FP32, x, d, t, e, g, stream);
}

namespace {
void pf_eval_async(const PFArgs& a, int prec, int pv, const void* x, const void* d, double t, bool wantE,
                   double* e_dev, void* g, void* stream) {
  int mode = (wantE ? 1 : 0) | (g ? 2 : 0) | (d ? 4 : 0);
  if (mode == 0) return;
  if ((mode & 4) && (mode & 2)) throw std::invalid_argument("PhaseField: gradients at a trial point are not supported.");
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto& ws = minim_amd::Workspace::current();
  minim_amd::pf_launch_eval(a, prec, pv, mode, x, d, t, g, ws.rs(), s);
  if (wantE && e_dev)
    MINIM_HIP_CHECK(hipMemcpyAsync(e_dev, ws.device_result(), sizeof(double), hipMemcpyDeviceToDevice, s));
}
}  // namespace

void PhaseField::evaluateAsync(int prec, const void* x, const void* d, double t, bool wantE, double* e_dev, void* g,
                               void* stream) const {
  pf_eval_async(args_of(args_), prec, prec, x, d, t, wantE, e_dev, g, stream);
}

void PhaseField::evaluateMixedAsync(const void* x, const void* d, double t, bool wantE, double* e_dev, void* g,
                                    void* stream) const {
  pf_eval_async(args_of(args_), FP64, FP32, x, d, t, wantE, e_dev, g, stream);
}

void PhaseField::applyConstraints(int prec, void* v, size_t n, void* stream) const {
  const PFArgs& a = args_of(args_);
  minim_amd::pf_launch_constrain(a, prec, v, n / a.nf, static_cast<hipStream_t>(stream));
}

SiteProjection PhaseField::projection() const {
  const PFArgs& a = args_of(args_);
  SiteProjection p;
  p.nf = a.nf;
  p.fixmask = a.fixmask;
  p.hard = (a.density == minim_amd::kDensityHard && a.nf > 1) ? 1 : 0;
  p.nvar = a.nvar;
  p.inv_nvar = a.inv_nvar;
  return p;
}

}  // namespace minim
