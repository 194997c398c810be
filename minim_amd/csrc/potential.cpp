// potential.cpp -- Potential base class: single-DOF constraints on device vectors.
#include <algorithm>
#include <stdexcept>

#include "minim/Potential.h"
#include "runtime.hpp"
#include "vec.hpp"

namespace minim {

using minim_amd::DeviceBuffer;

// Potential::setConstraints(vector<int> iFix) (src/Potential.cpp:80-85)
Potential& Potential::setConstraints(vector<int> iFix) {
  for (int i : iFix) fixedDofs.push_back(i);
  fixedIdxDev_.reset();
  nFixedDev_ = 0;
  return *this;
}

// Potential::isFixed (src/Potential.cpp:124-130)
bool Potential::isFixed(int index) const {
  return std::find(fixedDofs.begin(), fixedDofs.end(), index) != fixedDofs.end();
}

void Potential::uploadFixed(size_t ndof) {
  if (fixedDofs.empty()) {
    fixedIdxDev_.reset();
    nFixedDev_ = 0;
    return;
  }
  for (int i : fixedDofs)
    if (i < 0 || (size_t)i >= ndof) throw std::invalid_argument("Potential: constrained DOF index out of range.");
  auto buf = std::make_shared<DeviceBuffer>(sizeof(int) * fixedDofs.size());
  MINIM_HIP_CHECK(hipMemcpy(buf->get(), fixedDofs.data(), sizeof(int) * fixedDofs.size(), hipMemcpyHostToDevice));
  nFixedDev_ = fixedDofs.size();
  fixedIdxDev_ = std::shared_ptr<void>(buf, buf->get());
}

// Potential::applyConstraints for single-DOF constraints: step[i] = 0 (src/Potential.cpp:104)
void Potential::applyConstraints(int prec, void* v, size_t n, void* stream) const {
  (void)n;
  if (nFixedDev_)
    minim_amd::launch_zero_idx(prec, v, static_cast<const int*>(fixedIdxDev_.get()), nFixedDev_,
                               static_cast<hipStream_t>(stream));
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
  if (wantE && e_dev) {
    MINIM_HIP_CHECK(hipMemcpyAsync(e_dev, &e, sizeof(double), hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)));
    MINIM_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  }
}

}  // namespace minim
