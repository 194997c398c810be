// ref_gpu_binding.cpp -- the reference-side binding of this library's C ABI:
// a Potential subclass compiled with the reference (sjavis/minim) that keeps
// the reference's State and minimisers and moves PhaseField's energy and
// gradient to the GPU through minim_potential_eval.  The class GpuPhaseField
// is the one a maintainer would add to the reference; main() minimises the
// same ternary field with the reference's own PhaseField and with
// GpuPhaseField, both under the reference's Lbfgs, and compares them.
//
// Built against the reference's headers and library (oracle/Makefile, target
// _ref/ref_gpu_binding) and libminim_amd.so; run by tests/test_examples.py.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <vector>

#include "State.h"  // the reference's
#include "minim_hip.h"  // this library's C ABI
#include "minimisers/Lbfgs.h"
#include "potentials/PhaseField.h"

class GpuPhaseField : public minim::NewPotential<GpuPhaseField> {
 public:
  minim_pf_params p{};  // grid, n_fluid, surface tensions, interface sizes ... (PhaseField setters)
  std::vector<double> tension, width;
  GpuPhaseField() = default;
  GpuPhaseField(const GpuPhaseField& o)
      : minim::NewPotential<GpuPhaseField>(o), p(o.p), tension(o.tension), width(o.width) {
    point();  // clone(): a fresh handle, made by init
  }
  ~GpuPhaseField() override { release(); }
  int potentialType() const override { return minim::Potential::GRID; }

  void init(const std::vector<double>& coords) override {
    release();
    point();
    check(minim_pf_create(&p, &h_));
    check(minim_potential_init(h_, coords.data(), coords.size()));
    n_ = coords.size();
    if (hipMalloc(&x_, n_ * sizeof(double)) != hipSuccess || hipMalloc(&g_, n_ * sizeof(double)) != hipSuccess)
      throw std::runtime_error("hipMalloc failed");
    check(minim_potential_convergence(h_, &convergence));
    gridSize = {p.grid[0], p.grid[1], p.grid[2]};
    dofPerNode = p.n_fluid;
  }

  void energyGradient(const std::vector<double>& x, const minim::Communicator&, double* e,
                      std::vector<double>* g) const override {
    if (hipMemcpy(x_, x.data(), n_ * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      throw std::runtime_error("hipMemcpy failed");
    double ev = 0;
    check(minim_potential_eval(h_, MINIM_FP64, x_, n_, nullptr, 0.0, e ? &ev : nullptr, g ? g_ : nullptr, nullptr));
    if (e) *e += ev;  // *e is accumulated into (State.cpp:49)
    if (g) {
      g->assign(n_, 0.0);  // resized and zeroed by the callee (PhaseField.cpp:663)
      if (hipMemcpy(g->data(), g_, n_ * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("hipMemcpy failed");
    }
  }
  // The device gradient already carries PhaseField::specialisedConstraints
  // (the hard density projection), which is idempotent, so State::basicEG's
  // applyConstraints call leaves it unchanged.

 private:
  static void check(int rc) {
    if (rc == MINIM_EINVAL) throw std::invalid_argument(minim_last_error());
    if (rc != MINIM_OK) throw std::runtime_error(minim_last_error());
  }
  void point() {
    p.surface_tension = tension.data();
    p.n_surface_tension = (int)tension.size();
    p.interface_size = width.data();
    p.n_interface_size = (int)width.size();
  }
  void release() {
    if (h_) minim_potential_destroy(h_);
    if (x_) (void)hipFree(x_);
    if (g_) (void)hipFree(g_);
    h_ = nullptr;
    x_ = g_ = nullptr;
  }
  minim_potential* h_ = nullptr;
  double *x_ = nullptr, *g_ = nullptr;
  size_t n_ = 0;
};

int main() {
  const std::vector<int> grid = {12, 10, 8};
  const std::vector<double> tension = {1.0, 0.8, 1.2};
  const double width = 2.0;
  const int ns = grid[0] * grid[1] * grid[2];
  std::mt19937_64 gen(4242);
  std::vector<double> x0(3 * ns);
  for (int i = 0; i < ns; i++) {
    double u[3], s = 0;
    for (double& v : u) s += (v = (double)(gen() >> 11) / 9007199254740992.0);
    for (int k = 0; k < 3; k++) x0[3 * i + k] = u[k] / s;
  }

  minim::PhaseField cpu;
  cpu.setNFluid(3).setGridSize(grid).setSurfaceTension(tension).setInterfaceSize(width);
  GpuPhaseField gpu;
  gpu.tension = tension;
  gpu.width = {width};
  gpu.p.grid[0] = grid[0];
  gpu.p.grid[1] = grid[1];
  gpu.p.grid[2] = grid[2];
  gpu.p.n_fluid = 3;
  gpu.p.resolution = 1.0;

  minim::State sc(cpu, x0), sg(gpu, x0);
  const double ec0 = sc.energy(), eg0 = sg.energy();
  std::vector<double> gc = sc.gradient(), gg = sg.gradient();
  double dg = 0, ng = 0;
  for (size_t i = 0; i < gc.size(); i++) {
    dg += (gc[i] - gg[i]) * (gc[i] - gg[i]);
    ng += gc[i] * gc[i];
  }
  minim::Lbfgs lc, lg;
  lc.setMaxIter(300);
  lg.setMaxIter(300);
  lc.minimise(sc);
  lg.minimise(sg);
  const double ec = sc.energy(), eg = sg.energy();
  std::printf("initial-energy %.17g %.17g\n", ec0, eg0);
  std::printf("initial-gradient-relerr %.3e\n", std::sqrt(dg / ng));
  std::printf("lbfgs-300 %d %.17g %d %.17g\n", lc.iter, ec, lg.iter, eg);
  const bool ok = std::fabs(ec0 - eg0) <= 1e-12 * std::fabs(ec0) && std::sqrt(dg / ng) <= 1e-12 &&
                  std::fabs(ec - eg) <= 1e-8 * std::fabs(ec0);
  std::printf("%s\n", ok ? "binding ok" : "binding MISMATCH");
  return ok ? 0 : 1;
}
