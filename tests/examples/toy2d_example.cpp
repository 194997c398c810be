// toy2d_example.cpp -- a user program written against minim's C++ API, shaped
// like the reference's examples/toy2d/toy2d.cpp:7-19: a user-defined potential
// (energy and gradient overrides), newState, and the minimisers.  It compiles
// unchanged against the reference's headers (-I <reference>/include, linked
// with the reference library) and against this library's (-I include/minim,
// linked with libminim_amd.so); the two builds' outputs are compared by
// tests/test_examples.py.  Without the reference's utils/print.h and
// utils/mpi.h it prints with printf.  Every line: a label, the iterations,
// the energy and the coordinates.
#include <cmath>
#include <cstdio>
#include <vector>

#include "minim.h"

using namespace minim;

class Toy2d : public NewPotential<Toy2d> {
 public:
  double energy(const std::vector<double>& coords) const override {
    return coords[0] * coords[0] + coords[1] * coords[1] + coords[0] * coords[0] * coords[0] * coords[0] +
           coords[1] * coords[1] * coords[1] * coords[1];
  }

  std::vector<double> gradient(const std::vector<double>& coords) const override {
    std::vector<double> g(coords.size());
    g[0] = 2 * coords[0] + 4 * coords[0] * coords[0] * coords[0];
    g[1] = 2 * coords[1] + 4 * coords[1] * coords[1] * coords[1];
    return g;
  }
};

// The same function through the function-pointer constructor Potential(EGFunc).
static void toyEG(const std::vector<double>& c, double* e, std::vector<double>* g) {
  if (e) *e = (c[0] - 1) * (c[0] - 1) + 3 * (c[1] + 2) * (c[1] + 2) + c[0] * c[1];
  if (g) *g = {2 * (c[0] - 1) + c[1], 6 * (c[1] + 2) + c[0]};
}

// An UNSTRUCTURED user potential: springs between coordinate pairs given as
// energy elements (setElements), each with its stiffness and rest length.
class Springs : public NewPotential<Springs> {
 public:
  int potentialType() const override { return Potential::UNSTRUCTURED; }
  void elementEnergyGradient(const std::vector<double>& x, const Element& el, double* e,
                             std::vector<double>* g) const override {
    const double k = el.parameters[0], l0 = el.parameters[1];
    const double d = x[el.idof[1]] - x[el.idof[0]] - l0;
    if (e) *e += 0.5 * k * d * d;
    if (g) {
      (*g)[el.idof[1]] += k * d;
      (*g)[el.idof[0]] -= k * d;
    }
  }
};

static void report(const char* label, int iter, const State& state, const std::vector<double>& r) {
  std::printf("%s %d %.17g", label, iter, state.energy());
  for (double v : r) std::printf(" %.17g", v);
  std::printf("\n");
}

int main() {
  Toy2d potential = Toy2d();
  State state = potential.newState({1, 4});
  state.convergence = 1e-4;

  GradDescent gd = GradDescent();
  auto result = gd.minimise(state);
  report("toy2d-graddescent", gd.iter, state, result);

  state.coords({1, 4});
  Lbfgs lbfgs = Lbfgs().setMaxIter(100);
  result = lbfgs.minimise(state);
  report("toy2d-lbfgs", lbfgs.iter, state, result);

  state.coords({1, 4});
  Fire fire = Fire();
  fire.setMaxIter(3000);
  result = fire.minimise(state);
  report("toy2d-fire", fire.iter, state, result);

  // a constraint: steps stay orthogonal to (1, 1), so x0 + x1 keeps its value
  Toy2d constrained = Toy2d();
  constrained.setConstraints({{0, 1}}, {1.0, 1.0});
  State cs(constrained, {1, 4});
  cs.convergence = 1e-6;
  Lbfgs lb2 = Lbfgs().setMaxIter(200);
  result = lb2.minimise(cs);
  report("toy2d-normal-constraint", lb2.iter, cs, result);

  // function pointer potential and a single-DOF constraint
  Potential fp(toyEG);
  fp.setConstraints({1});
  State fs(fp, {3, 0.5});
  fs.convergence = 1e-8;
  Lbfgs lb3 = Lbfgs();
  result = lb3.minimise(fs);
  report("fnptr-fixed-dof", lb3.iter, fs, result);

  // energy elements
  Springs springs;
  springs.setElements({{0, 1}, {1, 2}, {2, 3}}, {0, 0, 1}, {{1.0, 1.0}, {2.0, 1.5}, {0.5, 0.7}});
  springs.setConstraints({0});
  State ss(springs, {0.0, 0.3, 0.9, 1.1});
  ss.convergence = 1e-9;
  Lbfgs lb4 = Lbfgs();
  result = lb4.minimise(ss);
  report("springs-elements", lb4.iter, ss, result);
  ss.coords({0.0, 1.3, 2.1, 3.4});
  std::printf("springs-components %.17g %.17g\n", ss.componentEnergy(0), ss.componentEnergy(1));

  // a State on a rank this process does not hold (ranks = {1}): unused here
  State other(potential, {1, 4}, {1});
  std::printf("unused-rank %d %.17g %zu\n", other.usesThisProc ? 1 : 0, other.energy(), other.coords().size());
  return 0;
}
