// lj3d_example.cpp -- a user program written against minim's C++ API, shaped
// like the reference's examples/lj3d/lj3d.cpp:14-24 and :43-52: four
// Lennard-Jones particles in 3-D minimised by L-BFGS, then repeated
// minimisations from other starting points through state.comm->bcast and
// state.coords.  It compiles unchanged against the reference's headers and
// against this library's; tests/test_examples.py compares the outputs.
// The starting points come from a fixed linear congruential sequence instead
// of rand(), so every build draws the same ones.
#include <cmath>
#include <cstdio>
#include <vector>

#include "minim.h"

using namespace minim;

static double sep(const std::vector<double>& r, int a, int b) {
  return std::sqrt(std::pow(r[3 * a] - r[3 * b], 2) + std::pow(r[3 * a + 1] - r[3 * b + 1], 2) +
                   std::pow(r[3 * a + 2] - r[3 * b + 2], 2));
}

static void report(const char* label, int iter, const State& state, const std::vector<double>& r) {
  std::printf("%s %d %.17g", label, iter, state.energy());
  for (int a = 0; a < 4; a++)
    for (int b = a + 1; b < 4; b++) std::printf(" %.17g", sep(r, a, b));
  std::printf("\n");
}

int main() {
  Lj3d potential = Lj3d();
  State state(potential, {0, 0, 0, 2, 0, 0, 1, 1, 0, 5, 1, 1});
  state.convergence = 1e-4;

  Lbfgs min = Lbfgs();
  auto result = min.minimise(state);
  report("lj3d", min.iter, state, result);
  std::printf("lj3d-component %.17g\n", state.componentEnergy(0));

  unsigned long long seed = 2;
  auto next = [&seed]() {
    seed = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)((seed >> 33) % 1000) / 100. - 5;
  };
  std::vector<double> init(12);
  for (int run = 0; run < 5; run++) {
    for (double& v : init) v = next();
    state.comm->bcast(init);
    state.coords(init);
    result = min.minimise(state);
    char label[32];
    std::snprintf(label, sizeof label, "lj3d-run%d", run);
    report(label, min.iter, state, result);
  }

  Fire fire = Fire();
  state.coords({0, 0, 0, 2, 0, 0, 1, 1, 0, 5, 1, 1});
  result = fire.minimise(state);
  report("lj3d-fire", fire.iter, state, result);
  return 0;
}
