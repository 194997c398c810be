// cpp_api_example.cpp -- a program written against the reference's C++ API
// (State / Potential / Minimiser, PhaseField and PhaseFieldUnstructured
// setters), compiled against include/minim/ and linked with libminim_amd.so.
// With "host" as its argument it runs only what needs no GPU (potential set-up
// on the host: coefficients, element counts, input errors); with "gpu" it also
// minimises on the device.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "minim/Minimiser.h"
#include "minim/PhaseField.h"
#include "minim/PhaseFieldUnstructured.h"
#include "minim/State.h"

static int fails = 0;
#define EXPECT(c)                                          \
  do {                                                     \
    if (!(c)) {                                            \
      std::printf("FAILED line %d: %s\n", __LINE__, #c);   \
      fails++;                                             \
    }                                                      \
  } while (0)

int main(int argc, char** argv) {
  const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
  // PhaseField_test.cpp-style set-up: three fluids, tensions in pair order
  minim::PhaseField pf;
  pf.setNFluid(3).setGridSize({4, 5, 6}).setSurfaceTension({1.0, 0.8, 1.2}).setInterfaceSize(2.0);
  std::vector<double> x(4 * 5 * 6 * 3, 1.0 / 3);
  pf.init(x);
  EXPECT(std::fabs(pf.kappa[0] - 0.9) < 1e-15 && std::fabs(pf.kappa[1] - 2.1) < 1e-15);
  EXPECT(std::fabs(pf.kappaP[2] - 6.0) < 1e-15);
  EXPECT(std::fabs(pf.convergence - 1e-8) < 1e-22);
  // the element form of the same physics
  minim::PhaseFieldUnstructured pfu;
  pfu.setNFluid(3).setGridSize({4, 5, 6}).setSurfaceTension({1.0, 0.8, 1.2}).setInterfaceSize(2.0);
  pfu.setPressure({0.0, 0.5, 0.0});
  pfu.init(x);
  EXPECT(pfu.nElements == (size_t)(4 * 5 * 6) * 3 + 4 * 5 * 6);  // 3 gradient + 1 pressure element per site
  EXPECT(std::fabs(pfu.kappa[2] - pf.kappa[2]) < 1e-15);
  // the reference's input errors
  bool threw = false;
  try {
    minim::PhaseField bad;
    bad.setNFluid(3).setGridSize({2, 2, 2}).setSurfaceTension({1.0, 2.0});
    bad.init(std::vector<double>(24, 0.0));
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  EXPECT(threw);
  if (gpu) {
    // the library's autotuner picks the stencil's launch shape for this field
    // (FP64); the State's clone of the potential launches with it
    pf.autotune(x, minim::FP64, {256, 512}, {0, 8}, 3);
    const std::pair<int, int> tuned = pf.tunedShape(minim::FP64);
    EXPECT(tuned.first == 256 || tuned.first == 512);
    minim::State state(pf, x);
    EXPECT(static_cast<const minim::PhaseField&>(*state.pot).tunedShape(minim::FP64) == tuned);
    minim::Lbfgs lbfgs;
    lbfgs.setM(5).setMaxIter(100);
    std::vector<double> r = lbfgs.minimise(state);
    EXPECT(r.size() == x.size());
    EXPECT(state.energy() <= 1e-12);
  }
  std::printf("%s\n", fails ? "cpp api example: FAILED" : "cpp api example: ok");
  return fails ? 1 : 0;
}
