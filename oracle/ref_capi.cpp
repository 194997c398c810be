// ref_capi.cpp -- a C-ABI shim over the reference library (sjavis/minim),
// compiled together with the reference's own sources from the reference source tree by
// oracle/Makefile into oracle/_ref/libminim_ref.so.
//
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle restatement and by
// bench.py's cpu_baseline leg (kind "reference").  The product never loads it.
//
// Every function drives the reference's public API exactly as a user program
// would (PhaseField/LjNd/BarAndHinge setters, State, Lbfgs/Fire/GradDescent).
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "State.h"
#include "minimisers/Fire.h"
#include "minimisers/GradDescent.h"
#include "minimisers/Lbfgs.h"
#include "potentials/BarAndHinge.h"
#include "potentials/LjNd.h"
#include "potentials/PhaseField.h"
#include "potentials/PhaseFieldUnstructured.h"

using namespace minim;
using std::vector;

namespace {

long g_nEnergy = 0, g_nGradient = 0, g_nBoth = 0;

// Counts calls the way a subclass-of-PhaseField probe does: by overriding the
// plugin virtual that State::basicEG calls (src/State.cpp:48-56).
class CountingPF : public PhaseField {
 public:
  CountingPF() = default;
  explicit CountingPF(const PhaseField& p) : PhaseField(p) {}
  std::unique_ptr<Potential> clone() const override { return std::make_unique<CountingPF>(*this); }
  void energyGradient(const vector<double>& coords, const Communicator& comm, double* e,
                      vector<double>* g) const override {
    if (e && g) g_nBoth++;
    else if (e) g_nEnergy++;
    else if (g) g_nGradient++;
    PhaseField::energyGradient(coords, comm, e, g);
  }
};

const char* density_name(int d) {
  switch (d) {
    case 0: return "hard";
    case 1: return "soft";
    case 2: return "none";
    case 3: return "fixed";
  }
  return "bad";
}

CountingPF make_pf(const int* grid, int nfluid, double res, const double* st, int nst,
                   const double* iw, int niw, int density) {
  CountingPF pot;
  pot.setNFluid(nfluid);
  pot.setGridSize({grid[0], grid[1], grid[2]});
  pot.setResolution(res);
  pot.setSurfaceTension(vector<double>(st, st + nst));
  // one value means the scalar setter (PhaseField.cpp:764-767), else the vector one
  if (niw == 1) pot.setInterfaceSize(iw[0]);
  else if (niw > 1) pot.setInterfaceSize(vector<double>(iw, iw + niw));
  pot.setDensityConstraint(density_name(density));
  return pot;
}

struct Result {
  int iterations;
  int converged;
  long n_energy, n_gradient;
  double energy;
  double grad_rms;
};

void fill_result(Result* r, const State& s, int iter, int maxIter, Minimiser& min) {
  (void)min;
  r->iterations = iter;
  r->converged = iter <= maxIter;
  r->n_energy = g_nEnergy;
  r->n_gradient = g_nGradient;
  r->energy = s.energy();
  vector<double> g = s.gradient();
  double gg = 0;
  for (double v : g) gg += v * v;
  r->grad_rms = std::sqrt(gg / g.size());
}

}  // namespace

extern "C" {

// ---- PhaseField: persistent State so that construction is timed separately ----
void* ref_pf_state_create(const int* grid, int nfluid, double res, const double* st, int nst,
                          const double* iw, int niw, int density, const double* x) {
  try {
    CountingPF pot = make_pf(grid, nfluid, res, st, nst, iw, niw, density);
    size_t n = (size_t)grid[0] * grid[1] * grid[2] * nfluid;
    return new State(pot, vector<double>(x, x + n));
  } catch (...) {
    return nullptr;
  }
}

void ref_pf_state_destroy(void* h) { delete static_cast<State*>(h); }

// e and/or g may be NULL; calls State::energy / State::gradient /
// State::energyGradient on the given coordinates, as the minimisers do.
int ref_state_eval(void* h, const double* x, double* e, double* g) {
  State* s = static_cast<State*>(h);
  vector<double> xv(x, x + s->ndof);
  try {
    if (e && g) {
      vector<double> gv;
      s->energyGradient(xv, e, &gv);
      std::memcpy(g, gv.data(), sizeof(double) * gv.size());
    } else if (e) {
      *e = s->energy(xv);
    } else if (g) {
      vector<double> gv = s->gradient(xv);
      std::memcpy(g, gv.data(), sizeof(double) * gv.size());
    }
  } catch (...) {
    return 1;
  }
  return 0;
}

int ref_pf_params(const int* grid, int nfluid, double res, const double* st, int nst,
                  const double* iw, int niw, int density, double* kappa, double* kappaP,
                  double* convergence) {
  try {
    CountingPF pot = make_pf(grid, nfluid, res, st, nst, iw, niw, density);
    size_t n = (size_t)grid[0] * grid[1] * grid[2] * nfluid;
    State s(pot, vector<double>(n, 0.0));
    const PhaseField& p = static_cast<const PhaseField&>(*s.pot);
    for (int i = 0; i < nfluid; i++) {
      kappa[i] = p.kappa[i];
      kappaP[i] = p.kappaP[i];
    }
    *convergence = s.convergence;
  } catch (...) {
    return 1;
  }
  return 0;
}

// minimiser: 0 = Lbfgs, 1 = Fire, 2 = GradDescent.  ls: 0 backtracking, 1 none.
// param_a: Lbfgs maxStep | Fire dtMax | GradDescent alpha.  param_b: Lbfgs initStep.
int ref_state_minimise(void* h, int minimiser, int m, int maxIter, double param_a, double param_b,
                       int ls, double convergence, double* x_out, Result* r) {
  State* s = static_cast<State*>(h);
  if (convergence > 0) s->convergence = convergence;
  g_nEnergy = g_nGradient = g_nBoth = 0;
  try {
    std::unique_ptr<Minimiser> min;
    if (minimiser == 0) {
      Lbfgs l;
      l.setM(m).setMaxIter(maxIter);
      l.setMaxStep(param_a).setInitStep(param_b);
      min = l.clone();
    } else if (minimiser == 1) {
      Fire f;
      f.setMaxIter(maxIter);
      f.setDtMax(param_a);
      min = f.clone();
    } else {
      GradDescent gd;
      gd.setMaxIter(maxIter);
      gd.setAlpha(param_a);
      min = gd.clone();
    }
    min->setLinesearch(ls == 0 ? "backtracking" : "none");
    vector<double> xr = min->minimise(*s);
    std::memcpy(x_out, xr.data(), sizeof(double) * xr.size());
    long ne = g_nEnergy, ng = g_nGradient;
    fill_result(r, *s, min->iter, maxIter, *min);
    r->n_energy = ne;
    r->n_gradient = ng;
  } catch (...) {
    return 1;
  }
  return 0;
}

// ---- LjNd (all pairs) ----
void* ref_lj_state_create(int ndim, double sigma, double epsilon, const double* x, int ndof) {
  try {
    LjNd pot(ndim, sigma, epsilon);
    return new State(pot, vector<double>(x, x + ndof));
  } catch (...) {
    return nullptr;
  }
}

// ---- BarAndHinge ----
// Builds the reference's element lists for a triangulation (bond and hinge
// node orders from computeBondList / computeHingeList) and returns them so
// that other implementations can be fed identical elements.
// bars_out[2*nb], bar_k/l0; hinges_out[4*nh], hinge_k/t0.  Returns counts via pointers.
void* ref_bah_state_create(const int* tri, int ntri, double kb, double kh, int set_l0, double l0,
                           int set_t0, double t0, const int* fixed_dofs, int nfixed,
                           const double* force3, const double* x, int ndof) {
  try {
    BarAndHinge pot;
    vector<vector<int>> tl(ntri);
    for (int i = 0; i < ntri; i++) tl[i] = {tri[3 * i], tri[3 * i + 1], tri[3 * i + 2]};
    pot.setTriangulation(tl);
    pot.setRigidity(kb, kh);
    if (set_l0) pot.setLength0(l0);
    if (set_t0) pot.setTheta0(t0);
    if (force3) pot.setForce(vector<double>(force3, force3 + 3));
    if (nfixed > 0) pot.setConstraints(vector<int>(fixed_dofs, fixed_dofs + nfixed));
    return new State(pot, vector<double>(x, x + ndof));
  } catch (...) {
    return nullptr;
  }
}

// Copies out the element list of a BarAndHinge state (types 0 and 1 only).
int ref_bah_elements(void* h, int* nbar, int* bars, double* bar_k, double* bar_l0, int* nhinge,
                     int* hinges, double* hinge_k, double* hinge_t0, int cap) {
  State* s = static_cast<State*>(h);
  int nb = 0, nh = 0;
  for (const auto& el : s->pot->elements) {
    if (el.type == 0) {
      if (nb >= cap) return 1;
      bars[2 * nb] = el.idof[0] / 3;
      bars[2 * nb + 1] = el.idof[3] / 3;
      bar_k[nb] = el.parameters[0];
      bar_l0[nb] = el.parameters[1];
      nb++;
    } else if (el.type == 1) {
      if (nh >= cap) return 1;
      for (int k = 0; k < 4; k++) hinges[4 * nh + k] = el.idof[3 * k] / 3;
      hinge_k[nh] = el.parameters[0];
      hinge_t0[nh] = el.parameters[1];
      nh++;
    }
  }
  *nbar = nb;
  *nhinge = nh;
  return 0;
}

// Overwrites hinge parameter 1 (theta0) of hinge element ih, as
// BarAndHinge_test.cpp:27 does.
int ref_bah_set_hinge_theta0(void* h, int ih, double t0) {
  State* s = static_cast<State*>(h);
  int k = 0;
  for (auto& el : s->pot->elements) {
    if (el.type != 1) continue;
    if (k == ih) {
      el.parameters[1] = t0;
      return 0;
    }
    k++;
  }
  return 1;
}

// ---- PhaseField with its remaining terms (solid, wetting, pressure, force,
// confinement, volume) set through the reference's own setters ----
struct PFExtOpts {
  const unsigned char* solid;  // ngrid, or NULL
  const double* contact_angle;
  int n_ca;
  const double* pressure;  // nfluid, or NULL
  const double* force3;    // 3, or NULL
  const int* force_fluids;
  int n_force_fluids;
  const double* confinement;  // nfluid, or NULL
  int volume_fixed;
  double vol_const;
  const double* volume;  // nfluid, or NULL
  const unsigned char* fix_fluid;  // nfluid, or NULL
  double density_const;
};

static CountingPF make_pf_ext(const int* grid, int nfluid, double res, const double* st, int nst, const double* iw,
                              int niw, int density, const PFExtOpts* o) {
  CountingPF pot = make_pf(grid, nfluid, res, st, nst, iw, niw, density);
  const size_t ng = (size_t)grid[0] * grid[1] * grid[2];
  pot.densityConst = o->density_const;
  if (o->solid) pot.setSolid(vector<char>(o->solid, o->solid + ng));
  if (o->contact_angle) pot.setContactAngle(vector<double>(o->contact_angle, o->contact_angle + o->n_ca));
  if (o->pressure) pot.setPressure(vector<double>(o->pressure, o->pressure + nfluid));
  if (o->force3)
    pot.setForce(vector<double>(o->force3, o->force3 + 3),
                 vector<int>(o->force_fluids, o->force_fluids + o->n_force_fluids));
  if (o->confinement) pot.setConfinement(vector<double>(o->confinement, o->confinement + nfluid));
  if (o->volume) pot.setVolume(vector<double>(o->volume, o->volume + nfluid), o->vol_const);
  else if (o->volume_fixed) pot.setVolumeFixed(true, o->vol_const);
  if (o->fix_fluid)
    for (int i = 0; i < nfluid; i++)
      if (o->fix_fluid[i]) pot.setFixFluid(i);
  return pot;
}

void* ref_pf_state_create_ex(const int* grid, int nfluid, double res, const double* st, int nst, const double* iw,
                             int niw, int density, const double* x, const PFExtOpts* o) {
  try {
    CountingPF pot = make_pf_ext(grid, nfluid, res, st, nst, iw, niw, density, o);
    size_t n = (size_t)grid[0] * grid[1] * grid[2] * nfluid;
    return new State(pot, vector<double>(x, x + n));
  } catch (...) {
    return nullptr;
  }
}

// PhaseField::energyComponents at the state's coordinates: out[2c], out[2c+1]
// = energy and gradient norm of fluid, surface, pressure, density constraint,
// force, confinement, volume constraint.
int ref_pf_components(void* h, double* out) {
  State* s = static_cast<State*>(h);
  try {
    const PhaseField& p = static_cast<const PhaseField&>(*s->pot);
    auto m = p.energyComponents(s->coords(), *s->comm);
    const char* names[7] = {"fluid", "surface", "pressure", "density constraint", "force", "confinement",
                            "volume constraint"};
    for (int c = 0; c < 7; c++) {
      out[2 * c] = m.at(names[c])[0];
      out[2 * c + 1] = m.at(names[c])[1];
    }
  } catch (...) {
    return 1;
  }
  return 0;
}

// Node volumes and surface areas after initLocal.
int ref_pf_geometry(void* h, double* node_vol, double* surface_area) {
  State* s = static_cast<State*>(h);
  const PhaseField& p = static_cast<const PhaseField&>(*s->pot);
  std::memcpy(node_vol, p.nodeVol.data(), sizeof(double) * p.nodeVol.size());
  std::memcpy(surface_area, p.surfaceArea.data(), sizeof(double) * p.surfaceArea.size());
  return 0;
}

// State::applyConstraints (src/State.cpp:321-323) on a vector of ndof entries.
int ref_state_constraints(void* h, double* data) {
  State* s = static_cast<State*>(h);
  try {
    vector<double> d(data, data + s->ndof);
    s->applyConstraints(d);
    std::memcpy(data, d.data(), sizeof(double) * d.size());
  } catch (...) {
    return 1;
  }
  return 0;
}

// PhaseField::diffuseSolid (PhaseField.cpp:1005-1035) of a potential built
// like ref_pf_state_create_ex; out has ngrid * nfluid entries.
int ref_pf_diffuse_solid(const int* grid, int nfluid, double res, const double* st, int nst, const double* iw,
                         int niw, int density, const PFExtOpts* o, const unsigned char* solid_in, int ifluid,
                         int two_step, double* out) {
  try {
    CountingPF pot = make_pf_ext(grid, nfluid, res, st, nst, iw, niw, density, o);
    const size_t ng = (size_t)grid[0] * grid[1] * grid[2];
    PhaseField base(pot);
    vector<double> r = base.diffuseSolid(vector<char>(solid_in, solid_in + ng), ifluid, two_step != 0);
    std::memcpy(out, r.data(), sizeof(double) * r.size());
  } catch (...) {
    return 1;
  }
  return 0;
}

// ---- PhaseFieldUnstructured through its own setters ----
void* ref_pfu_state_create_ex(const int* grid, int nfluid, double res, const double* st, int nst, const double* iw,
                              int niw, int density, const double* x, const PFExtOpts* o) {
  try {
    PhaseFieldUnstructured pot;
    pot.setNFluid(nfluid);
    pot.setGridSize({grid[0], grid[1], grid[2]});
    pot.setResolution(res);
    pot.setSurfaceTension(vector<double>(st, st + nst));
    if (niw == 1) pot.setInterfaceSize(iw[0]);
    else if (niw > 1) pot.setInterfaceSize(vector<double>(iw, iw + niw));
    pot.setDensityConstraint(density_name(density));
    const size_t ng = (size_t)grid[0] * grid[1] * grid[2];
    pot.densityConst = o->density_const;
    if (o->solid) pot.setSolid(vector<char>(o->solid, o->solid + ng));
    if (o->contact_angle) pot.setContactAngle(vector<double>(o->contact_angle, o->contact_angle + o->n_ca));
    if (o->pressure) pot.setPressure(vector<double>(o->pressure, o->pressure + nfluid));
    if (o->force3)
      pot.setForce(vector<double>(o->force3, o->force3 + 3),
                   vector<int>(o->force_fluids, o->force_fluids + o->n_force_fluids));
    if (o->confinement) pot.setConfinement(vector<double>(o->confinement, o->confinement + nfluid));
    if (o->volume) pot.setVolume(vector<double>(o->volume, o->volume + nfluid), o->vol_const);
    else if (o->volume_fixed) pot.setVolumeFixed(true, o->vol_const);
    if (o->fix_fluid)
      for (int i = 0; i < nfluid; i++)
        if (o->fix_fluid[i]) pot.setFixFluid(i);
    size_t n = ng * nfluid;
    return new State(pot, vector<double>(x, x + n));
  } catch (...) {
    return nullptr;
  }
}

// Number of energy elements of a state's potential.
long ref_state_nelements(void* h) { return (long)static_cast<State*>(h)->pot->elements.size(); }

void ref_state_destroy(void* h) { delete static_cast<State*>(h); }
long ref_state_ndof(void* h) { return (long)static_cast<State*>(h)->ndof; }

}  // extern "C"
