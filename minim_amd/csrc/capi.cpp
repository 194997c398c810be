// capi.cpp -- the extern "C" boundary declared in include/minim_hip.h.
// Maps the C++ classes' exceptions to status codes (std::invalid_argument ->
// MINIM_EINVAL, HIP failures -> MINIM_EHIP, other std::exception -> MINIM_ERUNTIME).
#include "minim_hip.h"

#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "minim/LjNd.h"
#include "minim/BarAndHinge.h"
#include "minim/Minimiser.h"
#include "minim/PhaseField.h"
#include "minim/PhaseFieldUnstructured.h"
#include "minim/State.h"
#include "autotune.hpp"
#include "runtime.hpp"
#include "vec.hpp"

struct minim_potential {
  std::unique_ptr<minim::Potential> p;
  bool initialised = false;
  size_t ndof = 0;
};
struct minim_state {
  std::unique_ptr<minim::State> s;
};
struct minim_minimiser {
  std::unique_ptr<minim::Minimiser> m;
};

namespace {

thread_local std::string g_err;

template <typename F>
int guard(F&& f) {
  try {
    f();
    return MINIM_OK;
  } catch (const std::invalid_argument& e) {
    g_err = e.what();
    return MINIM_EINVAL;
  } catch (const minim_amd::HipError& e) {
    g_err = e.what();
    return MINIM_EHIP;
  } catch (const std::bad_alloc& e) {
    g_err = "out of memory";
    return MINIM_ENOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return MINIM_ERUNTIME;
  } catch (...) {
    g_err = "unknown error";
    return MINIM_ERUNTIME;
  }
}

void need(const void* p, const char* what) {
  if (!p) throw std::invalid_argument(std::string(what) + " must not be NULL");
}

void check_prec(int prec) {
  if (prec != MINIM_FP64 && prec != MINIM_FP32) throw std::invalid_argument("precision must be MINIM_FP64 or MINIM_FP32");
}

// The setters both phase-field classes share (PhaseField.cpp:747-900,
// PhaseFieldUnstructured.cpp:727-856), called in the order a user would.
template <typename PF>
void apply_pf_params(const minim_pf_params* p, PF* pf) {
  pf->setNFluid(p->n_fluid);
  pf->setGridSize({p->grid[0], p->grid[1], p->grid[2]});
  pf->setResolution(p->resolution);
  if (p->n_surface_tension > 0) {
    need(p->surface_tension, "surface_tension");
    pf->setSurfaceTension(std::vector<double>(p->surface_tension, p->surface_tension + p->n_surface_tension));
  }
  if (p->n_interface_size == 1) {
    need(p->interface_size, "interface_size");
    pf->setInterfaceSize(p->interface_size[0]);
  } else if (p->n_interface_size > 1) {
    need(p->interface_size, "interface_size");
    pf->setInterfaceSize(std::vector<double>(p->interface_size, p->interface_size + p->n_interface_size));
  }
  static const char* names[] = {"hard", "soft", "none", "fixed"};
  if (p->density_constraint < 0 || p->density_constraint > 3)
    throw std::invalid_argument("Invalid density constraint.");
  pf->setDensityConstraint(names[p->density_constraint]);
  pf->densityConst = p->density_const;
  if (p->fix_fluid) {
    for (int k = 0; k < p->n_fluid; k++) pf->setFixFluid(k, p->fix_fluid[k] != 0);
  }
  const long ng = (long)p->grid[0] * p->grid[1] * p->grid[2];
  if (p->solid) pf->setSolid(std::vector<char>(p->solid, p->solid + ng));
  if (p->contact_angle) pf->setContactAngle(std::vector<double>(p->contact_angle, p->contact_angle + p->n_contact_angle));
  if (p->pressure) pf->setPressure(std::vector<double>(p->pressure, p->pressure + p->n_fluid));
  if (p->force) {
    std::vector<int> fl;
    if (p->n_force_fluids > 0) {
      need(p->force_fluids, "force_fluids");
      fl.assign(p->force_fluids, p->force_fluids + p->n_force_fluids);
    }
    pf->setForce(std::vector<double>(p->force, p->force + 3), fl);
  }
  if (p->confinement) pf->setConfinement(std::vector<double>(p->confinement, p->confinement + p->n_fluid));
  if (p->volume) pf->setVolume(std::vector<double>(p->volume, p->volume + p->n_fluid), p->vol_const);
  else if (p->volume_fixed) pf->setVolumeFixed(true, p->vol_const);
}

}  // namespace

extern "C" {

const char* minim_last_error(void) { return g_err.c_str(); }
int minim_abi_version(void) { return 10; }

// ---------------------------------------------------------------- potentials
int minim_pf_create(const minim_pf_params* p, minim_potential** out) {
  return guard([&] {
    need(p, "params");
    need(out, "out");
    auto pf = std::make_unique<minim::PhaseField>();
    apply_pf_params(p, pf.get());
    pf->stencilXChunk = p->stencil_xchunk;
    if (p->stencil_target_blocks > 0) pf->stencilTargetBlocks = p->stencil_target_blocks;
    pf->stencilKernel = p->stencil_kernel;
    auto* h = new minim_potential;
    h->p = std::move(pf);
    *out = h;
  });
}

int minim_pfu_create(const minim_pf_params* p, minim_potential** out) {
  return guard([&] {
    need(p, "params");
    need(out, "out");
    auto pf = std::make_unique<minim::PhaseFieldUnstructured>();
    apply_pf_params(p, pf.get());
    auto* h = new minim_potential;
    h->p = std::move(pf);
    *out = h;
  });
}

int minim_pfu_element_list(const minim_potential* pot, size_t* n_elements, size_t* device_bytes) {
  return guard([&] {
    need(pot, "pot");
    auto* pf = dynamic_cast<const minim::PhaseFieldUnstructured*>(pot->p.get());
    if (!pf) throw std::invalid_argument("not a PhaseFieldUnstructured potential");
    if (!pot->initialised) throw std::runtime_error("PhaseFieldUnstructured: call minim_potential_init first");
    if (n_elements) *n_elements = pf->nElements;
    if (device_bytes) *device_bytes = pf->listBytes;
  });
}

int minim_pf_coefficients(const minim_potential* pot, double* kappa, double* kappa_p) {
  return guard([&] {
    need(pot, "pot");
    const std::vector<double>* k = nullptr;
    const std::vector<double>* kp = nullptr;
    if (auto* pf = dynamic_cast<const minim::PhaseField*>(pot->p.get())) {
      k = &pf->kappa;
      kp = &pf->kappaP;
    } else if (auto* pu = dynamic_cast<const minim::PhaseFieldUnstructured*>(pot->p.get())) {
      k = &pu->kappa;
      kp = &pu->kappaP;
    } else {
      throw std::invalid_argument("not a PhaseField potential");
    }
    if (!pot->initialised) throw std::runtime_error("PhaseField: call minim_potential_init first");
    for (size_t i = 0; i < k->size(); i++) {
      if (kappa) kappa[i] = (*k)[i];
      if (kappa_p) kappa_p[i] = (*kp)[i];
    }
  });
}

namespace {
const minim::PhaseField& pf_of(const minim_potential* pot, bool initialised) {
  need(pot, "pot");
  auto* pf = dynamic_cast<const minim::PhaseField*>(pot->p.get());
  if (!pf) throw std::invalid_argument("not a PhaseField potential");
  if (initialised && !pot->initialised) throw std::runtime_error("PhaseField: call minim_potential_init first");
  return *pf;
}
}  // namespace

int minim_pf_energy_components(const minim_potential* pot, const double* coords, size_t ndof, double* out) {
  return guard([&] {
    const minim::PhaseField& pf = pf_of(pot, true);
    need(coords, "coords");
    need(out, "out");
    auto m = pf.energyComponents(std::vector<double>(coords, coords + ndof));
    static const char* names[7] = {"fluid", "surface", "pressure", "density constraint", "force", "confinement",
                                   "volume constraint"};
    for (int c = 0; c < 7; c++) {
      out[2 * c] = m.at(names[c])[0];
      out[2 * c + 1] = m.at(names[c])[1];
    }
  });
}

int minim_pf_node_volumes(const minim_potential* pot, double* node_vol, double* surface_area) {
  return guard([&] {
    need(pot, "pot");
    if (auto* pu = dynamic_cast<const minim::PhaseFieldUnstructured*>(pot->p.get())) {
      // the element form keeps node volumes per DOF and no surface areas
      // (PhaseFieldUnstructured.cpp:252-268)
      if (!pot->initialised) throw std::runtime_error("PhaseField: call minim_potential_init first");
      if (surface_area) throw std::invalid_argument("PhaseFieldUnstructured keeps no surface areas");
      for (size_t i = 0; i < (size_t)pu->nGrid; i++)
        if (node_vol) node_vol[i] = pu->nodeVol[i * pu->nFluid];
      return;
    }
    const minim::PhaseField& pf = pf_of(pot, true);
    const size_t ng = (size_t)pf.nGrid;
    for (size_t i = 0; i < ng; i++) {
      if (node_vol) node_vol[i] = pf.nodeVol.empty() ? std::pow(pf.resolution, 3) : pf.nodeVol[i];
      if (surface_area) surface_area[i] = pf.surfaceArea.empty() ? 0.0 : pf.surfaceArea[i];
    }
  });
}

int minim_pf_diffuse_solid(const minim_potential* pot, const unsigned char* solid, int i_fluid, int two_step,
                           double* out) {
  return guard([&] {
    need(pot, "pot");
    need(solid, "solid");
    need(out, "out");
    std::vector<double> r;
    if (auto* pu = dynamic_cast<const minim::PhaseFieldUnstructured*>(pot->p.get())) {
      if (pu->gridSize.size() != 3) throw std::invalid_argument("PhaseField: gridSize must have three entries.");
      const size_t ng = (size_t)pu->gridSize[0] * pu->gridSize[1] * pu->gridSize[2];
      r = minim::PhaseFieldUnstructured::diffuseSolid(std::vector<char>(solid, solid + ng), *pu, i_fluid,
                                                       two_step != 0);
    } else {
      const minim::PhaseField& pf = pf_of(pot, false);
      if (pf.gridSize.size() != 3) throw std::invalid_argument("PhaseField: gridSize must have three entries.");
      const size_t ng = (size_t)pf.gridSize[0] * pf.gridSize[1] * pf.gridSize[2];
      r = minim::PhaseField::diffuseSolid(std::vector<char>(solid, solid + ng), pf, i_fluid, two_step != 0);
    }
    std::memcpy(out, r.data(), sizeof(double) * r.size());
  });
}

int minim_pf_site_types(const int grid[3], const unsigned char* solid, int* type) {
  return guard([&] {
    need(grid, "grid");
    need(solid, "solid");
    need(type, "type");
    for (int d = 0; d < 3; d++)
      if (grid[d] < 1) throw std::invalid_argument("PhaseField: grid sizes must be positive.");
    const size_t ng = (size_t)grid[0] * grid[1] * grid[2];
    std::vector<char> s(solid, solid + ng);
    std::vector<int> t = minim::pf_site_types(grid, s.data());
    std::memcpy(type, t.data(), sizeof(int) * ng);
  });
}

int minim_lj_create(int ndim, double sigma, double epsilon, minim_potential** out) {
  return guard([&] {
    need(out, "out");
    auto* h = new minim_potential;
    h->p = std::make_unique<minim::LjNd>(ndim, sigma, epsilon);
    *out = h;
  });
}

int minim_lj_set_cutoff(minim_potential* pot, double rc, const double box[3], double skin) {
  return guard([&] {
    need(pot, "pot");
    need(box, "box");
    auto* lj = dynamic_cast<minim::LjNd*>(pot->p.get());
    if (!lj) throw std::invalid_argument("not a LjNd potential");
    lj->setCutoff(rc, {box[0], box[1], box[2]}, skin);
    pot->initialised = false;
  });
}

int minim_lj_list_builds(const minim_state* state, long* builds) {
  return guard([&] {
    need(state, "state");
    need(builds, "builds");
    auto* lj = dynamic_cast<const minim::LjNd*>(state->s->pot.get());
    if (!lj) throw std::invalid_argument("not a LjNd potential");
    *builds = lj->listBuilds();
  });
}

int minim_lj_max_neighbours(const minim_state* state, int* longest) {
  return guard([&] {
    need(state, "state");
    need(longest, "longest");
    auto* lj = dynamic_cast<const minim::LjNd*>(state->s->pot.get());
    if (!lj) throw std::invalid_argument("not a LjNd potential");
    *longest = lj->maxNeighbours();
  });
}

int minim_bah_create(const minim_bah_params* p, minim_potential** out) {
  return guard([&] {
    need(p, "params");
    need(out, "out");
    auto b = std::make_unique<minim::BarAndHinge>();
    if (p->n_bar > 0 || p->n_hinge > 0) {
      if ((p->n_bar && !p->bars) || (p->n_hinge && !p->hinges)) throw std::invalid_argument("bars/hinges must not be NULL");
      std::vector<std::vector<int>> bl(p->n_bar), hl(p->n_hinge);
      for (int i = 0; i < p->n_bar; i++) bl[i] = {p->bars[2 * i], p->bars[2 * i + 1]};
      for (int i = 0; i < p->n_hinge; i++)
        hl[i] = {p->hinges[4 * i], p->hinges[4 * i + 1], p->hinges[4 * i + 2], p->hinges[4 * i + 3]};
      b->setBondList(bl);
      b->setHingeList(hl);
    } else {
      if (p->n_triangle <= 0) throw std::invalid_argument("BarAndHinge: The triangulation has not been set.");
      need(p->triangles, "triangles");
      std::vector<std::vector<int>> tl(p->n_triangle);
      for (int i = 0; i < p->n_triangle; i++)
        tl[i] = {p->triangles[3 * i], p->triangles[3 * i + 1], p->triangles[3 * i + 2]};
      b->setTriangulation(tl);
    }
    if (p->bar_k || p->hinge_k) {
      if (!p->bar_k || !p->hinge_k) throw std::invalid_argument("bar_k and hinge_k must be given together");
      b->setRigidity(std::vector<double>(p->bar_k, p->bar_k + p->n_bar),
                     std::vector<double>(p->hinge_k, p->hinge_k + p->n_hinge));
    } else {
      b->setRigidity(p->k_bar, p->k_hinge);
    }
    if (p->bar_l0) b->setLength0(std::vector<double>(p->bar_l0, p->bar_l0 + p->n_bar));
    else if (!std::isnan(p->l0)) b->setLength0(p->l0);
    if (p->hinge_theta0) b->setTheta0(std::vector<double>(p->hinge_theta0, p->hinge_theta0 + p->n_hinge));
    else if (!std::isnan(p->theta0)) b->setTheta0(p->theta0);
    if (p->force) {
      if (p->force_per_node) b->setForce(std::vector<double>(p->force, p->force + 3 * (size_t)p->n_node));
      else b->setForce(std::vector<double>(p->force, p->force + 3));
    }
    b->nNodeHint = p->n_node;
    auto* h = new minim_potential;
    h->p = std::move(b);
    *out = h;
  });
}

int minim_potential_set_fixed(minim_potential* pot, const int* idof, size_t n) {
  return guard([&] {
    need(pot, "pot");
    if (n) need(idof, "idof");
    pot->p->setConstraints(std::vector<int>(idof, idof + n));
    pot->initialised = false;
  });
}

int minim_potential_init(minim_potential* pot, const double* coords, size_t ndof) {
  return guard([&] {
    need(pot, "pot");
    if (ndof) need(coords, "coords");
    pot->p->init(std::vector<double>(coords, coords + ndof));
    pot->initialised = true;
    pot->ndof = ndof;
  });
}

int minim_potential_convergence(const minim_potential* pot, double* out) {
  return guard([&] {
    need(pot, "pot");
    need(out, "out");
    *out = pot->p->convergence;
  });
}

int minim_potential_destroy(minim_potential* pot) {
  return guard([&] { delete pot; });
}

int minim_potential_eval(const minim_potential* pot, int precision, const void* x, size_t ndof, const void* d,
                         double t, double* e, void* g, void* stream) {
  return guard([&] {
    need(pot, "pot");
    need(x, "x");
    check_prec(precision);
    if (!pot->initialised) throw std::runtime_error("potential not initialised: call minim_potential_init");
    if (ndof != pot->ndof) throw std::invalid_argument("ndof does not match the initialised potential");
    if (d && g) throw std::invalid_argument("a gradient at a trial point (d != NULL) is not supported");
    pot->p->evaluate(precision, x, d, t, e, g, stream);
  });
}

int minim_potential_eval_async(const minim_potential* pot, int precision, const void* x, size_t ndof,
                               const void* d, double t, int want_energy, double* e_dev, void* g, void* stream) {
  return guard([&] {
    need(pot, "pot");
    need(x, "x");
    check_prec(precision);
    if (!pot->initialised) throw std::runtime_error("potential not initialised: call minim_potential_init");
    if (ndof != pot->ndof) throw std::invalid_argument("ndof does not match the initialised potential");
    if (d && g) throw std::invalid_argument("a gradient at a trial point (d != NULL) is not supported");
    pot->p->evaluateAsync(precision, x, d, t, want_energy != 0, e_dev, g, stream);
  });
}

int minim_potential_eval_async_noalloc(const minim_potential* pot, int precision, const void* x, size_t ndof,
                                       const void* d, double t, int want_energy, double* e_dev, void* g,
                                       void* stream) {
  return guard([&] {
    need(pot, "pot");
    need(x, "x");
    check_prec(precision);
    if (!pot->initialised) throw std::runtime_error("potential not initialised: call minim_potential_init");
    if (ndof != pot->ndof) throw std::invalid_argument("ndof does not match the initialised potential");
    if (d && g) throw std::invalid_argument("a gradient at a trial point (d != NULL) is not supported");
    auto* pf = dynamic_cast<const minim::PhaseField*>(pot->p.get());
    if (!pf) throw std::invalid_argument("the allocation-free evaluation supports PhaseField only");
    size_t nPartials = 0;
    double* scratch = minim_amd::noalloc_scratch(&nPartials);
    pf->evaluateAsyncScratch(precision, x, d, t, want_energy != 0, e_dev, g, scratch, nPartials, stream);
  });
}

int minim_pf_autotune(minim_potential* pot, int precision, const double* coords, size_t ndof, const int* targets,
                      int n_targets, const int* marches, int n_marches, int reps, double* times, int* best_target,
                      int* best_march) {
  return guard([&] {
    need(pot, "pot");
    need(coords, "coords");
    need(targets, "targets");
    need(marches, "marches");
    check_prec(precision);
    if (!pot->initialised) throw std::runtime_error("potential not initialised: call minim_potential_init");
    if (ndof != pot->ndof) throw std::invalid_argument("ndof does not match the initialised potential");
    if (n_targets < 1 || n_marches < 1) throw std::invalid_argument("no candidate shapes");
    auto* pf = dynamic_cast<minim::PhaseField*>(pot->p.get());
    if (!pf || dynamic_cast<minim::PhaseFieldUnstructured*>(pot->p.get()))
      throw std::invalid_argument("the launch-shape autotuner is for PhaseField");
    std::vector<int> tv(targets, targets + n_targets), mv(marches, marches + n_marches);
    const auto rows = pf->autotune(std::vector<double>(coords, coords + ndof), precision, tv, mv, reps);
    if (times) {
      for (int i = 0; i < n_targets * n_marches; i++) times[i] = -1.0;
      for (const auto& r : rows)
        for (int i = 0; i < n_targets; i++)
          for (int j = 0; j < n_marches; j++)
            if ((int)r[0] == targets[i] && (int)r[1] == marches[j]) times[i * n_marches + j] = r[2];
    }
    const auto best = pf->tunedShape(precision);
    if (best_target) *best_target = best.first;
    if (best_march) *best_march = best.second;
  });
}

int minim_pf_tuned_shape(const minim_potential* pot, int precision, int* target, int* march) {
  return guard([&] {
    need(pot, "pot");
    check_prec(precision);
    auto* pf = dynamic_cast<const minim::PhaseField*>(pot->p.get());
    if (!pf || dynamic_cast<const minim::PhaseFieldUnstructured*>(pot->p.get()))
      throw std::invalid_argument("launch shapes are tuned for PhaseField");
    if (!pot->initialised) throw std::runtime_error("potential not initialised: call minim_potential_init");
    const auto best = pf->tunedShape(precision);
    if (target) *target = best.first;
    if (march) *march = best.second;
  });
}

int minim_lbfgs_autotune(minim_minimiser* m, minim_state* s, const int* blocks, int n_blocks, int reps,
                         double* times, int* best_direction, int* best_gram) {
  return guard([&] {
    need(m, "minimiser");
    need(s, "state");
    need(blocks, "blocks");
    auto* lb = dynamic_cast<minim::Lbfgs*>(m->m.get());
    if (!lb) throw std::invalid_argument("the vector-pass autotuner is for Lbfgs");
    const auto rows = lb->autotune(*s->s, std::vector<int>(blocks, blocks + n_blocks), reps);
    if (times)
      for (size_t i = 0; i < rows.size(); i++) times[i] = rows[i][2];
    if (best_direction) *best_direction = lb->tunedBlocks().first;
    if (best_gram) *best_gram = lb->tunedBlocks().second;
  });
}

int minim_lbfgs_autotune_batch(minim_minimiser* m, minim_state* const* states, int n_states, const int* candidates,
                               int n_candidates, int replays, double* seconds, int* best) {
  return guard([&] {
    need(m, "minimiser");
    need(states, "states");
    need(candidates, "candidates");
    if (n_states < 1 || n_candidates < 1) throw std::invalid_argument("nothing to time");
    auto* lb = dynamic_cast<minim::Lbfgs*>(m->m.get());
    if (!lb) throw std::invalid_argument("the batch autotuner is for Lbfgs");
    std::vector<minim::State*> ss(n_states);
    for (int i = 0; i < n_states; i++) {
      need(states[i], "state");
      ss[i] = states[i]->s.get();
    }
    const auto rows = lb->autotuneBatch(ss, std::vector<int>(candidates, candidates + n_candidates), replays);
    int b = 0;
    double bs = 0;
    for (size_t i = 0; i < rows.size(); i++) {
      if (seconds) seconds[i] = rows[i][1];
      if (b == 0 || rows[i][1] < bs) {
        b = (int)rows[i][0];
        bs = rows[i][1];
      }
    }
    if (best) *best = b;
  });
}

int minim_autotune_clear(void) {
  return guard([&] { minim_amd::tune_clear(); });
}

int minim_potential_eval_mixed(const minim_potential* pot, const void* x, size_t ndof, const void* d, double t,
                               double* e, void* g, void* stream) {
  return guard([&] {
    need(pot, "pot");
    need(x, "x");
    if (!pot->initialised) throw std::runtime_error("potential not initialised: call minim_potential_init");
    if (ndof != pot->ndof) throw std::invalid_argument("ndof does not match the initialised potential");
    if (d && g) throw std::invalid_argument("a gradient at a trial point (d != NULL) is not supported");
    if (!pot->p->supportsMixed()) throw std::invalid_argument("this potential has no mixed FP64/FP32 evaluation");
    pot->p->evaluateMixed(x, d, t, e, g, stream);
  });
}

int minim_potential_eval_step(const minim_potential* pot, int precision_x, int precision_v, const void* x,
                              size_t ndof, const void* d, double t, void* x_out, double* e, void* g, void* stream) {
  return guard([&] {
    need(pot, "pot");
    need(x, "x");
    need(d, "d");
    need(x_out, "x_out");
    need(g, "g");
    check_prec(precision_x);
    check_prec(precision_v);
    if (!pot->initialised) throw std::runtime_error("potential not initialised: call minim_potential_init");
    if (ndof != pot->ndof) throw std::invalid_argument("ndof does not match the initialised potential");
    if (!pot->p->supportsStep()) throw std::invalid_argument("this potential has no fused step evaluation");
    hipStream_t s = static_cast<hipStream_t>(stream);
    pot->p->evaluateStepAsync(precision_x, precision_v, x, d, t, x_out, g, stream);
    if (e) {
      auto& ws = minim_amd::Workspace::current();
      ws.fetch(s, 1, e);
    } else {
      MINIM_HIP_CHECK(hipStreamSynchronize(s));
    }
  });
}

int minim_potential_time_kernel(const minim_potential* pot, int precision, const void* x, size_t ndof,
                                int want_energy, void* g, int reps, void* stream, double* ms_per_launch) {
  return guard([&] {
    need(pot, "pot");
    need(x, "x");
    need(ms_per_launch, "ms_per_launch");
    check_prec(precision);
    if (!pot->initialised) throw std::runtime_error("potential not initialised: call minim_potential_init");
    if (ndof != pot->ndof) throw std::invalid_argument("ndof does not match the initialised potential");
    *ms_per_launch = pot->p->timeKernel(precision, x, want_energy != 0, g, reps, stream);
  });
}

int minim_potential_time_steps(const minim_potential* pot, int precision, const void* x, size_t ndof, void* g,
                               double* e_dev, int steps, double warm_ms, void* stream, double* out) {
  return guard([&] {
    need(pot, "pot");
    need(x, "x");
    need(g, "g");
    need(e_dev, "e_dev");
    need(out, "out");
    check_prec(precision);
    if (!pot->initialised) throw std::runtime_error("potential not initialised: call minim_potential_init");
    if (ndof != pot->ndof) throw std::invalid_argument("ndof does not match the initialised potential");
    pot->p->timeSteps(precision, x, g, e_dev, steps, warm_ms, stream, out);
  });
}

int minim_potential_constrain(const minim_potential* pot, int precision, void* v, size_t n, void* stream) {
  return guard([&] {
    need(pot, "pot");
    need(v, "v");
    check_prec(precision);
    if (!pot->initialised) throw std::runtime_error("potential not initialised: call minim_potential_init");
    if (n != pot->ndof) throw std::invalid_argument("n does not match the initialised potential");
    pot->p->constrainDevice(precision, nullptr, v, n, stream);
  });
}

int minim_dot(int precision, const void* a, const void* b, size_t n, double* out, void* stream) {
  return guard([&] {
    need(a, "a");
    need(b, "b");
    need(out, "out");
    check_prec(precision);
    auto& ws = minim_amd::Workspace::current();
    const void* pa[1] = {a};
    const void* pb[1] = {b};
    minim_amd::launch_dots(precision, pa, pb, 1, n, ws.rs(), static_cast<hipStream_t>(stream));
    ws.fetch(static_cast<hipStream_t>(stream), 1, out);
  });
}

// ---------------------------------------------------------------- State
int minim_state_create(const minim_potential* pot, const double* coords, size_t ndof, int precision,
                       minim_state** out) {
  return guard([&] {
    need(pot, "pot");
    need(out, "out");
    if (ndof) need(coords, "coords");
    check_prec(precision);
    auto* h = new minim_state;
    try {
      h->s = std::make_unique<minim::State>(*pot->p, std::vector<double>(coords, coords + ndof),
                                             static_cast<minim::Precision>(precision));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int minim_state_copy(const minim_state* s, minim_state** out) {
  return guard([&] {
    need(s, "state");
    need(out, "out");
    auto* h = new minim_state;
    h->s = std::make_unique<minim::State>(*s->s);
    *out = h;
  });
}

int minim_state_destroy(minim_state* s) {
  return guard([&] { delete s; });
}

int minim_state_energy(minim_state* s, const double* coords, double* e) {
  return guard([&] {
    need(s, "state");
    need(e, "e");
    *e = 0;
    s->s->energyGradientRaw(coords, e, nullptr);
  });
}

int minim_state_gradient(minim_state* s, const double* coords, double* g) {
  return guard([&] {
    need(s, "state");
    need(g, "g");
    if (!s->s->usesThisProc) return;
    s->s->energyGradientRaw(coords, nullptr, g);
  });
}

int minim_state_energy_gradient(minim_state* s, const double* coords, double* e, double* g) {
  return guard([&] {
    need(s, "state");
    s->s->energyGradientRaw(coords, e, g);
  });
}

int minim_state_get_coords(minim_state* s, double* out) {
  return guard([&] {
    need(s, "state");
    need(out, "out");
    const auto& c = s->s->blockCoords();
    std::memcpy(out, c.data(), sizeof(double) * c.size());
  });
}

int minim_state_set_coords(minim_state* s, const double* in) {
  return guard([&] {
    need(s, "state");
    need(in, "in");
    s->s->coords(std::vector<double>(in, in + s->s->ndof));
  });
}

int minim_state_component_energy(minim_state* s, int component, double* e) {
  return guard([&] {
    need(s, "state");
    need(e, "e");
    *e = s->s->componentEnergy(component);
  });
}

int minim_batch_report(const minim_minimiser* m, int* iterations, long* replays, long* host_syncs,
                       long* graph_launches, long* captures, int* block) {
  return guard([&] {
    need(m, "minimiser");
    auto* lb = dynamic_cast<const minim::Lbfgs*>(m->m.get());
    if (!lb) throw std::invalid_argument("batch reports come from Lbfgs");
    const auto& r = lb->batchReport;
    if (iterations) *iterations = r.iterations;
    if (replays) *replays = r.replays;
    if (host_syncs) *host_syncs = r.hostSyncs;
    if (graph_launches) *graph_launches = r.graphLaunches;
    if (captures) *captures = r.captures;
    if (block) *block = r.block;
  });
}

int minim_state_apply_constraints(minim_state* s, double* data) {
  return guard([&] {
    need(s, "state");
    need(data, "data");
    std::vector<double> v(data, data + s->s->ndof);
    s->s->applyConstraints(v);
    std::memcpy(data, v.data(), sizeof(double) * v.size());
  });
}

int minim_state_failed(minim_state* s) {
  return guard([&] {
    need(s, "state");
    s->s->failed();
  });
}

int minim_state_info(const minim_state* s, size_t* ndof, int* precision, double* convergence, int* is_failed) {
  return guard([&] {
    need(s, "state");
    if (ndof) *ndof = s->s->ndof;
    if (precision) *precision = s->s->precision();
    if (convergence) *convergence = s->s->convergence;
    if (is_failed) *is_failed = s->s->isFailed ? 1 : 0;
  });
}

int minim_state_set_convergence(minim_state* s, double c) {
  return guard([&] {
    need(s, "state");
    s->s->convergence = c;
  });
}

int minim_state_set_precision(minim_state* s, int precision) {
  return guard([&] {
    need(s, "state");
    check_prec(precision);
    s->s->setPrecision(precision);
  });
}

int minim_state_set_stream(minim_state* s, void* stream) {
  return guard([&] {
    need(s, "state");
    s->s->setStream(stream);
  });
}

void* minim_state_device_coords(minim_state* s) { return s ? s->s->deviceCoords() : nullptr; }

int minim_state_mark_coords_changed(minim_state* s) {
  return guard([&] {
    need(s, "state");
    s->s->markCoordsChanged();
  });
}

// ---------------------------------------------------------------- minimisers
int minim_minimiser_create(const minim_minimiser_params* p, minim_minimiser** out) {
  return guard([&] {
    need(p, "params");
    need(out, "out");
    std::unique_ptr<minim::Minimiser> m;
    static const char* modes[] = {"state", "fp64", "fp32", "mixed"};
    if (p->precision_mode < 0 || p->precision_mode > 3) throw std::invalid_argument("invalid precision mode");
    if (p->kind == MINIM_LBFGS) {
      auto l = std::make_unique<minim::Lbfgs>();
      l->setM(p->m > 0 ? p->m : 5);
      l->setMaxStep(p->max_step);
      l->setInitStep(p->init_step);
      l->setPrecision(modes[p->precision_mode], p->switch_rms);
      static const char* controls[] = {"device", "graph", "host", "resident"};
      if (p->batch_control < 0 || p->batch_control > 3) throw std::invalid_argument("invalid batch control");
      l->setBatchControl(controls[p->batch_control]);
      l->setBatchBudget(p->batch_budget);
      l->setBatchBlock(p->batch_block);
      m = std::move(l);
    } else if (p->kind == MINIM_FIRE) {
      auto f = std::make_unique<minim::Fire>();
      f->setDtMax(p->dt_max);
      f->setDt0(p->dt0);
      f->setMaxForceConvergence(p->max_force_dim);
      m = std::move(f);
    } else if (p->kind == MINIM_GRADDESCENT) {
      auto gd = std::make_unique<minim::GradDescent>();
      gd->setAlpha(p->alpha);
      m = std::move(gd);
    } else {
      throw std::invalid_argument("unknown minimiser kind");
    }
    m->setMaxIter(p->max_iter);
    m->setLinesearch(p->linesearch == MINIM_LS_NONE ? "none" : "backtracking");
    auto* h = new minim_minimiser;
    h->m = std::move(m);
    *out = h;
  });
}

int minim_minimiser_destroy(minim_minimiser* m) {
  return guard([&] { delete m; });
}

static void fill_result(const minim_minimiser* m, minim_result* r) {
  if (!r) return;
  const auto& st = m->m->stats;
  r->iterations = m->m->iter;
  r->converged = st.converged ? 1 : 0;
  r->n_energy = st.nEnergy;
  r->n_gradient = st.nGradient;
  r->n_energy_reused = st.nEnergyReused;
  r->iter_fp32 = st.iterFP32;
  r->iter_fp64 = st.iterFP64;
  r->energy = st.energy;
  r->grad_rms = st.gradRms;
  auto* f = dynamic_cast<const minim::Fire*>(m->m.get());
  r->dt_max = f ? f->dtMax : 0.0;
  r->grad_max = st.gradMax;
  r->seconds = st.seconds;
  for (int k = 0; k < 8; k++) {
    r->cross_seconds[k] = st.crossSeconds[k];
    r->cross_iter[k] = st.crossIter[k];
  }
}

int minim_minimise(minim_minimiser* m, minim_state* s, minim_adjust_fn fn, void* user, double* coords_out,
                   minim_result* r) {
  return guard([&] {
    need(m, "minimiser");
    need(s, "state");
    std::function<void(int, minim::State&)> adj = nullptr;
    if (fn) adj = [fn, user, s](int it, minim::State&) { fn(it, s, user); };
    if (coords_out) {
      std::vector<double> x = m->m->minimise(*s->s, adj);
      std::memcpy(coords_out, x.data(), sizeof(double) * x.size());
    } else {
      m->m->run(*s->s, adj);  // result stays in the State's device coordinates
    }
    fill_result(m, r);
  });
}

int minim_minimise_batch(minim_minimiser* m, minim_state* const* states, int n_states, minim_result* results) {
  return guard([&] {
    need(m, "minimiser");
    if (n_states < 0) throw std::invalid_argument("n_states must not be negative");
    if (n_states > 0) need(states, "states");
    auto* lb = dynamic_cast<minim::Lbfgs*>(m->m.get());
    if (!lb) throw std::invalid_argument("batched minimisation needs an Lbfgs minimiser");
    std::vector<minim::State*> v(n_states);
    for (int i = 0; i < n_states; i++) {
      need(states[i], "state");
      v[i] = states[i]->s.get();
    }
    std::vector<minim::MinimiseStats> st = lb->minimiseBatch(v);
    if (!results) return;
    for (int i = 0; i < n_states; i++) {
      minim_result& r = results[i];
      r = minim_result{};
      const int runs = st[i].iterFP32 + st[i].iterFP64;
      r.iterations = st[i].converged ? runs - 1 : runs;  // Minimiser::iter at exit
      r.converged = st[i].converged ? 1 : 0;
      r.n_energy = st[i].nEnergy;
      r.n_gradient = st[i].nGradient;
      r.n_energy_reused = st[i].nEnergyReused;
      r.iter_fp32 = st[i].iterFP32;
      r.iter_fp64 = st[i].iterFP64;
      r.energy = st[i].energy;
      r.grad_rms = st[i].gradRms;
      r.seconds = st[i].seconds;
      r.start_seconds = st[i].startSeconds;
      for (int k = 0; k < 8; k++) {
        r.cross_seconds[k] = -1;
        r.cross_iter[k] = -1;
      }
    }
  });
}

int minim_minimise_log(minim_minimiser* m, minim_state* s, const char* log_type, double* coords_out,
                       minim_result* r) {
  return guard([&] {
    need(m, "minimiser");
    need(s, "state");
    need(log_type, "log_type");
    std::vector<double> x = m->m->minimise(*s->s, std::string(log_type));
    if (coords_out) std::memcpy(coords_out, x.data(), sizeof(double) * x.size());
    fill_result(m, r);
  });
}

}  // extern "C"
