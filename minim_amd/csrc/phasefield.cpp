// phasefield.cpp -- host side of the device PhaseField potential.
//
// Set-up mirrors PhaseField::init (src/potentials/PhaseField.cpp:137-158),
// initLocal on one rank (:161-235), setDefaults (:961-967), checkArraySizes
// (:970-995), assignFluidCoefficients (:78-134) and getType (:899-958);
// evaluation launches the tiled stencil of pf.hip, or the general kernel of
// pf_ext.hip when solid nodes or any of the terms beyond bulk, gradient and
// density are set.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "minim/Minimiser.h"
#include "minim/PhaseField.h"
#include "minim/State.h"
#include "autotune.hpp"
#include "pf.hpp"
#include "runtime.hpp"

namespace minim {

using minim_amd::PFArgs;

namespace {

// Launch arguments plus the device memory they point into.  init() fills the
// host tables only (it needs no GPU); the first launch uploads them once.
struct PFHolder {
  PFArgs a;
  minim_amd::PFExt ext_h{};
  std::vector<uint16_t> code_h;
  std::vector<uint32_t> wcode_h;
  std::vector<double> wet_h, bs_h, fs_h;
  std::vector<int8_t> csign_h;
  minim_amd::DeviceBuffer ext, code, wcode, wet, csign, bs, fs;
  std::once_flag uploaded;
  std::atomic<bool> resident{false};  // the tables above are in device memory
};

template <typename T>
const T* to_device(minim_amd::DeviceBuffer& buf, std::vector<T>& v) {
  if (v.empty()) return nullptr;
  buf.reset(sizeof(T) * v.size());
  MINIM_HIP_CHECK(hipMemcpy(buf.get(), v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  std::vector<T>().swap(v);  // the device copy is the one used from here on
  return static_cast<const T*>(buf.get());
}

const PFArgs& args_of(const std::shared_ptr<void>& p) {
  if (!p) throw std::runtime_error("PhaseField: not initialised (construct a State first).");
  PFHolder* h = static_cast<PFHolder*>(p.get());
  std::call_once(h->uploaded, [h] {
    h->ext.reset(sizeof(h->ext_h));
    MINIM_HIP_CHECK(hipMemcpy(h->ext.get(), &h->ext_h, sizeof(h->ext_h), hipMemcpyHostToDevice));
    h->a.ext = static_cast<const minim_amd::PFExt*>(h->ext.get());
    h->a.code = to_device(h->code, h->code_h);
    h->a.wcode = to_device(h->wcode, h->wcode_h);
    h->a.wet = to_device(h->wet, h->wet_h);
    h->a.csign = to_device(h->csign, h->csign_h);
    h->a.Bs = to_device(h->bs, h->bs_h);
    h->a.Fs = to_device(h->fs, h->fs_h);
    h->resident = true;
  });
  return h->a;
}
}  // namespace

// Surface type of every site (PhaseField::getType, PhaseField.cpp:899-958):
// -1 solid, 0 bulk, 1..7 by the solid faces, edges and corners among the 26
// periodic neighbours.  Throws std::runtime_error for the layouts the
// reference leaves undefined.
std::vector<int> pf_site_types(const int grid[3], const char* solid) {
  const int n0 = grid[0], n1 = grid[1], n2 = grid[2];
  std::vector<int> type((size_t)n0 * n1 * n2);
  for (int x = 0; x < n0; x++)
    for (int y = 0; y < n1; y++)
      for (int z = 0; z < n2; z++) {
        const size_t i = ((size_t)x * n1 + y) * n2 + z;
        if (solid[i]) {
          type[i] = -1;
          continue;
        }
        int faces = 0, edges = 0, corners = 0;
        bool faceAxis[3] = {false, false, false}, edgeAcross[3] = {false, false, false};
        for (int dx = -1; dx <= 1; dx++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dz = -1; dz <= 1; dz++) {
              const int nonzero = (dx != 0) + (dy != 0) + (dz != 0);
              if (nonzero == 0) continue;
              const size_t j = (((size_t)((x + dx + n0) % n0) * n1) + (y + dy + n1) % n1) * n2 + (z + dz + n2) % n2;
              if (!solid[j]) continue;
              if (nonzero == 1) {
                faces++;
                faceAxis[dx ? 0 : (dy ? 1 : 2)] = true;
              } else if (nonzero == 2) {
                edges++;
                edgeAcross[0] |= dx == 0;
                edgeAcross[1] |= dy == 0;
                edgeAcross[2] |= dz == 0;
              } else {
                corners++;
              }
            }
        int t = -2;
        if (faces + edges + corners == 0) t = 0;
        else if (faces == 0) t = (edges == 0 && corners == 1) ? 1 : (edges == 1 && corners <= 2) ? 2
                                : (edges == 2 && corners <= 3) ? 3 : -2;
        else if (faces == 1) t = ((faceAxis[0] && !edgeAcross[0]) || (faceAxis[1] && !edgeAcross[1]) ||
                                  (faceAxis[2] && !edgeAcross[2])) ? 4 : 5;
        else if (faces == 2) t = 6;
        else if (faces == 3) t = 7;
        if (t == -2) throw std::runtime_error("PhaseField: Undefined surface type");
        type[i] = t;
      }
  return type;
}

PhaseField& PhaseField::setNFluid(int n) {
  nFluid = n;
  dofPerNode = n;  // PhaseField.cpp:753-757
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

// PhaseField.cpp:779-797
PhaseField& PhaseField::setSurfaceTension(std::function<vector<double>(int, int, int)> fn) {
  if (gridSize.size() != 3) throw std::invalid_argument("PhaseField: set the grid size first.");
  const size_t np = fn(0, 0, 0).size();
  surfaceTension.assign((size_t)nGrid * np, 0.0);
  size_t i = 0;
  for (int x = 0; x < gridSize[0]; x++)
    for (int y = 0; y < gridSize[1]; y++)
      for (int z = 0; z < gridSize[2]; z++) {
        vector<double> st = fn(x, y, z);
        for (size_t k = 0; k < np; k++) surfaceTension[i++] = st.at(k);
      }
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

// PhaseField.cpp:818-833
PhaseField& PhaseField::setPressure(vector<double> p) {
  pressure = p;
  return *this;
}

PhaseField& PhaseField::setVolume(vector<double> v, double c) {
  volumeFixed = true;
  volConst = c;
  volume = v;
  return *this;
}

PhaseField& PhaseField::setVolumeFixed(bool fixed, double c) {
  volumeFixed = fixed;
  volConst = c;
  return *this;
}

// PhaseField.cpp:836-873
PhaseField& PhaseField::setSolid(vector<char> s) {
  solid = s;
  return *this;
}

PhaseField& PhaseField::setSolid(std::function<bool(int, int, int)> fn) {
  if (gridSize.size() != 3) throw std::invalid_argument("PhaseField: set the grid size first.");
  solid.assign(nGrid, 0);
  size_t i = 0;
  for (int x = 0; x < gridSize[0]; x++)
    for (int y = 0; y < gridSize[1]; y++)
      for (int z = 0; z < gridSize[2]; z++) solid[i++] = fn(x, y, z);
  return *this;
}

PhaseField& PhaseField::setContactAngle(double a) {
  contactAngle = vector<double>(nGrid, a);
  return *this;
}

PhaseField& PhaseField::setContactAngle(vector<double> a) {
  contactAngle = a;
  return *this;
}

PhaseField& PhaseField::setContactAngle(std::function<double(int, int, int)> fn) {
  if (gridSize.size() != 3) throw std::invalid_argument("PhaseField: set the grid size first.");
  contactAngle.assign(nGrid, 0.0);
  size_t i = 0;
  for (int x = 0; x < gridSize[0]; x++)
    for (int y = 0; y < gridSize[1]; y++)
      for (int z = 0; z < gridSize[2]; z++) contactAngle[i++] = fn(x, y, z);
  return *this;
}

// PhaseField.cpp:876-887: one force for every fluid (one fluid, or no list),
// else only for the listed fluids
PhaseField& PhaseField::setForce(vector<double> f, vector<int> iFluid) {
  if ((int)f.size() != 3) throw std::invalid_argument("PhaseField: Invalid size of force array.");
  if (nFluid == 1 || iFluid.empty()) {
    force = vector2d<double>(nFluid, f);
  } else {
    force = vector2d<double>(nFluid, {0, 0, 0});
    for (int i : iFluid) force.at(i) = f;
  }
  return *this;
}

// PhaseField::setFixFluid (PhaseField.cpp:891-895)
PhaseField& PhaseField::setFixFluid(int iFluid, bool fix) {
  if ((int)fixFluid.size() != nFluid) fixFluid = vector<char>(nFluid, false);
  fixFluid.at(iFluid) = fix;
  return *this;
}

PhaseField& PhaseField::setConfinement(vector<double> s) {
  confinementStrength = s;
  return *this;
}

void PhaseField::setDefaults() {
  if (interfaceSize.empty()) interfaceSize = vector<double>(nFluid, resolution);
  if (fixFluid.empty()) fixFluid = vector<char>(nFluid, false);
  if (pressure.empty()) pressure = vector<double>(nFluid, 0);
  if (confinementStrength.empty()) confinementStrength = vector<double>(nFluid, 0);
}

void PhaseField::checkArraySizes() {
  if ((int)interfaceSize.size() != nFluid) throw std::invalid_argument("PhaseField: Invalid size interfaceSize array.");
  if (!solid.empty() && (int)solid.size() != nGrid) throw std::invalid_argument("PhaseField: Invalid size of solid array.");
  if (!contactAngle.empty() && (long)contactAngle.size() != (long)nGrid * nFluid)
    throw std::invalid_argument("PhaseField: Invalid size of contactAngle array.");
  if (!force.empty() && (int)force.size() != nFluid) throw std::invalid_argument("PhaseField: Invalid size of force array.");
  if (!volume.empty() && (int)volume.size() != nFluid) throw std::invalid_argument("PhaseField: Invalid size of volume array.");
  if ((int)pressure.size() != nFluid) throw std::invalid_argument("PhaseField: Invalid size of pressure array.");
  if ((int)fixFluid.size() != nFluid) throw std::invalid_argument("PhaseField: Invalid size of fixed fluid array.");
  if ((int)confinementStrength.size() != nFluid)
    throw std::invalid_argument("PhaseField: Invalid size of confinement strength array.");
}

void PhaseField::assignFluidCoefficients() {
  // per-site tensions when the array holds nParams values per site (:79)
  const int nk = ((long)surfaceTension.size() == (long)nGrid * nParams) ? nGrid : 1;
  vector<double> st = surfaceTension;
  if ((int)st.size() == 1) {
    st = vector<double>((size_t)nk * nParams, st[0]);
  } else if ((long)st.size() != (long)nk * nParams) {
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

  kappa.assign((size_t)nk * nParams, 0.0);
  kappaP.assign((size_t)nk * nParams, 0.0);
  for (int ig = 0; ig < nk; ig++) {
    const double* s = st.data() + (size_t)ig * nParams;
    double* ka = kappa.data() + (size_t)ig * nParams;
    double* kp = kappaP.data() + (size_t)ig * nParams;
    if (nFluid <= 2) {
      for (int i = 0; i < nParams; i++) {
        ka[i] = (3 * s[i]) / iw[i];
        kp[i] = (3 * s[i]) * iw[i];
      }
    } else {
      // pairwise tensions (1-2, 1-3, 2-3, ...) to per-fluid coefficients
      vector<double> ks(nParams), kps(nParams);
      for (int i = 0; i < nParams; i++) {
        ks[i] = (6 * s[i]) / iw[i];
        kps[i] = (6 * s[i]) * iw[i];
      }
      const int l = nFluid - 1;
      ka[0] = 0.5 * (ks[0] + ks[1] - ks[l]);
      ka[1] = 0.5 * (ks[0] - ks[1] + ks[l]);
      ka[2] = 0.5 * (-ks[0] + ks[1] + ks[l]);
      kp[0] = 0.5 * (kps[0] + kps[1] - kps[l]);
      kp[1] = 0.5 * (kps[0] - kps[1] + kps[l]);
      kp[2] = 0.5 * (-kps[0] + kps[1] + kps[l]);
      for (int i = 3; i < nFluid; i++) {
        ka[i] = ks[i - 1] - ka[0];
        kp[i] = kps[i - 1] - kp[0];
      }
    }
  }
}

namespace {
// Tuned shapes are per kernel family: the plain tiled stencil, the general
// tiled kernel, and the untiled kernels (more than four fluids, per-site
// tensions), whose single launch shape serves both precisions.
bool pf_tiled(const PFArgs& a) { return a.use_ext ? (!a.Fs && a.nf <= 4) : a.nf <= 4; }

std::string tune_key_of(const PFArgs& a, int prec) {
  const char* fam = !pf_tiled(a) ? "pf-untiled" : a.use_ext ? "pf-ext-tile" : "pf-tile";
  return minim_amd::tune_key(fam, {a.X, a.Y, a.Z, a.nf}, pf_tiled(a) ? prec : 0);
}

void apply_shape(PFArgs& a, int prec, int target, int march) {
  a.tuned_target[prec] = target;
  a.tuned_xchunk[prec] = march;
  if (!pf_tiled(a) && a.target_blocks <= 0 && a.xchunk_override <= 0) {
    a.xchunk_override = march;
    minim_amd::pf_args_finalize(a, target);
    a.target_blocks = 0;  // still the heuristic's slot: the shape is in xchunk and nblocks
    a.xchunk_override = 0;
  }
}
}  // namespace

void PhaseField::init(const vector<double>& coords) {
  if (gridSize.size() != 3) throw std::invalid_argument("PhaseField: gridSize must have three entries.");
  for (int s : gridSize)
    if (s < 1) throw std::invalid_argument("PhaseField: grid sizes must be positive.");
  nGrid = gridSize[0] * gridSize[1] * gridSize[2];
  if (nFluid < 1 || nFluid > minim_amd::kPFMaxFluid)
    throw std::invalid_argument("PhaseField: nFluid must be between 1 and 16.");
  if ((long)coords.size() != (long)nGrid * nFluid)
    throw std::invalid_argument("PhaseField: Size of coordinates array does not match the grid size.");
  if (!constraints.empty())
    throw std::invalid_argument(
        "PhaseField: single-DOF and normal constraints are not supported; use setFixFluid.");
  nParams = nFluid;
  setDefaults();
  checkArraySizes();
  assignFluidCoefficients();
  convergence = 1e-8 * surfaceTensionMean * std::pow(resolution, 2);
  if (densityConstraint == minim_amd::kDensityFixed) fixFluid[nFluid - 1] = true;
  const int nF = nFluid;
  const size_t ng = (size_t)nGrid;

  // forces (:149-157)
  fMag.assign(nF, 0.0);
  fNorm.assign(nF, vector<double>(3, 0.0));
  bool anyForce = false;
  if (!force.empty()) {
    for (int i = 0; i < nF; i++) {
      double r = 0;
      for (int d = 0; d < 3; d++) r += force[i].at(d) * force[i].at(d);
      fMag[i] = std::sqrt(r);
      for (int d = 0; d < 3; d++) fNorm[i][d] = force[i][d] / fMag[i];
      if (fMag[i] != 0) anyForce = true;  // vec::any: NaN counts as set
    }
  }

  auto h = std::make_shared<PFHolder>();
  PFArgs& A = h->a;
  A = PFArgs{};
  // canonical grid: drop size-1 dimensions to the front (keeps the linear index)
  int cmap[3] = {-1, -1, -1}, nd = 0, dims[3];
  for (int i = 0; i < 3; i++)
    if (gridSize[i] > 1) dims[nd++] = i;
  for (int i = 0; i < nd; i++) cmap[3 - nd + i] = dims[i];
  int canon[3];
  for (int j = 0; j < 3; j++) canon[j] = cmap[j] >= 0 ? gridSize[cmap[j]] : 1;
  A.X = canon[0];
  A.Y = canon[1];
  A.Z = canon[2];
  A.nf = nF;
  const double res3 = std::pow(resolution, 3), res2 = std::pow(resolution, 2);
  const bool perSite = kappa.size() != (size_t)nParams;
  for (int k = 0; k < nF; k++) {
    const double kk = perSite ? 0.0 : kappa[k], kkp = perSite ? 0.0 : kappaP[k];
    A.fb[k] = (nF == 1) ? kk / 16 * (1 * res3) : 0.5 * kk * (1 * res3);
    double f = (nF == 1) ? 0.25 * kkp * (1 * res3) : 0.5 * kkp * (1 * res3);
    A.fg[k] = f / res2;
  }
  A.fixmask = 0;
  int nvar = 0;
  for (int k = 0; k < nF; k++) {
    if (fixFluid[k]) A.fixmask |= (1u << k);
    else nvar++;
  }
  A.nvar = nvar;
  A.inv_nvar = nvar > 0 ? 1.0 / nvar : 0.0;
  A.density = densityConstraint;
  A.coef = densityConst * surfaceTensionMean * res2;
  A.xchunk_override = stencilXChunk;
  minim_amd::pf_args_finalize(A, stencilTargetBlocks);

  // ---- node volumes, surface areas and bond weights (initLocal, :190-214)
  bool anySolid = false;
  for (char s : solid) anySolid |= s != 0;
  static const double kArea[8] = {0, 0.75, 1, 1.25, 1, 1.25, 1, 0.75};
  minim_amd::PFExt& E = h->ext_h;
  for (int kind = 0; kind < minim_amd::kPFKinds; kind++) {
    const int t = kind - 1;
    const double v = (t < 0) ? 0.0 : (t == 0 ? 1.0 : t / 8.0);
    E.vol[kind] = v * res3;
    E.area[kind] = (t < 0 ? 0.0 : kArea[t]) * res2;
    for (int k = 0; k < nF; k++) {
      const double kk = perSite ? 0.0 : kappa[k], kkp = perSite ? 0.0 : kappaP[k];
      E.B[k][kind] = (nF == 1) ? kk / 16 * E.vol[kind] : 0.5 * kk * E.vol[kind];
      E.F[k][kind] = ((nF == 1) ? 0.25 * kkp * E.vol[kind] : 0.5 * kkp * E.vol[kind]) / res2;
    }
  }
  std::vector<int> kinds;  // per site, when solid nodes exist
  nodeVol.clear();
  surfaceArea.clear();
  if (anySolid) {
    std::vector<int> type = pf_site_types(gridSize.data(), solid.data());
    kinds.resize(ng);
    nodeVol.resize(ng);
    surfaceArea.resize(ng);
    for (size_t i = 0; i < ng; i++) {
      kinds[i] = type[i] + 1;
      nodeVol[i] = E.vol[kinds[i]];
      surfaceArea[i] = E.area[kinds[i]];
    }
    // bond weights per canonical dimension: 0 towards a solid site, 1 towards
    // the fluid side when the other neighbour is solid, else one half
    std::vector<uint16_t> code(ng);
    const int n0 = gridSize[0], n1 = gridSize[1], n2 = gridSize[2];
    for (int x = 0; x < n0; x++)
      for (int y = 0; y < n1; y++)
        for (int z = 0; z < n2; z++) {
          const size_t i = ((size_t)x * n1 + y) * n2 + z;
          unsigned q = (unsigned)kinds[i];
          if (!solid[i]) {
            for (int j = 0; j < 3; j++) {
              const int dd = cmap[j];
              if (dd < 0) continue;
              int cm[3] = {x, y, z}, cp[3] = {x, y, z};
              cm[dd] = (cm[dd] - 1 + gridSize[dd]) % gridSize[dd];
              cp[dd] = (cp[dd] + 1) % gridSize[dd];
              const bool sm = solid[((size_t)cm[0] * n1 + cm[1]) * n2 + cm[2]] != 0;
              const bool sp = solid[((size_t)cp[0] * n1 + cp[1]) * n2 + cp[2]] != 0;
              const unsigned hm = sm ? 0u : (sp ? 2u : 1u);
              const unsigned hp = sp ? 0u : (sm ? 2u : 1u);
              q |= hm << (minim_amd::kPFCodeShift + 4 * j);
              q |= hp << (minim_amd::kPFCodeShift + 4 * j + 2);
            }
          }
          code[i] = (uint16_t)q;
        }
    // forward-bond weights for the tiled kernel: w(i, i+e) = (v_i h_i->i+e +
    // v_{i+e} h_{i+e}->i) / res^3 in sixteenths (v/res^3 in eighths, h in halves)
    std::vector<uint32_t> wcode(ng);
    auto v8 = [](unsigned q) { const unsigned k = q & 15u; return k >= 2 ? k - 1 : 8 * k; };
    auto hq = [](unsigned q, int j, int side) { return (q >> (minim_amd::kPFCodeShift + 4 * j + 2 * side)) & 3u; };
    for (int x = 0; x < n0; x++)
      for (int y = 0; y < n1; y++)
        for (int z = 0; z < n2; z++) {
          const size_t i = ((size_t)x * n1 + y) * n2 + z;
          uint32_t w = (uint32_t)kinds[i];
          for (int j = 0; j < 3; j++) {
            const int dd = cmap[j];
            if (dd < 0) continue;
            int cp[3] = {x, y, z};
            cp[dd] = (cp[dd] + 1) % gridSize[dd];
            const size_t ip = ((size_t)cp[0] * n1 + cp[1]) * n2 + cp[2];
            const unsigned w16 = v8(code[i]) * hq(code[i], j, 1) + v8(code[ip]) * hq(code[ip], j, 0);
            w |= w16 << (minim_amd::kPFWeightShift + minim_amd::kPFWeightBits * j);
          }
          wcode[i] = w;
        }
    h->code_h = std::move(code);
    h->wcode_h = std::move(wcode);
  }
  {
    unsigned q = 1;  // bulk site, both bonds one half along every dimension of size > 1
    uint32_t w = 1;  // and forward-bond weight one
    for (int j = 0; j < 3; j++)
      if (cmap[j] >= 0) {
        q |= (1u << (minim_amd::kPFCodeShift + 4 * j)) | (1u << (minim_amd::kPFCodeShift + 4 * j + 2));
        w |= 16u << (minim_amd::kPFWeightShift + minim_amd::kPFWeightBits * j);
      }
    A.ucode = q;
    A.wucode = w;
  }
  auto volAt = [&](size_t i) { return anySolid ? nodeVol[i] : E.vol[1]; };

  // per-site tensions: factors per DOF (fluidEnergy, :369-391)
  if (perSite) {
    std::vector<double> bs(ng * nF), fs(ng * nF);
    for (size_t i = 0; i < ng; i++)
      for (int k = 0; k < nF; k++) {
        const size_t iK = i * nF + k;
        const double v = volAt(i);
        bs[iK] = (nF == 1) ? kappa[iK] / 16 * v : 0.5 * kappa[iK] * v;
        fs[iK] = ((nF == 1) ? 0.25 * kappaP[iK] * v : 0.5 * kappaP[iK] * v) / res2;
      }
    h->bs_h = std::move(bs);
    h->fs_h = std::move(fs);
  }

  // wetting (surfaceEnergy, :476-487): binary fluids on sites next to a solid
  bool anyWet = false;
  if (!contactAngle.empty() && nF == 1 && anySolid) {
    std::vector<double> wet(ng, 0.0);
    for (size_t i = 0; i < ng; i++) {
      if (surfaceArea[i] == 0 || contactAngle[i] == 90) continue;
      wet[i] = 1 / std::sqrt(2.0) * std::cos(contactAngle[i] * 3.1415926536 / 180);
      anyWet = true;
    }
    if (anyWet) h->wet_h = std::move(wet);
  }

  // pressure (:396-421): skipped for zero pressure and fixed fluids
  for (int k = 0; k < nF; k++) {
    E.pres[k] = (pressure[k] == 0 || fixFluid[k]) ? 0.0 : pressure[k];
    if (E.pres[k] != 0) A.any_pres = 1;
  }
  // force (:490-518) in canonical dimensions
  A.any_force = anyForce ? 1 : 0;
  for (int j = 0; j < 3; j++) E.fc[j] = cmap[j] >= 0 ? (gridSize[cmap[j]] - 1) / 2.0 : 0.0;
  for (int k = 0; k < nF; k++) {
    E.fmag[k] = fMag[k];
    for (int j = 0; j < 3; j++) E.fn[k][j] = (cmap[j] >= 0 && fMag[k] != 0) ? fNorm[k][cmap[j]] : 0.0;
  }
  E.res = resolution;
  // confinement (:521-537): which side of 0.5 each DOF started on
  bool anyConf = false;
  for (int k = 0; k < nF; k++) {
    E.conf[k] = confinementStrength[k] * surfaceTensionMean * res2;
    if (confinementStrength[k] != 0) anyConf = true;
  }
  if (anyConf) {
    std::vector<int8_t> sg(ng * nF);
    for (size_t i = 0; i < ng * nF; i++) sg[i] = (int8_t)((coords[i] > 0.5) - (coords[i] < 0.5));
    h->csign_h = std::move(sg);
  }
  // fixed volume (:216-231): the target from the initial state when none is
  // given, and the sum of squared node volumes (which the reference leaves
  // unset when a target is given)
  totalVolume2 = 0;
  if (volumeFixed) {
    if (volume.empty()) {
      volume.assign(nF, 0.0);
      for (size_t i = 0; i < ng; i++)
        for (int k = 0; k < nF; k++) volume[k] += coords[k + nF * i] * volAt(i);
    }
    for (size_t i = 0; i < ng; i++) totalVolume2 += volAt(i) * volAt(i);
    A.volfix = 1;
    A.tv2 = totalVolume2;
  }
  A.ext_host = &h->ext_h;
  A.use_ext = (anySolid || anyWet || A.any_pres || anyForce || anyConf || perSite || stencilKernel == 1) ? 1 : 0;
  // launch shapes an autotuner chose earlier in this process (or in the file
  // named by MINIM_AUTOTUNE_CACHE) for a field of this shape and kernel
  A.tuned_target[0] = A.tuned_target[1] = A.tuned_xchunk[0] = A.tuned_xchunk[1] = 0;
  for (int prec = 0; prec < 2; prec++) {
    std::vector<int> v;
    if (minim_amd::tune_lookup(tune_key_of(A, prec), v) && v.size() == 2) apply_shape(A, prec, v[0], v[1]);
  }
  args_ = h;
}

// Launch-shape autotuner of the evaluation kernel this field uses (the plain
// tiled stencil, the general tiled kernel, or the untiled kernels): one
// energy+gradient launch (kernel only, HIP events, the average of reps after
// two warm-up launches) for every pair of target block count and x-march
// length (0 = automatic); the fastest is kept for this precision and stored in
// the process-wide cache, where the next field of the same shape finds it.
// The shape changes the tiling only: gradients are the same bit for bit.
std::vector<std::array<double, 3>> PhaseField::autotune(const vector<double>& coords, int prec, vector<int> targets,
                                                         vector<int> marches, int reps) {
  if (prec != FP64 && prec != FP32) throw std::invalid_argument("PhaseField::autotune: precision must be FP64 or FP32.");
  if (reps < 1 || targets.empty() || marches.empty()) throw std::invalid_argument("PhaseField::autotune: nothing to time");
  const PFArgs& base = args_of(args_);
  PFHolder* h = static_cast<PFHolder*>(args_.get());
  const size_t n = (size_t)nGrid * nFluid;
  if (coords.size() != n) throw std::invalid_argument("PhaseField::autotune: coordinates have the wrong size.");
  hipStream_t s = nullptr;
  const size_t bytes = n * minim_amd::prec_bytes(prec);
  minim_amd::DeviceBuffer xb(bytes), gb(bytes);
  minim_amd::upload_raw(coords.data(), n, prec, xb.get(), s);
  auto& ws = minim_amd::Workspace::current();
  hipEvent_t e0, e1;
  MINIM_HIP_CHECK(hipEventCreate(&e0));
  MINIM_HIP_CHECK(hipEventCreate(&e1));
  std::vector<std::array<double, 3>> rows;
  double bestMs = 0;
  int bestT = 0, bestM = 0;
  for (int t : targets) {
    for (int mch : marches) {
      PFArgs c = base;
      c.xchunk_override = mch;
      try {
        minim_amd::pf_args_finalize(c, t);
        for (int w = 0; w < 2; w++)
          minim_amd::pf_launch_eval(c, prec, prec, 3, xb.get(), nullptr, 0.0, gb.get(), ws.rs(), s, false);
      } catch (const std::invalid_argument&) {
        continue;  // a shape this grid cannot take
      }
      MINIM_HIP_CHECK(hipEventRecord(e0, s));
      for (int r = 0; r < reps; r++)
        minim_amd::pf_launch_eval(c, prec, prec, 3, xb.get(), nullptr, 0.0, gb.get(), ws.rs(), s, false);
      MINIM_HIP_CHECK(hipEventRecord(e1, s));
      MINIM_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0;
      MINIM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      const double avg = (double)ms / reps;
      rows.push_back({(double)t, (double)mch, avg});
      if (bestT == 0 || avg < bestMs) {
        bestMs = avg;
        bestT = t;
        bestM = mch;
      }
    }
  }
  MINIM_HIP_CHECK(hipEventDestroy(e0));
  MINIM_HIP_CHECK(hipEventDestroy(e1));
  if (bestT == 0) throw std::invalid_argument("PhaseField::autotune: no candidate shape fits this grid");
  apply_shape(h->a, prec, bestT, bestM);
  minim_amd::tune_store(tune_key_of(h->a, prec), {bestT, bestM});
  return rows;
}

std::pair<int, int> PhaseField::tunedShape(int prec) const {
  const PFArgs& a = args_of(args_);
  if (prec != FP64 && prec != FP32) throw std::invalid_argument("PhaseField: precision must be FP64 or FP32.");
  return {a.tuned_target[prec], a.tuned_xchunk[prec]};
}

bool PhaseField::supportsMixed() const {
  if (nFluid <= 4) return true;
  return args_ && args_of(args_).use_ext;
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

void PhaseField::timeSteps(int prec, const void* x, void* g, double* e_dev, int steps, double warmMs, void* stream,
                           double* out) const {
  const PFArgs& a = args_of(args_);
  if (steps < 1 || !g || !e_dev || !out) throw std::invalid_argument("PhaseField: nothing to time");
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto& ws = minim_amd::Workspace::current();
  auto step = [&](hipEvent_t b, hipEvent_t e) {
    minim_amd::pf_launch_eval(a, prec, prec, 3, x, nullptr, 0.0, g, ws.rs(), s, true, nullptr, b, e, e_dev);
  };
  // untimed evaluations until warmMs of host time have passed
  const auto t0 = std::chrono::steady_clock::now();
  do {
    for (int i = 0; i < 16; i++) step(nullptr, nullptr);
    MINIM_HIP_CHECK(hipStreamSynchronize(s));
  } while (std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() < warmMs);
  std::vector<hipEvent_t> ev(2 * (size_t)steps + 2);
  for (auto& e : ev) MINIM_HIP_CHECK(hipEventCreate(&e));
  MINIM_HIP_CHECK(hipEventRecord(ev[0], s));
  for (int i = 0; i < steps; i++) step(ev[2 + 2 * i], ev[3 + 2 * i]);
  MINIM_HIP_CHECK(hipEventRecord(ev[1], s));
  MINIM_HIP_CHECK(hipEventSynchronize(ev[1]));
  float ms = 0;
  MINIM_HIP_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
  out[0] = (double)ms / steps;
  std::vector<double> k(steps);
  double sum = 0;
  for (int i = 0; i < steps; i++) {
    MINIM_HIP_CHECK(hipEventElapsedTime(&ms, ev[2 + 2 * i], ev[3 + 2 * i]));
    k[i] = ms;
    sum += ms;
  }
  for (auto& e : ev) MINIM_HIP_CHECK(hipEventDestroy(e));
  std::sort(k.begin(), k.end());
  out[1] = (steps % 2) ? k[steps / 2] : 0.5 * (k[steps / 2 - 1] + k[steps / 2]);
  out[2] = k.front();
  out[3] = k.back();
  out[4] = sum / steps;
}

// evaluateAsync with the cross-block reduction in caller-provided device
// scratch: partials in scratch[0, nPartials), results in scratch[nPartials,
// nPartials + 96).  Allocates nothing: the potential's tables must already be
// resident (any earlier evaluation uploads them).
void PhaseField::evaluateAsyncScratch(int prec, const void* x, const void* d, double t, bool wantE, double* e_dev,
                                      void* g, double* scratch, size_t nPartials, void* stream) const {
  if (!args_) throw std::runtime_error("PhaseField: not initialised (construct a State first).");
  const PFHolder* h = static_cast<const PFHolder*>(args_.get());
  if (!h->resident)
    throw std::invalid_argument("PhaseField: device tables not resident yet; evaluate once through an allocating "
                                "entry point first");
  const PFArgs& a = h->a;
  // the widest reductions: the evaluation (2 values per block) and the fixed
  // volume projection (up to 16 fluid sums over at most 2048 blocks)
  const size_t need = std::max<size_t>(2 * (size_t)a.nblocks, a.volfix ? (size_t)minim_amd::kPFMaxFluid * 2048 : 0);
  if (need > nPartials)
    throw std::invalid_argument("PhaseField: the reduction needs " + std::to_string(need) +
                                " partial sums, more than the scratch holds (" + std::to_string(nPartials) + ")");
  const int mode = (wantE ? 1 : 0) | (g ? 2 : 0) | (d ? 4 : 0);
  if (mode == 0) return;
  if ((mode & 4) && (mode & 2)) throw std::invalid_argument("PhaseField: gradients at a trial point are not supported.");
  hipStream_t s = static_cast<hipStream_t>(stream);
  minim_amd::RedScratch rs{scratch, scratch + nPartials};
  minim_amd::pf_launch_eval(a, prec, prec, mode, x, d, t, g, rs, s, true, nullptr, nullptr, nullptr,
                            wantE ? e_dev : nullptr);
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
  minim_amd::pf_launch_eval(a, prec, pv, mode, x, d, t, g, ws.rs(), s, true, nullptr, nullptr, nullptr,
                            wantE ? e_dev : nullptr);
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

// The tiled kernel with the trial point written out (no general kernel, no
// fixed volumes: their projection needs the whole gradient first).
bool PhaseField::plainStencil(int canon[3], double& fb, double& fg) const {
  if (!args_) return false;
  const PFArgs& a = args_of(args_);
  if (a.use_ext || a.volfix || a.nf != 1 || a.fixmask) return false;
  canon[0] = a.X;
  canon[1] = a.Y;
  canon[2] = a.Z;
  fb = a.fb[0];
  fg = a.fg[0];
  return true;
}

bool PhaseField::supportsStep() const {
  if (!args_) return false;
  const PFArgs& a = args_of(args_);
  return !a.use_ext && !a.volfix && a.nf <= 4;
}

void PhaseField::evaluateStepAsync(int px, int pv, const void* x, const void* d, double t, void* xout, void* g,
                                   void* stream) const {
  if (!supportsStep()) throw std::invalid_argument("PhaseField: no fused step for this configuration");
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto& ws = minim_amd::Workspace::current();
  minim_amd::pf_launch_eval(args_of(args_), px, pv, 7, x, d, t, g, ws.rs(), s, true, xout);
}

void PhaseField::constrainDevice(int prec, const void* x, void* v, size_t n, void* stream) const {
  (void)x;
  const PFArgs& a = args_of(args_);
  auto& ws = minim_amd::Workspace::current();
  minim_amd::pf_launch_constrain(a, prec, v, n / a.nf, ws.rs(), static_cast<hipStream_t>(stream));
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

// PhaseField::energyComponents (PhaseField.cpp:706-744): one launch of the
// general kernel per term with the raw (unconstrained) gradient; the volume
// term from the fluid volumes (volumeConstraintEnergy, :540-577).
std::map<std::string, vector<double>> PhaseField::energyComponents(const vector<double>& coords) const {
  const PFArgs& a = args_of(args_);
  const size_t n = (size_t)nGrid * nFluid;
  if (coords.size() != n) throw std::invalid_argument("PhaseField: coordinates have the wrong size.");
  hipStream_t s = nullptr;
  auto& ws = minim_amd::Workspace::current();
  minim_amd::DVec x(n, FP64), g(n, FP64);
  minim_amd::upload(coords.data(), n, x, s);
  static const char* names[7] = {"fluid", "surface", "pressure", "density constraint", "force", "confinement",
                                 "volume constraint"};
  static const unsigned bits[6] = {minim_amd::kTermFluid,    minim_amd::kTermSurface, minim_amd::kTermPressure,
                                   minim_amd::kTermDensity,  minim_amd::kTermForce,   minim_amd::kTermConfinement};
  std::map<std::string, vector<double>> out;
  for (int c = 0; c < 6; c++) {
    minim_amd::PFCall cl{};
    cl.terms = bits[c];
    cl.raw = 1;
    const int nb = minim_amd::pf_launch_ext(a, cl, FP64, FP64, 3, x.data(), nullptr, 0.0, g.data(), ws.rs(), s);
    minim_amd::launch_finalize(ws.rs(), 2, nb, s);
    double r[2];
    ws.fetch(s, 2, r);
    out[names[c]] = {r[0], std::sqrt(r[1])};
  }
  vector<double> vc = {0.0, 0.0};
  if (volumeFixed) {
    minim_amd::pf_launch_volumes(a, FP64, x.data(), (size_t)nGrid, ws.rs(), s);
    vector<double> vol(nFluid);
    ws.fetch(s, nFluid, vol.data());
    const double coef = volConst * surfaceTensionMean / std::pow(resolution, 4);
    minim_amd::PFCall cl{};
    cl.terms = minim_amd::kTermVolume;
    cl.raw = 1;
    double e = 0;
    bool any = false;
    for (int k = 0; k < nFluid; k++) {
      if (fixFluid[k] || volume[k] < 0) continue;
      const double diff = vol[k] - volume[k];
      if (diff == 0) continue;
      e += coef * std::pow(diff, 2) / 1;
      cl.volg[k] = coef * diff;
      any = true;
    }
    double gg = 0;
    if (any) {
      const int nb = minim_amd::pf_launch_ext(a, cl, FP64, FP64, 2, x.data(), nullptr, 0.0, g.data(), ws.rs(), s);
      minim_amd::launch_finalize(ws.rs(), 2, nb, s);
      double r[2];
      ws.fetch(s, 2, r);
      gg = r[1];
    }
    vc = {e, std::sqrt(gg)};
  }
  out[names[6]] = vc;
  return out;
}

// PhaseField::diffuseSolid (PhaseField.cpp:998-1035): the solid indicator of
// fluid iFluid relaxed by 100 L-BFGS steps (no line search) under a strong
// confinement, without a density constraint.
vector<double> PhaseField::diffuseSolid(vector<char> s, int iFluid, bool twoStep) {
  return diffuseSolid(s, *this, iFluid, twoStep);
}

vector<double> PhaseField::diffuseSolid(vector<char> s, vector<int> grid, int nF, int iFluid, bool twoStep) {
  PhaseField p;
  p.setNFluid(nF);
  p.setGridSize(grid);
  return diffuseSolid(s, p, iFluid, twoStep);
}

vector<double> PhaseField::diffuseSolid(vector<char> s, PhaseField p, int iFluid, bool twoStep) {
  if (iFluid < 0 || iFluid >= p.nFluid) throw std::invalid_argument("PhaseField: invalid fluid index.");
  vector<double> conf(p.nFluid, 0.0);
  conf[iFluid] = 100;
  p.setConfinement(conf);
  p.setDensityConstraint("none");
  const int ng = p.gridSize.at(0) * p.gridSize.at(1) * p.gridSize.at(2);
  if ((int)s.size() != ng) throw std::invalid_argument("PhaseField: Invalid size of solid array.");
  vector<double> init((size_t)p.nFluid * ng, 0.0);
  for (int i = 0; i < ng; i++) init[(size_t)i * p.nFluid + iFluid] = (int)s[i];
  // the reference indexes surfaceTension by the fluid (:1024); a single value
  // stands for every fluid
  const double st = (int)p.surfaceTension.size() > iFluid ? p.surfaceTension[iFluid] : p.surfaceTension.at(0);
  Lbfgs min;
  min.setLinesearch("none");
  min.setMaxIter(100);
  vector<double> minimum;
  {
    State state(p, init);
    state.convergence = 1e-8 * st * std::pow(p.resolution, 2);
    minimum = min.minimise(state);
  }
  if (twoStep) {
    State state(p, minimum);
    minimum = min.minimise(state);
  }
  return minimum;
}

}  // namespace minim
