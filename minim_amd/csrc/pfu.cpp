// pfu.cpp -- host side of PhaseFieldUnstructured: the reference's setters,
// defaults and checks (src/potentials/PhaseFieldUnstructured.cpp:71-107,
// :110-268, :594-978), the element list built in the reference's order, and
// the launches of pfu.hip.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <vector>

#include "minim/Minimiser.h"
#include "minim/PhaseField.h"
#include "minim/PhaseFieldUnstructured.h"
#include "minim/State.h"
#include "pf.hpp"
#include "pfu.hpp"
#include "runtime.hpp"

namespace minim {

namespace {

// The element list in the reference's order (PhaseFieldUnstructured::init,
// :157-240): type, DOF indices (unused slots -1) and parameters.
struct RefElements {
  std::vector<int8_t> type;
  std::vector<int32_t> idof;  // kPFUNodes per element
  std::vector<double> par;    // 3 per element
  void push(int t, std::initializer_list<int32_t> ids, std::initializer_list<double> p) {
    type.push_back((int8_t)t);
    size_t k = 0;
    for (int32_t i : ids) {
      idof.push_back(i);
      k++;
    }
    for (; k < (size_t)minim_amd::kPFUNodes; k++) idof.push_back(-1);
    k = 0;
    for (double v : p) {
      par.push_back(v);
      k++;
    }
    for (; k < 3; k++) par.push_back(0.0);
  }
  size_t size() const { return type.size(); }
};

// The device form of the list, on the host until the first launch uploads it:
// elements sorted (stably) into chunks by the tile of their first node, each
// chunk's sorted table of the sites its elements touch (flagged where another
// chunk touches the site too), and 16-byte element records naming their
// nodes by table position.
struct PFUPacked {
  std::vector<uint32_t> rec;  // 4 words per element
  std::vector<uint32_t> pidx;
  std::vector<double> par;
  std::vector<uint32_t> chunk_el, chunk_site, sites;
  // sites not exclusive to one chunk, and the site-table positions of each
  std::vector<uint32_t> gsite, gpos_off, gpos;
};

// Largest tile of tx x ty x tz sites whose sites plus face neighbours hold at
// most kPFUChunkDofs DOFs (fewer counted in a direction the tile spans).
void choose_tile(const int n[3], int nf, int t[3]) {
  long best = 0, best_touch = 0;
  t[0] = t[1] = t[2] = 1;
  const long cap = minim_amd::kPFUChunkDofs / nf;
  for (int tz = 1; tz <= std::min(n[2], 64); tz++)
    for (int ty = 1; ty <= std::min(n[1], 64); ty++)
      for (int tx = 1; tx <= std::min(n[0], 64); tx++) {
        const long v = (long)tx * ty * tz;
        const long touch = v + (n[0] > tx ? 2L * ty * tz : 0) + (n[1] > ty ? 2L * tx * tz : 0) +
                           (n[2] > tz ? 2L * tx * ty : 0);
        if (touch > cap) break;
        if (v > best || (v == best && touch < best_touch)) {
          best = v;
          best_touch = touch;
          t[0] = tx;
          t[1] = ty;
          t[2] = tz;
        }
      }
}

// Packs the list for tiles of t[] sites; false when a chunk touches more than
// kPFUChunkDofs DOFs.
bool pack_elements(const RefElements& el, const int n[3], int nf, const int t[3], PFUPacked& out) {
  const size_t nel = el.size();
  const size_t ng = (size_t)n[0] * n[1] * n[2];
  const int nt[3] = {(n[0] + t[0] - 1) / t[0], (n[1] + t[1] - 1) / t[1], (n[2] + t[2] - 1) / t[2]};
  const size_t ntile = (size_t)nt[0] * nt[1] * nt[2];
  auto tile_of = [&](size_t site) {
    const size_t z = site % n[2], y = (site / n[2]) % n[1], x = site / ((size_t)n[1] * n[2]);
    return ((x / t[0]) * nt[1] + y / t[1]) * nt[2] + z / t[2];
  };
  // counting sort of the elements by the tile of their first node
  std::vector<uint32_t> start(ntile + 1, 0);
  std::vector<uint32_t> key(nel);
  for (size_t e = 0; e < nel; e++) {
    key[e] = (uint32_t)tile_of((size_t)el.idof[e * minim_amd::kPFUNodes] / nf);
    start[key[e] + 1]++;
  }
  for (size_t i = 0; i < ntile; i++) start[i + 1] += start[i];
  std::vector<uint32_t> order(nel);
  {
    std::vector<uint32_t> pos(start.begin(), start.end() - 1);
    for (size_t e = 0; e < nel; e++) order[pos[key[e]]++] = (uint32_t)e;
  }
  std::vector<uint32_t>().swap(key);
  std::vector<int32_t> stamp(ng, -1);
  std::vector<uint16_t> local(ng, 0);
  std::vector<uint8_t> touch(ng, 0);
  out = PFUPacked{};
  out.rec.resize(4 * nel);
  out.chunk_el.push_back(0);
  out.chunk_site.push_back(0);
  std::vector<uint32_t> list;
  auto nnodes = [&](int type) {
    return (type == minim_amd::kPFUFluid || type == minim_amd::kPFUFluidAll) ? minim_amd::kPFUNodes : 1;
  };
  size_t pos = 0;
  for (size_t tl = 0; tl < ntile; tl++) {
    if (start[tl] == start[tl + 1]) continue;
    const int32_t c = (int32_t)(out.chunk_el.size() - 1);
    list.clear();
    for (uint32_t k = start[tl]; k < start[tl + 1]; k++) {
      const size_t e = order[k];
      const int nn = nnodes(el.type[e]);
      for (int i = 0; i < nn; i++) {
        const size_t site = (size_t)el.idof[e * minim_amd::kPFUNodes + i] / nf;
        if (stamp[site] != c) {
          stamp[site] = c;
          list.push_back((uint32_t)site);
        }
      }
    }
    if (list.size() * (size_t)nf > (size_t)minim_amd::kPFUChunkDofs) return false;
    std::sort(list.begin(), list.end());
    for (size_t i = 0; i < list.size(); i++) {
      local[list[i]] = (uint16_t)i;
      if (touch[list[i]] < 2) touch[list[i]]++;
    }
    for (uint32_t k = start[tl]; k < start[tl + 1]; k++, pos++) {
      const size_t e = order[k];
      const int type = el.type[e];
      const int nn = nnodes(type);
      const int32_t* id = &el.idof[e * minim_amd::kPFUNodes];
      uint32_t l[minim_amd::kPFUNodes];
      for (int i = 0; i < minim_amd::kPFUNodes; i++) l[i] = local[(size_t)id[i < nn ? i : 0] / nf];
      const uint32_t fluid = (uint32_t)(id[0] % nf);
      uint32_t* r = &out.rec[4 * pos];
      r[0] = l[0] | (l[1] << 16);
      r[1] = l[2] | (l[3] << 16);
      r[2] = l[4] | (l[5] << 16);
      r[3] = l[6] | ((uint32_t)type << 16) | (fluid << 24);
      if (type == minim_amd::kPFUSurface || type == minim_amd::kPFUForce || type == minim_amd::kPFUConfinement) {
        if (out.pidx.empty()) out.pidx.assign(nel, 0);
        out.pidx[pos] = (uint32_t)(out.par.size() / 3);
        out.par.insert(out.par.end(), &el.par[3 * e], &el.par[3 * e] + 3);
      }
    }
    out.sites.insert(out.sites.end(), list.begin(), list.end());
    out.chunk_el.push_back((uint32_t)pos);
    out.chunk_site.push_back((uint32_t)out.sites.size());
  }
  for (uint32_t& en : out.sites)
    if (touch[en] > 1) en |= minim_amd::kPFUShared;
  // the gather pass: every site but the exclusive ones (untouched sites too,
  // whose gradient it sets to zero), with its table positions in chunk order
  std::vector<uint32_t>().swap(order);
  std::vector<int32_t>& gidx = stamp;  // reused: site -> gather index, or -1
  out.gpos_off.push_back(0);
  for (size_t site = 0; site < ng; site++) {
    gidx[site] = -1;
    if (touch[site] == 1) continue;
    gidx[site] = (int32_t)out.gsite.size();
    out.gsite.push_back((uint32_t)site);
    out.gpos_off.push_back(0);
  }
  for (uint32_t en : out.sites)
    if (en & minim_amd::kPFUShared) out.gpos_off[gidx[en & ~minim_amd::kPFUShared] + 1]++;
  for (size_t j = 0; j < out.gsite.size(); j++) out.gpos_off[j + 1] += out.gpos_off[j];
  out.gpos.resize(out.gpos_off.back());
  std::vector<uint32_t> fill(out.gpos_off.begin(), out.gpos_off.end() - 1);
  for (size_t p = 0; p < out.sites.size(); p++)
    if (out.sites[p] & minim_amd::kPFUShared) out.gpos[fill[gidx[out.sites[p] & ~minim_amd::kPFUShared]]++] = (uint32_t)p;
  return true;
}

// Launch arguments, the packed element list until the first launch uploads
// it, and the device memory.
struct PFUHolder {
  minim_amd::PFUArgs u{};
  minim_amd::PFArgs con{};  // the per-site projection (hard density, fixed fluids)
  PFUPacked pk;
  std::vector<double> vol_h, target_h;
  minim_amd::DeviceBuffer rec, pidx, par, chunk_el, chunk_site, sites, gsite, gpos_off, gpos, scratch, vol, target,
      volg;
  std::once_flag uploaded;
};

template <typename T>
const T* upload_vec(minim_amd::DeviceBuffer& buf, std::vector<T>& v) {
  if (v.empty()) return nullptr;
  buf.reset(sizeof(T) * v.size());
  MINIM_HIP_CHECK(hipMemcpy(buf.get(), v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  std::vector<T>().swap(v);
  return static_cast<const T*>(buf.get());
}

PFUHolder& holder_peek(const std::shared_ptr<void>& p) {
  if (!p) throw std::runtime_error("PhaseFieldUnstructured: not initialised (construct a State first).");
  return *static_cast<PFUHolder*>(p.get());
}

// The holder with its element list in device memory (uploaded on first use,
// so that init itself needs no GPU).
PFUHolder& holder_of(const std::shared_ptr<void>& p) {
  PFUHolder* h = &holder_peek(p);
  std::call_once(h->uploaded, [h] {
    h->u.rec = reinterpret_cast<const uint4*>(upload_vec(h->rec, h->pk.rec));
    h->u.pidx = upload_vec(h->pidx, h->pk.pidx);
    h->u.par = upload_vec(h->par, h->pk.par);
    h->u.chunk_el = upload_vec(h->chunk_el, h->pk.chunk_el);
    h->u.chunk_site = upload_vec(h->chunk_site, h->pk.chunk_site);
    const size_t nslot = h->pk.sites.size() * (size_t)h->u.nf;
    h->u.sites = upload_vec(h->sites, h->pk.sites);
    h->u.ngather = h->pk.gsite.size();
    h->u.gsite = upload_vec(h->gsite, h->pk.gsite);
    h->u.gpos_off = upload_vec(h->gpos_off, h->pk.gpos_off);
    h->u.gpos = upload_vec(h->gpos, h->pk.gpos);
    h->scratch.reset(sizeof(double) * (nslot ? nslot : 1));
    h->u.scratch = h->scratch.get();
    h->u.node_vol = upload_vec(h->vol, h->vol_h);
    h->u.volume = upload_vec(h->target, h->target_h);
    h->volg.reset(sizeof(double) * minim_amd::kPFMaxFluid);
    h->u.volg = static_cast<double*>(h->volg.get());
  });
  return *h;
}
}  // namespace

PhaseFieldUnstructured& PhaseFieldUnstructured::setNFluid(int n) {
  nFluid = n;
  dofPerNode = n;
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setGridSize(vector<int> g) {
  gridSize = g;
  nGrid = 1;
  for (int s : g) nGrid *= s;
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setResolution(double r) {
  resolution = r;
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setInterfaceSize(double w) {
  interfaceSize = vector<double>(nFluid, w);
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setInterfaceSize(vector<double> w) {
  interfaceSize = w;
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setSurfaceTension(double s) {
  surfaceTension = vector<double>(nFluid, s);
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setSurfaceTension(vector<double> s) {
  surfaceTension = s;
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setDensityConstraint(std::string method) {
  if (method == "gradient" || method == "hard") densityConstraint = minim_amd::kDensityHard;
  else if (method == "energy penalty" || method == "soft") densityConstraint = minim_amd::kDensitySoft;
  else if (method == "none") densityConstraint = minim_amd::kDensityNone;
  else if (method == "fixed") densityConstraint = minim_amd::kDensityFixed;
  else
    throw std::invalid_argument(
        "Invalid density constraint. Allowed methods are: fixed, none, gradient / hard, or energy penalty / soft");
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setPressure(vector<double> p) {
  pressure = p;
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setVolume(vector<double> v, double c) {
  volumeFixed = true;
  volConst = c;
  volume = v;
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setVolumeFixed(bool f, double c) {
  volumeFixed = f;
  volConst = c;
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setSolid(vector<char> s) {
  solid = s;
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setSolid(std::function<bool(int, int, int)> fn) {
  if (gridSize.size() != 3) throw std::invalid_argument("PhaseFieldUnstructured: set the grid size first.");
  solid.assign(nGrid, 0);
  size_t i = 0;
  for (int x = 0; x < gridSize[0]; x++)
    for (int y = 0; y < gridSize[1]; y++)
      for (int z = 0; z < gridSize[2]; z++) solid[i++] = fn(x, y, z);
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setContactAngle(vector<double> a) {
  contactAngle = a;
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setContactAngle(std::function<double(int, int, int)> fn) {
  if (gridSize.size() != 3) throw std::invalid_argument("PhaseFieldUnstructured: set the grid size first.");
  contactAngle.assign(nGrid, 0.0);
  size_t i = 0;
  for (int x = 0; x < gridSize[0]; x++)
    for (int y = 0; y < gridSize[1]; y++)
      for (int z = 0; z < gridSize[2]; z++) contactAngle[i++] = fn(x, y, z);
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setForce(vector<double> f, vector<int> iFluid) {
  if ((int)f.size() != 3) throw std::invalid_argument("Invalid size of force array.");
  if (nFluid == 1 || iFluid.empty()) {
    force = vector2d<double>(nFluid, f);
  } else {
    force = vector2d<double>(nFluid, {0, 0, 0});
    for (int i : iFluid) force.at(i) = f;
  }
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setFixFluid(int iFluid, bool fix) {
  if ((int)fixFluid.size() != nFluid) fixFluid = vector<char>(nFluid, false);
  fixFluid.at(iFluid) = fix;
  return *this;
}
PhaseFieldUnstructured& PhaseFieldUnstructured::setConfinement(vector<double> s) {
  confinementStrength = s;
  return *this;
}

void PhaseFieldUnstructured::setDefaults() {
  if (interfaceSize.empty()) interfaceSize = vector<double>(nFluid, resolution);
  if (solid.empty()) solid = vector<char>(nGrid, false);
  if (pressure.empty()) pressure = vector<double>(nFluid, 0);
  if (fixFluid.empty()) fixFluid = vector<char>(nFluid, false);
  if (confinementStrength.empty()) confinementStrength = vector<double>(nFluid, 0);
}

void PhaseFieldUnstructured::checkArraySizes() {
  if ((int)interfaceSize.size() != nFluid) throw std::invalid_argument("Invalid size interfaceSize array.");
  if ((int)solid.size() != nGrid) throw std::invalid_argument("Invalid size of solid array.");
  if ((long)contactAngle.size() != (long)nGrid * nFluid && !contactAngle.empty())
    throw std::invalid_argument("Invalid size of contactAngle array.");
  if ((int)force.size() != nFluid && !force.empty()) throw std::invalid_argument("Invalid size of force array.");
  if ((int)volume.size() != nFluid && !volume.empty()) throw std::invalid_argument("Invalid size of volume array.");
  if ((int)pressure.size() != nFluid) throw std::invalid_argument("Invalid size of pressure array.");
  if ((int)fixFluid.size() != nFluid) throw std::invalid_argument("Invalid size of fixed fluid array.");
  if ((int)confinementStrength.size() != nFluid)
    throw std::invalid_argument("Invalid size of confinement strength array.");
}

// PhaseFieldUnstructured.cpp:71-107 (a fourth and later fluid's coefficients
// follow :103-104, which differ from PhaseField's)
void PhaseFieldUnstructured::assignFluidCoefficients() {
  const int np = nFluid;
  if ((int)interfaceSize.size() == 1) interfaceSize = vector<double>(np, interfaceSize[0]);
  else if ((int)interfaceSize.size() != np) throw std::invalid_argument("Invalid size of interfaceSize array.");
  if ((int)surfaceTension.size() == 1) surfaceTension = vector<double>(np, surfaceTension[0]);
  else if ((int)surfaceTension.size() != np) throw std::invalid_argument("Invalid size of surfaceTension array.");
  double sum = 0;
  for (double v : surfaceTension) sum += v;
  surfaceTensionMean = sum / np;
  kappa.assign(np, 0.0);
  kappaP.assign(np, 0.0);
  const vector<double>& st = surfaceTension;
  const vector<double>& iw = interfaceSize;
  if (nFluid <= 2) {
    for (int i = 0; i < np; i++) {
      kappa[i] = 3 * st[i] / iw[i];
      kappaP[i] = 3 * st[i] * iw[i];
    }
  } else {
    vector<double> ks(np), kps(np);
    for (int i = 0; i < np; i++) {
      ks[i] = 6 * st[i] / iw[i];
      kps[i] = 6 * st[i] * iw[i];
    }
    const int l = nFluid - 1;
    kappa[0] = 0.5 * (ks[0] + ks[1] - ks[l]);
    kappa[1] = 0.5 * (ks[0] - ks[1] + ks[l]);
    kappa[2] = 0.5 * (-ks[0] + ks[1] + ks[l]);
    kappaP[0] = 0.5 * (kps[0] + kps[1] - kps[l]);
    kappaP[1] = 0.5 * (kps[0] - kps[1] + kps[l]);
    kappaP[2] = 0.5 * (-kps[0] + kps[1] + kps[l]);
    for (int i = 3; i < nFluid; i++) {
      kappa[i] = 0.5 * (-ks[0] - ks[1] + ks[l]) + ks[i - 1];
      kappaP[i] = 0.5 * (-kps[0] - kps[1] + kps[l]) + kps[i - 1];
    }
  }
}

// PhaseFieldUnstructured::init (:110-257) and initLocal on one rank (:260-277)
void PhaseFieldUnstructured::init(const vector<double>& coords) {
  if (gridSize.size() != 3) throw std::invalid_argument("PhaseFieldUnstructured: gridSize must have three entries.");
  for (int s : gridSize)
    if (s < 1) throw std::invalid_argument("PhaseFieldUnstructured: grid sizes must be positive.");
  nGrid = gridSize[0] * gridSize[1] * gridSize[2];
  if (nFluid < 1 || nFluid > minim_amd::kPFMaxFluid)
    throw std::invalid_argument("PhaseFieldUnstructured: nFluid must be between 1 and 16.");
  if ((long)coords.size() != (long)nGrid * nFluid)
    throw std::invalid_argument("Size of coordinates array does not match the grid size.");
  if (!constraints.empty())
    throw std::invalid_argument(
        "PhaseFieldUnstructured: single-DOF and normal constraints are not supported; use setFixFluid.");
  setDefaults();
  checkArraySizes();
  assignFluidCoefficients();
  convergence = 1e-8 * surfaceTensionMean * std::pow(resolution, 2);
  if (densityConstraint == minim_amd::kDensityFixed) fixFluid[nFluid - 1] = true;
  const int nF = nFluid;
  // the soft constraint's element reads three DOFs of its site (:631-642)
  if (nF == 2 && densityConstraint == minim_amd::kDensitySoft)
    throw std::invalid_argument(
        "PhaseFieldUnstructured: the soft density constraint needs at least three fluids (its element reads three "
        "concentrations per site)");

  vector<double> fMag(nF, 0.0);
  vector2d<double> fNorm(nF, vector<double>(3, 0.0));
  if (!force.empty()) {
    for (int i = 0; i < nF; i++) {
      double r = 0;
      for (int d = 0; d < 3; d++) r += force[i][d] * force[i][d];
      fMag[i] = std::sqrt(r);
      for (int d = 0; d < 3; d++) fNorm[i][d] = force[i][d] / fMag[i];
    }
  }

  const size_t ng = (size_t)nGrid;
  std::vector<int> type = pf_site_types(gridSize.data(), solid.data());
  static const double kArea[8] = {0, 0.75, 1, 1.25, 1, 1.25, 1, 0.75};
  vector<double> vol(ng), area(ng);
  for (size_t i = 0; i < ng; i++) {
    const int t = type[i];
    const double v = (t < 0) ? 0.0 : (t == 0 ? 1.0 : t / 8.0);
    vol[i] = v * std::pow(resolution, 3);
    area[i] = (t < 0 ? 0.0 : kArea[t]) * std::pow(resolution, 2);
  }
  if (volume.empty() && volumeFixed) {
    volume.assign(nF, 0.0);
    for (size_t i = 0; i < ng; i++)
      for (int k = 0; k < nF; k++) volume[k] += coords[k + nF * i] * vol[i];
  }

  auto h = std::make_shared<PFUHolder>();
  RefElements el;
  auto push = [&el](int t, std::initializer_list<int32_t> ids, std::initializer_list<double> par) {
    el.push(t, ids, par);
  };
  const int n0 = gridSize[0], n1 = gridSize[1], n2 = gridSize[2];
  static const int off[6][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}};
  for (int x = 0; x < n0; x++)
    for (int y = 0; y < n1; y++)
      for (int z = 0; z < n2; z++) {
        const int iGrid = (x * n1 + y) * n2 + z;
        if (solid[iGrid]) continue;
        int nodes[7];
        nodes[0] = iGrid;
        for (int k = 0; k < 6; k++) {
          const int j =
              (((x + off[k][0] + n0) % n0) * n1 + (y + off[k][1] + n1) % n1) * n2 + (z + off[k][2] + n2) % n2;
          nodes[k + 1] = solid[j] ? iGrid : j;
        }
        const double v = vol[iGrid];
        if (densityConstraint == minim_amd::kDensityFixed) {
          push(minim_amd::kPFUFluidAll,
               {nodes[0] * nF, nodes[1] * nF, nodes[2] * nF, nodes[3] * nF, nodes[4] * nF, nodes[5] * nF,
                nodes[6] * nF},
               {v});
        } else {
          for (int f = 0; f < nF; f++)
            push(minim_amd::kPFUFluid,
                 {nodes[0] * nF + f, nodes[1] * nF + f, nodes[2] * nF + f, nodes[3] * nF + f, nodes[4] * nF + f,
                  nodes[5] * nF + f, nodes[6] * nF + f},
                 {v, (double)f});
        }
        for (int f = 0; f < nF; f++)
          if (pressure[f] != 0) push(minim_amd::kPFUPressure, {iGrid * nF + f}, {});
        if (nF > 1 && densityConstraint == minim_amd::kDensitySoft)
          push(minim_amd::kPFUDensity, {iGrid * nF, iGrid * nF + 1, iGrid * nF + 2}, {});
        if (area[iGrid] > 0 && !contactAngle.empty() && contactAngle[iGrid] != 90) {
          const double w = 1 / std::sqrt(2.0) * std::cos(contactAngle[iGrid] * 3.1415926536 / 180);
          for (int f = 0; f < nF; f++) push(minim_amd::kPFUSurface, {iGrid * nF + f}, {area[iGrid], w});
        }
        for (int f = 0; f < nF; f++) {
          if (!(fMag[f] > 0)) continue;
          const double coord[3] = {x - (n0 - 1) / 2.0, y - (n1 - 1) / 2.0, z - (n2 - 1) / 2.0};
          double dot = 0;
          for (int d = 0; d < 3; d++) dot += coord[d] * fNorm[f][d];
          const double hh = -dot * resolution;
          push(minim_amd::kPFUForce, {iGrid * nF + f}, {v, fMag[f], hh});
        }
        for (int f = 0; f < nF; f++)
          if (confinementStrength[f] != 0)
            push(minim_amd::kPFUConfinement, {iGrid * nF + f}, {confinementStrength[f], coords[iGrid * nF + f]});
      }
  nElements = el.size();
  if (ng * (size_t)nF >= ((size_t)1 << 31)) throw std::invalid_argument("PhaseFieldUnstructured: more than 2^31 DOFs");
  if (nElements >= ((size_t)1 << 32)) throw std::invalid_argument("PhaseFieldUnstructured: more than 2^32 elements");
  {
    // the device list: chunks of tiles, shrunk until every chunk fits its LDS accumulator
    int t[3];
    choose_tile(gridSize.data(), nF, t);
    while (!pack_elements(el, gridSize.data(), nF, t, h->pk)) {
      int* big = std::max_element(t, t + 3);
      if (*big == 1) throw std::runtime_error("PhaseFieldUnstructured: an element touches too many sites");
      *big = (*big + 1) / 2;
    }
    h->u.nchunk = (int)(h->pk.chunk_el.size() - 1);
    const PFUPacked& pk = h->pk;
    listBytes = 4 * (pk.rec.size() + pk.pidx.size() + pk.chunk_el.size() + pk.chunk_site.size() + pk.sites.size()) +
                8 * pk.par.size();
  }

  // initLocal: node volume and fluid of every DOF
  nodeVol.assign(ng * nF, 0.0);
  fluidType.assign(ng * nF, 0);
  for (size_t i = 0; i < ng; i++)
    for (int k = 0; k < nF; k++) {
      nodeVol[i * nF + k] = vol[i];
      fluidType[i * nF + k] = k;
    }

  minim_amd::PFUArgs& u = h->u;
  u.nel = nElements;
  u.nf = nF;
  u.ngrid = ng;
  for (int k = 0; k < nF; k++) {
    u.kappa[k] = kappa[k];
    u.kappaP[k] = kappaP[k];
    u.pressure[k] = pressure[k];
  }
  u.res2 = std::pow(resolution, 2);
  u.coef_density = densityConst * surfaceTensionMean * std::pow(resolution, 2);
  u.st_mean = surfaceTensionMean;
  u.volfix = volumeFixed ? 1 : 0;
  u.vol_coef = volConst * surfaceTensionMean / std::pow(resolution, 4);
  h->vol_h = vol;
  if (volumeFixed) h->target_h = volume;

  // constraints (:242-256): the variable fluids' sum per site, then the fixed fluids
  minim_amd::PFArgs& a = h->con;
  a.nf = nF;
  int nvar = 0;
  for (int k = 0; k < nF; k++) {
    if (fixFluid[k]) a.fixmask |= (1u << k);
    else nvar++;
  }
  a.nvar = nvar;
  a.inv_nvar = nvar > 0 ? 1.0 / nvar : 0.0;
  a.density = densityConstraint;
  dev_ = h;
}

void PhaseFieldUnstructured::evaluate(int prec, const void* x, const void* d, double t, double* e, void* g,
                                      void* stream) const {
  PFUHolder& h = holder_of(dev_);
  const int mode = (e ? 1 : 0) | (g ? 2 : 0) | (d ? 4 : 0);
  if (mode == 0) return;
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto& ws = minim_amd::Workspace::current();
  minim_amd::pfu_launch_eval(h.u, prec, mode, x, d, t, g, ws.rs(), s);
  if (g) minim_amd::pf_launch_constrain(h.con, prec, g, h.u.ngrid, ws.rs(), s);
  if (e) {
    double r[1];
    ws.fetch(s, 1, r);
    *e = r[0];
  }
}

void PhaseFieldUnstructured::evaluateAsync(int prec, const void* x, const void* d, double t, bool wantE, double* e_dev,
                                           void* g, void* stream) const {
  PFUHolder& h = holder_of(dev_);
  const int mode = (wantE ? 1 : 0) | (g ? 2 : 0) | (d ? 4 : 0);
  if (mode == 0) return;
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto& ws = minim_amd::Workspace::current();
  minim_amd::pfu_launch_eval(h.u, prec, mode, x, d, t, g, ws.rs(), s);
  if (g) minim_amd::pf_launch_constrain(h.con, prec, g, h.u.ngrid, ws.rs(), s);
  if (wantE && e_dev)
    MINIM_HIP_CHECK(hipMemcpyAsync(e_dev, ws.device_result(), sizeof(double), hipMemcpyDeviceToDevice, s));
}

double PhaseFieldUnstructured::componentEnergy(int, const vector<double>&) const {
  throw std::invalid_argument(
      "PhaseFieldUnstructured: energies per element type are not available (the element list lives on the device).");
}

double PhaseFieldUnstructured::timeKernel(int prec, const void* x, bool wantE, void* g, int reps, void* stream) const {
  PFUHolder& h = holder_of(dev_);
  if (!g) throw std::invalid_argument("PhaseFieldUnstructured: timing needs a gradient buffer");
  auto& ws = minim_amd::Workspace::current();
  return minim_amd::pfu_time_elements(h.u, prec, wantE ? 3 : 2, x, g, reps, ws.rs(), static_cast<hipStream_t>(stream));
}

void PhaseFieldUnstructured::constrainDevice(int prec, const void* x, void* v, size_t n, void* stream) const {
  (void)x;
  PFUHolder& h = holder_of(dev_);
  auto& ws = minim_amd::Workspace::current();
  minim_amd::pf_launch_constrain(h.con, prec, v, n / h.u.nf, ws.rs(), static_cast<hipStream_t>(stream));
}

SiteProjection PhaseFieldUnstructured::projection() const {
  const PFUHolder& h = holder_peek(dev_);
  SiteProjection p;
  p.nf = h.con.nf;
  p.fixmask = h.con.fixmask;
  p.hard = (h.con.density == minim_amd::kDensityHard && h.con.nf > 1) ? 1 : 0;
  p.nvar = h.con.nvar;
  p.inv_nvar = h.con.inv_nvar;
  return p;
}

// PhaseFieldUnstructured::diffuseSolid (:957-994), the same procedure as PhaseField's
vector<double> PhaseFieldUnstructured::diffuseSolid(vector<char> s, int iFluid, bool twoStep) {
  return diffuseSolid(s, *this, iFluid, twoStep);
}

vector<double> PhaseFieldUnstructured::diffuseSolid(vector<char> s, vector<int> grid, int nF, int iFluid,
                                                    bool twoStep) {
  PhaseFieldUnstructured p;
  p.setNFluid(nF);
  p.setGridSize(grid);
  return diffuseSolid(s, p, iFluid, twoStep);
}

vector<double> PhaseFieldUnstructured::diffuseSolid(vector<char> s, PhaseFieldUnstructured p, int iFluid,
                                                    bool twoStep) {
  if (iFluid < 0 || iFluid >= p.nFluid) throw std::invalid_argument("PhaseFieldUnstructured: invalid fluid index.");
  vector<double> conf(p.nFluid, 0.0);
  conf[iFluid] = 100;
  p.setConfinement(conf);
  p.setDensityConstraint("none");
  const int ng = p.gridSize.at(0) * p.gridSize.at(1) * p.gridSize.at(2);
  if ((int)s.size() != ng) throw std::invalid_argument("Invalid size of solid array.");
  vector<double> init((size_t)p.nFluid * ng, 0.0);
  for (int i = 0; i < ng; i++) init[(size_t)i * p.nFluid + iFluid] = s[i];
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
