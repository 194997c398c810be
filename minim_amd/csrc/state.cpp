// state.cpp -- State with device-resident coordinates.
//
// Mirrors src/State.cpp: the constructor clones and initialises the potential,
// sets up its communicator and initialises it locally (:12-26), copies
// re-clone both (:28-45), energy()/gradient() evaluate the potential with
// constraints applied to gradients (basicEG / elementEG, :48-79), and the
// coordinate accessors gather/scatter (:282-297), which on one GPU are a
// device<->host copy through the pinned staging buffers.  A State whose rank
// list leaves this process out answers like the reference's unused ranks:
// zero energies, empty vectors.
#include <stdexcept>

#include "minim/State.h"
#include "runtime.hpp"
#include "vec.hpp"

namespace minim {

using minim_amd::DVec;

namespace {
DVec& dv(const std::shared_ptr<void>& p) { return *static_cast<DVec*>(p.get()); }
std::shared_ptr<void> make_dvec(size_t n, int prec) {
  auto v = std::make_shared<DVec>(n, prec);
  return std::static_pointer_cast<void>(v);
}
}  // namespace

// ---- HostCoords
const vector<double>& HostCoords::get() const { return s_->blockCoords(); }

HostCoords& HostCoords::operator=(const vector<double>& v) {
  s_->blockCoords(v);
  return *this;
}

HostCoords& HostCoords::operator=(const HostCoords& o) { return *this = o.get(); }

// ---- construction
State::State(const Potential& p, const vector<double>& coords, const vector<int>& ranks)
    : ndof(coords.size()), pot(p.clone()) {
  construct(p, coords, FP64, ranks);
}

State::State(const Potential& p, const vector<double>& coords, Precision precision, const vector<int>& ranks)
    : ndof(coords.size()), pot(p.clone()) {
  construct(p, coords, static_cast<int>(precision), ranks);
}

void State::construct(const Potential&, const vector<double>& coords, int precision, const vector<int>& ranks) {
  if (precision != FP64 && precision != FP32) throw std::invalid_argument("State: precision must be FP64 or FP32.");
  prec_ = precision;
  pot->ndofHint = ndof;
  pot->init(coords);
  convergence = pot->convergence;
  comm = pot->newComm();
  comm->setup(*pot, ndof, ranks);
  usesThisProc = comm->usesThisProc;
  bindPotential();
  pot->initLocal(coords, *comm);
  x_ = make_dvec(ndof, prec_);
  minim_amd::upload(coords.data(), ndof, dv(x_), static_cast<hipStream_t>(stream_));
  host_ = coords;
  hostValid_ = (prec_ == FP64);
}

void State::bindPotential() {
  pot->ndofHint = ndof;
  pot->stateComm = comm.get();
}

State::State(const State& o)
    : ndof(o.ndof),
      convergence(o.convergence),
      pot(o.pot->clone()),
      comm(o.comm->clone()),
      usesThisProc(o.usesThisProc),
      isFailed(o.isFailed),
      prec_(o.prec_),
      stream_(o.stream_) {
  bindPotential();
  x_ = make_dvec(ndof, prec_);
  minim_amd::device_copy(dv(x_).data(), dv(o.x_).data(), ndof * minim_amd::prec_bytes(prec_),
                         static_cast<hipStream_t>(stream_));
  MINIM_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream_)));
  host_ = o.host_;
  hostValid_ = o.hostValid_;
}

State& State::operator=(const State& o) {
  if (this == &o) return *this;
  ndof = o.ndof;
  convergence = o.convergence;
  pot = o.pot->clone();
  comm = o.comm->clone();
  usesThisProc = o.usesThisProc;
  isFailed = o.isFailed;
  prec_ = o.prec_;
  stream_ = o.stream_;
  bindPotential();
  x_ = make_dvec(ndof, prec_);
  minim_amd::device_copy(dv(x_).data(), dv(o.x_).data(), ndof * minim_amd::prec_bytes(prec_),
                         static_cast<hipStream_t>(stream_));
  MINIM_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream_)));
  host_ = o.host_;
  hostValid_ = o.hostValid_;
  version_++;
  return *this;
}

State::~State() = default;

void* State::deviceCoords() const { return dv(x_).data(); }

std::shared_ptr<void> State::newCoordsBuffer() const { return make_dvec(ndof, prec_); }

void State::swapDeviceCoords(std::shared_ptr<void>& buffer) {
  if (!buffer || dv(buffer).n != dv(x_).n || dv(buffer).prec != dv(x_).prec)
    throw std::invalid_argument("State: coordinate buffer of another size or precision");
  std::swap(x_, buffer);
  markCoordsChanged();
}

void State::setPrecision(int precision) {
  if (precision == prec_) return;
  if (precision != FP64 && precision != FP32) throw std::invalid_argument("State: precision must be FP64 or FP32.");
  auto nx = make_dvec(ndof, precision);
  minim_amd::launch_convert(precision, dv(nx).data(), prec_, dv(x_).data(), ndof, static_cast<hipStream_t>(stream_));
  MINIM_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream_)));
  x_ = nx;
  prec_ = precision;
  markCoordsChanged();
}

// ---- energy / gradient
namespace {
void check_size(size_t have, size_t want) {
  if (have != want) throw std::invalid_argument("State: coordinates have the wrong size.");
}
}  // namespace

double State::energy() const {
  if (!usesThisProc) return 0;
  double e = 0;
  pot->evaluate(prec_, deviceCoords(), nullptr, 0.0, &e, nullptr, stream_);
  return e;
}

double State::energy(const vector<double>& c) const {
  if (!usesThisProc) return 0;
  check_size(c.size(), ndof);
  hipStream_t s = static_cast<hipStream_t>(stream_);
  void* x = minim_amd::Staging::current().device(2, ndof * minim_amd::prec_bytes(prec_));
  minim_amd::upload_raw(c.data(), ndof, prec_, x, s);
  double e = 0;
  pot->evaluate(prec_, x, nullptr, 0.0, &e, nullptr, stream_);
  return e;
}

vector<double> State::gradient() const {
  if (!usesThisProc) return vector<double>();
  vector<double> out(ndof);
  energyGradientRaw(nullptr, nullptr, out.data());
  return out;
}

vector<double> State::gradient(const vector<double>& c) const {
  if (!usesThisProc) return vector<double>();
  check_size(c.size(), ndof);
  vector<double> out(ndof);
  energyGradientRaw(c.data(), nullptr, out.data());
  return out;
}

void State::energyGradient(double* e, vector<double>* g) const {
  if (!usesThisProc) return;
  if (g) g->resize(ndof);
  energyGradientRaw(nullptr, e, g ? g->data() : nullptr);
}

void State::energyGradient(const vector<double>& c, double* e, vector<double>* g) const {
  if (!usesThisProc) return;
  check_size(c.size(), ndof);
  if (g) g->resize(ndof);
  energyGradientRaw(c.data(), e, g ? g->data() : nullptr);
}

// Host coordinates go up through the pinned staging chunks (converted to the
// State's precision on the way), the gradient comes down the same way.
void State::energyGradientRaw(const double* c, double* e, double* g) const {
  if (!usesThisProc || (!e && !g)) return;
  hipStream_t s = static_cast<hipStream_t>(stream_);
  auto& st = minim_amd::Staging::current();
  const size_t bytes = ndof * minim_amd::prec_bytes(prec_);
  const void* x = deviceCoords();
  if (c) {
    void* xs = st.device(2, bytes);
    minim_amd::upload_raw(c, ndof, prec_, xs, s);
    x = xs;
  }
  void* gd = g ? st.device(3, bytes) : nullptr;
  pot->evaluate(prec_, x, nullptr, 0.0, e, gd, stream_);
  if (g) minim_amd::download_raw(gd, prec_, ndof, g, s);
}

double State::componentEnergy(int component) const {
  if (!usesThisProc) return 0;
  return pot->componentEnergy(component, blockCoords());
}

// ---- coordinates
double State::operator[](int i) { return comm->get(blockCoords(), i); }

vector<double> State::coords() const {
  if (!usesThisProc) return vector<double>();
  return blockCoords();
}

void State::coords(const vector<double>& in) {
  if (!usesThisProc) return;
  check_size(in.size(), ndof);
  minim_amd::upload(in.data(), ndof, dv(x_), static_cast<hipStream_t>(stream_));
  markCoordsChanged();
  host_ = in;
  hostValid_ = (prec_ == FP64);
}

const vector<double>& State::blockCoords() const {
  if (!hostValid_) {
    host_.resize(ndof);
    minim_amd::download(dv(x_), host_.data(), static_cast<hipStream_t>(stream_));
    hostValid_ = true;
  }
  return host_;
}

void State::blockCoords(const vector<double>& in) { coords(in); }

void State::applyConstraints(vector<double>& data) const {
  if (data.size() != ndof) throw std::invalid_argument("State: vector has the wrong size.");
  if (!pot->deviceNative()) {
    pot->applyConstraints(blockCoords(), *comm, data);
    return;
  }
  hipStream_t s = static_cast<hipStream_t>(stream_);
  void* v = minim_amd::Staging::current().device(2, ndof * sizeof(double));
  minim_amd::upload_raw(data.data(), ndof, FP64, v, s);
  const void* x = deviceCoords();
  if (prec_ != FP64) {
    // the device projection works on one precision: the coordinates widened
    void* xw = minim_amd::Staging::current().device(3, ndof * sizeof(double));
    minim_amd::launch_convert(FP64, xw, prec_, x, ndof, s);
    x = xw;
  }
  pot->constrainDevice(FP64, x, v, ndof, stream_);
  minim_amd::download_raw(v, FP64, ndof, data.data(), s);
}

void State::failed() { isFailed = true; }

}  // namespace minim
