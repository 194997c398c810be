// state.cpp -- State with device-resident coordinates.
//
// Mirrors src/State.cpp: the constructor clones and initialises the potential
// (:12-26), copies re-clone it (:28-45), energy()/gradient() evaluate the
// potential with constraints applied to gradients (basicEG, :48-56), and the
// coordinate accessors gather/scatter (:282-297), which on one GPU are a
// device<->host copy.
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

State::State(const Potential& p, const vector<double>& coords, int precision)
    : ndof(coords.size()), pot(p.clone()), prec_(precision) {
  if (precision != FP64 && precision != FP32) throw std::invalid_argument("State: precision must be FP64 or FP32.");
  pot->init(coords);
  convergence = pot->convergence;
  x_ = make_dvec(ndof, prec_);
  minim_amd::upload(coords.data(), ndof, dv(x_), static_cast<hipStream_t>(stream_));
  host_ = coords;
  hostValid_ = (prec_ == FP64);
}

State::State(const State& o)
    : ndof(o.ndof), convergence(o.convergence), pot(o.pot->clone()), usesThisProc(o.usesThisProc),
      isFailed(o.isFailed), prec_(o.prec_), stream_(o.stream_) {
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
  usesThisProc = o.usesThisProc;
  isFailed = o.isFailed;
  prec_ = o.prec_;
  stream_ = o.stream_;
  x_ = make_dvec(ndof, prec_);
  minim_amd::device_copy(dv(x_).data(), dv(o.x_).data(), ndof * minim_amd::prec_bytes(prec_),
                         static_cast<hipStream_t>(stream_));
  MINIM_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream_)));
  host_ = o.host_;
  hostValid_ = o.hostValid_;
  return *this;
}

State::~State() = default;

void* State::deviceCoords() const { return dv(x_).data(); }

std::shared_ptr<void> State::newCoordsBuffer() const { return make_dvec(ndof, prec_); }

void State::swapDeviceCoords(std::shared_ptr<void>& buffer) {
  if (!buffer || dv(buffer).n != dv(x_).n || dv(buffer).prec != dv(x_).prec)
    throw std::invalid_argument("State: coordinate buffer of another size or precision");
  std::swap(x_, buffer);
  hostValid_ = false;
}

void State::setPrecision(int precision) {
  if (precision == prec_) return;
  if (precision != FP64 && precision != FP32) throw std::invalid_argument("State: precision must be FP64 or FP32.");
  auto nx = make_dvec(ndof, precision);
  minim_amd::launch_convert(precision, dv(nx).data(), prec_, dv(x_).data(), ndof, static_cast<hipStream_t>(stream_));
  MINIM_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream_)));
  x_ = nx;
  prec_ = precision;
  hostValid_ = false;
  version_++;
}

// ---- energy / gradient ----
double State::energy() const {
  double e = 0;
  pot->evaluate(prec_, deviceCoords(), nullptr, 0.0, &e, nullptr, stream_);
  return e;
}

double State::energy(const vector<double>& c) const {
  if (c.size() != ndof) throw std::invalid_argument("State: coordinates have the wrong size.");
  DVec tmp(ndof, prec_);
  minim_amd::upload(c.data(), ndof, tmp, static_cast<hipStream_t>(stream_));
  double e = 0;
  pot->evaluate(prec_, tmp.data(), nullptr, 0.0, &e, nullptr, stream_);
  return e;
}

vector<double> State::gradient() const {
  DVec g(ndof, prec_);
  pot->evaluate(prec_, deviceCoords(), nullptr, 0.0, nullptr, g.data(), stream_);
  vector<double> out(ndof);
  minim_amd::download(g, out.data(), static_cast<hipStream_t>(stream_));
  return out;
}

vector<double> State::gradient(const vector<double>& c) const {
  if (c.size() != ndof) throw std::invalid_argument("State: coordinates have the wrong size.");
  DVec tmp(ndof, prec_), g(ndof, prec_);
  minim_amd::upload(c.data(), ndof, tmp, static_cast<hipStream_t>(stream_));
  pot->evaluate(prec_, tmp.data(), nullptr, 0.0, nullptr, g.data(), stream_);
  vector<double> out(ndof);
  minim_amd::download(g, out.data(), static_cast<hipStream_t>(stream_));
  return out;
}

void State::energyGradient(double* e, vector<double>* g) const {
  if (!g) {
    if (e) *e = energy();
    return;
  }
  DVec gd(ndof, prec_);
  pot->evaluate(prec_, deviceCoords(), nullptr, 0.0, e, gd.data(), stream_);
  g->resize(ndof);
  minim_amd::download(gd, g->data(), static_cast<hipStream_t>(stream_));
}

void State::energyGradient(const vector<double>& c, double* e, vector<double>* g) const {
  if (c.size() != ndof) throw std::invalid_argument("State: coordinates have the wrong size.");
  DVec tmp(ndof, prec_);
  minim_amd::upload(c.data(), ndof, tmp, static_cast<hipStream_t>(stream_));
  if (!g) {
    if (e) pot->evaluate(prec_, tmp.data(), nullptr, 0.0, e, nullptr, stream_);
    return;
  }
  DVec gd(ndof, prec_);
  pot->evaluate(prec_, tmp.data(), nullptr, 0.0, e, gd.data(), stream_);
  g->resize(ndof);
  minim_amd::download(gd, g->data(), static_cast<hipStream_t>(stream_));
}

// ---- coordinates ----
double State::operator[](int i) { return blockCoords().at(i); }

vector<double> State::coords() const { return blockCoords(); }

void State::coords(const vector<double>& in) {
  if (in.size() != ndof) throw std::invalid_argument("State: coordinates have the wrong size.");
  minim_amd::upload(in.data(), ndof, dv(x_), static_cast<hipStream_t>(stream_));
  host_ = in;
  hostValid_ = (prec_ == FP64);
  version_++;
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
  DVec tmp(ndof, FP64);
  minim_amd::upload(data.data(), ndof, tmp, static_cast<hipStream_t>(stream_));
  pot->applyConstraints(FP64, tmp.data(), ndof, stream_);
  minim_amd::download(tmp, data.data(), static_cast<hipStream_t>(stream_));
}

void State::failed() { isFailed = true; }

}  // namespace minim
