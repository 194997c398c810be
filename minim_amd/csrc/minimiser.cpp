// minimiser.cpp -- device-resident minimisers with the reference's semantics.
//
//   Minimiser::minimise  src/Minimiser.cpp:27-38 (callback, failure flag,
//                        iter = 0..maxIter inclusive), log variant :41-69
//   Lbfgs                src/minimisers/Lbfgs.cpp:41-133
//   line search          src/minimisers/linesearch.cpp:8-29
//   Fire                 src/minimisers/Fire.cpp:27-90
//   GradDescent          src/minimisers/GradDescent.cpp:27-46
//
// Every vector lives on the GPU; the host only holds scalars.  Departures
// from the reference's *implementation* (not its results):
//  * L-BFGS keeps the (2m+1)^2 Gram matrix of {g, s_i, y_i} on the host and
//    runs the two-loop recursion on coefficients, so a direction costs one
//    fused pass over the basis instead of 2m dot/axpy sweeps;
//  * the line search's e0 (recomputed by the reference on every call,
//    linesearch.cpp:13) is taken from the previous accepted trial, which is the
//    energy of exactly the current coordinates; it is re-evaluated whenever
//    that is not the case (after the unevaluated 10-failure fallback, a step
//    without line search, or coordinates changed by the caller);
//  * L-BFGS's first trial at x + d is, where the potential supports it, one
//    pass that also yields the gradient there and x + d itself, used as is
//    when the trial is accepted (the same bits as trial, axpy and gradient);
//  * Fire reads the scalars an iteration needs (g.g, the next v.g and v.v,
//    the largest force, the potential's validity flag) in one host
//    synchronisation.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <deque>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include "minim/Minimiser.h"
#include "minim/State.h"
#include "autotune.hpp"
#include "runtime.hpp"
#include "vec.hpp"

namespace minim {

using minim_amd::DVec;
using minim_amd::Workspace;

// ------------------------------------------------------------------ Minimiser
Minimiser& Minimiser::setMaxIter(int m) {
  maxIter = m;
  return *this;
}

Minimiser& Minimiser::setLinesearch(std::string method) {
  if (method != "backtracking" && method != "none") throw std::invalid_argument("Invalid line search method.");
  linesearch = method;
  return *this;
}

std::vector<double> Minimiser::minimise(State& state, std::function<void(int, State&)> adjustState) {
  if (!state.usesThisProc) return std::vector<double>();
  run(state, adjustState);
  return state.coords();
}

void Minimiser::run(State& state, std::function<void(int, State&)> adjustState) {
  if (!state.usesThisProc) return;
  stats = MinimiseStats{};
  const auto t0 = std::chrono::steady_clock::now();
  init(state);
  for (iter = 0; iter <= maxIter; iter++) {
    if (adjustState) adjustState(iter, state);
    if (state.isFailed) break;
    iteration(state);
    const bool conv = checkConvergence(state);
    // first crossings of the decades of the convergence measure
    const double measure = stats.gradMax > 0 ? stats.gradMax : stats.gradRms;
    for (int k = 0; k < 8; k++) {
      if (stats.crossIter[k] >= 0 || !(measure < std::pow(10.0, -(k + 1)))) continue;
      stats.crossIter[k] = iter;
      stats.crossSeconds[k] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (conv) {
      stats.converged = true;
      break;
    }
  }
  finish(state);
  state.invalidateHost();
  stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

std::vector<double> Minimiser::minimise(State& state, std::string logType) {
  std::function<void(int, State&)> logFn = nullptr;
  size_t divPos = logType.find("-");
  if (divPos < logType.length()) {
    std::string logFields = logType.substr(0, divPos);
    int logIter = std::stoi(logType.substr(divPos + 1));
    if (logIter > 0) {
      int iterDigits = (int)std::to_string(maxIter).length();
      logFn = [logIter, logFields, iterDigits](int i, State& s) {
        if (i % logIter != 0) return;
        bool logE = logFields.find('e') < logFields.length();
        bool logG = logFields.find('g') < logFields.length();
        double e = 0;
        std::vector<double> g;
        s.invalidateHost();
        s.energyGradient(logE ? &e : nullptr, logG ? &g : nullptr);
        std::ostringstream log;
        log << "I: " << std::setw(iterDigits) << i;
        if (logE) log << "  E: " << e;
        if (logG) {
          double gg = 0;
          for (double v : g) gg += v * v;
          log << "  G: " << std::sqrt(gg / g.size());
        }
        std::printf("%s\n", log.str().c_str());
        std::fflush(stdout);
      };
    }
  }
  return minimise(state, logFn);
}

namespace {

hipStream_t hs(const State& s) { return static_cast<hipStream_t>(s.stream()); }

// Energy cache of the coordinates a minimiser last left in the State.
struct EnergyCache {
  bool valid = false;
  double e = 0;
  unsigned long version = 0;
  void set(const State& s, double v) {
    valid = true;
    e = v;
    version = s.coordsVersion();
  }
  bool ok(const State& s) const { return valid && version == s.coordsVersion(); }
};

// Evaluation at x (+ t*d) with the vectors d and g in precision pv: the
// State's own precision, or FP32 vectors against FP64 coordinates.
void eval_at(State& s, int pv, const void* d, double t, double* e, void* g) {
  if (pv == s.precision())
    s.pot->evaluate(pv, s.deviceCoords(), d, t, e, g, s.stream());
  else
    s.pot->evaluateMixed(s.deviceCoords(), d, t, e, g, s.stream());
}

// Energy at x + t*d launched without a host synchronisation; the value is
// left in the calling thread's reduction result slot 0.
void energy_async_at(State& s, int pv, const void* d, double t) {
  if (pv == s.precision())
    s.pot->evaluateAsync(pv, s.deviceCoords(), d, t, true, nullptr, nullptr, s.stream());
  else
    s.pot->evaluateMixedAsync(s.deviceCoords(), d, t, true, nullptr, nullptr, s.stream());
}

// backtrackingLinesearch (linesearch.cpp:8-29) along step = tau0 * d, with d a
// device vector of precision pv.  Trials are evaluated at x + tau*d without
// materialising them; tau halves exactly like step *= 0.5.  firstTrial, when
// given, is the energy already evaluated at x + tau0*d; xTrial, when given,
// holds that point itself (written by a fused evaluation): if the first trial
// is accepted it becomes the coordinates instead of an axpy, and *usedTrial
// says so.  Returns the tau applied to x.
double backtrack(State& s, int pv, const void* d, double tau0, double gs, EnergyCache& ec, MinimiseStats& st,
                 bool* acceptedOut = nullptr, const double* firstTrial = nullptr,
                 std::shared_ptr<void>* xTrial = nullptr, bool* usedTrial = nullptr) {
  const double c = 0.5, shrink = 0.5;
  double t = -c * gs;
  double e0;
  if (ec.ok(s)) {
    e0 = ec.e;
    st.nEnergyReused++;
  } else {
    s.pot->evaluate(s.precision(), s.deviceCoords(), nullptr, 0.0, &e0, nullptr, s.stream());
    st.nEnergy++;
  }
  double tau = tau0;
  bool accepted = false, fromTrial = false;
  double e = 0;
  for (int i = 0; i < 10; i++) {
    if (i == 0 && firstTrial) {
      e = *firstTrial;
    } else {
      eval_at(s, pv, d, tau, &e, nullptr);
      st.nEnergy++;
    }
    if (e0 - e >= t) {
      accepted = true;
      fromTrial = i == 0 && firstTrial && xTrial;
      break;
    }
    tau = tau * shrink;
    t = t * shrink;
  }
  if (fromTrial)
    s.swapDeviceCoords(*xTrial);  // x + tau0*d, bit for bit what the axpy below computes
  else
    minim_amd::launch_axpy_mixed(s.precision(), s.deviceCoords(), pv, d, tau, s.ndof, hs(s));
  if (usedTrial) *usedTrial = fromTrial;
  s.invalidateHost();
  if (accepted) {
    ec.set(s, e);
    st.energy = e;
  } else {
    ec.valid = false;  // the reference's unevaluated fallback point
  }
  if (acceptedOut) *acceptedOut = accepted;
  return tau;
}

void gradient_into(State& s, int pv, void* g, MinimiseStats& st) {
  eval_at(s, pv, nullptr, 0.0, nullptr, g);
  st.nGradient++;
}

double dot1(int prec, const void* a, const void* b, size_t n, hipStream_t st) {
  auto& ws = Workspace::current();
  const void* pa[1] = {a};
  const void* pb[1] = {b};
  minim_amd::launch_dots(prec, pa, pb, 1, n, ws.rs(), st);
  double r;
  ws.fetch(st, 1, &r);
  return r;
}

}  // namespace

// ------------------------------------------------------------------ L-BFGS
struct Lbfgs::Impl {
  int prec = FP64;  // precision of every vector below (g, D, Y)
  size_t n = 0;
  int m = 5;
  DVec g, gnew;
  std::vector<DVec> D, Y;  // m+1 slots: directions d (s = tau*d) and y
  struct Pair {
    int slot;
    double tau;
    double rho;
  };
  std::deque<Pair> pairs;  // oldest .. newest
  std::vector<char> used;
  // Gram data indexed by slot
  double gg = 0;
  std::vector<double> dg, yg;                  // d_a.g, y_a.g (current g)
  std::vector<std::vector<double>> dy, yy;     // d_a.y_b, y_a.y_b
  EnergyCache ec;
  bool fresh = true;  // next iteration starts from a new gradient
  int stall = 0;
  std::shared_ptr<void> xtrial;  // coordinates-sized buffer for the fused first trial
  int nbDir = 0, nbGram = 0;     // launch sizes of the two passes (0: the heuristic)

  // Empty history; device buffers are kept from the previous call when they
  // are large enough and of the same precision (a minimise on the same or a
  // smaller problem allocates nothing; vector lengths come from n, not from
  // the buffer sizes).
  void alloc(int prec_, size_t n_, int m_) {
    const bool reuse = prec_ == prec && m_ == m && (int)D.size() == m_ + 1 && g.n >= n_ && g.prec == prec_;
    prec = prec_;
    n = n_;
    m = m_;
    if (!reuse) {
      D.clear();
      Y.clear();
      g = DVec();
      gnew = DVec();
      g = DVec(n, prec);
      gnew = DVec(n, prec);
      for (int i = 0; i <= m; i++) {
        D.emplace_back(n, prec);
        Y.emplace_back(n, prec);
      }
    }
    used.assign(m + 1, 0);
    dg.assign(m + 1, 0.0);
    yg.assign(m + 1, 0.0);
    dy.assign(m + 1, std::vector<double>(m + 1, 0.0));
    yy.assign(m + 1, std::vector<double>(m + 1, 0.0));
    pairs.clear();
    fresh = true;
    stall = 0;
    ec.valid = false;
  }

  int spare() const {
    for (int i = 0; i <= m; i++)
      if (!used[i]) return i;
    throw std::logic_error("Lbfgs: no free history slot");
  }

  // Promote every device vector to FP64 in place (exact widening); the Gram
  // data, accumulated in FP64 from the same values, stays valid.
  void promote(hipStream_t s) {
    if (prec == FP64) return;
    auto widen = [&](DVec& v) {
      DVec w(n, FP64);
      minim_amd::launch_convert(FP64, w.data(), FP32, v.data(), n, s);
      MINIM_HIP_CHECK(hipStreamSynchronize(s));
      v = std::move(w);
    };
    widen(g);
    gnew = DVec();
    gnew = DVec(n, FP64);
    for (auto& v : D) widen(v);
    for (auto& v : Y) widen(v);
    prec = FP64;
  }
};

Lbfgs::Lbfgs() : impl_(new Impl) {}
Lbfgs::Lbfgs(const Lbfgs& o)
    : NewMinimiser<Lbfgs>(o), _m(o._m), _maxStep(o._maxStep), _initStep(o._initStep), _precMode(o._precMode),
      _switchRms(o._switchRms), _batchControl(o._batchControl), _batchBudget(o._batchBudget),
      _batchBlock(o._batchBlock), impl_(new Impl) {}
Lbfgs& Lbfgs::operator=(const Lbfgs& o) {
  Minimiser::operator=(o);
  _m = o._m;
  _maxStep = o._maxStep;
  _initStep = o._initStep;
  _precMode = o._precMode;
  _switchRms = o._switchRms;
  _batchControl = o._batchControl;
  _batchBudget = o._batchBudget;
  _batchBlock = o._batchBlock;
  impl_.reset(new Impl);
  return *this;
}
Lbfgs::~Lbfgs() = default;

Lbfgs& Lbfgs::setM(int m) {
  if (m < 1 || m > minim_amd::kMaxM) throw std::invalid_argument("Lbfgs: m must be between 1 and 15.");
  _m = m;
  return *this;
}
Lbfgs& Lbfgs::setMaxIter(int mi) {
  Minimiser::setMaxIter(mi);
  return *this;
}
Lbfgs& Lbfgs::setMaxStep(double v) {
  _maxStep = v;
  return *this;
}
Lbfgs& Lbfgs::setInitStep(double v) {
  _initStep = v;
  return *this;
}
Lbfgs& Lbfgs::setBatchControl(std::string mode) {
  if (mode != "graph" && mode != "device" && mode != "host" && mode != "resident")
    throw std::invalid_argument("Lbfgs: batch control must be graph, device, host or resident.");
  _batchControl = mode;
  return *this;
}
Lbfgs& Lbfgs::setBatchBlock(int e) {
  if (e != 0 && e != 4 && e != 8 && e != 16)
    throw std::invalid_argument("Lbfgs: the batch block must be 4, 8 or 16 elements per thread (or 0).");
  _batchBlock = e;
  return *this;
}
Lbfgs& Lbfgs::setBatchBudget(long sites) {
  if (sites < 0) throw std::invalid_argument("Lbfgs: the batch budget must be non-negative.");
  _batchBudget = sites;
  return *this;
}

Lbfgs& Lbfgs::setPrecision(std::string mode, double switchRms) {
  if (mode != "state" && mode != "fp64" && mode != "fp32" && mode != "mixed")
    throw std::invalid_argument("Lbfgs: precision must be state, fp64, fp32 or mixed.");
  _precMode = mode;
  _switchRms = switchRms;
  return *this;
}

namespace {
// Cache keys of the vector passes' launch sizes: the pass, the vector length,
// the sites' component count and the number of stored pairs.
std::string vec_key(const char* pass, size_t n, int nf, int m, int prec) {
  return minim_amd::tune_key(pass, {(long)n, (long)nf, (long)m}, prec);
}
}  // namespace

// Lbfgs::init (Lbfgs.cpp:34-38): empty history.
void Lbfgs::init(State& state) {
  if (_precMode == "fp64" || _precMode == "mixed") state.setPrecision(_precMode == "mixed" ? FP32 : FP64);
  if (_precMode == "fp32") state.setPrecision(FP32);
  Impl& I = *impl_;
  I.alloc(state.precision(), state.ndof, _m);
  // launch sizes an autotuner chose for vectors of this shape (Lbfgs::autotune)
  const int nf = state.pot->projection().nf;
  std::vector<int> v;
  I.nbDir = minim_amd::tune_lookup(vec_key("lbfgs-dir", I.n, nf, _m, I.prec), v) && v.size() == 1 ? v[0] : 0;
  I.nbGram = minim_amd::tune_lookup(vec_key("lbfgs-gram", I.n, nf, _m, I.prec), v) && v.size() == 1 ? v[0] : 0;
}

// The vector passes' launch sizes timed on this State's problem (the
// direction pass and the Gram pass with m stored pairs, the steady state of
// an iteration), with the minimiser's own history buffers; the fastest of
// each is kept and stored in the process-wide cache, where the next
// minimisation of the same shape finds it.  The launch size changes which
// block sums which elements: results agree to the order of the FP64 sums.
std::vector<std::array<double, 3>> Lbfgs::autotune(State& state, std::vector<int> blocks, int reps) {
  if (reps < 1 || blocks.empty()) throw std::invalid_argument("Lbfgs::autotune: nothing to time");
  for (int b : blocks)
    if (b < 1 || b > minim_amd::kMaxRedBlocks) throw std::invalid_argument("Lbfgs::autotune: block count out of range");
  init(state);
  Impl& I = *impl_;
  hipStream_t s = hs(state);
  auto& ws = Workspace::current();
  const size_t n = I.n;
  const int P = I.m;
  // any values: the coordinates widened or rounded into every vector
  auto fill = [&](DVec& v) { minim_amd::launch_convert(I.prec, v.data(), state.precision(), state.deviceCoords(), n, s); };
  fill(I.g);
  fill(I.gnew);
  for (auto& v : I.D) fill(v);
  for (auto& v : I.Y) fill(v);
  minim_amd::DirArgs a{};
  a.out = I.D[P].data();
  a.n = n;
  SiteProjection sp = state.pot->projection();
  a.proj = minim_amd::SiteProj{sp.nf, sp.fixmask, sp.hard, sp.nvar, sp.inv_nvar};
  a.basis[0] = I.g.data();
  a.coef[0] = -1.0;
  a.nb = 1;
  for (int q = 0; q < P; q++) {
    a.basis[a.nb] = I.D[q].data();
    a.coef[a.nb++] = 0.01;
  }
  for (int q = 0; q < P; q++) {
    a.basis[a.nb] = I.Y[q].data();
    a.coef[a.nb++] = 0.01;
  }
  a.dotv[0] = I.g.data();
  for (int q = 0; q < P; q++) a.dotv[1 + q] = I.Y[q].data();
  a.ndot = 1 + P;
  a.pairs = P;
  minim_amd::GramArgs ga{};
  ga.gnew = I.gnew.data();
  ga.gold = I.g.data();
  ga.y = I.Y[P].data();
  ga.n = n;
  for (int q = 0; q <= P; q++) ga.D[q] = I.D[q].data();
  ga.nd = P + 1;
  for (int q = 0; q < P; q++) ga.Y[q] = I.Y[q].data();
  ga.ny = P;
  hipEvent_t e0, e1;
  MINIM_HIP_CHECK(hipEventCreate(&e0));
  MINIM_HIP_CHECK(hipEventCreate(&e1));
  auto timeit = [&](auto&& launch) {
    launch();
    launch();
    MINIM_HIP_CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++) launch();
    MINIM_HIP_CHECK(hipEventRecord(e1, s));
    MINIM_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    MINIM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    return (double)ms / reps;
  };
  std::vector<std::array<double, 3>> rows;  // (pass 0 direction / 1 Gram, blocks, ms)
  double bestD = 0, bestG = 0;
  int nbD = 0, nbG = 0;
  for (int b : blocks) {
    a.blocks = b;
    const double md = timeit([&] { minim_amd::launch_direction(a, I.prec, ws.rs(), s); });
    ga.blocks = b;
    const double mg = timeit([&] { minim_amd::launch_gram(ga, I.prec, ws.rs(), s); });
    rows.push_back({0.0, (double)b, md});
    rows.push_back({1.0, (double)b, mg});
    if (nbD == 0 || md < bestD) bestD = md, nbD = b;
    if (nbG == 0 || mg < bestG) bestG = mg, nbG = b;
  }
  MINIM_HIP_CHECK(hipEventDestroy(e0));
  MINIM_HIP_CHECK(hipEventDestroy(e1));
  I.nbDir = nbD;
  I.nbGram = nbG;
  minim_amd::tune_store(vec_key("lbfgs-dir", n, sp.nf, _m, I.prec), {nbD});
  minim_amd::tune_store(vec_key("lbfgs-gram", n, sp.nf, _m, I.prec), {nbG});
  I.alloc(I.prec, n, _m);  // an empty history again
  return rows;
}

std::pair<int, int> Lbfgs::tunedBlocks() const { return {impl_->nbDir, impl_->nbGram}; }

// Lbfgs::iteration (Lbfgs.cpp:41-81) with getDirection (:84-127).
void Lbfgs::iteration(State& state) {
  Impl& I = *impl_;
  hipStream_t s = hs(state);
  auto& ws = Workspace::current();
  const int prec = I.prec;
  const size_t n = I.n;

  if (I.fresh) {
    gradient_into(state, prec, I.g.data(), stats);
    I.gg = dot1(prec, I.g.data(), I.g.data(), n, s);
    if (iter == 0) I.pairs.clear();
    // new gradient: refresh its dots with every stored pair
    if (!I.pairs.empty()) {
      minim_amd::DirArgs a{};
      a.basis[0] = I.g.data();
      a.coef[0] = 1.0;
      a.nb = 1;
      int k = 0;
      for (auto& p : I.pairs) a.dotv[k++] = I.D[p.slot].data();
      for (auto& p : I.pairs) a.dotv[k++] = I.Y[p.slot].data();
      a.ndot = k;
      a.out = I.gnew.data();
      a.n = n;
      a.proj = minim_amd::SiteProj{1, 0, 0, 1, 1.0};
      minim_amd::launch_direction(a, prec, ws.rs(), s);
      std::vector<double> r(1 + k);
      ws.fetch(s, 1 + k, r.data());
      const int P = (int)I.pairs.size();
      for (int i = 0; i < P; i++) {
        I.dg[I.pairs[i].slot] = r[1 + i];
        I.yg[I.pairs[i].slot] = r[1 + P + i];
      }
    }
    I.fresh = false;
  }

  // ---- direction: coefficients from the two-loop recursion on the Gram matrix
  const int P = (int)I.pairs.size();
  const int dslot = I.spare();
  std::vector<int> oldSlots(P);
  for (int i = 0; i < P; i++) oldSlots[i] = I.pairs[i].slot;
  minim_amd::DirArgs a{};
  a.out = I.D[dslot].data();
  a.n = n;
  a.blocks = I.nbDir;
  {
    SiteProjection sp = state.pot->projection();
    a.proj = minim_amd::SiteProj{sp.nf, sp.fixmask, sp.hard, sp.nvar, sp.inv_nvar};
  }
  a.dotv[0] = I.g.data();
  for (int i = 0; i < P; i++) a.dotv[1 + i] = I.Y[I.pairs[i].slot].data();
  a.ndot = 1 + P;
  a.pairs = P;  // basis = [g, d_0.., y_0..], dotv = [g, y_0..]
  a.basis[0] = I.g.data();
  a.nb = 1;
  if (P == 0) {
    // first step: -g * initStep/|g| (Lbfgs.cpp:91-96)
    double gnorm = std::sqrt(I.gg);
    a.coef[0] = (gnorm > 0) ? -(_initStep / gnorm) : -1.0;
  } else {
    std::vector<double> alpha(P), ag_s(P, 0.0), ay(P, 0.0);
    double ag = -1.0;  // q = -g
    auto sdotg = [&](int i) { return I.pairs[i].tau * I.dg[I.pairs[i].slot]; };
    auto sdoty = [&](int i, int j) { return I.pairs[i].tau * I.dy[I.pairs[i].slot][I.pairs[j].slot]; };
    auto ydoty = [&](int i, int j) { return I.yy[I.pairs[i].slot][I.pairs[j].slot]; };
    for (int k = P - 1; k >= 0; k--) {  // newest -> oldest
      double sq = ag * sdotg(k);
      for (int q = 0; q < P; q++) sq += ay[q] * sdoty(k, q);
      alpha[k] = I.pairs[k].rho * sq;
      ay[k] -= alpha[k];
    }
    const int nw = P - 1;
    double gamma = 1 / (I.pairs[nw].rho * ydoty(nw, nw));
    ag *= gamma;
    for (int q = 0; q < P; q++) ay[q] *= gamma;
    for (int k = 0; k < P; k++) {  // oldest -> newest
      double yr = ag * I.yg[I.pairs[k].slot];
      for (int q = 0; q < P; q++) yr += ag_s[q] * sdoty(q, k) + ay[q] * ydoty(k, q);
      double beta = I.pairs[k].rho * yr;
      ag_s[k] += alpha[k] - beta;
    }
    a.coef[0] = ag;
    for (int q = 0; q < P; q++) {
      a.basis[a.nb] = I.D[I.pairs[q].slot].data();
      a.coef[a.nb++] = ag_s[q] * I.pairs[q].tau;
    }
    for (int q = 0; q < P; q++) {
      a.basis[a.nb] = I.Y[I.pairs[q].slot].data();
      a.coef[a.nb++] = ay[q];
    }
  }
  // Constraints that couple all DOFs (fixed volumes) act on the step too
  // (Lbfgs.cpp:51); d lies in their subspace up to rounding, so the direction
  // pass's dots stay valid.  The projection leaves result[0, 80) untouched.
  // Normal constraints and a host potential's constraints can move d: then
  // the dots are taken again after the projection, except d.d, which the
  // maxStep cap reads before it (Lbfgs.cpp:121-124 precede :51).
  const bool moves = state.pot->constraintsMoveSteps();
  // Without a maxStep cap the line search starts at x + d (tau = 1) unless d
  // must be flipped (rare).  That trial energy is launched right after the
  // direction pass, whose dots then go to result slots [48, 48 + 1 + ndot)
  // (the trial's energy takes slot 0), so that one host synchronisation
  // returns both.  When the potential can, that trial is the fused step:
  // energy and gradient at x + d, and x + d itself into a spare buffer, in one
  // pass; if the trial is accepted the spare becomes the coordinates and the
  // gradient is there.
  const bool spec = linesearch == "backtracking" && (P == 0 || _maxStep == 0) && !moves;
  const bool fused = spec && state.pot->supportsStep();
  const bool defer = fused && state.pot->stepDefersGradient();
  constexpr int kStash = 48;
  minim_amd::RedScratch rd = ws.rs();
  if (spec) rd.result += kStash;
  minim_amd::launch_direction(a, prec, rd, s);
  std::vector<double> dres(1 + a.ndot);
  double ddPre = 0;
  if (moves) {
    ws.fetch(s, 1 + a.ndot, dres.data());
    ddPre = dres[0];
  }
  if (state.pot->hasGlobalConstraints())
    state.pot->constrainDevice(prec, prec == state.precision() ? state.deviceCoords() : nullptr, a.out, n, s);
  if (moves) {
    // dres = [d.d, d.g, d.y_0, ...] of the projected d, up to 8 dots a launch
    std::vector<const void*> pb(1 + a.ndot);
    pb[0] = a.out;
    for (int i = 0; i < a.ndot; i++) pb[1 + i] = a.dotv[i];
    const void* pa[8] = {a.out, a.out, a.out, a.out, a.out, a.out, a.out, a.out};
    for (int j0 = 0; j0 <= a.ndot; j0 += 8) {
      const int k = std::min(8, a.ndot + 1 - j0);
      minim_amd::launch_dots(prec, pa, pb.data() + j0, k, n, ws.rs(), s);
      ws.fetch(s, k, dres.data() + j0);
    }
  }
  double eSpec = 0;
  if (spec) {
    if (fused) {
      const auto* xt = static_cast<const DVec*>(I.xtrial.get());
      if (!xt || xt->n != n || xt->prec != state.precision()) {
        I.xtrial = state.newCoordsBuffer();
        xt = static_cast<const DVec*>(I.xtrial.get());
      }
      state.pot->evaluateStepAsync(state.precision(), prec, state.deviceCoords(), a.out, 1.0, xt->data(),
                                   defer ? nullptr : I.gnew.data(), s);
    } else {
      energy_async_at(state, prec, a.out, 1.0);
    }
    std::vector<double> r(kStash + 1 + a.ndot);
    ws.fetch(s, kStash + 1 + a.ndot, r.data());
    eSpec = r[0];
    for (int i = 0; i <= a.ndot; i++) dres[i] = r[kStash + i];
    stats.nEnergy++;
  } else if (!moves) {
    ws.fetch(s, 1 + a.ndot, dres.data());
  }
  const double dd = moves ? ddPre : dres[0], dgv = dres[1];

  double tau0 = 1.0;
  if (P > 0 && _maxStep != 0) {  // Lbfgs.cpp:121-124
    double ss = std::sqrt(dd);
    if (ss > _maxStep) tau0 = _maxStep / ss;
  }
  double gs = tau0 * dgv;  // Lbfgs.cpp:53-57
  if (gs > 0) {
    gs = -gs;
    tau0 = -tau0;
  }

  double tau;
  bool lsFailed = false, haveGradient = false;
  if (linesearch == "backtracking") {
    bool acc = true;
    const bool first = spec && tau0 == 1.0;
    tau = backtrack(state, prec, a.out, tau0, gs, I.ec, stats, &acc, first ? &eSpec : nullptr,
                    (first && fused) ? &I.xtrial : nullptr, &haveGradient);
    lsFailed = !acc;
    if (defer) haveGradient = false;  // the gradient at the accepted trial is still to come
  } else {
    minim_amd::launch_axpy_mixed(state.precision(), state.deviceCoords(), prec, a.out, tau0, n, s);
    state.invalidateHost();
    I.ec.valid = false;
    tau = tau0;
  }

  // ---- new gradient and Gram update (Lbfgs.cpp:67-80)
  if (haveGradient)
    stats.nGradient++;  // computed with the accepted trial point
  else
    gradient_into(state, prec, I.gnew.data(), stats);
  minim_amd::GramArgs ga{};
  ga.gnew = I.gnew.data();
  ga.gold = I.g.data();
  ga.y = I.Y[dslot].data();
  ga.n = n;
  ga.blocks = I.nbGram;
  for (int i = 0; i < P; i++) ga.D[i] = I.D[I.pairs[i].slot].data();
  ga.D[P] = I.D[dslot].data();
  ga.nd = P + 1;
  for (int i = 0; i < P; i++) ga.Y[i] = I.Y[I.pairs[i].slot].data();
  ga.ny = P;
  minim_amd::launch_gram(ga, prec, ws.rs(), s);
  const int KB = minim_amd::gram_kb(ga.nd, ga.ny);
  std::vector<double> r(4 * KB + 3);
  ws.fetch(s, 4 * KB + 3, r.data());
  auto gD = [&](int j) { return r[j]; };
  auto gY = [&](int j) { return r[KB + j]; };
  const double gng = r[2 * KB], gny = r[2 * KB + 1];
  auto yD = [&](int j) { return r[2 * KB + 2 + j]; };
  auto yY = [&](int j) { return r[3 * KB + 2 + j]; };
  const double yyv = r[4 * KB + 2];

  const double sy = tau * yD(P);
  if (sy != 0) {
    // store the pair (slot dslot); d_new.y_old comes from the direction pass
    for (int i = 0; i < P; i++) {
      const int sl = I.pairs[i].slot;
      I.dy[dslot][sl] = dres[2 + i];
      I.dy[sl][dslot] = yD(i);
      I.yy[dslot][sl] = yY(i);
      I.yy[sl][dslot] = yY(i);
    }
    I.dy[dslot][dslot] = yD(P);
    I.yy[dslot][dslot] = yyv;
    I.dg[dslot] = gD(P);
    I.yg[dslot] = gny;
    I.used[dslot] = 1;
    I.pairs.push_back({dslot, tau, 1 / sy});
    if ((int)I.pairs.size() > I.m) {
      I.used[I.pairs.front().slot] = 0;
      I.pairs.pop_front();
    }
  }
  // dots of the old pairs with the new gradient
  for (int i = 0; i < P; i++) {
    I.dg[oldSlots[i]] = gD(i);
    I.yg[oldSlots[i]] = gY(i);
  }
  I.gg = gng;
  std::swap(I.g, I.gnew);

  // ---- mixed precision: leave FP32 coordinates once they have done their job
  if (state.precision() == FP32 && _precMode == "mixed") {
    double rms = std::sqrt(I.gg / n);
    double sw = _switchRms > 0 ? _switchRms : 1e3 * state.convergence;
    if (lsFailed) I.stall++;
    else I.stall = 0;
    if (rms < sw || I.stall >= 2) {
      // x widens exactly to FP64; the vectors stay FP32 when the potential
      // evaluates FP64 coordinates against them, else widen too.  Both keep
      // the Gram data valid.
      if (!state.pot->supportsMixed()) I.promote(s);
      state.setPrecision(FP64);
      I.ec.valid = false;
      I.fresh = true;  // recompute g from FP64 coordinates at the same point; history kept
      stats.iterFP32++;
      return;
    }
  }
  if (state.precision() == FP32) stats.iterFP32++;
  else stats.iterFP64++;
}

// Lbfgs::checkConvergence (Lbfgs.cpp:130-133)
bool Lbfgs::checkConvergence(const State& state) {
  const Impl& I = *impl_;
  double rms = std::sqrt(I.gg / state.ndof);
  stats.gradRms = rms;
  if (I.fresh) return false;  // a precision switch is pending
  return rms < state.convergence;
}

void Lbfgs::finish(State& state) {
  if (!impl_->ec.ok(state)) {
    double e = 0;
    state.pot->evaluate(state.precision(), state.deviceCoords(), nullptr, 0.0, &e, nullptr, state.stream());
    stats.energy = e;
  }
}

// ------------------------------------------------------------------ FIRE
struct Fire::Impl {
  int prec = FP64;
  size_t n = 0;
  DVec g, v, step;
  double gNorm = 0, gMax = 0, a = 0.1, dt = 0;
  int nSteps = 0;
  EnergyCache ec;
  // v.g and v.v for the next iteration, fetched with the previous one's g.g
  bool haveDots = false;
  double vg = 0, vv = 0;
};

Fire::Fire() : impl_(new Impl) {}
Fire::Fire(double dtMax_) : dtMax(dtMax_), impl_(new Impl) { linesearch = "none"; }  // Fire.cpp:11-13
Fire::Fire(const Fire& o)
    : NewMinimiser<Fire>(o), dtMax(o.dtMax), _dt0(o._dt0), _maxForceDim(o._maxForceDim), _nMin(o._nMin),
      _fInc(o._fInc), _fDec(o._fDec), _fA(o._fA), _aStart(o._aStart), impl_(new Impl) {}
Fire& Fire::operator=(const Fire& o) {
  Minimiser::operator=(o);
  dtMax = o.dtMax;
  _dt0 = o._dt0;
  _maxForceDim = o._maxForceDim;
  impl_.reset(new Impl);
  return *this;
}
Fire::~Fire() = default;
Fire& Fire::setMaxIter(int mi) {
  Minimiser::setMaxIter(mi);
  return *this;
}
Fire& Fire::setDtMax(double d) {
  dtMax = d;
  return *this;
}
Fire& Fire::setDt0(double d) {
  if (d < 0) throw std::invalid_argument("Fire: dt0 must not be negative.");
  _dt0 = d;
  return *this;
}
Fire& Fire::setMaxForceConvergence(int dim) {
  if (dim < 0) throw std::invalid_argument("Fire: the force group size must not be negative.");
  _maxForceDim = dim;
  return *this;
}

namespace {
double group_norm_max(int prec, const void* v, size_t n, int dim, hipStream_t s) {
  auto& ws = Workspace::current();
  minim_amd::launch_group_norm_max(prec, v, n, dim, ws.rs(), s);
  double r;
  ws.fetch(s, 1, &r);
  return r;
}
}  // namespace

// Fire::init (Fire.cpp:27-38)
void Fire::init(State& state) {
  Impl& I = *impl_;
  I.prec = state.precision();
  I.n = state.ndof;
  I.g = DVec(I.n, I.prec);
  I.v = DVec(I.n, I.prec);
  I.step = DVec(I.n, I.prec);
  MINIM_HIP_CHECK(hipMemsetAsync(I.v.data(), 0, I.n * minim_amd::prec_bytes(I.prec), hs(state)));
  I.ec.valid = false;
  I.haveDots = false;
  I.a = _aStart;
  I.nSteps = 0;
  if (dtMax != 0) return;
  gradient_into(state, I.prec, I.g.data(), stats);
  I.gNorm = std::sqrt(dot1(I.prec, I.g.data(), I.g.data(), I.n, hs(state)));
  if (I.gNorm != 0) {
    dtMax = 0.1 / std::sqrt(I.gNorm);
  } else {
    std::printf("Warning: Unable to estimate dtMax parameter for FIRE.\n");
    dtMax = 1;
  }
}

// Fire::iteration (Fire.cpp:41-84)
void Fire::iteration(State& state) {
  Impl& I = *impl_;
  hipStream_t s = hs(state);
  auto& ws = Workspace::current();
  if (iter == 0) {
    I.dt = _dt0 > 0 ? _dt0 : dtMax;
    gradient_into(state, I.prec, I.g.data(), stats);
    I.gNorm = std::sqrt(dot1(I.prec, I.g.data(), I.g.data(), I.n, s));
  }
  double r[2];
  if (I.haveDots) {
    r[0] = I.vg;
    r[1] = I.vv;
  } else {
    const void* pa[2] = {I.v.data(), I.v.data()};
    const void* pb[2] = {I.g.data(), I.v.data()};
    minim_amd::launch_dots(I.prec, pa, pb, 2, I.n, ws.rs(), s);
    ws.fetch(s, 2, r);
  }
  I.haveDots = false;
  const double p = -r[0];
  // velocity update with the current a and dt (Fire.cpp:50-57) ...
  int mode;
  double f1 = 0, f2;
  if (p > 0) {
    double vNorm = std::sqrt(r[1]);
    mode = 1;
    f1 = 1 - I.a;
    f2 = I.a * vNorm / I.gNorm + I.dt;
    I.nSteps++;
  } else {
    mode = 0;
    f2 = -I.dt;
    I.nSteps = 0;
  }
  // ... then the dt / a adaptation (:60-67), and step = dt*v with the new dt (:70)
  if ((p > 0) && (I.nSteps > _nMin)) {
    I.a = I.a * _fA;
    I.dt = I.dt * _fInc;
    if (I.dt > dtMax) I.dt = dtMax;
  } else if (p <= 0) {
    I.a = _aStart;
    I.dt = I.dt * _fDec;
  }
  minim_amd::launch_fire_update(I.prec, mode, I.v.data(), I.g.data(), I.step.data(), f1, f2, I.dt, I.n, s);
  state.pot->constrainDevice(I.prec, I.prec == state.precision() ? state.deviceCoords() : nullptr, I.step.data(),
                             I.n, state.stream());
  if (linesearch == "backtracking") {
    double gs = dot1(I.prec, I.g.data(), I.step.data(), I.n, s);
    backtrack(state, I.prec, I.step.data(), 1.0, gs, I.ec, stats);
  } else {
    minim_amd::launch_add(I.prec, state.deviceCoords(), I.step.data(), I.n, s);
    state.invalidateHost();
    I.ec.valid = false;
  }
  // the new gradient (its validity check, if any, read with the dots below),
  // g.g now and the next iteration's v.g and v.v (v is already updated), with
  // the largest per-particle force, in one host synchronisation
  int check = state.pot->gradientDeferred(I.prec, state.deviceCoords(), I.g.data(), state.stream());
  stats.nGradient++;
  for (;;) {
    const void* pa[3] = {I.g.data(), I.v.data(), I.v.data()};
    const void* pb[3] = {I.g.data(), I.g.data(), I.v.data()};
    constexpr int kStash = 48;  // with the largest force in slot 0, the dots go to [48, 51)
    minim_amd::RedScratch rd = ws.rs();
    if (_maxForceDim > 0) rd.result += kStash;
    minim_amd::launch_dots(I.prec, pa, pb, 3, I.n, rd, s);
    std::array<double, minim_amd::kMaxDots> q;
    int nfetch = 3;
    if (_maxForceDim > 0) {
      minim_amd::launch_group_norm_max(I.prec, I.g.data(), I.n, _maxForceDim, ws.rs(), s);
      nfetch = kStash + 3;
    }
    if (check >= 0 && check + 1 > nfetch) nfetch = check + 1;
    ws.fetch(s, nfetch, q.data());
    if (check >= 0 && q[check] > 0) {  // the gradient was evaluated with a stale list
      state.pot->redoGradient(I.prec, state.deviceCoords(), I.g.data(), state.stream());
      check = -1;
      continue;
    }
    const double* d3 = _maxForceDim > 0 ? q.data() + kStash : q.data();
    I.gNorm = std::sqrt(d3[0]);
    I.vg = d3[1];
    I.vv = d3[2];
    I.haveDots = true;
    if (_maxForceDim > 0) I.gMax = q[0];
    break;
  }
  if (I.prec == FP32) stats.iterFP32++;
  else stats.iterFP64++;
}

// Fire::checkConvergence (Fire.cpp:87-90), or the largest per-particle force
bool Fire::checkConvergence(const State& state) {
  double rms = impl_->gNorm / std::sqrt((double)state.ndof);
  stats.gradRms = rms;
  if (_maxForceDim > 0) {
    stats.gradMax = impl_->gMax;
    return impl_->gMax < state.convergence;
  }
  return rms < state.convergence;
}

// ------------------------------------------------------------------ GradDescent
struct GradDescent::Impl {
  int prec = FP64;
  size_t n = 0;
  DVec g, step;
  double gg = 0;
  EnergyCache ec;
};

GradDescent::GradDescent() : impl_(new Impl) {}
GradDescent::GradDescent(const GradDescent& o) : NewMinimiser<GradDescent>(o), _alpha(o._alpha), impl_(new Impl) {}
GradDescent& GradDescent::operator=(const GradDescent& o) {
  Minimiser::operator=(o);
  _alpha = o._alpha;
  impl_.reset(new Impl);
  return *this;
}
GradDescent::~GradDescent() = default;
GradDescent& GradDescent::setAlpha(double al) {
  _alpha = al;
  return *this;
}
GradDescent& GradDescent::setMaxIter(int mi) {
  Minimiser::setMaxIter(mi);
  return *this;
}

void GradDescent::init(State& state) {
  Impl& I = *impl_;
  I.prec = state.precision();
  I.n = state.ndof;
  I.g = DVec(I.n, I.prec);
  I.step = DVec(I.n, I.prec);
  I.ec.valid = false;
}

// GradDescent::iteration (GradDescent.cpp:27-39)
void GradDescent::iteration(State& state) {
  Impl& I = *impl_;
  hipStream_t s = hs(state);
  gradient_into(state, I.prec, I.g.data(), stats);
  // step = -alpha * g (explicit vector, rounded as the reference's)
  minim_amd::launch_scale_copy(I.prec, I.step.data(), I.g.data(), -_alpha, I.n, s);
  const void* pa[2] = {I.g.data(), I.g.data()};
  const void* pb[2] = {I.step.data(), I.g.data()};
  auto& ws = Workspace::current();
  minim_amd::launch_dots(I.prec, pa, pb, 2, I.n, ws.rs(), s);
  double r[2];
  ws.fetch(s, 2, r);
  I.gg = r[1];
  if (linesearch == "backtracking") {
    backtrack(state, I.prec, I.step.data(), 1.0, r[0], I.ec, stats);
  } else {
    minim_amd::launch_add(I.prec, state.deviceCoords(), I.step.data(), I.n, s);
    state.invalidateHost();
    I.ec.valid = false;
  }
}

// GradDescent::checkConvergence (GradDescent.cpp:42-46)
bool GradDescent::checkConvergence(const State& state) {
  double rms = std::sqrt(impl_->gg / state.ndof);
  stats.gradRms = rms;
  return rms < state.convergence;
}

}  // namespace minim
