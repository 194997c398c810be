// Minimiser.h -- device-resident minimisers, mirroring the reference's
// Minimiser (include/Minimiser.h:13-33), Lbfgs (include/minimisers/Lbfgs.h:14-38),
// Fire (include/minimisers/Fire.h:9-36) and GradDescent
// (include/minimisers/GradDescent.h:9-23).
#ifndef MINIM_AMD_MINIMISER_H
#define MINIM_AMD_MINIMISER_H

#include <array>
#include <functional>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace minim {
class State;

// Counters and timings filled by every minimise() call.
struct MinimiseStats {
  long nEnergy = 0;      // energy-only passes (line-search trials)
  long nGradient = 0;    // gradient passes
  long nEnergyReused = 0;  // e0 values reused instead of re-evaluated
  int iterFP32 = 0;      // iterations run with FP32 storage (mixed schedule)
  int iterFP64 = 0;
  bool converged = false;
  double gradRms = 0;    // sqrt(g.g/ndof) of the last gradient
  double gradMax = 0;    // Fire with setMaxForceConvergence: largest per-particle force norm
  double seconds = 0;    // wall time of the minimisation (host clock); minimiseBatch: from the
                         // start of the call until this instance finished (device clock)
  double startSeconds = 0;  // minimiseBatch: from the start of the call until this instance was
                            // admitted (0 without a budget)
  double energy = 0;     // last accepted energy (FP64 accumulation)
  // Wall time (s) and iteration at which the convergence measure (gradRms,
  // or gradMax for Fire::setMaxForceConvergence) first fell below
  // 10^-(k+1), k = 0..7; -1 where it never did.
  double crossSeconds[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
  int crossIter[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
};

class Minimiser {
 public:
  int maxIter = 100000;
  std::string linesearch = "backtracking";
  int iter = 0;

  virtual ~Minimiser() = default;
  virtual std::unique_ptr<Minimiser> clone() const = 0;

  virtual Minimiser& setMaxIter(int maxIter);
  Minimiser& setLinesearch(std::string method);

  std::vector<double> minimise(State& state, std::function<void(int, State&)> adjustState = nullptr);
  std::vector<double> minimise(State& state, std::string logType);
  // minimise() without the final download: the result stays in the State's
  // device coordinates.
  void run(State& state, std::function<void(int, State&)> adjustState = nullptr);

  virtual void init(State& state) {}
  virtual void iteration(State& state) = 0;
  virtual bool checkConvergence(const State& state) { return false; }
  virtual void finish(State& state) {}

  MinimiseStats stats;
};

template <typename Derived>
class NewMinimiser : public Minimiser {
 public:
  std::unique_ptr<Minimiser> clone() const override {
    return std::make_unique<Derived>(static_cast<const Derived&>(*this));
  }
};

// L-BFGS with the reference's semantics (src/minimisers/Lbfgs.cpp:41-133 and
// src/minimisers/linesearch.cpp:8-29), evaluated as a Gram-matrix recursion
// over device vectors.
class Lbfgs : public NewMinimiser<Lbfgs> {
 public:
  Lbfgs();
  Lbfgs(const Lbfgs& o);
  Lbfgs& operator=(const Lbfgs& o);
  ~Lbfgs() override;

  Lbfgs& setM(int m);
  Lbfgs& setMaxIter(int maxIter) override;
  Lbfgs& setMaxStep(double maxStep);
  Lbfgs& setInitStep(double initStep);
  // "state" (use the State's precision), "fp64", "fp32" or "mixed": FP32
  // storage until sqrt(g.g/ndof) < switchRms (or the FP32 line search stalls),
  // then FP64 coordinates and arithmetic to the requested convergence.  In the
  // second phase the gradient, direction and history vectors stay in FP32 when
  // the potential supportsMixed(), else they are widened to FP64.
  Lbfgs& setPrecision(std::string mode, double switchRms = 0);

  void init(State& state) override;
  void iteration(State& state) override;
  bool checkConvergence(const State& state) override;
  void finish(State& state) override;

  int m() const { return _m; }

  // Launch sizes of the direction and Gram passes timed on this State's
  // problem (m stored pairs, the minimiser's own buffers, HIP events, the
  // average of reps per candidate block count); the fastest of each is kept
  // and stored in the process-wide cache, where later minimisations of the
  // same shape find it.  Returns rows (pass 0 direction / 1 Gram, blocks, ms).
  std::vector<std::array<double, 3>> autotune(State& state, std::vector<int> blocks = {256, 512, 1024, 2048, 4096},
                                              int reps = 10);
  // (direction, Gram) launch sizes in use; 0 = the heuristic.
  std::pair<int, int> tunedBlocks() const;

  // Independent minimisations of many States at once, each with the
  // semantics of minimise(state) under this minimiser's settings: one launch
  // per phase (direction, line-search trials, step with gradient and Gram
  // update) covers every instance still running.  Every State must hold an
  // initialised one-fluid
  // PhaseField without solid nodes, extra terms or fixed fluid
  // (PhaseField::plainStencil), in FP64 ("state" or "fp64" precision), m <= 10
  // and no adjustState callback.  The result is left in each State's
  // coordinates; one MinimiseStats per State, in order.  The reference runs
  // such instances as States on disjoint rank subsets (include/State.h:23,
  // src/Communicator.cpp:219-240).
  std::vector<MinimiseStats> minimiseBatch(const std::vector<State*>& states);
  // How minimiseBatch drives the iterations.  "device" (the default): the
  // per-instance decisions (two-loop recursion, Armijo tests, Gram update,
  // convergence) run as device kernels, one iteration is a fixed sequence of
  // launches with per-instance phase masks, and the host reads the live
  // flags every 8 iterations; "graph": the same sequence captured once as a
  // HIP graph and replayed (measured 2-6 % slower than launching it);
  // "host": the decisions on the host, one host synchronisation per phase.
  // All three give the same results bit for bit.  "resident" (FP64 schedule only): one workgroup per
  // instance runs that instance's whole minimisation in a single launch, the
  // line search included, with no launches or host synchronisations between
  // iterations; its sums are split differently, so its results agree with
  // the others to rounding.
  Lbfgs& setBatchControl(std::string mode);
  // Under device control ("graph" or "device"): instances are admitted
  // smallest first while the running ones hold at most this many elements
  // (sites), until half of the instances have finished; then every remaining
  // one.  Small instances finish sooner, at some cost in total time.  0 (the
  // default): every instance from the start.  An instance's result does not
  // depend on it.
  Lbfgs& setBatchBudget(long sites);
  // Elements per thread of the batch's streaming passes (a block covers 256
  // times as many elements of one instance): 4, 8 or 16; 0 (the default) the
  // size autotuneBatch stored for this batch's shape, else 8 (4 below 2^20
  // elements in all, where the passes are bound by latency).  The block size
  // sets how an instance's sums are split over blocks, so results for
  // different sizes agree to rounding, not bit for bit.
  Lbfgs& setBatchBlock(int elementsPerThread);
  // Times the first `replays` iterations of minimiseBatch(states) for each
  // candidate block size (coordinates are read, not changed), keeps the
  // fastest in the process-wide cache for batches of the same shape (number
  // of instances, total elements, m, precision schedule) and returns rows
  // (elements per thread, seconds).
  std::vector<std::array<double, 2>> autotuneBatch(const std::vector<State*>& states,
                                                   std::vector<int> candidates = {4, 8, 16}, int replays = 32);
  struct BatchReport {
    std::string control;
    int iterations = 0;       // the most iterations any instance ran
    long replays = 0;         // iterations executed by the batch
    long hostSyncs = 0;       // host synchronisations waiting for results
    long graphLaunches = 0;   // HIP graph launches
    long captures = 0;        // graph captures (one more each time the grid shrank)
    int block = 0;            // elements per thread
  };
  BatchReport batchReport;  // of the last minimiseBatch

 private:
  int _m = 5;
  double _maxStep = 0;
  double _initStep = 1;
  std::string _precMode = "state";
  double _switchRms = 0;
  std::string _batchControl = "device";
  long _batchBudget = 0;
  int _batchBlock = 0;
  // minimiseBatch with a given block size; with replayLimit > 0 it stops after
  // that many iterations and, like writeBack false, leaves the States as
  // they were.
  std::vector<MinimiseStats> batchRun(const std::vector<State*>& states, int block, long replayLimit,
                                      bool writeBack);
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

class Fire : public NewMinimiser<Fire> {
 public:
  Fire();
  explicit Fire(double dtMax);
  Fire(const Fire& o);
  Fire& operator=(const Fire& o);
  ~Fire() override;

  double dtMax = 0;
  Fire& setMaxIter(int maxIter) override;
  Fire& setDtMax(double dtMax);
  // Options without a reference counterpart (both off reproduce the
  // reference): the initial time step (0 = dtMax, Fire.cpp:43), and
  // convergence on the largest per-particle force norm over groups of dim
  // DOFs (0 = RMS gradient, Fire.cpp:87-90).
  Fire& setDt0(double dt0);
  Fire& setMaxForceConvergence(int dim);

  void init(State& state) override;
  void iteration(State& state) override;
  bool checkConvergence(const State& state) override;

 private:
  double _dt0 = 0;
  int _maxForceDim = 0;
  int _nMin = 5;
  double _fInc = 1.1;
  double _fDec = 0.5;
  double _fA = 0.99;
  double _aStart = 0.1;
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

class GradDescent : public NewMinimiser<GradDescent> {
 public:
  GradDescent();
  GradDescent(const GradDescent& o);
  GradDescent& operator=(const GradDescent& o);
  ~GradDescent() override;

  GradDescent& setAlpha(double alpha);
  GradDescent& setMaxIter(int maxIter) override;

  void init(State& state) override;
  void iteration(State& state) override;
  bool checkConvergence(const State& state) override;

 private:
  double _alpha = 1e-1;
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

}  // namespace minim

#endif
