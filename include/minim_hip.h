/*
 * minim_hip.h -- C ABI of the MI355X-native minim hot path (libminim_amd.so).
 *
 * This is the drop-in boundary: plain C types, device pointers and sizes, no
 * HIP or torch types.  Each entry point names the reference interface
 * (sjavis/minim, file:line) whose role it takes over.  Every function returns
 * MINIM_OK or an error code; minim_last_error() gives the message.  The codes
 * map the reference's exceptions: std::invalid_argument -> MINIM_EINVAL,
 * std::runtime_error -> MINIM_ERUNTIME.
 *
 * Device vectors are flat arrays of double (MINIM_FP64) or float (MINIM_FP32)
 * in GPU memory.  Streams are hipStream_t passed as void* (NULL = default).
 * Calls that reduce (energies, dots, minimisers) use scratch memory private to
 * the calling host thread and device: independent problems may run
 * concurrently from several host threads, each on its own stream.
 */
#ifndef MINIM_HIP_H
#define MINIM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { MINIM_OK = 0, MINIM_EINVAL = 1, MINIM_ERUNTIME = 2, MINIM_EHIP = 3, MINIM_ENOMEM = 4 };
enum { MINIM_FP64 = 0, MINIM_FP32 = 1 };

/* Message of the last failed call on this thread. */
const char* minim_last_error(void);
int minim_abi_version(void);

typedef struct minim_potential minim_potential;
typedef struct minim_state minim_state;
typedef struct minim_minimiser minim_minimiser;

/* ------------------------------------------------------------------ potentials */

/* PhaseField setters (include/potentials/PhaseField.h:20-65, src/potentials/PhaseField.cpp:747-900).
 * n_interface_size: 0 = default (resolution), 1 = setInterfaceSize(double),
 * n_fluid = setInterfaceSize(vector).  density_constraint: 0 hard, 1 soft,
 * 2 none, 3 fixed (PhaseField.cpp:802-815). */
typedef struct minim_pf_params {
  int grid[3];
  int n_fluid;
  double resolution;
  const double* surface_tension;
  int n_surface_tension;
  const double* interface_size;
  int n_interface_size;
  int density_constraint;
  double density_const;
  const unsigned char* fix_fluid; /* NULL or n_fluid flags (setFixFluid) */
  int stencil_xchunk;             /* 0 = automatic x-march length */
  int stencil_target_blocks;      /* 0 = automatic */
  /* The remaining terms (PhaseField.h:36-68, PhaseField.cpp:818-895); NULL / 0
   * leaves a setter uncalled.  n_surface_tension may also be
   * ngrid * n_fluid: per-site tensions. */
  const unsigned char* solid;     /* ngrid flags: setSolid */
  const double* contact_angle;    /* setContactAngle(vector) */
  size_t n_contact_angle;
  const double* pressure;         /* n_fluid: setPressure */
  const double* force;            /* 3: setForce(force, force_fluids) */
  const int* force_fluids;
  int n_force_fluids;
  const double* confinement;      /* n_fluid: setConfinement */
  int volume_fixed;               /* setVolumeFixed(true, vol_const) ... */
  double vol_const;
  const double* volume;           /* ... or setVolume(volume, vol_const): n_fluid */
  int stencil_kernel;             /* 0 automatic, 1 the general (solid-aware) kernel */
} minim_pf_params;
int minim_pf_create(const minim_pf_params* p, minim_potential** out);

/* PhaseField::energyComponents (PhaseField.cpp:706-744) at host coordinates
 * (initialised potential): out[2c], out[2c+1] = energy and unconstrained
 * gradient norm of c = fluid, surface, pressure, density constraint, force,
 * confinement, volume constraint. */
int minim_pf_energy_components(const minim_potential* pot, const double* coords, size_t ndof, double* out);
/* nodeVol / surfaceArea after init (PhaseField.cpp:190-214), ngrid each (host). */
int minim_pf_node_volumes(const minim_potential* pot, double* node_vol, double* surface_area);
/* PhaseField::diffuseSolid(solid, iFluid, twoStep) (PhaseField.cpp:998-1035)
 * on a copy of the configured potential; out: ngrid * n_fluid host doubles. */
int minim_pf_diffuse_solid(const minim_potential* pot, const unsigned char* solid, int i_fluid, int two_step,
                           double* out);
/* PhaseField::getType (PhaseField.cpp:899-958) of every site: type[ngrid]
 * (-1 solid, 0 bulk, 1..7).  Host only. */
int minim_pf_site_types(const int grid[3], const unsigned char* solid, int* type);

/* kappa / kappaP per fluid after init (PhaseField::assignFluidCoefficients, PhaseField.cpp:78-134;
 * PhaseFieldUnstructured::assignFluidCoefficients, PhaseFieldUnstructured.cpp:71-107). */
int minim_pf_coefficients(const minim_potential* pot, double* kappa, double* kappa_p);

/* Launch-shape autotuner of a PhaseField's evaluation kernel (no reference
 * counterpart; the reference has no launch shapes).  For every pair of target
 * block count targets[i] and x-march length marches[j] (0 = automatic) one
 * energy+gradient launch of the kernel this field uses (the plain tiled
 * stencil, the general tiled kernel or the untiled kernels) is timed with HIP
 * events (average of reps after two warm-up launches) at the host coordinates
 * coords (ndof values), in device buffers of the library's own.  The fastest
 * shape is kept for this precision and stored in the process-wide cache
 * (appended to the file named by MINIM_AUTOTUNE_CACHE, when set), where every
 * later PhaseField of the same shape and kernel finds it at init.  times
 * (may be NULL) receives n_targets * n_marches milliseconds, row-major
 * (negative: a shape the grid cannot take).  The shape changes the tiling
 * only: gradients are the same bit for bit. */
int minim_pf_autotune(minim_potential* pot, int precision, const double* coords, size_t ndof, const int* targets,
                      int n_targets, const int* marches, int n_marches, int reps, double* times, int* best_target,
                      int* best_march);
/* The launch shape a PhaseField uses for a precision after autotuning (0, 0:
 * the heuristic). */
int minim_pf_tuned_shape(const minim_potential* pot, int precision, int* target, int* march);
/* Launch sizes of an L-BFGS minimiser's direction and Gram passes on this
 * State's problem (no reference counterpart): each candidate block count is
 * timed (HIP events, average of reps) with the minimiser's own history
 * buffers and m stored pairs; the fastest of each pass is kept and stored in
 * the process-wide cache.  times (may be NULL): 2 * n_blocks milliseconds,
 * direction then Gram per candidate. */
int minim_lbfgs_autotune(minim_minimiser* m, minim_state* s, const int* blocks, int n_blocks, int reps,
                         double* times, int* best_direction, int* best_gram);
/* Forget every launch shape the autotuners stored in this process. */
int minim_autotune_clear(void);

/* PhaseFieldUnstructured (include/potentials/PhaseFieldUnstructured.h:13-104):
 * the element form of the same physics, built from the same parameters
 * (PhaseFieldUnstructured.cpp:110-277, setters :727-856).  Differences from
 * minim_pf_create: surface tensions are 1 or n_fluid values; the stencil_*
 * fields are ignored; the fixed volume is an energy term, not a projection;
 * the soft density constraint needs at least three fluids (its element reads
 * three concentrations per site, :528-541).  minim_pf_coefficients,
 * minim_pf_node_volumes (node volumes only) and minim_pf_diffuse_solid accept
 * the handle; minim_pf_energy_components does not (the class has none). */
int minim_pfu_create(const minim_pf_params* p, minim_potential** out);
/* Number of energy elements init built (the size of the reference's
 * Potential::elements, include/Potential.h:90) and the device bytes of the
 * list as this library stores it (16-byte records, parameters, chunk and
 * site tables); either pointer may be NULL. */
int minim_pfu_element_list(const minim_potential* pot, size_t* n_elements, size_t* device_bytes);

/* LjNd(nDim, sigma, epsilon) (include/potentials/LjNd.h:12-44): all pairs, no cutoff. */
int minim_lj_create(int ndim, double sigma, double epsilon, minim_potential** out);
/* Truncated-and-shifted cutoff rc in a periodic box with a cell list (no
 * reference counterpart: the reference's LjNd has neither, LjNd.cpp:29-53).
 * ndim must be 3; skin is the neighbour-list margin (rebuild when a particle
 * has moved more than skin/2). */
int minim_lj_set_cutoff(minim_potential* pot, double rc, const double box[3], double skin);
/* Neighbour-list builds so far of the LjNd (cutoff mode) behind a State -- a
 * State holds its own copy of the potential (LjNd::listBuilds; no reference
 * counterpart). */
int minim_lj_list_builds(const minim_state* state, long* builds);
/* The longest neighbour list of the last build of the LjNd behind a State
 * (LjNd::maxNeighbours; 0 before the first build; waits for the device). */
int minim_lj_max_neighbours(const minim_state* state, int* longest);

/* BarAndHinge (BarAndHinge::setTriangulation / setBondList / setHingeList /
 * setRigidity / setLength0 / setTheta0 / setForce,
 * include/potentials/BarAndHinge.h:17-43).  Either a triangulation
 * (3 node ids per triangle; bars and hinges are derived in the reference's
 * order, BarAndHinge.cpp:70-147) or explicit lists (n_bar > 0 or n_hinge > 0:
 * 2 node ids per bar, 4 per hinge).  Per-element arrays may be NULL, in which
 * case the scalars apply; a NaN l0 / theta0 means "from the initial
 * coordinates" (computeLength0 / computeTheta0, BarAndHinge.cpp:213-248).
 * force: NULL, a 3-vector, or 3 values per node (force_per_node). */
typedef struct minim_bah_params {
  int n_node;
  const int* triangles;
  int n_triangle;
  int n_bar;
  const int* bars;
  int n_hinge;
  const int* hinges;
  const double* bar_k;
  const double* hinge_k;
  double k_bar;
  double k_hinge;
  const double* bar_l0;
  const double* hinge_theta0;
  double l0;
  double theta0;
  const double* force;
  int force_per_node;
} minim_bah_params;
int minim_bah_create(const minim_bah_params* p, minim_potential** out);

/* Potential::setConstraints(vector<int> iFix) (src/Potential.cpp:80-85). */
int minim_potential_set_fixed(minim_potential* pot, const int* idof, size_t n);
/* Potential::init / initLocal on the initial coordinates (include/Potential.h:62-63). */
int minim_potential_init(minim_potential* pot, const double* coords, size_t ndof);
/* Potential::convergence after init (include/Potential.h:25, PhaseField.cpp:146). */
int minim_potential_convergence(const minim_potential* pot, double* out);
int minim_potential_destroy(minim_potential* pot);

/* Energy and/or gradient at x (or x + t*d when d != NULL) on device vectors:
 * Potential::energyGradient followed, for the gradient, by
 * Potential::applyConstraints, as State::basicEG / elementEG do
 * (src/State.cpp:48-79; PhaseField.cpp:662-703; LjNd.cpp:29-53;
 * BarAndHinge.cpp:251-369).  e: host double or NULL; g: device or NULL.
 * A gradient together with d != NULL is not supported. */
int minim_potential_eval(const minim_potential* pot, int precision, const void* x, size_t ndof,
                         const void* d, double t, double* e, void* g, void* stream);
/* As minim_potential_eval, without a host synchronisation: the energy (when
 * want_energy) is written to the device double e_dev, in stream order. */
int minim_potential_eval_async(const minim_potential* pot, int precision, const void* x, size_t ndof,
                               const void* d, double t, int want_energy, double* e_dev, void* g, void* stream);
/* The allocation-free form of minim_potential_eval_async.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  PhaseField only (MINIM_EINVAL otherwise).  The cross-block
 * reduction runs in a fixed device array of this library, 32768 partial sums
 * plus 96 results (263 KB per GPU, held until the process ends), instead of
 * the per-thread workspace; every PhaseField evaluation fits in it (at most
 * 8192 blocks x 2 values, or 16 fluid sums x 2048 blocks for fixed volumes).
 * The potential's device tables must already be resident: evaluate it once
 * through an allocating entry point first, else MINIM_EINVAL.  Energy, when
 * want_energy, goes to the device double e_dev in stream order.  Every call
 * on one GPU shares that array: calls that overlap on one GPU (several
 * streams or host threads) race; serialise them. */
int minim_potential_eval_async_noalloc(const minim_potential* pot, int precision, const void* x, size_t ndof,
                                       const void* d, double t, int want_energy, double* e_dev, void* g,
                                       void* stream);
/* As minim_potential_eval with FP64 coordinates x and FP64 arithmetic, but the
 * direction d and the gradient g in FP32 (the second phase of the mixed L-BFGS
 * schedule).  MINIM_ERR_INVALID for potentials without mixed support
 * (PhaseField with nFluid <= 4 has it). */
int minim_potential_eval_mixed(const minim_potential* pot, const void* x, size_t ndof, const void* d, double t,
                               double* e, void* g, void* stream);
/* The fused accepted step of this library's L-BFGS line search (no
 * reference counterpart: it replaces the reference's separate trial energy,
 * `x += step` and gradient, linesearch.cpp:13-25 and Lbfgs.cpp:67): energy
 * (host, may be NULL) and constrained gradient g at x + t*d, with x + t*d
 * written to x_out (precision_x, like x; d and g in precision_v).
 * MINIM_EINVAL for potentials without it (PhaseField on the tiled kernel has
 * it: nFluid <= 4, no solid nodes or extra terms; BarAndHinge has it, with
 * precision_v != precision_x only without normal constraints).  Synchronises
 * the stream. */
int minim_potential_eval_step(const minim_potential* pot, int precision_x, int precision_v, const void* x,
                              size_t ndof, const void* d, double t, void* x_out, double* e, void* g, void* stream);
/* Benchmark support (no reference counterpart): average duration in ms,
 * measured with HIP events on stream, of reps back-to-back launches of the
 * evaluation kernels (energy when want_energy, gradient into g when g !=
 * NULL).  PhaseField: the stencil kernel without the cross-block reduction;
 * PhaseFieldUnstructured: the element kernel (and the gather of shared sites
 * with g); LjNd: the padding and force passes (cutoff mode) or the all-pairs
 * kernel with its reduction; BarAndHinge: the energy pass with its reduction
 * and the gradient pass.  MINIM_EINVAL for other potentials. */
int minim_potential_time_kernel(const minim_potential* pot, int precision, const void* x, size_t ndof,
                                int want_energy, void* g, int reps, void* stream, double* ms_per_launch);
/* Benchmark support (no reference counterpart): steps whole energy+gradient
 * evaluations at x (energy copied to the device double e_dev) on stream,
 * after untimed evaluations for warm_ms of host time.  HIP events bracket the
 * loop and each launch of the evaluation kernel inside it.  out[0] = ms per
 * evaluation over the loop; out[1..4] = median, min, max, mean of the kernel's
 * per-launch durations (ms).  PhaseField only; MINIM_EINVAL for the others. */
int minim_potential_time_steps(const minim_potential* pot, int precision, const void* x, size_t ndof, void* g,
                               double* e_dev, int steps, double warm_ms, void* stream, double* out);
/* Potential::applyConstraints on a device vector (src/Potential.cpp:102-115,
 * PhaseField::specialisedConstraints PhaseField.cpp:580-612). */
int minim_potential_constrain(const minim_potential* pot, int precision, void* v, size_t n, void* stream);

/* CommGrid::dotProduct / Communicator::dotProduct on one GPU
 * (src/communicators/CommGrid.cpp:187-194, src/Communicator.cpp:170-173): FP64 sum. */
int minim_dot(int precision, const void* a, const void* b, size_t n, double* out, void* stream);

/* ------------------------------------------------------------------ State */

/* State(pot, coords) (src/State.cpp:12-26): clones and initialises pot, uploads coords. */
int minim_state_create(const minim_potential* pot, const double* coords, size_t ndof, int precision,
                       minim_state** out);
/* State copy constructor (src/State.cpp:28-35). */
int minim_state_copy(const minim_state* s, minim_state** out);
int minim_state_destroy(minim_state* s);
/* State::energy() / energy(coords) (src/State.cpp:84-87, 134-152); coords NULL = own. */
int minim_state_energy(minim_state* s, const double* coords, double* e);
/* State::gradient() / gradient(coords) (src/State.cpp:89-92, 154-172) into host g. */
int minim_state_gradient(minim_state* s, const double* coords, double* g);
/* State::energyGradient (src/State.cpp:94-97, 174-192); e or g may be NULL. */
int minim_state_energy_gradient(minim_state* s, const double* coords, double* e, double* g);
/* State::coords() get / set (src/State.cpp:282-288). */
int minim_state_get_coords(minim_state* s, double* out);
int minim_state_set_coords(minim_state* s, const double* in);
/* State::componentEnergy (src/State.cpp:307-312): the energy of the
 * potential's elements of one type at the State's coordinates
 * (BarAndHinge: 0 bars, 1 hinges, 2 external force; LjNd: 0 pairs). */
int minim_state_component_energy(minim_state* s, int component, double* e);
/* State::applyConstraints on a host vector (src/State.cpp:321-323). */
int minim_state_apply_constraints(minim_state* s, double* data);
/* State::failed (src/State.cpp:326-328) and the public fields of include/State.h:15-19,72. */
int minim_state_failed(minim_state* s);
int minim_state_info(const minim_state* s, size_t* ndof, int* precision, double* convergence, int* is_failed);
int minim_state_set_convergence(minim_state* s, double convergence);
int minim_state_set_precision(minim_state* s, int precision);
int minim_state_set_stream(minim_state* s, void* stream);
/* Device address of the coordinates (ndof entries of the State's precision).
 * Valid until the next minimise or precision change on this state: L-BFGS's
 * fused accepted step exchanges the coordinates with a buffer of its own.
 * After writing through it, call minim_state_mark_coords_changed. */
void* minim_state_device_coords(minim_state* s);
/* Declares that the device coordinates were changed by the caller (no
 * reference counterpart: the reference's State::coords(in) setter,
 * src/State.cpp:286-288, is the only writer there): the host copy is
 * refreshed on its next read, and a minimiser that continues from this State
 * (e.g. after an adjustState callback) evaluates the energy afresh instead of
 * reusing the one it cached for the old coordinates. */
int minim_state_mark_coords_changed(minim_state* s);

/* ------------------------------------------------------------------ minimisers */

/* Lbfgs (include/minimisers/Lbfgs.h:14-38), Fire (include/minimisers/Fire.h:9-36),
 * GradDescent (include/minimisers/GradDescent.h:9-23) with
 * Minimiser::setMaxIter / setLinesearch (src/Minimiser.cpp:12-24). */
enum { MINIM_LBFGS = 0, MINIM_FIRE = 1, MINIM_GRADDESCENT = 2 };
enum { MINIM_LS_BACKTRACKING = 0, MINIM_LS_NONE = 1 };
enum { MINIM_PREC_STATE = 0, MINIM_PREC_FP64 = 1, MINIM_PREC_FP32 = 2, MINIM_PREC_MIXED = 3 };
typedef struct minim_minimiser_params {
  int kind;
  int max_iter;
  int linesearch;
  int m;            /* Lbfgs::setM */
  double max_step;  /* Lbfgs::setMaxStep */
  double init_step; /* Lbfgs::setInitStep */
  int precision_mode;
  double switch_rms; /* mixed: gradient RMS at which FP32 storage is left (0 = 1e3*convergence) */
  double dt_max;     /* Fire::setDtMax (0 = estimate, Fire.cpp:29-37) */
  double alpha;      /* GradDescent::setAlpha */
  double dt0;        /* Fire::setDt0: initial time step (0 = dt_max, the reference) */
  int max_force_dim; /* Fire::setMaxForceConvergence: converge on the largest norm over groups of
                        this many DOFs (0 = RMS gradient, the reference) */
  int batch_control; /* Lbfgs::setBatchControl for minim_minimise_batch: 0 device, 1 graph, 2 host, 3 resident */
  long batch_budget; /* Lbfgs::setBatchBudget: sites of running instances under device control (0 = all) */
  int batch_block;   /* Lbfgs::setBatchBlock: elements per thread of the batch passes, 4, 8 or 16 (0 = tuned, else 8, or 4 below 2^20 elements) */
} minim_minimiser_params;

typedef struct minim_result {
  int iterations; /* Minimiser::iter at exit */
  int converged;
  long n_energy;
  long n_gradient;
  long n_energy_reused;
  int iter_fp32;
  int iter_fp64;
  double energy;
  double grad_rms;
  double dt_max; /* Fire: dtMax after init */
  double grad_max; /* Fire with max_force_dim: largest per-group norm of the last gradient */
  double seconds;  /* wall time from the start of the call until this minimisation ended */
  double cross_seconds[8]; /* wall time and iteration at which the convergence measure (grad_rms, or */
  int cross_iter[8];       /* grad_max with max_force_dim) first fell below 10^-(k+1); -1: never */
  double start_seconds;    /* minim_minimise_batch: wall time from the start of the call until this
                              instance was admitted (0 without a budget); 0 for a single minimisation */
} minim_result;

typedef void (*minim_adjust_fn)(int iter, minim_state* state, void* user);

int minim_minimiser_create(const minim_minimiser_params* p, minim_minimiser** out);
int minim_minimiser_destroy(minim_minimiser* m);
/* Minimiser::minimise(state, adjustState) (src/Minimiser.cpp:27-38); coords_out
 * (host, ndof) receives state.coords().  With coords_out NULL nothing is
 * copied to the host: the result stays in the State's device coordinates. */
int minim_minimise(minim_minimiser* m, minim_state* s, minim_adjust_fn fn, void* user, double* coords_out,
                   minim_result* r);
/* Minimiser::minimise(state, "<fields>-<N>") (src/Minimiser.cpp:41-69). */
int minim_minimise_log(minim_minimiser* m, minim_state* s, const char* log_type, double* coords_out,
                       minim_result* r);
/* Independent L-BFGS minimisations of n_states States at once (Lbfgs::minimiseBatch):
 * each with the semantics of minim_minimise on its own State (Lbfgs.cpp:41-133,
 * linesearch.cpp:8-29, Minimiser.cpp:27-38), one launch per phase over every
 * instance still running (batch_control and batch_budget in the parameters).
 * Replaces the reference's States on disjoint rank subsets
 * (include/State.h:23, src/Communicator.cpp:219-240).  Every State holds an
 * initialised one-fluid PhaseField without solid nodes, extra terms or fixed
 * fluid, in FP64; m <= 10.  Results stay in each State's device coordinates;
 * results[i] (may be NULL) describes states[i], with start_seconds the time
 * at which it was admitted. */
int minim_minimise_batch(minim_minimiser* m, minim_state* const* states, int n_states, minim_result* results);
/* How the last minim_minimise_batch ran (no reference counterpart): the most
 * iterations of any instance, the iterations executed, the host
 * synchronisations that waited for results, the HIP graph launches and the
 * graph captures (one more each time the launch grid shrank to the instances
 * still running). */
int minim_batch_report(const minim_minimiser* m, int* iterations, long* replays, long* host_syncs,
                       long* graph_launches, long* captures, int* block);
/* Lbfgs::autotuneBatch (no reference counterpart): the first `replays`
 * iterations of minim_minimise_batch(states) timed for each candidate block
 * size (4, 8 or 16 elements per thread; the States are not changed); the
 * fastest is stored for batches of the same shape.  seconds[i] (may be NULL)
 * receives candidate i's time, *best (may be NULL) the chosen size. */
int minim_lbfgs_autotune_batch(minim_minimiser* m, minim_state* const* states, int n_states, const int* candidates,
                               int n_candidates, int replays, double* seconds, int* best);

#ifdef __cplusplus
}
#endif

#endif
