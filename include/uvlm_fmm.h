/* uvlm_fmm.h -- C ABI of the MI355X-native FMM vortex-particle solver (libuvlm_fmm.so).
 *
 * This is the drop-in boundary for the hot path of kursatyurt/UVLM: the work that
 * Vortex::FMMCalculator (src/fmmCalculator.hpp:25-1006) and the exafmm free functions
 * (src/tree.hpp:101, src/traverse.hpp:32, src/traverse.hpp:88) do for one particle set.
 * Each entry point below names the reference interface it replaces.
 *
 * Conventions
 *  - Every function returns 0 (UVLM_OK) or a positive UVLM_ERR_* code; the message of the
 *    last failure on the calling thread is uvlm_last_error().  No exception crosses the ABI.
 *  - A handle owns its device buffers and one HIP stream.  Calls on one handle are
 *    stream-ordered and not re-entrant; separate handles are independent.
 *  - Host arrays belong to the caller.  Positions, strengths and velocities are [n][3]
 *    row-major doubles, J is [n][3][3] with J[a][b] = d u_a / d x_b.
 *  - Bodies may be passed as exafmm::Body records (37 doubles, 296 bytes,
 *    src/exafmmTypes.hpp:18-32) with a byte stride, or as separate arrays.
 *  - Like the reference, building the tree PERMUTES the particle set into tree order;
 *    getters return the bodies in the current tree order, and uvlm_get_permutation maps
 *    each tree position to the index the caller gave in the last uvlm_set_bodies_*.
 */
#ifndef UVLM_FMM_H
#define UVLM_FMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UVLM_ABI_VERSION 1

enum uvlm_status {
  UVLM_OK           = 0,
  UVLM_ERR_ARG      = 1, /* bad argument (null pointer, size mismatch, unsupported P...) */
  UVLM_ERR_DEPTH    = 2, /* more than ncrit coincident particles: octree deeper than 21 levels */
  UVLM_ERR_STATE    = 3, /* call out of order (e.g. eval before build) */
  UVLM_ERR_HIP      = 4, /* HIP runtime error */
  UVLM_ERR_NOMEM    = 5, /* device allocation failed */
  UVLM_ERR_INTERNAL = 6
};

enum uvlm_precision { UVLM_FP64 = 0, UVLM_MIXED = 1 };            /* MIXED: FP32 P2P, FP64 expansions */
enum uvlm_stretching { UVLM_TRANSPOSED = 0, UVLM_CLASSIC = 1 };   /* fmmCalculator.hpp:22-23 */
enum uvlm_formulation { UVLM_CVPM = 0, UVLM_RVPM = 1 };           /* fmmCalculator.hpp:19-20 */
enum uvlm_scheme { UVLM_EULER = 0, UVLM_RK2 = 1, UVLM_RK3 = 2, UVLM_RK4 = 3 }; /* fmmCalculator.hpp:14-17 */

typedef struct uvlm_params {
  int     P;           /* expansion terms: degrees 0..P-1 (exafmmTypes.hpp:12); 2 <= P <= 16 */
  double  theta;       /* multipole acceptance criterion (exafmmTypes.hpp:15) */
  int64_t ncrit;       /* max bodies per leaf (exafmmTypes.hpp:14) */
  int     precision;   /* uvlm_precision */
  int     stretching;  /* uvlm_stretching */
  int     formulation; /* uvlm_formulation */
  int     scheme;      /* uvlm_scheme, used by uvlm_advance */
  int     device;      /* HIP device ordinal */
} uvlm_params;

/* Per-phase device times of the most recent uvlm_build / uvlm_eval, in milliseconds,
 * measured with HIP events on the handle's stream, and the sizes they worked on. */
typedef struct uvlm_timing {
  double tree_ms;     /* bounds, keys, sort, cells, permutation */
  double lists_ms;    /* dual-tree traversal, list sort, CSR */
  double upward_ms;   /* reset, P2M, M2M */
  double m2l_ms;
  double p2p_ms;
  double downward_ms; /* L2L, L2P, stretching, rVPM */
  double eval_ms;     /* whole uvlm_eval */
  int64_t n_bodies, n_cells, n_leaves, n_p2p_pairs, n_m2l_pairs, n_levels;
  int64_t n_interactions; /* sum over P2P pairs of n_target * n_source */
  double  sensor_ms;      /* last uvlm_sensors_aos: probe tree, lists, M2L, P2P, L2L, L2P */
  double  adjoint_ms;     /* last uvlm_probe_adjoint, same stages transposed */
  int64_t ncrit;          /* leaf size in use (uvlm_autotune may change it) */
  int64_t m2l_slots;      /* M2L accumulator slots per lane of the plan in use */
  double  vjp_ms;         /* last uvlm_eval_vjp: the adjoint pass (forward eval excluded) */
} uvlm_timing;

typedef struct uvlm_handle_s *uvlm_handle;

/* Thread-local message of the last failure ("" if none). */
const char *uvlm_last_error(void);
int         uvlm_abi_version(void);

/* Lifecycle: replaces constructing Vortex::FMMCalculator<...> (fmmCalculator.hpp:25-26);
 * the template tags become uvlm_params fields, and P / theta / ncrit become runtime
 * values instead of exafmmTypes.hpp:12-15 constants. */
int uvlm_create(const uvlm_params *params, uvlm_handle *out);
int uvlm_destroy(uvlm_handle h);

/* Upload a particle set: replaces copyBodies2Device (fmmCalculator.hpp:214-223).
 * `bodies` points at n records of `stride` bytes laid out like exafmm::Body. */
int uvlm_set_bodies_aos(uvlm_handle h, const void *bodies, int64_t n, int64_t stride);
/* Download in tree order: replaces copyBodies2Host (fmmCalculator.hpp:235-243). */
int uvlm_get_bodies_aos(uvlm_handle h, void *bodies, int64_t n, int64_t stride);
/* Array form of set/get.  Any output pointer may be NULL. */
int uvlm_set_bodies_soa(uvlm_handle h, int64_t n, const double *X, const double *alpha, const double *sigma);
int uvlm_get_bodies_soa(uvlm_handle h, int64_t n, double *X, double *alpha, double *sigma, double *velocity,
                        double *J, double *dadt, double *drdt);
/* perm[i] = index, in the last uvlm_set_bodies_* call, of the body now at tree position i. */
int uvlm_get_permutation(uvlm_handle h, int64_t n, int64_t *perm);

/* Tree + interaction lists on the device: replaces FMMCalculator::init / createTree
 * (fmmCalculator.hpp:71-115, 282-288), i.e. exafmm::buildTree (tree.hpp:101-113) and
 * exafmm::getList(root, root) (traverse.hpp:32-85).  Permutes the device particle set. */
int uvlm_build(uvlm_handle h);
/* One FMM evaluation of velocity, velocity gradient J, stretching dadt (and drdt for rVPM):
 * replaces FMMCalculator::calculate (fmmCalculator.hpp:420-467). */
int uvlm_eval(uvlm_handle h);
/* One RK stage update kernel (stage 0..stages-1 of params.scheme, after uvlm_eval):
 * replaces stepEuler / stepRK2x / stepRK3x / stepRK4x (fmmCalculator.hpp:493-711).
 * store_initial marks the first stage of RK2/RK4 (storeInitialState, 493-509). */
int uvlm_stage(uvlm_handle h, int stage, double dt);
/* A full time step with a tree rebuild before every stage, device-resident:
 * replaces FMMCalculator::advance (fmmCalculator.hpp:919-1005) minus the host copies. */
int uvlm_advance(uvlm_handle h, double dt);

/* Tree and list inspection (for bit-exact tests against the reference semantics).
 * ints: [C][5] = NCHILD, NBODY, ChildOffset, BodyOffset, level; reals: [C][4] = X, R. */
int uvlm_num_cells(uvlm_handle h, int64_t *n_cells);
int uvlm_get_tree(uvlm_handle h, int64_t n_cells, int64_t *ints, double *reals);
/* Pair lists as (target cell, source cell), sorted by target then source.  Pass NULL
 * arrays to query the sizes. */
int uvlm_get_lists(uvlm_handle h, int64_t *n_p2p, int64_t *n_m2l, int64_t *p2p, int64_t *m2l);
/* exafmm::getList(targetRoot, sourceRoot) (traverse.hpp:32-85) on two trees given by the
 * caller (e.g. exafmm::Cells from a build): the dual-tree traversal with the near test
 * |X_t - X_s|^2 theta^2 < (R_t + R_s)^2 and the split rule of traverse.hpp:74, run by the device
 * traversal on `device`.  Trees are [C][5] ints (NCHILD, NBODY, ChildOffset, BodyOffset, level)
 * and [C][4] reals (X, R), as uvlm_get_tree returns them; the children of a cell must follow it
 * (ChildOffset > its index), as in the reference's numbering.  *n_p2p and *n_m2l receive the list
 * sizes; p2p / m2l ([n][2] (target, source), sorted by target then source) are written only when
 * non-NULL and at least that long (cap_p2p / cap_m2l pairs), so a caller can ask for the sizes
 * first.  The reference returns the pairs in its stack order; as sets they are the same. */
int uvlm_get_list(int device, double theta, int64_t ct, const int64_t *t_ints, const double *t_reals, int64_t cs,
                  const int64_t *s_ints, const double *s_reals, int64_t cap_p2p, int64_t cap_m2l, int64_t *n_p2p,
                  int64_t *n_m2l, int64_t *p2p, int64_t *m2l);
/* Multipole and local expansions after uvlm_eval: [C][3*P(P+1)/2] complex doubles
 * (re, im interleaved), index 3*(n(n+1)/2 + m) + d as in multipoleKernels.hpp:468-471. */
int uvlm_get_expansions(uvlm_handle h, double *M, double *L);

/* Velocity (and far-field J) at probe points from the last uvlm_eval's multipoles:
 * replaces FMMCalculator::getSensorData (fmmCalculator.hpp:726-731).  Like the
 * reference, the probes are permuted into their own tree order in place (AoS records),
 * velocity and dadt are reset, and J accumulates the far field only
 * (multipoleKernels.hpp:303-341 adds no near-field J). */
int uvlm_sensors_aos(uvlm_handle h, void *probes, int64_t m, int64_t stride);

/* FP64 O(m*N) direct sum from the handle's current bodies (sources) to m target points
 * xt [m][3]: u [m][3] and, if non-NULL, J [m][9]; with target strengths at [m][3], dadt
 * [m][3] = J.at (classic stretching).  Replaces exafmm::direct (traverse.hpp:88-98) with
 * the device P2P formula (multipoleKernels.hpp:223-301); coincident points are skipped. */
int uvlm_direct(uvlm_handle h, int64_t m, const double *xt, const double *at, double *u, double *J, double *dadt);

/* uvlm_direct without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernel and results as uvlm_direct;
 * the targets and outputs are staged in a static device array of the library instead of a
 * per-call buffer, so m <= UVLM_DIRECT_NOALLOC_MAX (larger m fails with UVLM_ERR_ARG).  The
 * array is 2.75 MB, one copy per GPU the library runs on, held until the process ends; its
 * address is taken when the first handle on that GPU is created.  All calls on one GPU share
 * it: calls that overlap in time on one GPU (from several threads or handles) race. */
#define UVLM_DIRECT_NOALLOC_MAX 16384
int uvlm_direct_noalloc(uvlm_handle h, int64_t m, const double *xt, const double *at, double *u, double *J,
                        double *dadt);

/* Adjoint of the probe objective J = sum_p w_p . u(x_p), with u the velocity that
 * uvlm_sensors_aos computes at probe p from the current particles (the same interaction
 * lists, transposed).  Outputs, in the current tree order of the particles:
 *   dJ_dalpha [n][3] = dJ / d alpha_j   and   dJ_dx [n][3] = dJ / d x_j.
 * No counterpart in the reference (its solver has no adjoint): it is evaluated as one FMM
 * pass from probe "vortices" w_p onto the particles with the particle's own sigma, whose
 * field is dJ/d alpha and whose gradient gives dJ/dx = (grad field)^T alpha.  Needs a
 * uvlm_build; uses none of the forward results. */
int uvlm_probe_adjoint(uvlm_handle h, int64_t m, const double *probes, const double *weights, double *dJ_dalpha,
                       double *dJ_dx);
/* Vector-Jacobian product of one evaluation (cVPM, inviscid; both stretching forms): with
 * cotangents ubar [n][3] of the velocity and abar [n][3] of dadt, returns
 *   xbar = dS/dx and alphabar = dS/dalpha of S = sum_i ubar_i . u_i + abar_i . dadt_i,
 * for the tree and lists of the last uvlm_build held fixed (the exact gradient of the FMM
 * approximation, expansions included), all in the CURRENT tree order.  Evaluates the rates
 * first (uvlm_eval), then runs the adjoint in FP64: the P2P pair VJP on the forward and the
 * transposed P2P lists, the far field of adjoint vortices ubar and vortex dipoles
 * Jbar = alpha abar^T (classic: abar alpha^T) through P2M, M2M, the transposed M2L list, L2L
 * and L2P, and a degree-3 L2P of the forward local expansions for the target-side terms.
 * ubar or abar may be NULL (zero).  No counterpart in the reference (its solver has no
 * adjoint); replaces differentiating FMMCalculator::calculate (fmmCalculator.hpp:420-467). */
int uvlm_eval_vjp(uvlm_handle h, int64_t n, const double *ubar, const double *abar, double *xbar, double *alphabar);
/* Vector-Jacobian product of one uvlm_advance(dt) (the handle's scheme, cVPM, inviscid) from the
 * current particle state: on entry xbar/alphabar [n][3] hold the cotangent of the state after
 * the step, on exit that of the state before it.  Both are indexed by the ORIGINAL body order
 * (the ids uvlm_get_permutation reports), because a step reorders the bodies.  The stage
 * states are recomputed; each stage's tree is rebuilt as in the forward step and held fixed
 * in its VJP.  The particle state is left at the step's start (tree not built).  No
 * counterpart in the reference; differentiates FMMCalculator::advance (fmmCalculator.hpp:
 * 919-1005). */
int uvlm_step_vjp(uvlm_handle h, double dt, int64_t n, double *xbar, double *alphabar);
/* Checkpointed trajectory: uvlm_trajectory_forward advances `steps` steps of dt from the current
 * state, keeping (X, alpha, sigma) in device memory every `every` steps; uvlm_trajectory_vjp then
 * maps the cotangent of the final state (xbar/alphabar [n][3], original body order, in place)
 * to that of the initial state, recomputing each segment from its checkpoint.  The particles
 * end at the final state again.  No counterpart in the reference. */
int uvlm_trajectory_forward(uvlm_handle h, int64_t steps, double dt, int64_t every);
int uvlm_trajectory_vjp(uvlm_handle h, int64_t n, double *xbar, double *alphabar);
/* Number of particles on the handle (it grows with uvlm_append_bodies_soa and shedding). */
int uvlm_num_bodies(uvlm_handle h, int64_t *n);
/* Append n particles (X [n][3], alpha [n][3], sigma [n]) after the current ones; they take the
 * next original indices.  The tree is rebuilt by the next build/advance.  Replaces
 * particles.push_back (main.cpp:288) on a device-resident particle set. */
int uvlm_append_bodies_soa(uvlm_handle h, int64_t n, const double *X, const double *alpha, const double *sigma);

/* ---- Vortex-lattice wing (src/main.cpp:11-295): the caller that drives the particle wake ----
 * A lattice of vertices [nchord+1][nspan+1][3] (row 0 the leading edge); nchord = 1 is the
 * reference's Wing of LE/TE vertex couples (main.cpp:27-74).  Panel p = r * nspan + c carries
 * the ring of getPanelVortexLine (60-73).  The system A gamma = -U.n - v.n is solved with
 * A[i][j] = (velocity of ring j at control point i) . n_i (the reference fills AIC(j, i) with this
 * value, 198-205; on its wing, a lattice of congruent parallelograms, A is symmetric, so the
 * two agree), v the particle-induced velocity at the control points (getSensorData, 196-197). */
#define UVLM_WING_MAX_PANELS 2048
typedef struct uvlm_wing_s *uvlm_wing;
/* Geometry (calculateTopology, 45-58), ring influences, system matrix and its LU (one workgroup,
 * partial pivoting, in place of Eigen's fullPivLu, 214) on `device`.  Fails with UVLM_ERR_ARG
 * where the reference exits: a control point on a vortex line (19-21). */
int uvlm_wing_create(int device, int64_t nspan, int64_t nchord, const double *vertices, uvlm_wing *out);
int uvlm_wing_destroy(uvlm_wing w);
/* Control points [n][3], unit normals [n][3], areas [n], and the LU factors [n][n] (any NULL). */
int uvlm_wing_get(uvlm_wing w, double *cp, double *normal, double *area, double *aic_lu);
/* Rigid translation of the wing by d (the matrix and its factors are unchanged). */
int uvlm_wing_translate(uvlm_wing w, const double d[3]);
/* Circulations gamma [n] and the Kutta-Joukowski force [3] (223-244: sum over panels of the net
 * circulation of the leading-edge segment times u x dl, u = rings + U + v at the control point)
 * at the wing's current position.  particles may be NULL or empty; otherwise they are built and
 * evaluated at their current state for v.  gamma and force may be NULL. */
int uvlm_wing_solve(uvlm_wing w, uvlm_handle particles, const double U[3], double *gamma, double force[3]);
/* Shed one particle per trailing-edge panel from the last solve and append it (246-289).
 * reference_offset != 0 offsets z by U[1] dt as main.cpp:254 writes it, else by U[2] dt. */
int uvlm_wing_shed(uvlm_wing w, uvlm_handle particles, const double U[3], double dt, int reference_offset);
/* One pass of the reference's time loop (main.cpp:168-295): solve, shed, advance the particles
 * by dt with the handle's scheme (its viscous scheme applied like uvlm_advance), then translate
 * the wing by -U dt (its vertices are at -dt * step * U at step `step`, 182-184). */
int uvlm_wing_step(uvlm_wing w, uvlm_handle particles, const double U[3], double dt, int reference_offset,
                   double *gamma, double force[3]);

/* Adjoint of the lift coefficient through the loop (the build's dC_L adjoint through time; no
 * counterpart in the reference): from the current particles (n of them) and wing position,
 * `steps` passes of uvlm_wing_step, then CL = e . F / (|U|^2 S / 2) (S the wing area, e a unit
 * lift direction) at the final position; returns CL and its gradient with respect to the
 * particles' positions (xbar [n][3]) and strengths (alphabar [n][3]) at the start, ORIGINAL
 * order.  The chain: Kutta-Joukowski force -> circulations (A^{-T}) -> the particle velocity at
 * the control points (probe adjoint) and, back through each step, the RK step VJP of the wake
 * (uvlm_step_vjp) and the shedding map.  Trees are held fixed per evaluation (each rebuilt as in
 * the forward pass).  The particles and the wing end at the final state.  cVPM, inviscid. */
int uvlm_wing_lift_adjoint(uvlm_wing w, uvlm_handle particles, const double U[3], double dt, int64_t steps,
                           int reference_offset, const double e[3], double *CL, int64_t n, double *xbar,
                           double *alphabar);

/* Design gradient of the lift coefficient through a window of the loop (no counterpart in the
 * reference).  As uvlm_wing_lift_adjoint, from the current particles (n of them) and wing
 * position: `steps` passes of uvlm_wing_step, CL at the final position, and its gradient with
 * respect to the particles at the start (xbar, alphabar [n][3], original order) and to the wing's
 * vertices (vertex_bar [(nchord+1)(nspan+1)][3]: the same lattice, moved rigidly by -U dt per
 * step, so one set of vertices fixes the geometry of every step: the normals, control points,
 * areas, system matrix, right-hand sides, force segments and shed particles all follow it).  The
 * particle state at the window start is held fixed (the window's sensitivity).  The particle
 * state and the wing position are checkpointed every `every` steps and each segment is
 * recomputed from its checkpoint in the backward pass.  A design variable (angle of attack,
 * twist, ...) that moves the vertices by dV gives dCL = sum vertex_bar . dV.  Any of CL, xbar,
 * alphabar, vertex_bar may be NULL.  The particles and the wing end at the final state.  cVPM,
 * inviscid. */
int uvlm_wing_window_adjoint(uvlm_wing w, uvlm_handle particles, const double U[3], double dt, int64_t steps,
                             int64_t every, int reference_offset, const double e[3], double *CL, int64_t n,
                             double *xbar, double *alphabar, double *vertex_bar);

/* Overwrite positions and strengths [n][3] in the CURRENT tree order without rebuilding
 * the tree or lists (frozen-tree perturbations, e.g. finite differences).  No counterpart
 * in the reference. */
int uvlm_update_bodies_soa(uvlm_handle h, int64_t n, const double *X, const double *alpha);

/* Pedrizzetti relaxation of the strengths toward the vorticity w = curl u of the last
 * uvlm_eval: alpha <- (1-f) alpha + f |alpha| w / |w| (bodies with w = 0 unchanged).
 * Replaces the pseudo-code of src/relaxation.hpp:8-18, which does not compile. */
int uvlm_relax(uvlm_handle h, double f);
/* Invariants of the current particle set: out[0..2] linear impulse 1/2 sum x x alpha,
 * out[3..5] angular impulse 1/3 sum x x (x x alpha), out[6] energy sum u . (x x alpha)
 * (needs the velocity of a uvlm_eval), out[7..9] total strength sum alpha.  No reference
 * counterpart (the reference prints only strength extrema, src/main.cpp:219-221). */
int uvlm_diagnostics(uvlm_handle h, double *out);

/* Near-field option (no counterpart in the reference, off by default): with c > 0, a cell pair
 * whose bodies can come closer than c * (the largest core size sigma of the particles) is treated
 * as near (P2P with the regularised kernel) even where the acceptance test of traverse.hpp:56-57
 * would take it as far: |X_t - X_s| < c sigma_max + sqrt(3) (R_t + R_s).  The reference's far field
 * applies the singular kernel, which differs from the regularised one by a factor
 * 1 - erf(s) + 2 s e^{-s^2} / sqrt(pi), s = R / (sqrt(2) sigma) (about 1e-6 at R = 6 sigma), so c
 * bounds the far field's regularisation error; the expansion order P and theta bound its
 * truncation.  Applies from the next build; c = 0 restores the reference's lists exactly. */
int uvlm_set_near_field(uvlm_handle h, double c);

/* Viscous schemes (the ViscousScheme tag and _nu of fmmCalculator.hpp:9-13, 56; _nu is never
 * initialised there).  Once set, uvlm_advance applies the scheme after every stage's eval, as
 * advance() calls calculateViscosity() (fmmCalculator.hpp:919-1005):
 *   UVLM_PSE  dadt_i += nu/(sqrt2 pi^1.5) sum_j (vol_i a_j - vol_j a_i) e^{-R^2/(2 s_ij^2)} / s_ij^5
 *             over the P2P lists (multipoleKernels.hpp:392-450);
 *   UVLM_CSM  drdt_i += (sqrt(sigma_i^2 + 2 nu dt) - sigma_i) / dt (fmmCalculator.hpp:334-337). */
enum { UVLM_INVISCID = 0, UVLM_PSE = 1, UVLM_CSM = 2 };
int uvlm_set_viscosity(uvlm_handle h, int scheme, double nu);
/* calculateViscosity() (fmmCalculator.hpp:713-724) on the rates of the last uvlm_eval; dt is
 * the step CSM spreads over. */
int uvlm_viscosity(uvlm_handle h, double dt);
/* Gaussian vorticity of the particles at m probes over the near (P2P) pairs of the probe tree
 * against the particle tree: velocity_old = -sum_j a_j exp(-R^2/(2 sigma_j^2)) /
 * ((2 pi)^1.5 sigma_j^3) (calculateVorticityDualTree, fmmCalculator.hpp:746-753;
 * multipoleKernels.hpp:343-370).  Needs a uvlm_build; the probes come back in tree order. */
int uvlm_vorticity_aos(uvlm_handle h, void *probes, int64_t m, int64_t stride);
/* RBF re-initialisation (fmmCalculator.hpp:855-917): the particles' vorticity at the m new
 * points, the points below 1e-6 in every component dropped, and the strengths of the rest fitted
 * by the reference's conjugate gradients on the Gaussian near-field operator of their own tree
 * (at most 100 iterations, squared residual < 1e-10 per component).  points[0..*kept) receive
 * the kept points in tree order with the fitted alpha; *iterations and residuals[3] report the
 * CG (may be NULL).  Needs a uvlm_build of the particles. */
int uvlm_rbf_aos(uvlm_handle h, void *points, int64_t m, int64_t stride, int64_t *kept, int *iterations,
                 double *residuals);

int uvlm_timings(uvlm_handle h, uvlm_timing *out);

/* Chooses ncrit and the M2L plan size for the bodies currently on the handle by timing
 * uvlm_build + uvlm_eval: each candidate ncrit[0..n) with each accumulator-slot budget in
 * {5, 6, 8} (budgets that give the same plan are timed once), one warm-up and `reps` timed
 * passes each.  The fastest pair is kept on the handle for later builds; *best_ncrit and
 * *best_ms (mean milliseconds of build + eval) report it.  The bodies are left permuted by
 * the last trial build, and the handle needs a uvlm_build before its next uvlm_eval.
 * No counterpart in the reference, where ncrit is a compile-time constant
 * (src/exafmmTypes.hpp:14). */
int uvlm_autotune(uvlm_handle h, const int64_t *ncrit, int n, int reps, int64_t *best_ncrit, double *best_ms);

/* Device synchronisation of the handle's stream. */
int uvlm_sync(uvlm_handle h);

#ifdef __cplusplus
}
#endif
#endif /* UVLM_FMM_H */
