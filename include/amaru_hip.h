/*
 * amaru_hip.h — C ABI of libamaru_hip.so, the MI355X (gfx950) implementation
 * of Amaru.jl's static FE assembly and solve path.
 *
 * Plain C types only.  Every function returns AMR_OK (0) or a negative error
 * code; amr_last_error() gives the message of the calling thread's last
 * failure.  The caller owns every host buffer; the library owns the device
 * memory behind an amr_model handle.  One handle is used by one host thread
 * at a time; distinct handles are independent.
 *
 * Numbering: node ids, element connectivity and DOF equation ids are 0-based
 * (a Julia caller subtracts 1).  eq_id[3*n+d] is the equation id of
 * direction d (ux,uy,uz) of node n as numbered by configure_dofs!
 * (src/bc.jl:198-233: unknowns first); -1 marks a node without DOFs.
 * IP state is 14 doubles per integration point, AoS on the host:
 * sig[6] eps[6] epa dlam (Mandel notation), IPs numbered element by element
 * in quadrature order.
 */
#ifndef AMARU_HIP_H
#define AMARU_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    AMR_OK = 0,
    AMR_ERR_ARG = -1,      /* bad argument or size */
    AMR_ERR_HIP = -2,      /* HIP runtime error */
    AMR_ERR_JACOBIAN = -3, /* non-positive Jacobian determinant */
    AMR_ERR_MATERIAL = -4, /* constitutive update failed */
    AMR_ERR_SOLVE = -5,    /* linear solve failed (no convergence, breakdown, |U|>1e8) */
    AMR_ERR_NAN = -6,      /* NaN in internal forces */
    AMR_ERR_NOMEM = -7     /* device memory exhausted */
};

enum { AMR_HEX8 = 1, AMR_HEX20 = 2, AMR_TET10 = 3 };
enum { AMR_MAT_LE = 0, AMR_MAT_VM = 1, AMR_MAT_DP = 2 };
enum { AMR_PREC_FP64 = 0, AMR_PREC_MIXED = 1, AMR_PREC_MIXED_F64ACC = 2 };
/* Preconditioners: Jacobi; 3x3 nodal block Jacobi; two-level = Jacobi plus
 * a coarse correction on rigid-body modes of node aggregates (additive,
 * Phi Kc^-1 Phi^T with Kc = Phi^T K Phi), which removes most of the
 * iterations slender models need. */
enum { AMR_PRECOND_JACOBI = 0, AMR_PRECOND_BLOCK_JACOBI = 1, AMR_PRECOND_TWO_LEVEL = 2 };

typedef struct amr_model amr_model;

typedef struct {
    int precision;      /* AMR_PREC_FP64: FP64 Jacobi-PCG.  AMR_PREC_MIXED: FP32 matrix,
                           FP32 search direction and FP32 accumulation in the SpMV,
                           FP32 preconditioner, FP64 vectors/dots/residual.
                           AMR_PREC_MIXED_F64ACC: FP32 matrix gathered against the FP64
                           search direction with FP64 accumulation */
    int precond;        /* AMR_PRECOND_* */
    int maxiter;        /* cap on PCG iterations (all refinement steps together) */
    int check_every;    /* iterations between host convergence checks */
    double rtol;        /* stop when ||b - K11 x||_2 <= rtol ||b||_2 (FP64 residual) */
    double inner_rtol;  /* reserved */
    double replace_ratio; /* mixed: recompute the FP64 residual each time the recursive
                             residual drops by this factor (0 = default 0.1) */
    double drift_limit;   /* mixed: fall back to the FP64 operator when the recursive
                             residual drifts from the true one by more than this
                             fraction (0 = default 0.3) */
} amr_solve_opts;

typedef struct {
    int iterations;     /* PCG iterations performed (inner ones for mixed) */
    int refinements;    /* FP64 residual recomputations (residual replacements) */
    int converged;      /* 1 if rtol (or the rounding floor of the residual) was met */
    int fallback;       /* mixed: 1 if the solve finished on the FP64 operator */
    double relres;      /* final ||b - K11 x|| / ||b|| evaluated in FP64 */
} amr_solve_info;

/* Message of this thread's last error ("" if none). */
const char *amr_last_error(void);

/* Number of visible HIP devices. */
int amr_device_count(int *n);

/*
 * Build a device model (replaces the per-call element loop of mount_K,
 * src/mech/mech-solver.jl:78-110, with a persistent handle).
 *   etype/nips: AMR_HEX8|AMR_HEX20|AMR_TET10 and the IP rule (0 = the
 *     shape's default, 8 or 27 for hex, 4 for tet) — setquadrature!,
 *     src/element.jl:119-166.
 *   coords: nnodes×3; conn: ne×nn node ids in the reference's node order.
 *   mat_kind/mat_params: nmats materials, 6 params each (LE: E nu;
 *     VM: E nu fy H; DP: E nu alpha kappa H); emat: material of each element.
 *   eq_id/ndofs/nu: DOF numbering (see above).
 * The sparsity structure, the element->K and element->force maps and all
 * device buffers are built once here.  IP states start at zero.
 */
int amr_model_create(amr_model **out, int device, int etype, int nips, int64_t nnodes, const double *coords,
                     int64_t ne, const int32_t *conn, int nmats, const int32_t *mat_kind,
                     const double *mat_params, const int32_t *emat, const int32_t *eq_id, int64_t ndofs,
                     int64_t nu);
int amr_model_destroy(amr_model *m);

/* New DOF numbering (a new stage's configure_dofs!). */
int amr_model_set_dofs(amr_model *m, const int32_t *eq_id, int64_t ndofs, int64_t nu);

/* out[0..7] = nnodes, ne, nips, nblocks, max scalar nnz (9*nblocks),
 * element->K contributions, ndofs, nu */
int amr_model_info(amr_model *m, int64_t *out);

/* IP states (host AoS nip_total×14).  set writes both the committed and the
 * current copy; get reads the current (committed=0) or committed (1) copy. */
int amr_state_set(amr_model *m, const double *st);
int amr_state_get(amr_model *m, double *st, int committed);
/* copyto!.(StateBk, State) / copyto!.(State, StateBk)
 * (src/mech/mech-solver.jl:391, :333, :461). */
int amr_state_commit(amr_model *m);
int amr_state_restore(amr_model *m);

/*
 * mount_K (src/mech/mech-solver.jl:78-110): element stiffness from the
 * current IP states (elem_stiffness, src/mech/elem/mech-solid.jl:124-166)
 * and assembly into the device K, dropping |Ke_ij| < eps and summing the
 * survivors in element order.  K stays resident on the device.
 */
int amr_assemble(amr_model *m);

/* Element matrices of the last amr_assemble, ne × nd × nd row-major. */
int amr_get_Ke(amr_model *m, double *Ke);
/*
 * mount_M (src/mech/dyn-solver.jl:72-104): element mass matrices (elem_mass,
 * src/mech/elem/mech-solid.jl:169-205, thickness 1) with rho[e] per element
 * (MechSolidProps.rho), assembled with mount_K's rule (|Me_ij| < eps
 * dropped, survivors summed in element order) into a device matrix of its
 * own, kept beside K.  Errors as amr_assemble.
 */
int amr_assemble_M(amr_model *m, const double *rho);
/* CSC of the last amr_assemble_M, as amr_get_K_csc. */
int amr_get_M_csc(amr_model *m, int64_t *nnz, int64_t *colptr, int32_t *rowval, double *nzval);
/*
 * The dynamic solver's use of K and M (src/mech/dyn-solver.jl:326-345):
 * amr_combine_KM replaces the device K by a K + b M (the pseudo-stiffness
 * Kp = K + (4/dt^2) M + (2/dt) C with C = alpha M + beta K is
 * a = 1 + 2 beta/dt, b = 4/dt^2 + 2 alpha/dt), pattern = union of K's and
 * M's, Jacobi preconditioners refreshed, so that amr_solve_system and
 * amr_get_K_csc work on it; a two-level coarse operator is refreshed by its
 * own policy.  amr_matvec: y = K x (which 0: the device K, i.e. Kp after a
 * combination), M x (which 1) or K x with K as it was before the last
 * amr_combine_KM (which 2), x and y in eq order (ndofs), for the products
 * M*(...) and C*(...) of the Newmark right-hand side and internal force.
 */
int amr_combine_KM(amr_model *m, double a, double b);
int amr_matvec(amr_model *m, int which, const double *x, double *y);

/*
 * The assembled K as the reference's SparseMatrixCSC (0-based): column
 * pointers (ndofs+1), row ids (sorted in each column) and values, with the
 * reference's pattern (entries with at least one surviving contribution).
 * With colptr == NULL only *nnz is written.
 */
int amr_get_K_csc(amr_model *m, int64_t *nnz, int64_t *colptr, int32_t *rowval, double *nzval);

/*
 * solve_system! (src/solver.jl:5-79) on the assembled K: U and F are
 * ndofs-long host vectors in eq order.  Solves K11 U1 = F1 - K12 U2, writes
 * U[0:nu] and the reactions F[nu:] = K21 U1 + K22 U2; fails like the
 * reference if max|U1| > 1e8 or the solve does not converge.  opts may be
 * NULL (FP64 Jacobi-PCG, rtol 1e-12); info may be NULL.
 */
int amr_solve_system(amr_model *m, double *U, double *F, const amr_solve_opts *opts, amr_solve_info *info);

/*
 * update_state!(elems, dUt, t) (src/mech/mech-solver.jl:124-144): restores
 * nothing itself — the trial state is computed from the committed one —
 * runs the per-IP constitutive update from the committed state with the
 * strain of the total increment dUt (ndofs, eq order) and writes the
 * internal-force increment dFin (ndofs) gathered in element order.
 */
int amr_update_state(amr_model *m, const double *dUt, double *dFin);

/* ------------------------------------------------------------------------
 * Device-resident stage driver.  The vectors of mech_stage_solver!
 * (src/mech/mech-solver.jl:258-264: U, F, R, ΔUa, ΔUi, ΔFin, Rc) live on the
 * GPU; the host keeps the reference's control flow (increments, NR loop,
 * auto-increment) and reads back two scalars per NR iteration.
 * ------------------------------------------------------------------------ */
typedef struct {
    double res;          /* ||R||_inf after the iteration (R zero at prescribed dofs) */
    double err;          /* ||ΔUi||_inf / ||ΔUa||_inf */
    double relres;       /* final relative residual of the linear solve */
    int pcg_iterations;  /* PCG iterations of this NR iteration */
    int status;          /* AMR_OK, or the error code of a failed iteration */
    int replacements;    /* FP64 residual recomputations in the linear solve */
    int fallback;        /* mixed: 1 if the solve finished on the FP64 operator */
} amr_nr_info;

/* Uex, Fex (ndofs, eq order): essential values and natural loads of the
 * stage (get_bc_vals, src/bc.jl:237-249).  Zeroes U, F and Rc. */
int amr_stage_begin(amr_model *m, const double *Uex, const double *Fex);

/* ΔUex = dT Uex; ΔFex = dT Fex (+ acr Rc if add_residual); R = ΔFex;
 * ΔUa = 0; ΔUi = ΔUex  (src/mech/mech-solver.jl:301-309). */
int amr_increment_begin(amr_model *m, double dT, double acr, int add_residual);

/* One Newton-Raphson iteration it (1-based) of src/mech/mech-solver.jl:
 * 319-367 with scheme :FE: ΔUi = 0 if it > 1; mount_K; solve_system!;
 * update_state! from the committed state with ΔUa + ΔUi; ΔUa += ΔUi;
 * R = ΔFex - ΔFin with R zero at prescribed dofs.  A failed linear solve or
 * constitutive update is reported in info->status (and the return code). */
int amr_nr_iteration(amr_model *m, int it, const amr_solve_opts *opts, amr_nr_info *info);

/* Predictor-corrector scheme of amr_nr_iteration, mech_stage_solver!'s
 * `scheme` keyword (src/mech/mech-solver.jl:280-288, 325-349).  :FE (the
 * default): one mount_K, solve and update per iteration.  :ME, :BE,
 * :Ralston, with (p1, q11, a1, a2) = (1, 1, 1/2, 1/2), (1, 1, 0, 1),
 * (2/3, 2/3, 1/4, 3/4): solve_system!(K, p1 ΔUi, q11 R) and update_state!
 * at ΔUa + that solution (the predictor), K2 = mount_K at the predicted
 * state, K = a1 K + a2 K2, solve_system!(K, ΔUi, R) and update_state! at
 * ΔUa + ΔUi (the corrector).  info->pcg_iterations counts both solves.
 * Pools (amr_pool_add) run :FE only. */
enum { AMR_SCHEME_FE = 0, AMR_SCHEME_ME = 1, AMR_SCHEME_BE = 2, AMR_SCHEME_RALSTON = 3 };
int amr_set_scheme(amr_model *m, int scheme);

/* Converged increment: U += ΔUa; F += ΔFin; Rc = (1-acr) Rc + R; state
 * commit (src/mech/mech-solver.jl:386-391). */
int amr_increment_commit(amr_model *m);

/* Failed increment: state restore (src/mech/mech-solver.jl:461). */
int amr_increment_reject(amr_model *m);

/* Stage totals U and F (ndofs, eq order); either pointer may be NULL.
 * With last != 0 the last increment's ΔUa and ΔFin instead. */
int amr_stage_get(amr_model *m, double *U, double *F, int last);

/* SpMV lanes per block row (8, 16, 32 or 64) for the FP64 and FP32
 * operators; 0 restores the default choice, made on the next solve from the
 * matrix size and blocks per row.  Every choice gives bit-identical
 * results. */
int amr_set_spmv_lanes(amr_model *m, int lanes_fp64, int lanes_fp32);
/* Times every lanes choice on the model's assembled matrix (a few dozen SpMV
 * launches and waits) and keeps the fastest for each operand kind. */
int amr_tune_spmv_lanes(amr_model *m);
/* out[0..2]: lanes chosen for the FP64, FP32 and FP32-storage/FP64-accumulation
 * SpMVs (0 = not chosen yet). */
int amr_get_spmv_lanes(amr_model *m, int *out);

/* Element tiles of the stiffness and update kernels: ach = row nodes per
 * lane of the element stiffness (HEX8 1/2/4, HEX20 7/10/20, TET10 2/5/10),
 * epw = elements per update workgroup (HEX8 32/16/8, HEX20 27 IP 9/4/2,
 * HEX20 8 IP 16/8/32, TET10 32/16/64).  Every tiling gives bit-identical
 * K, states and forces.  amr_set_elem_tiles: 0 keeps the tuned (or default,
 * the first) choice.  amr_tune_elem_tiles times the three tilings of each
 * kernel on the model (stiffness from the current states; update from the
 * committed states with a zero increment, which it leaves as the current
 * states: call it before an analysis or between increments), keeps the
 * fastest for the model and, for every later model of the same element
 * shape in the process, and returns them. */
int amr_set_elem_tiles(amr_model *m, int ach, int epw);
/* Element stiffness of the Newton-Raphson iterations (amr_nr_prepare,
 * amr_nr_iteration; elem_stiffness, src/mech/elem/mech-solid.jl:124-166):
 * AMR_KE_EXACT (the default) follows the oracle's operation order bit for
 * bit; AMR_KE_MFMA uses the FP64 matrix cores, equal to the exact path to
 * rounding (the summation order differs): models whose materials are all
 * LinearElastic form Ke_ab = lambda S_ab + mu S_ab' + mu tr(S_ab) I from
 * S_ab = sum_q coef_q g_a g_b' (every shape and rule), other HEX20 models
 * Ke = sum_q Bq' (coef D) Bq; the remaining cases keep the exact path.
 * amr_assemble (mount_K with its pattern) always uses the exact path. */
enum { AMR_KE_EXACT = 0, AMR_KE_MFMA = 1 };
int amr_set_ke_path(amr_model *m, int path);
int amr_get_elem_tiles(amr_model *m, int *ach, int *epw);
int amr_tune_elem_tiles(amr_model *m, int *ach, int *epw);

/* Kernel timing of the PCG SpMV with HIP events on the model's stream:
 * enable = 1 starts (and clears) the counters, 0 stops.  Returns the
 * accumulated milliseconds and launch count so far. */
int amr_timing(amr_model *m, int enable, double *spmv_ms, int64_t *spmv_launches);

/* The two halves of amr_nr_iteration, for callers that solve many models'
 * linear systems together in a pool.  amr_nr_prepare: ΔUi = 0 if it > 1;
 * mount_K; the right-hand side of solve_system! (src/mech/mech-solver.jl:
 * 319-331), all queued on the model's stream without waiting.
 * amr_nr_finish: given the outcome of the linear solve (si, from
 * amr_pool_step; NULL means the model has no unknowns), combines ΔUi,
 * runs update_state! and the residual as amr_nr_iteration does
 * (src/mech/mech-solver.jl:331-367). */
int amr_nr_prepare(amr_model *m, int it);
int amr_nr_finish(amr_model *m, const amr_solve_info *si, amr_nr_info *info);
/* amr_nr_finish in two halves: _begin queues the work and the read-back of
 * the residual norms, _end waits for them and fills info (the return code as
 * amr_nr_finish's).  Finishing many models of one pool: begin all, then end
 * all, so they share one wait. */
int amr_nr_finish_begin(amr_model *m, const amr_solve_info *si);
int amr_nr_finish_end(amr_model *m, amr_nr_info *info);
/* The same for n models at once (replacing the per-element loops of
 * mount_K and update_state!, src/mech/mech-solver.jl:78-144, for every
 * model): amr_nr_prepare_many equals amr_nr_prepare(ms[i], its[i]) for each
 * i, amr_nr_finish_begin_many equals amr_nr_finish_begin(ms[i], &si[i])
 * (si may be NULL when no model has unknowns), with the element kernels of
 * all n models in one launch per kernel family (element stiffness per kernel
 * variant, assembly, Jacobi inverses; IP updates per tiling, force gather).
 * Results are the same bits.  The models are on one device and none is
 * solving in a pool; each appears once. */
int amr_nr_prepare_many(amr_model **ms, const int *its, int n);
int amr_nr_finish_begin_many(amr_model **ms, const amr_solve_info *si, int n);

/* ------------------------------------------------------------------------
 * Pools: the linear solves of many models advanced together, one launch
 * per PCG phase for all of them (batching), with models joining and leaving
 * between steps (continuous batching).  Each model's PCG iterates are
 * bit-identical to those of solving it alone with amr_solve_system or
 * amr_nr_iteration (same kernels, per-model reductions in a fixed order).
 * No reference counterpart: the reference solves one model at a time
 * (src/solver.jl:5); these entry points serve the batched and served
 * forms of solve_system!.
 * ------------------------------------------------------------------------ */
typedef struct amr_pool amr_pool;

/* A pool on device for up to max_members models solving at once; every
 * step runs check_every PCG iterations (0: 32) before the host reads the
 * members' residuals. */
int amr_pool_create(amr_pool **out, int device, int max_members, int check_every);
/* Detaches every attached model (they keep working on their own streams). */
int amr_pool_destroy(amr_pool *p);

/* Registers the model with the pool.  While it is a member (from
 * amr_pool_add until a step returns it) its device work runs on the pool's
 * stream, otherwise on its own; both switches are ordered with events, so
 * neither waits on the host.  amr_model_destroy detaches a model by
 * itself. */
int amr_pool_attach(amr_pool *p, amr_model *m);
int amr_pool_detach(amr_model *m);

/* After amr_nr_prepare: the model's system K11 ΔUi = R - K12 ΔUi2 joins
 * the pool and starts at the next step.  opts may be NULL. */
int amr_pool_add(amr_pool *p, amr_model *m, const amr_solve_opts *opts);

/* One step for every member.  Members whose solve ended (converged or
 * failed, see info->converged) leave the pool and are written to
 * models[0..*nfinished) and infos (arrays of at least capacity entries;
 * capacity >= the pool's member count). */
int amr_pool_step(amr_pool *p, int capacity, int *nfinished, amr_model **models, amr_solve_info *infos);

/* amr_pool_step in two halves, so that the host can work on other models
 * while the device runs a step: step_begin queues the step's iterations
 * and the read-back of the members' residuals and returns; step_end waits
 * for them, decides, and returns the members that left, as amr_pool_step
 * (capacity >= the members of that step).  Models added in between start
 * at the next step; the finished ones' work continues on their own streams
 * after the step, without waiting for the next one. */
int amr_pool_step_begin(amr_pool *p);
int amr_pool_step_end(amr_pool *p, int capacity, int *nfinished, amr_model **models, amr_solve_info *infos);
/* *done = 1 when no step is in flight or the one in flight has finished on
 * the device (amr_pool_step_end would not wait), else 0.  For hosts that
 * drive several pools from one thread. */
int amr_pool_step_done(amr_pool *p, int *done);
/* The models that left the pool at its last step_end (their solves
 * finished) run their following work on the pool's stream until the next
 * step is queued, so that the work queued for them in between (the NR
 * finish) runs ahead of that step.  amr_pool_release_left moves them back to
 * their own streams now, after the work queued for them so far: a host that
 * continues such a model on another thread calls it first, on the thread
 * that drives the pool.  A pool call of its own: the reference has no
 * counterpart. */
int amr_pool_release_left(amr_pool *p);

/* Members currently solving. */
int amr_pool_size(amr_pool *p, int *members);

/* SpMV kernel timing over the pool's steps, as amr_timing. */
int amr_pool_timing(amr_pool *p, int enable, double *spmv_ms, int64_t *spmv_launches);

/* With enable = 1 a step's check_every iterations of every PCG phase run
 * as one HIP graph launch: captured once per launch shape and replayed
 * while members join and leave (the work lists are padded to size classes
 * and re-uploaded each step; padding entries do nothing).  enable = 0 (the
 * default, also set by the environment variable AMR_POOL_GRAPHS) launches
 * them kernel by kernel: on MI355X the replays ran the pooled workloads
 * 4-12% slower.  -1 leaves it; *captures and *replays count the graphs
 * captured and replayed so far (either may be NULL).  Results are the same
 * bits either way; with amr_pool_timing on, steps launch kernel by kernel
 * (the SpMVs between event records). */
int amr_pool_set_graphs(amr_pool *p, int enable, int64_t *captures, int64_t *replays);
/* Counters of the pool so far, out[0..4]:
 *   [0] batched formations of two-level coarse operators: they are formed
 *       when their models start solving, for all models that start at one
 *       step together, one launch per phase (Galerkin product, each
 *       Gauss-Jordan block step) for all of them;
 *   [1] models those formations formed;
 *   [2] kernels of the steps' PCG iterations (every phase of every
 *       iteration; each serves all members);
 *   [3] host submissions that queued them (one per kernel when a step runs
 *       kernel by kernel, one per graph launch when it is captured);
 *   [4] kernels the same iterations would take launched member by member
 *       (SpMV, update and direction per member and iteration, two-level
 *       members also restrict and coarse product).
 * No reference counterpart: the reference solves each system on its own
 * (src/solver.jl:5-79). */
int amr_pool_stats(amr_pool *p, int64_t *out);

/* Aggregate size of the two-level preconditioner: about target_nodes nodes
 * per aggregate, fewer aggregates than max_aggregates (defaults 256, 170).
 * Builds the aggregates and their index structures now (host work and one
 * wait on the model's stream), so that the first two-level solve does not;
 * without this call the first two-level solve builds them with the
 * defaults.  A two-level solve takes at most 2,048 coarse dofs (341
 * aggregates); a larger coarse space is refused with AMR_ERR_ARG when the
 * model joins a pool. */
int amr_set_coarse(amr_model *m, int target_nodes, int max_aggregates);
/* When the coarse operator Kc = Phi^T K Phi and its inverse are formed
 * again (at the first two-level solve after an assembly): once per DOF map
 * (the default: it only preconditions, and the elastic one served plastic
 * increments with the same or 1% more iterations), at the first solve of
 * every load increment for models with non-elastic materials, or after
 * every assembly. */
enum { AMR_COARSE_PER_DOF_MAP = 0, AMR_COARSE_PER_INCREMENT = 1, AMR_COARSE_PER_ASSEMBLY = 2 };
int amr_set_coarse_refresh(amr_model *m, int policy);
/* out[0..1]: aggregates and coarse dofs of the built coarse space (0 if not built). */
int amr_get_coarse(amr_model *m, int64_t *out);
/* The built coarse space: each node's aggregate (nnodes) and Kc^-1 as of
 * the last two-level solve (coarse dofs squared, row-major); either pointer
 * may be NULL. */
int amr_get_coarse_space(amr_model *m, int32_t *agg, double *kc_inv);

/* Bytes per block of the column indices the PCG SpMV reads: 2 when every
 * block's column minus its row fits 16 bits (offsets stored beside the
 * 32-bit columns at model creation), else 4. */
int amr_spmv_col_bytes(amr_model *m, int *bytes);
/* Chooses the column indices the PCG SpMV reads: 4 drops the 16-bit
 * offsets (the SpMV then reads the 32-bit columns, as on meshes whose
 * offsets do not fit), 2 builds them again (AMR_ERR_ARG if one does not
 * fit).  Both give bit-identical results.  Not while the model is solving in
 * a pool. */
int amr_set_spmv_col_bytes(amr_model *m, int bytes);

/* Device memory held by the model (bytes), and the device's free and total
 * memory, for admission by HBM budget.  Model memory comes from a per-device
 * cache: a destroyed model's blocks are kept for the next models (at most
 * AMR_ARENA_BYTES per device, environment, default 32 GiB), so that a
 * serving loop does not call hipMalloc/hipFree; amr_mem_info counts the
 * cached blocks as free, and amr_mem_trim returns them to the device.  No
 * counterpart in the reference (Julia's GC owns its arrays). */
int amr_model_device_bytes(amr_model *m, int64_t *bytes);
int amr_mem_info(int device, int64_t *free_bytes, int64_t *total_bytes);
int amr_mem_trim(int device);

/*
 * Distributed loads on the device, for many cells at once: the per-cell
 * force vectors of mech_boundary_forces (tractions tx/ty/tz/tn on facets,
 * ndim 2; line loads on edges, ndim 1; src/mech/elem/distributed.jl:75-150)
 * and of the solid body forces (wx/wy/wz on elements, ndim 3; :155-210).
 * C: ncells x npts x 3 node coordinates; N (nq x npts), D (nq x npts x
 * ndim) and w (nq): the cell shape's functions, derivatives and weights at
 * its integration points; dir 0..2 a global direction, -1 the unit normal
 * (facets).  The load value is a postfix program over x, y, z, t: nops
 * (operation, argument) pairs, the argument indexing consts for
 * AMR_OP_CONST; at most AMR_EXPR_STACK values deep (checked before any
 * launch).  F: ncells x npts x 3.
 */
/* The operations are those of the reference's op_dict (src/tools/expr.jl:
 * 98-125) that give numbers; AMR_OP_IDIV is Julia's div (quotient truncated
 * toward zero). */
enum {
    AMR_OP_CONST = 0, AMR_OP_X, AMR_OP_Y, AMR_OP_Z, AMR_OP_T,                             /* push */
    AMR_OP_NEG, AMR_OP_SQRT, AMR_OP_ABS, AMR_OP_SIN, AMR_OP_COS, AMR_OP_EXP, AMR_OP_LOG,  /* unary */
    AMR_OP_TAN, AMR_OP_LOG10,
    AMR_OP_ADD, AMR_OP_SUB, AMR_OP_MUL, AMR_OP_DIV, AMR_OP_POW, AMR_OP_MAX, AMR_OP_MIN,   /* binary */
    AMR_OP_IDIV
};
#define AMR_EXPR_STACK 16
int amr_check_expr(const int32_t *ops, int nops, int nconsts);
int amr_cell_loads(int device, int64_t ncells, int npts, int ndim, int nq, const double *C, const double *N,
                   const double *D, const double *w, int dir, const int32_t *ops, const double *consts, int nops,
                   int nconsts, double t, double *F);

/*
 * Structured block mesh on the device: split_block! of one Block
 * (src/mesh/structured.jl:384-552) with the synchronize! numbering
 * (src/mesh/mesh.jl:348-350).  corners: the block's 8 corners (8x3, HEX8
 * order); nx, ny, nz cells per axis.  Outputs: node coordinates (nnodes x 3,
 * rounded to 8 decimals as src/node.jl:58-60) and connectivity (ne x nen,
 * 0-based; TET10: six tetrahedra per cell).  With coords and conn NULL only
 * nnodes and ne are returned.
 */
int amr_block_mesh(int device, int etype, const double *corners, int64_t nx, int64_t ny, int64_t nz,
                   int64_t *nnodes, int64_t *ne, double *coords, int32_t *conn);
/*
 * DOF numbering of configure_dofs! (src/bc.jl:198-233) on the device: the
 * DOFs (node, ux/uy/uz) of active nodes in node order, unknowns first, then
 * prescribed, each in DOF order.  active (nnodes, or NULL: all) and presc
 * (nnodes x 3) are 0/1 flags; eq (nnodes x 3) receives the 0-based eq ids,
 * -1 for DOFs of inactive nodes.
 */
int amr_dof_map(int device, int64_t nnodes, const uint8_t *active, const uint8_t *presc, int32_t *eq,
                int64_t *ndofs, int64_t *nu);
/*
 * amr_dof_map that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that keep their buffers, harnesses that forbid
 * hipMalloc).  active (nnodes flags, or NULL: all), presc (nnodes x 3
 * flags), eq (nnodes x 3 eq ids out) and sizes (2 int64 out: ndofs, nu) are
 * DEVICE pointers; stream is a hipStream_t (NULL: the default stream).  The
 * call queues three kernels on the stream and returns; results are those of
 * amr_dof_map.  At most AMR_DOFMAP_NOALLOC_MAX_DOFS dofs (3 x nnodes), else
 * AMR_ERR_ARG.  The kernels' work space (per 256-dof tile counts and
 * offsets, 1.5 MiB) is one static device array per GPU, held until the
 * process ends and shared by every call on that GPU: calls that overlap on
 * one GPU race.
 */
#define AMR_DOFMAP_NOALLOC_MAX_DOFS (1 << 24)
int amr_dof_map_noalloc(int device, int64_t nnodes, const uint8_t *active, const uint8_t *presc, int32_t *eq,
                        int64_t *sizes, void *stream);

/*
 * Output post-processing.  Values per IP are ip_state_vals of the solid
 * materials (src/mech/mat/linear-elastic.jl:140-142, von-mises.jl:159-168,
 * drucker-prager.jl:152-163) over stress_strain_dict
 * (src/tools/tensors.jl:162-215, 3D), in a fixed table of AMR_NFIELDS
 * columns: sxx syy szz syz sxz sxy svm s1 s2 s3 exx eyy ezz eyz exz exy
 * (shear = Mandel / sqrt 2; s1 >= s2 >= s3 the principal stresses), then
 * column 16 = ep (von Mises) or epa (Drucker-Prager), 17 = j1 and
 * 18 = srj2d (Drucker-Prager); unused columns hold 0.  committed selects the
 * committed (1) or current (0) IP states.
 */
#define AMR_NFIELDS 19
/* out: nip_total x AMR_NFIELDS, IP-major (element e, IP q at row e*nips+q). */
int amr_ip_values(amr_model *m, int committed, double *out);
/*
 * nodal_patch_recovery (src/fe-model.jl:506-692): least-squares polynomial
 * fits over the patches of elements around corner nodes, evaluated at the
 * patches' nodes and averaged per node (patch contributions summed in
 * patch order, as one reference thread does); nodes no patch reaches get 0.
 * out: nnodes x AMR_NFIELDS; reps (optional, nnodes): patches reaching each
 * node.  The patches are built on the first call (host work, kept with the
 * model).  Errors: AMR_ERR_ARG when the elements use different material
 * kinds or a patch has over 256 integration points; AMR_ERR_SOLVE when a
 * patch's least-squares basis is rank-deficient.
 */
int amr_nodal_recovery(amr_model *m, int committed, double *out, int32_t *reps);

#ifdef __cplusplus
}
#endif

#endif /* AMARU_HIP_H */
