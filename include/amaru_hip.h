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
enum { AMR_PRECOND_JACOBI = 0, AMR_PRECOND_BLOCK_JACOBI = 1 };

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

/* Converged increment: U += ΔUa; F += ΔFin; Rc = (1-acr) Rc + R; state
 * commit (src/mech/mech-solver.jl:386-391). */
int amr_increment_commit(amr_model *m);

/* Failed increment: state restore (src/mech/mech-solver.jl:461). */
int amr_increment_reject(amr_model *m);

/* Stage totals U and F (ndofs, eq order); either pointer may be NULL.
 * With last != 0 the last increment's ΔUa and ΔFin instead. */
int amr_stage_get(amr_model *m, double *U, double *F, int last);

/* SpMV lanes per block row (8, 16, 32 or 64) for the FP64 and FP32
 * operators; 0 lets the autotuner choose on the next solve.  Every choice
 * gives bit-identical results. */
int amr_set_spmv_lanes(amr_model *m, int lanes_fp64, int lanes_fp32);
/* out[0..2]: lanes chosen for the FP64, FP32 and FP32-storage/FP64-accumulation
 * SpMVs (0 = not tuned yet). */
int amr_get_spmv_lanes(amr_model *m, int *out);

/* Kernel timing of the PCG SpMV with HIP events on the model's stream:
 * enable = 1 starts (and clears) the counters, 0 stops.  Returns the
 * accumulated milliseconds and launch count so far. */
int amr_timing(amr_model *m, int enable, double *spmv_ms, int64_t *spmv_launches);

#ifdef __cplusplus
}
#endif

#endif /* AMARU_HIP_H */
