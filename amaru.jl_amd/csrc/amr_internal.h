// Internal declarations shared by the HIP library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/amaru_hip.h"

#define AMR_NSTATE 14
#define AMR_MAXNN 20
#define AMR_MAXIP 27

// material table entry (device and host)
struct AmrMat {
    int kind;  // AMR_MAT_*
    int pad;
    double p[6];  // LE: E nu; VM: E nu fy H; DP: E nu alpha kappa H
};

void amr_set_error(const std::string &msg);

// K block values in chunks of 64 consecutive blocks.  Inside a chunk the
// nine values (row-major 3x3 position s) of a block are stored as four pairs
// (s = 0,1 | 2,3 | 4,5 | 6,7) and a single (s = 8): pair p of all 64 blocks is
// one contiguous run of 64 two-element vectors, then the run of s = 8.  A
// lane reading block k fetches its values with four 2-wide loads and one
// scalar load, and the lanes of a wave reading consecutive blocks touch one
// contiguous run per load.
__host__ __device__ inline int64_t kv_at(int64_t k, int s)
{
    const int64_t base = (k >> 6) * 576;
    return s < 8 ? base + (int64_t)(s >> 1) * 128 + (k & 63) * 2 + (s & 1) : base + 512 + (k & 63);
}
inline int64_t kv_size(int64_t nblk) { return ((nblk + 63) >> 6) * 576; }
// The FP32 copy uses the same 64-block chunks with four-value groups
// (s = 0..3 | 4..7, then s = 8): a lane reads a block with two 16-byte
// loads and one scalar load.
__host__ __device__ inline int64_t kv32_at(int64_t k, int s)
{
    const int64_t base = (k >> 6) * 576;
    return s < 8 ? base + (int64_t)(s >> 2) * 256 + (k & 63) * 4 + (s & 3) : base + 512 + (k & 63);
}

#define AMR_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) {                                                        \
            amr_set_error(std::string("HIP error: ") + hipGetErrorString(_e) + " at " + \
                          __FILE__ + ":" + std::to_string(__LINE__));                  \
            return AMR_ERR_HIP;                                                        \
        }                                                                              \
    } while (0)

// Host-built sparsity structure (node-block BSR over all nodes).
struct AmrPattern {
    int64_t nnodes = 0, nblk = 0, ncontrib = 0;
    std::vector<int32_t> rowptr;   // nnodes+1
    std::vector<int32_t> col;      // nblk
    std::vector<int32_t> diag;     // nnodes: block index of (n,n)
    std::vector<int64_t> cptr;     // nblk+1: contributions of each block
    std::vector<int32_t> csrc;     // ncontrib: e*nn*nn + la*nn + lb (ascending e per block)
    std::vector<int32_t> fptr;     // nnodes+1: force-gather contributions per node
    std::vector<int32_t> fsrc;     // e*nn + la (ascending e per node)
};

int amr_build_pattern(int nn, int64_t nnodes, int64_t ne, const int32_t *conn, AmrPattern &P);

// Doubles of a model's reduction-partials buffer (m->part; layout in
// kernels_solve.hip).
constexpr int AMR_PART_DBL = 46080;

// Device-built pattern (pattern.hip): the same arrays as amr_build_pattern,
// made on the device from the model's conn.  Returns AMR_PATTERN_ON_HOST when
// a node's contributions exceed what one workgroup sorts (the host builder is
// used then).  The host copies in m->pat (rowptr, col, diag) are downloaded
// on first use by amr_pattern_host.
constexpr int AMR_PATTERN_ON_HOST = 1;
struct amr_model;
int amr_build_pattern_device(amr_model *m);
int amr_pattern_host(amr_model *m);

// Device memory of a model, counted in m->dev_bytes; amr_release gives it back.
int amr_dalloc(amr_model *m, void **p, size_t bytes);
void amr_release(amr_model *m, void *p);

// Host-built coarse space of the two-level preconditioner (model_build.cpp).
struct AmrCoarse {
    int32_t nagg = 0;
    std::vector<int32_t> agg;       // nnodes: aggregate of each node
    std::vector<int32_t> aggptr;    // nagg+1
    std::vector<int32_t> aggnodes;  // nnodes: nodes of each aggregate, ascending
    std::vector<double> cent;       // 3*nagg: aggregate centroids
    std::vector<int32_t> pptr;      // nnodes+1: (row, aggregate) pairs of each block row
    std::vector<int32_t> pagg;      // npairs: the pair's aggregate
    std::vector<int32_t> prow;      // npairs: the pair's block row
    std::vector<int32_t> cba, cbb;  // coarse blocks (a, b), ascending
    std::vector<int32_t> cbptr;     // ncb+1
    std::vector<int32_t> cbslot;    // npairs: pair slots of each coarse block, ascending row
};
int amr_build_coarse(int64_t nnodes, const double *X, const int32_t *rowptr, const int32_t *col, int target,
                     AmrCoarse &C);

struct amr_pool;

// Per-model device resources.
struct amr_model {
    int etype = 0, nn = 0, nips = 0, device = 0;
    int ncu = 256;  // compute units of the device
    int64_t nnodes = 0, ne = 0, ndofs = 0, nu = 0, nip_total = 0;
    AmrPattern pat;
    std::vector<int32_t> eq;      // nnodes*3 eq ids (-1: no dof)
    std::vector<uint8_t> free_h;  // nnodes*3: 1 if the dof is unknown

    // device buffers
    double *X = nullptr;          // nnodes*3
    int32_t *conn = nullptr;      // ne*nn
    int32_t *emat = nullptr;      // ne
    AmrMat *mats = nullptr;       // nmats
    int nmats = 0;
    bool all_elastic = true;
    double *dNdR = nullptr;       // nips*nn*3
    double *wip = nullptr;        // nips
    double *st_cur = nullptr;     // 14*nip_total SoA (trial/current)
    double *st_com = nullptr;     // 14*nip_total SoA (committed)
    double *Ke = nullptr;         // ne*nn*nn*9 (block layout)
    int32_t *rowptr = nullptr, *col = nullptr, *diag = nullptr;
    int16_t *col16 = nullptr;     // nblk column - row offsets (null when one does not fit 16 bits)
    int64_t *cptr = nullptr;
    int32_t *csrc = nullptr;
    int32_t *fptr = nullptr, *fsrc = nullptr;
    double *Kv = nullptr;         // FP64 block values, kv_at(k, s) layout
    float *Kv32 = nullptr;        // nblk*9 FP32 values
    uint8_t *Kflag = nullptr;     // nblk*9 survivor flags (exact pattern)
    double *dinv = nullptr;       // 3*nnodes Jacobi inverse (0 at prescribed)
    float *dinv32 = nullptr;
    double *binv = nullptr;       // 9*nnodes nodal block-Jacobi inverses
    float *binv32 = nullptr;
    double *mask = nullptr;       // 3*nnodes: 1 free, 0 prescribed
    int32_t *eqd = nullptr;       // 3*nnodes eq ids (device)
    double *dF = nullptr;         // ne*nn*3 element internal forces
    int32_t *estat = nullptr;     // ne status codes
    int32_t *err = nullptr;       // device error word
    // solver workspace (3*nnodes each)
    double *v[12] = {};
    float *f[8] = {};
    double *v_block = nullptr;  // v and f are slices of these
    float *f_block = nullptr;
    double *part = nullptr;       // reduction partials
    double *scal = nullptr;       // device scalars
    int nparts = 0;
    hipStream_t stream = nullptr;      // where the model's work runs (a pool's stream while attached)
    hipStream_t own_stream = nullptr;  // the stream created with the model
    amr_pool *self_pool = nullptr;     // pool of one for amr_pcg_solve and amr_tune_lanes
    amr_pool *pool = nullptr;          // shared pool the model is attached to
    int64_t dev_bytes = 0;             // device memory allocated for the model
    std::vector<double> Xh;            // host copy of the coordinates
    int64_t k_version = 0;             // bumped by every assembly
    int64_t dof_version = 0;           // bumped by every DOF map upload
    int64_t inc_version = 0;           // bumped by every amr_increment_begin
    // two-level preconditioner: aggregates of about coarse_target nodes (more
    // when the model would need over coarse_max aggregates), rigid-body modes
    // per aggregate, Kc = Phi^T K Phi inverted on the device
    int coarse_target = 256, coarse_max = 170;
    bool coarse_built = false;
    int64_t coarse_version = -1;       // k_version the coarse operator was formed from
    int64_t coarse_dofs = -1, coarse_inc = -1;
    int coarse_refresh = AMR_COARSE_PER_DOF_MAP;  // when the coarse operator is formed again
    int32_t nagg = 0, ncoarse = 0, ncb = 0;
    int64_t npairs = 0;
    int32_t *c_agg = nullptr, *c_aggptr = nullptr, *c_aggnodes = nullptr, *c_pptr = nullptr, *c_pagg = nullptr;
    int32_t *c_cba = nullptr, *c_cbb = nullptr, *c_cbptr = nullptr, *c_cbslot = nullptr, *c_info = nullptr;
    double *c_cent = nullptr, *c_pair = nullptr, *Kinv = nullptr, *c_vec = nullptr, *y_vec = nullptr;
    double *c_work = nullptr;  // workspace of the coarse inversion
    // split NR finish: pinned landing area of the residual scalars and error
    // words, and the state between the two halves
    struct FinHost {
        double s[16];
        int32_t c[16];
        int32_t err[4];
    } *fin_h = nullptr;
    bool fin_pending = false;
    int fin_rc = 0;
    amr_nr_info fin_info{};
    // stage_begin's host vectors stay alive until their copies are done
    double *sb_pin = nullptr;  // pinned host staging of the stage targets (2 x 3 nnodes)
    size_t sb_pin_bytes = 0;
    bool sb_fresh = false;  // sb_event not recorded yet
    hipEvent_t sb_event = nullptr;
    hipEvent_t ev_batch = nullptr;  // the model's queued work, for a batched launch to wait on
    hipEvent_t ev_handoff = nullptr;   // orders the switch between its own and the pool's stream
    hipEvent_t fin_event = nullptr;    // end of the queued NR finish (its read-back)
    bool assembled = false;
    // device-resident stage vectors (node order): Uex Fex U F dFex R Rc dUa dUi dFin dUt x
    double *sv[12] = {};
    double *sv_block = nullptr;  // sv are slices of it
    double acr = 0.0;
    // SpMV timing with events on the model stream
    bool timing = false;
    std::vector<hipEvent_t> ev;
    int ev_used = 0;
    double t_spmv_ms = 0.0;
    int64_t n_spmv = 0;
    // SpMV lanes per block row (0 = not chosen yet)
    int lanes64 = 0, lanes32 = 0, lanes32d = 0;
    // output post-processing (recovery.hip)
    std::vector<double> Nip;        // nips*nn shape functions at the IPs
    int elem_kind = -1;             // material kind of every element, -2 if they differ
    std::vector<int8_t> ekind;      // material kind per element (models that mix kinds)
    struct AmrRecovery *rec = nullptr;
    // mass matrix (mount_M): one value per element node pair (the block is
    // that value times I3), assembled values and survivor flags in the kv_at
    // layout of K, shape functions at the IPs, densities
    double *Me = nullptr, *Mv = nullptr, *Nd = nullptr, *rho_d = nullptr;
    uint8_t *Mflag = nullptr;
    bool m_assembled = false;
    double *mvx = nullptr, *mvy = nullptr;   // 3*nnodes operands of amr_matvec
    double *Ksave = nullptr;                 // K as it was before the last amr_combine_KM
    int64_t ksave_version = -1;
    // predictor-corrector scheme of the NR iterations (AMR_SCHEME_*) and the
    // predictor's tangent, kept while the corrector's is assembled
    int scheme = AMR_SCHEME_FE;
    double *Kpred = nullptr;
    // element tilings (index into the shape's table; -1: the tuned or default one)
    int ke_var = -1, up_var = -1;
    int ke_path = AMR_KE_EXACT;   // element stiffness of the NR iterations (amr_set_ke_path)
};

enum { SV_UEX = 0, SV_FEX, SV_U, SV_F, SV_DFEX, SV_R, SV_RC, SV_DUA, SV_DUI, SV_DFIN, SV_DUT, SV_X };

// stage vector kernels (kernels_solve.hip)
int amr_vec_inc_begin(amr_model *m, double dT, double acr, int add);
int amr_vec_fill(amr_model *m, double *x, double v);
int amr_vec_fill_n(amr_model *m, double *x, int64_t n, double v);  // n values from x
int amr_vec_solve_prep(amr_model *m, const double *dUi, const double *R, double *u2, double *ky, double *b);
int amr_vec_combine(amr_model *m, const double *x, double *dUi);
int amr_vec_add(amr_model *m, const double *a, const double *b, double *out);
int amr_vec_scale(amr_model *m, double a, const double *x, double *out);
int amr_vec_nan_flag(amr_model *m, const double *v, int32_t *flag);  // *flag = any NaN in v
int amr_vec_nr_post(amr_model *m, double *res, double *dui_max, double *dua_max, int *nan_flag);
int amr_vec_commit(amr_model *m);
int amr_vec_nr_post_async(amr_model *m, double *hs, int32_t *hc);
int amr_nr_post_slots(int *res, int *errn, int *errd, int *nan);

// kernels (kernels_elem.hip)
int amr_launch_ke(amr_model *m);                 // the model's path (amr_set_ke_path)
int amr_launch_ke_path(amr_model *m, int path);  // AMR_KE_EXACT or AMR_KE_MFMA
int amr_elem_tiles(amr_model *m, int *ach, int *epw);
int amr_elem_tiles_set(amr_model *m, int ach, int epw);
int amr_elem_tiles_tune(amr_model *m, int *ach, int *epw);
int amr_launch_assemble(amr_model *m, bool flags);
int amr_launch_update(amr_model *m, const double *dU_node /*device, node order*/, double *dFin_node);
int amr_launch_mass(amr_model *m);  // Me from Nd, rho_d; then assembly into Mv, Mflag
int amr_launch_combine_km(amr_model *m, double a, double b);  // K := a K + b M, preconditioners refreshed
int amr_launch_combine_kk(amr_model *m, double a1, const double *K1, double a2);  // K := a1 K1 + a2 K

// Pinned host buffer and its device copy for the uploads of one stream: the
// host side is rewritten only after the previous upload from it completed.
struct AmrStaging {
    char *d = nullptr, *h = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;  // recorded after the last upload
    bool pending = false;
};
void amr_staging_free(AmrStaging &s);
int amr_staging_reserve(AmrStaging &s, hipStream_t st, size_t bytes);  // waits for the last upload; may grow

// The element phase of several models in one launch per kernel family (the
// same arithmetic as launching each model alone), queued on stream s:
// mount_K (error words reset, element stiffness by each model's path and
// tiles, assembly, Jacobi and nodal block-Jacobi inverses) and update_state!
// (IP updates by each model's tiles, then the internal-force gather)
int amr_batch_mount_K(amr_model *const *ms, int n, hipStream_t s, AmrStaging &stg);
int amr_batch_update(amr_model *const *ms, const double *const *dU, double *const *dFin, int n, hipStream_t s,
                     AmrStaging &stg);

// two-level coarse space (coarse.hip): the index structures are built on
// first use (host work, one wait); Kc^-1 is formed again when
// amr_coarse_due says so, for any number of models at once
// (amr_coarse_form: one launch per phase and block step for all of them,
// queued on stream s, after which their coarse operators are current)
int amr_coarse_build(amr_model *m);
bool amr_coarse_due(const amr_model *m);
int amr_coarse_form(hipStream_t s, amr_model *const *ms, int n, AmrStaging &stg);
void amr_coarse_free(amr_model *m);

// output post-processing (recovery.hip)
void amr_recovery_free(amr_model *m);

// solver (kernels_solve.hip)
int amr_pcg_solve(amr_model *m, const double *b, double *x, const amr_solve_opts *o, amr_solve_info *info);

// PCG pools (kernels_solve.hip)
struct amr_pool_result {
    amr_model *m;
    amr_solve_info info;
};
int amr_pool_new(amr_pool **out, int device, int cap, int check, hipStream_t stream /* null: own stream */);
void amr_pool_free(amr_pool *p);
int amr_pool_attach_model(amr_pool *p, amr_model *m);
int amr_pool_detach_model(amr_model *m);
int amr_pool_count(const amr_pool *p);
bool amr_pool_is_member(const amr_model *m);  // solving in its shared pool right now
int amr_pool_join(amr_pool *p, amr_model *m, const double *b, double *x, const amr_solve_opts *o);
int amr_pool_run(amr_pool *p, std::vector<amr_pool_result> &out);
int amr_pool_step_launch(amr_pool *p);
int amr_pool_step_collect(amr_pool *p, std::vector<amr_pool_result> &out);
int amr_pool_step_members(const amr_pool *p);  // members of the step in flight (0: none)
int amr_pool_step_query(amr_pool *p, int *done);
int amr_pool_hand_back(amr_pool *p);
int amr_pool_set_timing(amr_pool *p, int enable, double *ms, int64_t *n);
int amr_pool_graphs(amr_pool *p, int enable, int64_t *captures, int64_t *replays);
int amr_pool_counts(amr_pool *p, int64_t *out);
int amr_spmv_full(amr_model *m, const double *x, double *y, int mask_mode);
int amr_spmv_vals(amr_model *m, const double *vals /* kv_at layout */, const double *x, double *y, int mask_mode);
int amr_choose_lanes(amr_model *m);
int amr_tune_lanes(amr_model *m);
