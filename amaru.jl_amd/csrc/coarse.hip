// Coarse space of the two-level preconditioner (AMR_PRECOND_TWO_LEVEL).
//
// The nodes are split into aggregates (recursive coordinate bisection,
// model_build.cpp).  The coarse basis Phi holds, per aggregate, the six
// rigid-body modes of its nodes about the aggregate centroid (restricted to
// the free dofs: rows of prescribed dofs are zero).  The preconditioner is
// additive:  M^-1 r = D^-1 r + Phi Kc^-1 Phi^T r  with  Kc = Phi^T K Phi.
//
// Kc is formed on the device without atomics, in a fixed order:
//   pass 1, one thread per (block row i, aggregate b) pair:
//       U_ib = R_i^T  sum_{j in row i, agg(j) = b}  K_ij R_j       (6x6)
//     (R_j: 3x6 rigid-mode matrix of node j, masked), blocks in column order;
//   pass 2, one thread per (coarse block (a,b), entry): Kc_ab = sum over the
//     pairs (i in a, b) in ascending i.
// Kc (dense, n = 6 nagg <= 2048) is then inverted in place by block
// Gauss-Jordan elimination without pivoting (Kc is symmetric positive
// definite, so every pivot block is), 64x64 pivot blocks, three kernels per
// block step, every sum in a fixed order: the inverse is bit-reproducible
// whatever else runs on the device.  Every kernel works through a list of
// (model, index) entries, so the coarse operators of all models joining a
// pool at one step are formed by one launch per phase and block step.  It is symmetrised by mirroring its upper
// triangle; each PCG iteration applies Kc^-1 as a dense product.  A coarse
// matrix with a pivot that is not positive (degenerate aggregates) is
// replaced by zero, which leaves plain Jacobi.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "amr_internal.h"

namespace {

// 3x6 rigid-body-mode matrix of a node at offset d from the centroid, rows
// scaled by the node's free-dof mask
__device__ inline void rigid_modes(const double d[3], const double mk[3], double R[3][6])
{
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 6; q++) R[r][q] = 0.0;
    R[0][0] = R[1][1] = R[2][2] = 1.0;
    R[1][3] = -d[2];
    R[2][3] = d[1];
    R[0][4] = d[2];
    R[2][4] = -d[0];
    R[0][5] = -d[1];
    R[1][5] = d[0];
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 6; q++) R[r][q] *= mk[r];
}

// One model's coarse formation, as the batched kernels see it.  Every
// kernel below takes the jobs and a work list of (job, index) entries, so
// one launch serves any number of models; each entry does the same
// arithmetic as a launch for its model alone would.
struct CoarseJob {
    int32_t nc, nt, ncb;
    int64_t npairs, nrows;
    const int32_t *pptr, *pagg, *agg, *rowptr, *col, *cba, *cbb, *cbptr, *cbslot;
    const double *cent, *X, *mask, *Kv;
    double *pair, *Kc, *P, *R, *C;
    int32_t *info;
};

// Pointers taken from a job are generic; the buffers are all in the global
// address space, and saying so lets the compiler use global (not flat)
// memory instructions.
template <typename T>
__device__ inline const __attribute__((address_space(1))) T *cg(const T *p)
{
    return (const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ inline __attribute__((address_space(1))) T *cgw(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

// entry (job, g): block rows 16 g .. 16 g + 15, four consecutive rows per
// wave.  Pair s = (row i, aggregate
// b) of the row's pairs: U = R_i^T sum_{j in row i, agg(j) = b} K_ij R_j,
// 36 values.  Lane l takes the row's blocks l, l + 64, ... (in that order)
// and the wave sums the lanes' 3 x 6 products by a fixed tree.
__global__ void __launch_bounds__(256) galerkin_pairs_kernel(const CoarseJob *__restrict__ J,
                                                             const int2 *__restrict__ wl)
{
    const int2 e = wl[blockIdx.x];
    const CoarseJob &q = J[e.x];
    const int lane = threadIdx.x % 64;
    const auto *__restrict__ agg = cg(q.agg);
    const auto *__restrict__ X = cg(q.X), *__restrict__ cent = cg(q.cent), *__restrict__ mask = cg(q.mask);
    const auto *__restrict__ rowptr = cg(q.rowptr), *__restrict__ col = cg(q.col);
    const auto *__restrict__ Kv = cg(q.Kv);
    for (int64_t i = (int64_t)e.y * 16 + (threadIdx.x / 64) * 4, i1 = i + 4; i < i1 && i < q.nrows; i++) {
        const int32_t k0 = rowptr[i], k1 = rowptr[i + 1], s0 = cg(q.pptr)[i], s1 = cg(q.pptr)[i + 1];
        for (int32_t s = s0; s < s1; s++) {
            const int32_t b = cg(q.pagg)[s];
            double W[3][6];
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 6; c++) W[r][c] = 0.0;
            for (int32_t k = k0 + lane; k < k1; k += 64) {
                const int32_t j = col[k];
                if (agg[j] != b) continue;
                double d[3], mk[3], R[3][6], Kb[9];
                for (int c = 0; c < 3; c++) {
                    d[c] = X[(int64_t)j * 3 + c] - cent[(int64_t)b * 3 + c];
                    mk[c] = mask[(int64_t)j * 3 + c];
                }
                rigid_modes(d, mk, R);
                for (int t = 0; t < 9; t++) Kb[t] = Kv[kv_at(k, t)];
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 6; c++)
                        W[r][c] = fma(Kb[r * 3 + 2], R[2][c],
                                      fma(Kb[r * 3 + 1], R[1][c], fma(Kb[r * 3], R[0][c], W[r][c])));
            }
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 6; c++)
                    for (int off = 32; off > 0; off >>= 1) W[r][c] += __shfl_xor(W[r][c], off, 64);
            if (lane == 0) {
                const int32_t a = agg[i];
                double d[3], mk[3], R[3][6];
                for (int c = 0; c < 3; c++) {
                    d[c] = X[i * 3 + c] - cent[(int64_t)a * 3 + c];
                    mk[c] = mask[i * 3 + c];
                }
                rigid_modes(d, mk, R);
                for (int p = 0; p < 6; p++)
                    for (int c = 0; c < 6; c++)
                        cgw(q.pair)[(int64_t)s * 36 + p * 6 + c] =
                            fma(R[2][p], W[2][c], fma(R[1][p], W[1][c], R[0][p] * W[0][c]));
            }
        }
    }
}

// entry (job, g): coarse blocks cb = 4 g + wave.  Kc_ab is the sum of the
// block's pairs: lane l takes pairs l, l + 64, ... of the block's list (in
// that order), the wave sums the lanes' 36 values by a fixed tree, and
// lane c < 36 writes entry c.
__global__ void __launch_bounds__(256) galerkin_blocks_kernel(const CoarseJob *__restrict__ J,
                                                              const int2 *__restrict__ wl)
{
    const int2 e = wl[blockIdx.x];
    const CoarseJob &q = J[e.x];
    const int32_t cb = e.y * 4 + (int32_t)(threadIdx.x / 64);
    const int lane = threadIdx.x % 64;
    if (cb >= q.ncb) return;  // the whole wave
    const auto *__restrict__ cbptr = cg(q.cbptr), *__restrict__ cbslot = cg(q.cbslot);
    const auto *__restrict__ pair = cg(q.pair);
    double acc[36];
    for (int c = 0; c < 36; c++) acc[c] = 0.0;
    for (int32_t k = cbptr[cb] + lane; k < cbptr[cb + 1]; k += 64) {
        const auto *pv = pair + (int64_t)cbslot[k] * 36;
        for (int c = 0; c < 36; c++) acc[c] += pv[c];
    }
    double out = 0.0;
    for (int c = 0; c < 36; c++) {
        for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
        if (lane == c) out = acc[c];
    }
    if (lane < 36)
        cgw(q.Kc)[(int64_t)(6 * cg(q.cba)[cb] + lane / 6) * q.nc + 6 * cg(q.cbb)[cb] + lane % 6] = out;
}

// entry (job, b): Kc[4096 b .. 4096 (b + 1)) = 0
__global__ void __launch_bounds__(256) zero_kernel(const CoarseJob *__restrict__ J, const int2 *__restrict__ wl)
{
    const int2 e = wl[blockIdx.x];
    const CoarseJob &q = J[e.x];
    const int64_t n = (int64_t)q.nc * q.nc, b0 = (int64_t)e.y * 4096;
    const int64_t b1 = n < b0 + 4096 ? n : b0 + 4096;
    auto *__restrict__ Kc = cgw(q.Kc);
    for (int64_t i = b0 + threadIdx.x; i < b1; i += 256) Kc[i] = 0.0;
}

// entry (job, b): coarse dof k = 256 b + thread.  Coarse dofs whose basis
// vector vanishes (all dofs of the aggregate prescribed) get a unit
// diagonal; their right-hand side is always zero
__global__ void __launch_bounds__(256) fix_diag_kernel(const CoarseJob *__restrict__ J, const int2 *__restrict__ wl)
{
    const int2 e = wl[blockIdx.x];
    const CoarseJob &q = J[e.x];
    const int32_t k = e.y * 256 + (int32_t)threadIdx.x;
    auto *__restrict__ Kc = cgw(q.Kc);
    if (k == 0) cgw(q.info)[0] = cgw(q.info)[1] = 0;
    if (k < q.nc && Kc[(int64_t)k * q.nc + k] == 0.0) Kc[(int64_t)k * q.nc + k] = 1.0;
}

// entry (job, b): entry t = 256 b + thread of the inverse.  Mirror the upper
// triangle onto the lower one (the elimination leaves them equal up to
// rounding), or clear everything if a pivot failed
__global__ void __launch_bounds__(256) symmetrize_kernel(const CoarseJob *__restrict__ J, const int2 *__restrict__ wl)
{
    const int2 e = wl[blockIdx.x];
    const CoarseJob &q = J[e.x];
    const int32_t nc = q.nc;
    const int64_t t = (int64_t)e.y * 256 + threadIdx.x;
    if (t >= (int64_t)nc * nc) return;
    const int32_t i = (int32_t)(t / nc), j = (int32_t)(t % nc);
    auto *__restrict__ Kc = cgw(q.Kc);
    if (cg(q.info)[0] != 0) {
        Kc[t] = 0.0;
        return;
    }
    if (i < j) Kc[(int64_t)j * nc + i] = Kc[(int64_t)i * nc + j];
}

// ---- block Gauss-Jordan inversion (in place, row-major nc x nc) ----------
// Step k, pivot block K = [k0, k0 + b):
//   P    = A_KK^-1                              (gj_pivot, one workgroup)
//   R    = P A_Kj for j outside K; C = A_iK     (gj_panel, old column panel)
//   A_ij -= C_i R_j,  A_Kj = R_j,  A_iK = -C_i P,  A_KK = P      (gj_update)
// A launch does step k for every listed model with more than k block rows.
constexpr int GJB = 64;

// entry (job, -): the pivot block of step k
__global__ void __launch_bounds__(1024) gj_pivot_kernel(const CoarseJob *__restrict__ J, const int2 *__restrict__ wl,
                                                        int32_t k)
{
    const CoarseJob &q = J[wl[blockIdx.x].x];
    const int32_t nc = q.nc, k0 = k * GJB, b = nc - k0 < GJB ? nc - k0 : GJB;
    const auto *__restrict__ A = cg(q.Kc);
    auto *__restrict__ P = cgw(q.P);
    // the pivot block in LDS; thread (i, j0) updates entries (i, j0..j0+3):
    // each step reads the pivot row and column, then (after a barrier)
    // writes the new values
    __shared__ double M[GJB][GJB + 1];
    const int i = threadIdx.x >> 4, j0 = (threadIdx.x & 15) * 4;
    for (int c = 0; c < 4; c++) M[i][j0 + c] = (i < b && j0 + c < b) ? A[(int64_t)(k0 + i) * nc + k0 + j0 + c] : 0.0;
    __syncthreads();
    for (int p = 0; p < b; p++) {
        const double piv = M[p][p];
        const bool ok = piv > 0.0 && std::isfinite(piv);
        if (!ok && threadIdx.x == 0) cgw(q.info)[0] = 1;
        const double ip = ok ? 1.0 / piv : 0.0;
        const double ci = M[i][p];
        double nv[4];
        for (int c = 0; c < 4; c++) {
            const int j = j0 + c;
            const double rj = M[p][j], v = M[i][j];
            if (i == p && j == p) nv[c] = ip;
            else if (i == p) nv[c] = rj * ip;
            else if (j == p) nv[c] = -(ci * ip);
            else nv[c] = v - ci * (rj * ip);
        }
        __syncthreads();
        for (int c = 0; c < 4; c++) M[i][j0 + c] = nv[c];
        __syncthreads();
    }
    for (int c = 0; c < 4; c++)
        if (i < b && j0 + c < b) P[i * GJB + j0 + c] = M[i][j0 + c];
}

// entry (job, t), one per 64-column tile t: R[:, tile] = P A_K,tile (t != k),
// and C[tile rows, :] = A_tile,K
__global__ void __launch_bounds__(256) gj_panel_kernel(const CoarseJob *__restrict__ J, const int2 *__restrict__ wl,
                                                       int32_t k)
{
    const int2 e = wl[blockIdx.x];
    const CoarseJob &q = J[e.x];
    const int32_t nc = q.nc, t = e.y, k0 = k * GJB, c0 = t * GJB, b = nc - k0 < GJB ? nc - k0 : GJB;
    const auto *__restrict__ A = cg(q.Kc), *__restrict__ P = cg(q.P);
    auto *__restrict__ R = cgw(q.R), *__restrict__ C = cgw(q.C);
    __shared__ double Ps[GJB][GJB + 1];
    __shared__ double As[GJB][GJB + 1];
    const int w = std::min(GJB, nc - c0);  // columns (and rows) of this tile
    const int r = threadIdx.x >> 2, q0 = (threadIdx.x & 3) * 16;  // 64 rows x 4 quarters
    if (r < w)
        for (int c = q0; c < q0 + 16 && c < b; c++) C[(int64_t)(c0 + r) * GJB + c] = A[(int64_t)(c0 + r) * nc + k0 + c];
    if (t == k) return;
    if (r < b)
        for (int c = q0; c < q0 + 16; c++) {
            if (c < b) Ps[r][c] = P[r * GJB + c];
            if (c < w) As[r][c] = A[(int64_t)(k0 + r) * nc + c0 + c];
        }
    __syncthreads();
    if (r < b)
        for (int c = q0; c < q0 + 16 && c < w; c++) {
            double acc = 0.0;
            for (int l = 0; l < b; l++) acc = fma(Ps[r][l], As[l][c], acc);
            R[(int64_t)r * nc + c0 + c] = acc;
        }
}

// entry (job, ti * nt + tj), one per 64x64 tile (ti, tj) of A, 4x4 entries
// per thread
__global__ void __launch_bounds__(256) gj_update_kernel(const CoarseJob *__restrict__ J, const int2 *__restrict__ wl,
                                                        int32_t k)
{
    const int2 e = wl[blockIdx.x];
    const CoarseJob &q = J[e.x];
    const int32_t nc = q.nc, ti = e.y / q.nt, tj = e.y % q.nt, k0 = k * GJB;
    const int32_t b = nc - k0 < GJB ? nc - k0 : GJB;
    auto *__restrict__ A = cgw(q.Kc);
    const auto *__restrict__ P = cg(q.P), *__restrict__ R = cg(q.R), *__restrict__ C = cg(q.C);
    __shared__ double Cs[GJB][GJB + 1];
    __shared__ double Bs[GJB][GJB + 1];
    const int r0 = ti * GJB, c0 = tj * GJB;
    const int hr = std::min(GJB, nc - r0), hc = std::min(GJB, nc - c0);
    const int rr = 4 * (threadIdx.x / 16), cc = 4 * (threadIdx.x % 16);
    const int lr = threadIdx.x >> 2, lq = (threadIdx.x & 3) * 16;  // load mapping: 64 rows x 4 quarters
    if (ti == k) {
        // pivot block row: A_KK = P, A_Kj = R_j
        if (lr < b)
            for (int c = lq; c < lq + 16 && c < hc; c++)
                A[(int64_t)(k0 + lr) * nc + c0 + c] = tj == k ? P[lr * GJB + c] : R[(int64_t)lr * nc + c0 + c];
        return;
    }
    if (lr < hr)
        for (int c = lq; c < lq + 16 && c < b; c++) Cs[lr][c] = C[(int64_t)(r0 + lr) * GJB + c];
    if (lr < b)
        for (int c = lq; c < lq + 16 && c < hc; c++)
            Bs[lr][c] = tj == k ? P[lr * GJB + c] : R[(int64_t)lr * nc + c0 + c];
    __syncthreads();
    double acc[4][4] = {};
    for (int l = 0; l < b; l++) {
        double cv[4], bv[4];
        for (int s = 0; s < 4; s++) {
            cv[s] = Cs[rr + s][l];
            bv[s] = Bs[l][cc + s];
        }
        for (int x = 0; x < 4; x++)
            for (int y = 0; y < 4; y++) acc[x][y] = fma(cv[x], bv[y], acc[x][y]);
    }
    for (int x = 0; x < 4; x++)
        for (int y = 0; y < 4; y++) {
            const int r = rr + x, c = cc + y;
            if (r >= hr || c >= hc) continue;
            auto *a = A + (int64_t)(r0 + r) * nc + c0 + c;
            *a = tj == k ? -acc[x][y] : *a - acc[x][y];
        }
}

template <typename T>
int dev_upload(amr_model *m, T **d, const std::vector<T> &h)
{
    const size_t n = std::max<size_t>(h.size(), 1);
    if (int rc = amr_dalloc(m, (void **)d, n * sizeof(T))) return rc;
    if (!h.empty()) AMR_HIP(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, m->stream));
    return AMR_OK;
}

template <typename T>
int dev_alloc(amr_model *m, T **d, size_t n)
{
    return amr_dalloc(m, (void **)d, std::max<size_t>(n, 1) * sizeof(T));
}

int build(amr_model *m)
{
    const int64_t n = m->nnodes;
    const int target = (int)std::max<int64_t>(m->coarse_target, (n + m->coarse_max - 1) / m->coarse_max);
    AmrCoarse C;
    if (int rc = amr_pattern_host(m)) return rc;
    if (int rc = amr_build_coarse(n, m->Xh.data(), m->pat.rowptr.data(), m->pat.col.data(), target, C)) return rc;
    m->nagg = C.nagg;
    m->ncoarse = 6 * C.nagg;
    m->npairs = (int64_t)C.pagg.size();
    m->ncb = (int32_t)C.cba.size();
    const size_t nc = (size_t)m->ncoarse;
    if (int rc = dev_upload(m, &m->c_agg, C.agg)) return rc;
    if (int rc = dev_upload(m, &m->c_aggptr, C.aggptr)) return rc;
    if (int rc = dev_upload(m, &m->c_aggnodes, C.aggnodes)) return rc;
    if (int rc = dev_upload(m, &m->c_pptr, C.pptr)) return rc;
    if (int rc = dev_upload(m, &m->c_pagg, C.pagg)) return rc;
    if (int rc = dev_upload(m, &m->c_cba, C.cba)) return rc;
    if (int rc = dev_upload(m, &m->c_cbb, C.cbb)) return rc;
    if (int rc = dev_upload(m, &m->c_cbptr, C.cbptr)) return rc;
    if (int rc = dev_upload(m, &m->c_cbslot, C.cbslot)) return rc;
    if (int rc = dev_upload(m, &m->c_cent, C.cent)) return rc;
    // written in full by the setup and solver kernels before any read
    if (int rc = dev_alloc(m, &m->c_pair, (size_t)m->npairs * 36)) return rc;
    if (int rc = dev_alloc(m, &m->Kinv, nc * nc)) return rc;
    if (int rc = dev_alloc(m, &m->c_vec, nc)) return rc;
    if (int rc = dev_alloc(m, &m->y_vec, nc)) return rc;
    if (int rc = dev_alloc(m, &m->c_info, 2)) return rc;
    // Gauss-Jordan workspace: P (64 x 64), R (64 x nc), C (nc x 64)
    if (int rc = dev_alloc(m, &m->c_work, (size_t)GJB * GJB + 2 * (size_t)GJB * nc)) return rc;
    // C's vectors are released on return; the copies read them
    // asynchronously, so wait here
    AMR_HIP(hipStreamSynchronize(m->stream));
    m->coarse_built = true;
    return AMR_OK;
}

}  // namespace

int amr_coarse_build(amr_model *m)
{
    if (m->coarse_built) return AMR_OK;
    return build(m);
}

bool amr_coarse_due(const amr_model *m)
{
    // The coarse operator only preconditions: it is formed from the K of the
    // first assembly of each DOF map, and for models with plastic materials
    // again at the first solve of each load increment (a linear elastic K
    // does not change between assemblies)
    if (m->coarse_version == m->k_version) return false;
    const bool same_inc = m->all_elastic || m->coarse_inc == m->inc_version;
    return !(m->coarse_version >= 0 && m->coarse_dofs == m->dof_version &&
             (m->coarse_refresh == AMR_COARSE_PER_DOF_MAP ||
              (m->coarse_refresh == AMR_COARSE_PER_INCREMENT && same_inc)));
}

void amr_staging_free(AmrStaging &s)
{
    if (s.pending && s.ev) (void)hipEventSynchronize(s.ev);
    if (s.d) (void)hipFree(s.d);
    if (s.h) (void)hipHostFree(s.h);
    if (s.ev) (void)hipEventDestroy(s.ev);
    s = AmrStaging();
}

// room for `bytes` in the staging buffers of stream st (the host side free
// to write)
int amr_staging_reserve(AmrStaging &s, hipStream_t st, size_t bytes)
{
    if (!s.ev) AMR_HIP(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
    if (s.pending) {
        AMR_HIP(hipEventSynchronize(s.ev));
        s.pending = false;
    }
    if (bytes <= s.cap) return AMR_OK;
    // kernels queued earlier on st may still read the device buffer
    AMR_HIP(hipStreamSynchronize(st));
    if (s.d) AMR_HIP(hipFree(s.d));
    if (s.h) AMR_HIP(hipHostFree(s.h));
    s.d = s.h = nullptr;
    s.cap = 0;
    const size_t cap = std::max<size_t>(bytes + bytes / 2, 1 << 16);
    AMR_HIP(hipMalloc(&s.d, cap));
    AMR_HIP(hipHostMalloc(&s.h, cap, hipHostMallocDefault));
    s.cap = cap;
    return AMR_OK;
}

int amr_coarse_form(hipStream_t s, amr_model *const *ms, int n, AmrStaging &stg)
{
    if (n <= 0) return AMR_OK;
    std::vector<CoarseJob> jobs(n);
    std::vector<int2> zl, pl, bl, fl, sl;
    int maxnt = 0;
    for (int j = 0; j < n; j++) {
        amr_model *m = ms[j];
        if (!m->coarse_built)
            if (int rc = build(m)) return rc;
        CoarseJob &q = jobs[j];
        q.nc = m->ncoarse;
        q.nt = (q.nc + GJB - 1) / GJB;
        q.ncb = m->ncb;
        q.npairs = m->npairs;
        q.nrows = m->nnodes;
        q.pptr = m->c_pptr;
        q.pagg = m->c_pagg;
        q.agg = m->c_agg;
        q.rowptr = m->rowptr;
        q.col = m->col;
        q.cba = m->c_cba;
        q.cbb = m->c_cbb;
        q.cbptr = m->c_cbptr;
        q.cbslot = m->c_cbslot;
        q.cent = m->c_cent;
        q.X = m->X;
        q.mask = m->mask;
        q.Kv = m->Kv;
        q.pair = m->c_pair;
        q.Kc = m->Kinv;
        q.P = m->c_work;
        q.R = q.P + GJB * GJB;
        q.C = q.R + (size_t)GJB * q.nc;
        q.info = m->c_info;
        maxnt = std::max(maxnt, q.nt);
        const int64_t nn = (int64_t)q.nc * q.nc;
        for (int64_t b = 0; b < (nn + 4095) / 4096; b++) zl.push_back(make_int2(j, (int)b));
        for (int64_t g = 0; g < (q.nrows + 15) / 16; g++) pl.push_back(make_int2(j, (int)g));
        for (int64_t g = 0; g < ((int64_t)q.ncb + 3) / 4; g++) bl.push_back(make_int2(j, (int)g));
        for (int64_t b = 0; b < (q.nc + 255) / 256; b++) fl.push_back(make_int2(j, (int)b));
        for (int64_t b = 0; b < (nn + 255) / 256; b++) sl.push_back(make_int2(j, (int)b));
    }
    // block step k: the pivots, panels and tiles of the models with more
    // than k block rows
    std::vector<std::vector<int2>> pv(maxnt), pn(maxnt), up(maxnt);
    for (int k = 0; k < maxnt; k++)
        for (int j = 0; j < n; j++) {
            const int nt = jobs[j].nt;
            if (k >= nt) continue;
            pv[k].push_back(make_int2(j, 0));
            for (int t = 0; t < nt; t++) pn[k].push_back(make_int2(j, t));
            for (int t = 0; t < nt * nt; t++) up[k].push_back(make_int2(j, t));
        }
    // one upload: the jobs, then every list (16-byte aligned offsets)
    size_t off = 0;
    auto place = [&off](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 15) & ~(size_t)15;
        return o;
    };
    const size_t o_jobs = place(jobs.size() * sizeof(CoarseJob));
    const size_t o_z = place(zl.size() * sizeof(int2)), o_p = place(pl.size() * sizeof(int2)),
                 o_b = place(bl.size() * sizeof(int2)), o_f = place(fl.size() * sizeof(int2)),
                 o_s = place(sl.size() * sizeof(int2));
    std::vector<size_t> o_pv(maxnt), o_pn(maxnt), o_up(maxnt);
    for (int k = 0; k < maxnt; k++) {
        o_pv[k] = place(pv[k].size() * sizeof(int2));
        o_pn[k] = place(pn[k].size() * sizeof(int2));
        o_up[k] = place(up[k].size() * sizeof(int2));
    }
    if (int rc = amr_staging_reserve(stg, s, off)) return rc;
    auto put = [&stg](size_t o, const void *src, size_t bytes) {
        if (bytes) std::memcpy(stg.h + o, src, bytes);
    };
    put(o_jobs, jobs.data(), jobs.size() * sizeof(CoarseJob));
    put(o_z, zl.data(), zl.size() * sizeof(int2));
    put(o_p, pl.data(), pl.size() * sizeof(int2));
    put(o_b, bl.data(), bl.size() * sizeof(int2));
    put(o_f, fl.data(), fl.size() * sizeof(int2));
    put(o_s, sl.data(), sl.size() * sizeof(int2));
    for (int k = 0; k < maxnt; k++) {
        put(o_pv[k], pv[k].data(), pv[k].size() * sizeof(int2));
        put(o_pn[k], pn[k].data(), pn[k].size() * sizeof(int2));
        put(o_up[k], up[k].data(), up[k].size() * sizeof(int2));
    }
    AMR_HIP(hipMemcpyAsync(stg.d, stg.h, off, hipMemcpyHostToDevice, s));
    AMR_HIP(hipEventRecord(stg.ev, s));
    stg.pending = true;
    const CoarseJob *dj = (const CoarseJob *)(stg.d + o_jobs);
    auto wl = [&stg](size_t o) { return (const int2 *)(stg.d + o); };
    auto grid = [](size_t n_) { return dim3((unsigned)n_); };
    if (!zl.empty()) hipLaunchKernelGGL(zero_kernel, grid(zl.size()), dim3(256), 0, s, dj, wl(o_z));
    if (!pl.empty()) hipLaunchKernelGGL(galerkin_pairs_kernel, grid(pl.size()), dim3(256), 0, s, dj, wl(o_p));
    if (!bl.empty()) hipLaunchKernelGGL(galerkin_blocks_kernel, grid(bl.size()), dim3(256), 0, s, dj, wl(o_b));
    if (!fl.empty()) hipLaunchKernelGGL(fix_diag_kernel, grid(fl.size()), dim3(256), 0, s, dj, wl(o_f));
    AMR_HIP(hipGetLastError());
    for (int k = 0; k < maxnt; k++) {
        hipLaunchKernelGGL(gj_pivot_kernel, grid(pv[k].size()), dim3(1024), 0, s, dj, wl(o_pv[k]), k);
        hipLaunchKernelGGL(gj_panel_kernel, grid(pn[k].size()), dim3(256), 0, s, dj, wl(o_pn[k]), k);
        hipLaunchKernelGGL(gj_update_kernel, grid(up[k].size()), dim3(256), 0, s, dj, wl(o_up[k]), k);
    }
    AMR_HIP(hipGetLastError());
    if (!sl.empty()) hipLaunchKernelGGL(symmetrize_kernel, grid(sl.size()), dim3(256), 0, s, dj, wl(o_s));
    AMR_HIP(hipGetLastError());
    for (int j = 0; j < n; j++) {
        amr_model *m = ms[j];
        m->coarse_version = m->k_version;
        m->coarse_dofs = m->dof_version;
        m->coarse_inc = m->inc_version;
    }
    return AMR_OK;
}

void amr_coarse_free(amr_model *m)
{
    void **ptrs[] = {(void **)&m->c_agg,  (void **)&m->c_aggptr, (void **)&m->c_aggnodes, (void **)&m->c_pptr,
                     (void **)&m->c_pagg, (void **)&m->c_cba,    (void **)&m->c_cbb,      (void **)&m->c_cbptr,
                     (void **)&m->c_cbslot, (void **)&m->c_info, (void **)&m->c_cent,     (void **)&m->c_pair,
                     (void **)&m->Kinv,   (void **)&m->c_vec,    (void **)&m->y_vec,      (void **)&m->c_work};
    for (void **p : ptrs) {
        amr_release(m, *p);
        *p = nullptr;
    }
    m->coarse_built = false;
    m->coarse_version = -1;
}
