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
// whatever else runs on the device.  It is symmetrised by mirroring its upper
// triangle; each PCG iteration applies Kc^-1 as a dense product.  A coarse
// matrix with a pivot that is not positive (degenerate aggregates) is
// replaced by zero, which leaves plain Jacobi.
#include <algorithm>
#include <cmath>

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

__global__ void __launch_bounds__(256) galerkin_pairs_kernel(int64_t npairs, const int32_t *__restrict__ prow,
                                                             const int32_t *__restrict__ pagg,
                                                             const int32_t *__restrict__ agg,
                                                             const double *__restrict__ cent,
                                                             const double *__restrict__ X,
                                                             const double *__restrict__ mask,
                                                             const int32_t *__restrict__ rowptr,
                                                             const int32_t *__restrict__ col,
                                                             const double *__restrict__ Kv, double *__restrict__ out)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npairs) return;
    const int32_t i = prow[s], b = pagg[s];
    double W[3][6];
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 6; q++) W[r][q] = 0.0;
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
        const int32_t j = col[k];
        if (agg[j] != b) continue;
        double d[3], mk[3], R[3][6], Kb[9];
        for (int c = 0; c < 3; c++) {
            d[c] = X[(int64_t)j * 3 + c] - cent[(int64_t)b * 3 + c];
            mk[c] = mask[(int64_t)j * 3 + c];
        }
        rigid_modes(d, mk, R);
        for (int e = 0; e < 9; e++) Kb[e] = Kv[kv_at(k, e)];
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 6; q++)
                W[r][q] = fma(Kb[r * 3 + 2], R[2][q], fma(Kb[r * 3 + 1], R[1][q], fma(Kb[r * 3], R[0][q], W[r][q])));
    }
    const int32_t a = agg[i];
    double d[3], mk[3], R[3][6];
    for (int c = 0; c < 3; c++) {
        d[c] = X[(int64_t)i * 3 + c] - cent[(int64_t)a * 3 + c];
        mk[c] = mask[(int64_t)i * 3 + c];
    }
    rigid_modes(d, mk, R);
    for (int p = 0; p < 6; p++)
        for (int q = 0; q < 6; q++)
            out[s * 36 + p * 6 + q] = fma(R[2][p], W[2][q], fma(R[1][p], W[1][q], R[0][p] * W[0][q]));
}

__global__ void __launch_bounds__(256) galerkin_blocks_kernel(int32_t ncb, int32_t nc, const int32_t *__restrict__ cba,
                                                              const int32_t *__restrict__ cbb,
                                                              const int32_t *__restrict__ cbptr,
                                                              const int32_t *__restrict__ cbslot,
                                                              const double *__restrict__ pair, double *__restrict__ Kc)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)ncb * 36) return;
    const int32_t cb = (int32_t)(t / 36), e = (int32_t)(t % 36);
    double acc = 0.0;
    for (int32_t k = cbptr[cb]; k < cbptr[cb + 1]; k++) acc += pair[(int64_t)cbslot[k] * 36 + e];
    Kc[(int64_t)(6 * cba[cb] + e / 6) * nc + 6 * cbb[cb] + e % 6] = acc;
}

__global__ void zero_kernel(int64_t n, double *__restrict__ x)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = 0.0;
}

// coarse dofs whose basis vector vanishes (all dofs of the aggregate
// prescribed) get a unit diagonal; their right-hand side is always zero
__global__ void fix_diag_kernel(int32_t nc, double *__restrict__ Kc, int32_t *__restrict__ info)
{
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) info[0] = info[1] = 0;
    if (k < nc && Kc[(int64_t)k * nc + k] == 0.0) Kc[(int64_t)k * nc + k] = 1.0;
}

// Mirror the upper triangle of the inverse onto the lower one (the
// elimination leaves them equal up to rounding), or clear everything if a
// pivot failed
__global__ void symmetrize_kernel(int32_t nc, double *__restrict__ A, const int32_t *__restrict__ info)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nc * nc) return;
    const int32_t i = (int32_t)(t / nc), j = (int32_t)(t % nc);
    if (info[0] != 0) {
        A[t] = 0.0;
        return;
    }
    if (i < j) A[(int64_t)j * nc + i] = A[(int64_t)i * nc + j];
}

// ---- block Gauss-Jordan inversion (in place, row-major nc x nc) ----------
// Step k, pivot block K = [k0, k0 + b):
//   P    = A_KK^-1                              (gj_pivot, one workgroup)
//   R    = P A_Kj for j outside K; C = A_iK     (gj_panel, old column panel)
//   A_ij -= C_i R_j,  A_Kj = R_j,  A_iK = -C_i P,  A_KK = P      (gj_update)
constexpr int GJB = 64;

__global__ void __launch_bounds__(1024) gj_pivot_kernel(int32_t nc, int32_t k0, int32_t b,
                                                        const double *__restrict__ A, double *__restrict__ P,
                                                        int32_t *__restrict__ info)
{
    // the pivot block in LDS; thread (i, j0) updates entries (i, j0..j0+3):
    // each step reads the pivot row and column, then (after a barrier)
    // writes the new values
    __shared__ double M[GJB][GJB + 1];
    const int i = threadIdx.x >> 4, j0 = (threadIdx.x & 15) * 4;
    for (int k = 0; k < 4; k++) M[i][j0 + k] = (i < b && j0 + k < b) ? A[(int64_t)(k0 + i) * nc + k0 + j0 + k] : 0.0;
    __syncthreads();
    for (int p = 0; p < b; p++) {
        const double piv = M[p][p];
        const bool ok = piv > 0.0 && std::isfinite(piv);
        if (!ok && threadIdx.x == 0) info[0] = 1;
        const double ip = ok ? 1.0 / piv : 0.0;
        const double ci = M[i][p];
        double nv[4];
        for (int k = 0; k < 4; k++) {
            const int j = j0 + k;
            const double rj = M[p][j], v = M[i][j];
            if (i == p && j == p) nv[k] = ip;
            else if (i == p) nv[k] = rj * ip;
            else if (j == p) nv[k] = -(ci * ip);
            else nv[k] = v - ci * (rj * ip);
        }
        __syncthreads();
        for (int k = 0; k < 4; k++) M[i][j0 + k] = nv[k];
        __syncthreads();
    }
    for (int k = 0; k < 4; k++)
        if (i < b && j0 + k < b) P[i * GJB + j0 + k] = M[i][j0 + k];
}

// one workgroup per 64-column tile t: R[:, tile] = P A_K,tile (t != k), and
// C[tile rows, :] = A_tile,K
__global__ void __launch_bounds__(256) gj_panel_kernel(int32_t nc, int32_t k, int32_t b,
                                                       const double *__restrict__ A, const double *__restrict__ P,
                                                       double *__restrict__ R, double *__restrict__ C)
{
    __shared__ double Ps[GJB][GJB + 1];
    __shared__ double As[GJB][GJB + 1];
    const int t = blockIdx.x, k0 = k * GJB, c0 = t * GJB;
    const int w = std::min(GJB, nc - c0);  // columns (and rows) of this tile
    const int r = threadIdx.x >> 2, q0 = (threadIdx.x & 3) * 16;  // 64 rows x 4 quarters
    if (r < w)
        for (int q = q0; q < q0 + 16 && q < b; q++) C[(int64_t)(c0 + r) * GJB + q] = A[(int64_t)(c0 + r) * nc + k0 + q];
    if (t == k) return;
    if (r < b)
        for (int q = q0; q < q0 + 16; q++) {
            if (q < b) Ps[r][q] = P[r * GJB + q];
            if (q < w) As[r][q] = A[(int64_t)(k0 + r) * nc + c0 + q];
        }
    __syncthreads();
    if (r < b)
        for (int q = q0; q < q0 + 16 && q < w; q++) {
            double acc = 0.0;
            for (int l = 0; l < b; l++) acc = fma(Ps[r][l], As[l][q], acc);
            R[(int64_t)r * nc + c0 + q] = acc;
        }
}

// one workgroup per 64x64 tile (ti, tj) of A, 4x4 entries per thread
__global__ void __launch_bounds__(256) gj_update_kernel(int32_t nc, int32_t k, int32_t b, double *__restrict__ A,
                                                        const double *__restrict__ P, const double *__restrict__ R,
                                                        const double *__restrict__ C)
{
    __shared__ double Cs[GJB][GJB + 1];
    __shared__ double Bs[GJB][GJB + 1];
    const int tj = blockIdx.x, ti = blockIdx.y, k0 = k * GJB;
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
        for (int q = 0; q < 4; q++) {
            cv[q] = Cs[rr + q][l];
            bv[q] = Bs[l][cc + q];
        }
        for (int x = 0; x < 4; x++)
            for (int y = 0; y < 4; y++) acc[x][y] = fma(cv[x], bv[y], acc[x][y]);
    }
    for (int x = 0; x < 4; x++)
        for (int y = 0; y < 4; y++) {
            const int r = rr + x, c = cc + y;
            if (r >= hr || c >= hc) continue;
            double *a = A + (int64_t)(r0 + r) * nc + c0 + c;
            *a = tj == k ? -acc[x][y] : *a - acc[x][y];
        }
}

// One rocBLAS handle per device, shared by every model.  The handle's
// device workspace is shared too, so its uses are chained on the device
// with an event: each factorisation waits for the previous one, whatever
// stream either ran on.
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
    if (int rc = dev_upload(m, &m->c_prow, C.prow)) return rc;
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

int amr_coarse_setup(amr_model *m)
{
    if (!m->coarse_built)
        if (int rc = build(m)) return rc;
    // The coarse operator only preconditions: it is formed from the K of the
    // first assembly of each DOF map, and for models with plastic materials
    // again at the first solve of each load increment (a linear elastic K
    // does not change between assemblies)
    if (m->coarse_version == m->k_version) return AMR_OK;
    const bool same_inc = m->all_elastic || m->coarse_inc == m->inc_version;
    if (m->coarse_version >= 0 && m->coarse_dofs == m->dof_version &&
        (m->coarse_refresh == AMR_COARSE_PER_DOF_MAP ||
         (m->coarse_refresh == AMR_COARSE_PER_INCREMENT && same_inc)))
        return AMR_OK;
    hipStream_t s = m->stream;
    const int32_t nc = m->ncoarse;
    const int64_t nn = (int64_t)nc * nc;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)std::min<int64_t>((nn + 255) / 256, 4096)), dim3(256), 0, s, nn,
                       m->Kinv);
    if (m->npairs > 0)
        hipLaunchKernelGGL(galerkin_pairs_kernel, dim3((unsigned)((m->npairs + 255) / 256)), dim3(256), 0, s,
                           m->npairs, m->c_prow, m->c_pagg, m->c_agg, m->c_cent, m->X, m->mask, m->rowptr, m->col,
                           m->Kv, m->c_pair);
    if (m->ncb > 0)
        hipLaunchKernelGGL(galerkin_blocks_kernel, dim3((unsigned)(((int64_t)m->ncb * 36 + 255) / 256)), dim3(256),
                           0, s, m->ncb, nc, m->c_cba, m->c_cbb, m->c_cbptr, m->c_cbslot, m->c_pair, m->Kinv);
    hipLaunchKernelGGL(fix_diag_kernel, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, nc, m->Kinv, m->c_info);
    AMR_HIP(hipGetLastError());
    {
        double *P = m->c_work, *R = P + GJB * GJB, *C = R + (size_t)GJB * nc;
        const int nt = (nc + GJB - 1) / GJB;
        for (int k = 0; k < nt; k++) {
            const int b = std::min(GJB, nc - k * GJB);
            hipLaunchKernelGGL(gj_pivot_kernel, dim3(1), dim3(1024), 0, s, nc, k * GJB, b, m->Kinv, P, m->c_info);
            hipLaunchKernelGGL(gj_panel_kernel, dim3((unsigned)nt), dim3(256), 0, s, nc, k, b, m->Kinv, P, R, C);
            hipLaunchKernelGGL(gj_update_kernel, dim3((unsigned)nt, (unsigned)nt), dim3(256), 0, s, nc, k, b, m->Kinv,
                               P, R, C);
        }
        AMR_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(symmetrize_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, nc, m->Kinv,
                       m->c_info);
    AMR_HIP(hipGetLastError());
    m->coarse_version = m->k_version;
    m->coarse_dofs = m->dof_version;
    m->coarse_inc = m->inc_version;
    return AMR_OK;
}

void amr_coarse_free(amr_model *m)
{
    void **ptrs[] = {(void **)&m->c_agg,  (void **)&m->c_aggptr, (void **)&m->c_aggnodes, (void **)&m->c_prow,
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
