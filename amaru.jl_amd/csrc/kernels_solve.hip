// Solver kernels for gfx950: node-block (3x3) SpMV and Jacobi-PCG, FP64 and
// FP32, replacing the UMFPACK solve of solve_system! (src/solver.jl:42-43).
//
// Vectors live in node order (3 dofs per node).  The operator is K with the
// prescribed dofs masked out: y = mask .* (K x).  With x zero at prescribed
// dofs this is K11 x on the unknowns, so one resident copy of K serves K11,
// K12 U2 (right-hand side) and K21 U1 + K22 U2 (reactions).
//
// Reductions: every reducing kernel runs on a fixed grid, writes one partial
// per workgroup, and the last workgroup to finish (ticket counter) sums the
// partials in index order and publishes the scalar.  Results are therefore
// deterministic for a given model and launch shape.
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "amr_internal.h"

namespace {

constexpr int WG = 256;

enum { S_RZ = 0, S_ALPHA, S_BETA, S_RR, S_BB, S_THR, S_PQ, S_RZ0, S_RES, S_ERRN, S_ERRD, S_KX, S_NSCAL = 16 };
enum { C_DONE = 0, C_ITERS, C_BREAK, C_T0, C_T1, C_T2, C_T3, C_T4, C_NAN, C_NCTL = 16 };

template <typename T>
__device__ inline T wave_sum(T v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// block sum of up to 2 values; result valid in thread 0
__device__ inline void block_sum2(double &a, double &b)
{
    __shared__ double sa[WG / 64], sb[WG / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) {
        sa[w] = a;
        sb[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0.0, y = 0.0;
        for (int i = 0; i < WG / 64; i++) {
            x += sa[i];
            y += sb[i];
        }
        a = x;
        b = y;
    }
    __syncthreads();
}

// Publish this workgroup's partials; returns true in thread 0 of the last
// workgroup, which then holds the totals in (ta, tb).
__device__ inline bool last_block_sum(double a, double b, double *pa, double *pb, int32_t *ticket,
                                      double &ta, double &tb)
{
    __shared__ bool last;
    if (threadIdx.x == 0) {
        pa[blockIdx.x] = a;
        pb[blockIdx.x] = b;
        __threadfence();
        int t = atomicAdd(ticket, 1);
        last = (t == (int)gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    // deterministic: every thread sums a fixed strided subset, then a fixed tree
    double x = 0.0, y = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += WG) {
        x += pa[i];
        y += pb[i];
    }
    block_sum2(x, y);
    if (threadIdx.x == 0) {
        ta = x;
        tb = y;
        *ticket = 0;
    }
    return threadIdx.x == 0;
}

// ------------------------------------------------------------------ SpMV
//
// L lanes per block row (node).  Each lane walks the row's blocks with
// stride L, multiplies the 3x3 block by x[col], and the L partial 3-vectors
// are reduced with cross-lane shuffles.  mode: 0 = plain y = K x,
// 1 = y = mask .* (K x).
// |K||x| row sums (for the attainable-accuracy floor of the residual test)
template <int L>
__device__ inline void absrow(int64_t row, int sub, const int32_t *__restrict__ rowptr,
                              const int32_t *__restrict__ col, const double *__restrict__ vals,
                              const double *__restrict__ x, double *a)
{
    a[0] = a[1] = a[2] = 0.0;
    for (int32_t k = rowptr[row] + sub; k < rowptr[row + 1]; k += L) {
        const int64_t c = col[k];
        const double *v = vals + (int64_t)k * 9;
        for (int r = 0; r < 3; r++)
            for (int j = 0; j < 3; j++) a[r] += fabs(v[r * 3 + j] * x[c * 3 + j]);
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1)
        for (int r = 0; r < 3; r++) a[r] += __shfl_xor(a[r], off, L);
}

template <int L, typename VT, typename XT, typename AT>
__device__ inline void spmv_row(int64_t row, int sub, const int32_t *__restrict__ rowptr,
                                const int32_t *__restrict__ col, const VT *__restrict__ vals,
                                const XT *__restrict__ x, AT &y0, AT &y1, AT &y2)
{
    y0 = 0;
    y1 = 0;
    y2 = 0;
    const int32_t rb = rowptr[row], re = rowptr[row + 1];
    for (int32_t k = rb + sub; k < re; k += L) {
        const int64_t c = col[k];
        const VT *v = vals + (int64_t)k * 9;
        const AT x0 = (AT)x[c * 3], x1 = (AT)x[c * 3 + 1], x2 = (AT)x[c * 3 + 2];
        y0 = fma((AT)v[2], x2, fma((AT)v[1], x1, fma((AT)v[0], x0, y0)));
        y1 = fma((AT)v[5], x2, fma((AT)v[4], x1, fma((AT)v[3], x0, y1)));
        y2 = fma((AT)v[8], x2, fma((AT)v[7], x1, fma((AT)v[6], x0, y2)));
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) {
        y0 += __shfl_xor(y0, off, L);
        y1 += __shfl_xor(y1, off, L);
        y2 += __shfl_xor(y2, off, L);
    }
}

// y = [mask .*] K x, optionally r = mask .* (b - K x) with rr partials
template <int L, typename VT, typename XT, typename AT>
__global__ void __launch_bounds__(WG) spmv_kernel(int64_t nrows, const int32_t *__restrict__ rowptr,
                                                  const int32_t *__restrict__ col, const VT *__restrict__ vals,
                                                  const XT *__restrict__ x, XT *__restrict__ y,
                                                  const double *__restrict__ mask, int mode)
{
    constexpr int RPW = WG / L;
    const int sub = threadIdx.x % L;
    for (int64_t row = (int64_t)blockIdx.x * RPW + threadIdx.x / L; row < nrows; row += (int64_t)gridDim.x * RPW) {
        AT y0, y1, y2;
        spmv_row<L, VT, XT, AT>(row, sub, rowptr, col, vals, x, y0, y1, y2);
        if (sub == 0) {
            if (mode == 1) {
                y0 = mask[row * 3] != 0.0 ? y0 : (AT)0;
                y1 = mask[row * 3 + 1] != 0.0 ? y1 : (AT)0;
                y2 = mask[row * 3 + 2] != 0.0 ? y2 : (AT)0;
            }
            y[row * 3] = (XT)y0;
            y[row * 3 + 1] = (XT)y1;
            y[row * 3 + 2] = (XT)y2;
        }
    }
}

// r = mask .* (b - K x) in FP64, with ||r||^2 and ||b||^2 (masked)
template <int L>
__global__ void __launch_bounds__(WG) residual_kernel(int64_t nrows, const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ col,
                                                      const double *__restrict__ vals,
                                                      const double *__restrict__ x, const double *__restrict__ b,
                                                      double *__restrict__ r, const double *__restrict__ mask,
                                                      double *__restrict__ pa, double *__restrict__ pb,
                                                      int32_t *__restrict__ ctl, double *__restrict__ scal)
{
    constexpr int RPW = WG / L;
    const int sub = threadIdx.x % L;
    double rr = 0.0, bb = 0.0, kx = 0.0;
    for (int64_t row = (int64_t)blockIdx.x * RPW + threadIdx.x / L; row < nrows; row += (int64_t)gridDim.x * RPW) {
        double y[3], a[3];
        spmv_row<L, double, double, double>(row, sub, rowptr, col, vals, x, y[0], y[1], y[2]);
        absrow<L>(row, sub, rowptr, col, vals, x, a);
        if (sub == 0) {
            for (int d = 0; d < 3; d++) {
                int64_t i = row * 3 + d;
                double m = mask[i];
                double ri = m != 0.0 ? b[i] - y[d] : 0.0;
                r[i] = ri;
                rr = fma(ri, ri, rr);
                bb = fma(m * b[i], m * b[i], bb);
                double ai = m != 0.0 ? a[d] + fabs(b[i]) : 0.0;
                kx = fma(ai, ai, kx);
            }
        }
    }
    block_sum2(rr, bb);
    double dummy = 0.0;
    block_sum2(kx, dummy);
    double trr, tbb, tkx, t2;
    __shared__ bool last3;
    if (threadIdx.x == 0) pa[8192 + blockIdx.x] = kx;
    if (last_block_sum(rr, bb, pa, pb, ctl + C_T0, trr, tbb)) last3 = true;
    else if (threadIdx.x == 0) last3 = false;
    __syncthreads();
    if (!last3) return;
    // the last block: the kx partials were written before every block's ticket
    double x3 = 0.0, y3 = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += WG) x3 += pa[8192 + i];
    block_sum2(x3, y3);
    if (threadIdx.x == 0) {
        scal[S_RR] = trr;
        scal[S_BB] = tbb;
        scal[S_KX] = x3;
    }
}

// ------------------------------------------------------------------- PCG
//
// Iteration k:  q = A p, pq = p.q          (pcg_spmv_kernel)
//               a = rz/pq; x += a p; r -= a q; z = Minv r;
//               rz' = r.z, rr = r.r, b = rz'/rz, convergence test  (pcg_update_kernel)
//               p = z + b p                (pcg_dir_kernel)
// Dot products accumulate in FP64 for both precisions.
template <int L, typename VT, typename XT>
__global__ void __launch_bounds__(WG) pcg_spmv_kernel(int64_t nrows, const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ col, const VT *__restrict__ vals,
                                                      const XT *__restrict__ p, XT *__restrict__ q,
                                                      const double *__restrict__ mask, double *__restrict__ pa,
                                                      double *__restrict__ pb, int32_t *__restrict__ ctl,
                                                      double *__restrict__ scal)
{
    if (ctl[C_DONE]) return;
    using AT = XT;
    constexpr int RPW = WG / L;
    const int sub = threadIdx.x % L;
    double pq = 0.0;
    for (int64_t row = (int64_t)blockIdx.x * RPW + threadIdx.x / L; row < nrows; row += (int64_t)gridDim.x * RPW) {
        AT y0, y1, y2;
        spmv_row<L, VT, XT, AT>(row, sub, rowptr, col, vals, p, y0, y1, y2);
        if (sub == 0) {
            AT yy[3] = {y0, y1, y2};
            for (int d = 0; d < 3; d++) {
                int64_t i = row * 3 + d;
                XT qi = mask[i] != 0.0 ? (XT)yy[d] : (XT)0;
                q[i] = qi;
                pq = fma((double)p[i], (double)qi, pq);
            }
        }
    }
    double dummy = 0.0;
    block_sum2(pq, dummy);
    double tpq, t2;
    if (last_block_sum(pq, 0.0, pa, pb, ctl + C_T1, tpq, t2)) {
        scal[S_PQ] = tpq;
        if (!(tpq > 0.0) || !isfinite(tpq)) {
            ctl[C_BREAK] = 1;
            ctl[C_DONE] = 1;
            scal[S_ALPHA] = 0.0;
        } else {
            scal[S_ALPHA] = scal[S_RZ] / tpq;
        }
    }
}

template <typename XT>
__global__ void __launch_bounds__(WG) pcg_update_kernel(int64_t n, XT *__restrict__ x, XT *__restrict__ r,
                                                        XT *__restrict__ z, const XT *__restrict__ p,
                                                        const XT *__restrict__ q, const XT *__restrict__ dinv,
                                                        double *__restrict__ pa, double *__restrict__ pb,
                                                        int32_t *__restrict__ ctl, double *__restrict__ scal,
                                                        int maxiter)
{
    if (ctl[C_DONE]) return;
    const double alpha = scal[S_ALPHA];
    const XT a = (XT)alpha;
    double rz = 0.0, rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) {
        XT xi = fma(a, p[i], x[i]);
        XT ri = fma(-a, q[i], r[i]);
        XT zi = dinv[i] * ri;
        x[i] = xi;
        r[i] = ri;
        z[i] = zi;
        rz = fma((double)ri, (double)zi, rz);
        rr = fma((double)ri, (double)ri, rr);
    }
    block_sum2(rz, rr);
    double trz, trr;
    if (last_block_sum(rz, rr, pa, pb, ctl + C_T2, trz, trr)) {
        double beta = trz / scal[S_RZ];
        scal[S_BETA] = beta;
        scal[S_RZ] = trz;
        scal[S_RR] = trr;
        int it = ctl[C_ITERS] + 1;
        ctl[C_ITERS] = it;
        if (trr <= scal[S_THR] || it >= maxiter || !isfinite(trr)) ctl[C_DONE] = 1;
    }
}

template <typename XT>
__global__ void __launch_bounds__(WG) pcg_dir_kernel(int64_t n, XT *__restrict__ p, const XT *__restrict__ z,
                                                     const int32_t *__restrict__ ctl,
                                                     const double *__restrict__ scal)
{
    if (ctl[C_DONE]) return;
    const XT b = (XT)scal[S_BETA];
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG)
        p[i] = fma(b, p[i], z[i]);
}

// x = 0, r = rhs (already masked), z = Minv r, p = z, rz = r.z, rr = r.r
template <typename XT>
__global__ void __launch_bounds__(WG) pcg_init_kernel(int64_t n, const XT *__restrict__ rhs, XT *__restrict__ x,
                                                      XT *__restrict__ r, XT *__restrict__ z, XT *__restrict__ p,
                                                      const XT *__restrict__ dinv, double *__restrict__ pa,
                                                      double *__restrict__ pb, int32_t *__restrict__ ctl,
                                                      double *__restrict__ scal, double rtol2)
{
    double rz = 0.0, rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) {
        XT ri = rhs[i];
        XT zi = dinv[i] * ri;
        x[i] = (XT)0;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        rz = fma((double)ri, (double)zi, rz);
        rr = fma((double)ri, (double)ri, rr);
    }
    block_sum2(rz, rr);
    double trz, trr;
    if (last_block_sum(rz, rr, pa, pb, ctl + C_T3, trz, trr)) {
        scal[S_RZ] = trz;
        scal[S_RZ0] = trz;
        scal[S_RR] = trr;
        scal[S_THR] = rtol2 * trr;
        ctl[C_ITERS] = 0;
        ctl[C_BREAK] = 0;
        ctl[C_DONE] = (trr == 0.0) ? 1 : 0;
    }
}

// elementwise helpers
__global__ void to_f32_kernel(int64_t n, const double *__restrict__ a, float *__restrict__ b)
{
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) b[i] = (float)a[i];
}

__global__ void add_f32_kernel(int64_t n, double *__restrict__ x, const float *__restrict__ d)
{
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG)
        x[i] += (double)d[i];
}

__global__ void add_f64_kernel(int64_t n, double *__restrict__ x, const double *__restrict__ d)
{
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) x[i] += d[i];
}

__global__ void fill_kernel(int64_t n, double *__restrict__ x, double v)
{
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) x[i] = v;
}

struct Launch {
    int lanes;     // lanes per block row
    unsigned grid_rows, grid_vec;
};

Launch pick_launch(const amr_model *m)
{
    Launch L;
    double avg = m->nnodes ? (double)m->pat.nblk / (double)m->nnodes : 0.0;
    L.lanes = avg > 40 ? 64 : (avg > 20 ? 32 : 16);
    int64_t rows_per_wg = WG / L.lanes;
    int64_t need = (m->nnodes + rows_per_wg - 1) / rows_per_wg;
    L.grid_rows = (unsigned)std::max<int64_t>(1, std::min<int64_t>(need, 4096));
    int64_t nv = (m->nnodes * 3 + WG - 1) / WG;
    L.grid_vec = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nv, 2048));
    return L;
}

template <typename VT, typename XT>
void launch_pcg_spmv(amr_model *m, const Launch &L, const VT *vals, const XT *p, XT *q, double *pa, double *pb,
                     int32_t *ctl, double *scal)
{
    switch (L.lanes) {
    case 64:
        hipLaunchKernelGGL((pcg_spmv_kernel<64, VT, XT>), dim3(L.grid_rows), dim3(WG), 0, m->stream, m->nnodes,
                           m->rowptr, m->col, vals, p, q, m->mask, pa, pb, ctl, scal);
        break;
    case 32:
        hipLaunchKernelGGL((pcg_spmv_kernel<32, VT, XT>), dim3(L.grid_rows), dim3(WG), 0, m->stream, m->nnodes,
                           m->rowptr, m->col, vals, p, q, m->mask, pa, pb, ctl, scal);
        break;
    default:
        hipLaunchKernelGGL((pcg_spmv_kernel<16, VT, XT>), dim3(L.grid_rows), dim3(WG), 0, m->stream, m->nnodes,
                           m->rowptr, m->col, vals, p, q, m->mask, pa, pb, ctl, scal);
        break;
    }
}

void launch_residual(amr_model *m, const Launch &L, const double *x, const double *b, double *r, double *pa,
                     double *pb, int32_t *ctl, double *scal)
{
    switch (L.lanes) {
    case 64:
        hipLaunchKernelGGL((residual_kernel<64>), dim3(L.grid_rows), dim3(WG), 0, m->stream, m->nnodes, m->rowptr,
                           m->col, m->Kv, x, b, r, m->mask, pa, pb, ctl, scal);
        break;
    case 32:
        hipLaunchKernelGGL((residual_kernel<32>), dim3(L.grid_rows), dim3(WG), 0, m->stream, m->nnodes, m->rowptr,
                           m->col, m->Kv, x, b, r, m->mask, pa, pb, ctl, scal);
        break;
    default:
        hipLaunchKernelGGL((residual_kernel<16>), dim3(L.grid_rows), dim3(WG), 0, m->stream, m->nnodes, m->rowptr,
                           m->col, m->Kv, x, b, r, m->mask, pa, pb, ctl, scal);
        break;
    }
}

// Runs PCG iterations on (x, r, z, p, q) until the device flag says done.
template <typename VT, typename XT>
int run_pcg(amr_model *m, const Launch &L, const VT *vals, const XT *dinv, const XT *rhs, XT *x, XT *r, XT *z,
            XT *p, XT *q, double rtol, int maxiter, int check_every, int32_t *ctl, double *scal, int *iters,
            int *broke, double *rr_out, double *rr0_out)
{
    const int64_t n = m->nnodes * 3;
    double *pa = m->part, *pb = m->part + 4096;
    hipLaunchKernelGGL((pcg_init_kernel<XT>), dim3(L.grid_vec), dim3(WG), 0, m->stream, n, rhs, x, r, z, p, dinv, pa,
                       pb, ctl, scal, rtol * rtol);
    AMR_HIP(hipGetLastError());
    int32_t hctl[C_NCTL];
    double hscal[S_NSCAL];
    int prev_iters = 0;
    if (m->timing && (int)m->ev.size() < 2 * check_every) {
        for (hipEvent_t e : m->ev) hipEventDestroy(e);
        m->ev.assign(2 * check_every, nullptr);
        for (hipEvent_t &e : m->ev) AMR_HIP(hipEventCreate(&e));
    }
    for (;;) {
        for (int k = 0; k < check_every; k++) {
            if (m->timing) AMR_HIP(hipEventRecord(m->ev[2 * k], m->stream));
            launch_pcg_spmv<VT, XT>(m, L, vals, p, q, pa, pb, ctl, scal);
            if (m->timing) AMR_HIP(hipEventRecord(m->ev[2 * k + 1], m->stream));
            hipLaunchKernelGGL((pcg_update_kernel<XT>), dim3(L.grid_vec), dim3(WG), 0, m->stream, n, x, r, z, p, q,
                               dinv, pa, pb, ctl, scal, maxiter);
            hipLaunchKernelGGL((pcg_dir_kernel<XT>), dim3(L.grid_vec), dim3(WG), 0, m->stream, n, p, z, ctl, scal);
        }
        AMR_HIP(hipGetLastError());
        AMR_HIP(hipMemcpyAsync(hctl, ctl, sizeof hctl, hipMemcpyDeviceToHost, m->stream));
        AMR_HIP(hipMemcpyAsync(hscal, scal, sizeof hscal, hipMemcpyDeviceToHost, m->stream));
        AMR_HIP(hipStreamSynchronize(m->stream));
        if (m->timing) {
            // the SpMVs of iterations executed in this interval ran in full;
            // later launches returned early on the done flag
            int real = std::min(hctl[C_ITERS] - prev_iters, check_every);
            for (int k = 0; k < real; k++) {
                float ms = 0.f;
                AMR_HIP(hipEventElapsedTime(&ms, m->ev[2 * k], m->ev[2 * k + 1]));
                m->t_spmv_ms += ms;
                m->n_spmv++;
            }
        }
        prev_iters = hctl[C_ITERS];
        if (hctl[C_DONE]) break;
    }
    *iters = hctl[C_ITERS];
    *broke = hctl[C_BREAK];
    *rr_out = hscal[S_RR];
    *rr0_out = hscal[S_THR] / (rtol * rtol);
    return AMR_OK;
}

}  // namespace

// y = K x (mask_mode 0) or mask .* (K x) (mask_mode 1), FP64, node order.
int amr_spmv_full(amr_model *m, const double *x, double *y, int mask_mode)
{
    Launch L = pick_launch(m);
    switch (L.lanes) {
    case 64:
        hipLaunchKernelGGL((spmv_kernel<64, double, double, double>), dim3(L.grid_rows), dim3(WG), 0, m->stream,
                           m->nnodes, m->rowptr, m->col, m->Kv, x, y, m->mask, mask_mode);
        break;
    case 32:
        hipLaunchKernelGGL((spmv_kernel<32, double, double, double>), dim3(L.grid_rows), dim3(WG), 0, m->stream,
                           m->nnodes, m->rowptr, m->col, m->Kv, x, y, m->mask, mask_mode);
        break;
    default:
        hipLaunchKernelGGL((spmv_kernel<16, double, double, double>), dim3(L.grid_rows), dim3(WG), 0, m->stream,
                           m->nnodes, m->rowptr, m->col, m->Kv, x, y, m->mask, mask_mode);
        break;
    }
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

// Solve K11 x = b (node order, b zero at prescribed dofs).  x is FP64.
// Outer loop: FP64 residual r = b - K x and its norm; stop at rtol.  Inner
// solve of K d = r: FP64 PCG to the remaining tolerance (plain path) or FP32
// PCG to inner_rtol (mixed path); then x += d in FP64.
int amr_pcg_solve(amr_model *m, const double *b, double *x, const amr_solve_opts *o, amr_solve_info *info)
{
    const int64_t n = m->nnodes * 3;
    Launch L = pick_launch(m);
    int32_t *ctl = (int32_t *)(m->scal + S_NSCAL);
    double *scal = m->scal;
    double *pa = m->part, *pb = m->part + 4096;
    int total = 0, refinements = 0, broke = 0, converged = 0;
    double relres = 0.0;
    const int check = o->check_every > 0 ? o->check_every : 16;
    double *rres = m->v[0];
    double rr_prev = INFINITY;
    const double kEps = 2.220446049250313e-16, kFloorC = 64.0;

    hipLaunchKernelGGL(fill_kernel, dim3(L.grid_vec), dim3(WG), 0, m->stream, n, x, 0.0);
    AMR_HIP(hipGetLastError());
    for (;;) {
        launch_residual(m, L, x, b, rres, pa, pb, ctl, scal);
        AMR_HIP(hipGetLastError());
        double hs[S_NSCAL];
        AMR_HIP(hipMemcpyAsync(hs, scal, sizeof hs, hipMemcpyDeviceToHost, m->stream));
        AMR_HIP(hipStreamSynchronize(m->stream));
        const double bb = hs[S_BB], rr = hs[S_RR];
        relres = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
        // attainable accuracy: a residual at the rounding level of computing
        // b - K x itself (c * eps * || |K||x| + |b| ||) cannot be improved
        const double floor2 = kFloorC * kFloorC * kEps * kEps * hs[S_KX];
        if (bb == 0.0 || rr <= o->rtol * o->rtol * bb || (refinements > 0 && rr <= floor2)) {
            converged = 1;
            break;
        }
        if (refinements > 0 && rr > 0.25 * rr_prev) break;  // refinement stagnated above both bounds
        rr_prev = rr;
        if (total >= o->maxiter || refinements >= 64 || !std::isfinite(rr) || broke) break;
        int it = 0;
        double rri, rr0i;
        if (o->precision == AMR_PREC_FP64) {
            double *r = m->v[1], *z = m->v[2], *p = m->v[3], *q = m->v[4], *d = m->v[5];
            double inner = std::min(0.5 * o->rtol * std::sqrt(bb / rr), 0.5);
            int rc = run_pcg<double, double>(m, L, m->Kv, m->dinv, rres, d, r, z, p, q, inner, o->maxiter - total,
                                             check, ctl, scal, &it, &broke, &rri, &rr0i);
            if (rc) return rc;
            hipLaunchKernelGGL(add_f64_kernel, dim3(L.grid_vec), dim3(WG), 0, m->stream, n, x, d);
        } else {
            float *rf = m->f[0], *d = m->f[1], *ri = m->f[2], *zf = m->f[3], *pf = m->f[4], *qf = m->f[5];
            hipLaunchKernelGGL(to_f32_kernel, dim3(L.grid_vec), dim3(WG), 0, m->stream, n, rres, rf);
            int rc = run_pcg<float, float>(m, L, m->Kv32, m->dinv32, rf, d, ri, zf, pf, qf, o->inner_rtol,
                                           o->maxiter - total, check, ctl, scal, &it, &broke, &rri, &rr0i);
            if (rc) return rc;
            hipLaunchKernelGGL(add_f32_kernel, dim3(L.grid_vec), dim3(WG), 0, m->stream, n, x, d);
        }
        AMR_HIP(hipGetLastError());
        total += it;
        refinements++;
    }
    if (info) {
        info->iterations = total;
        info->refinements = refinements;
        info->converged = converged;
        info->relres = relres;
    }
    if (!converged) {
        char buf[256];
        snprintf(buf, sizeof buf, "PCG did not reach rtol=%.3e (relres %.3e after %d iterations, %d refinements%s)",
                 o->rtol, relres, total, refinements, broke ? ", breakdown" : "");
        amr_set_error(buf);
        return AMR_ERR_SOLVE;
    }
    return AMR_OK;
}

// ----------------------------------------------------------- stage vectors
namespace {

__global__ void inc_begin_kernel(int64_t n, double dT, double acr, int add, const double *__restrict__ Uex,
                                 const double *__restrict__ Fex, const double *__restrict__ Rc,
                                 double *__restrict__ dFex, double *__restrict__ R, double *__restrict__ dUa,
                                 double *__restrict__ dUi)
{
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) {
        double f = dT * Fex[i];
        if (add) f = f + acr * Rc[i];
        dFex[i] = f;
        R[i] = f;
        dUa[i] = 0.0;
        dUi[i] = dT * Uex[i];
    }
}

__global__ void prescribed_part_kernel(int64_t n, const double *__restrict__ mask, const double *__restrict__ u,
                                       double *__restrict__ u2)
{
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG)
        u2[i] = mask[i] != 0.0 ? 0.0 : u[i];
}

__global__ void masked_diff_kernel(int64_t n, const double *__restrict__ mask, const double *__restrict__ a,
                                   const double *__restrict__ b, double *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG)
        out[i] = mask[i] != 0.0 ? a[i] - b[i] : 0.0;
}

__global__ void combine_kernel(int64_t n, const double *__restrict__ mask, const double *__restrict__ x,
                               double *__restrict__ u)
{
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG)
        if (mask[i] != 0.0) u[i] = x[i];
}

__global__ void vadd_kernel(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                            double *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) out[i] = a[i] + b[i];
}

__device__ inline double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    return v;
}

// ΔUa += ΔUi; R = mask .* (ΔFex - ΔFin); max|R|, max|ΔUi|, max|ΔUa|, NaN in ΔFin
__global__ void __launch_bounds__(WG) nr_post_kernel(int64_t n, const double *__restrict__ mask,
                                                     const double *__restrict__ dFex,
                                                     const double *__restrict__ dFin,
                                                     const double *__restrict__ dUi, double *__restrict__ dUa,
                                                     double *__restrict__ R, double *__restrict__ part,
                                                     int32_t *__restrict__ ctl, double *__restrict__ scal)
{
    double mr = 0.0, mi = 0.0, ma = 0.0, bad = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) {
        double ua = dUa[i] + dUi[i];
        dUa[i] = ua;
        double fi = dFin[i];
        double r = mask[i] != 0.0 ? dFex[i] - fi : 0.0;
        R[i] = r;
        if (fi != fi) bad = 1.0;
        mr = fmax(mr, fabs(r));
        mi = fmax(mi, fabs(dUi[i]));
        ma = fmax(ma, fabs(ua));
    }
    __shared__ double sm[4][WG / 64];
    __shared__ bool last;
    mr = wave_max(mr);
    mi = wave_max(mi);
    ma = wave_max(ma);
    bad = wave_max(bad);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) {
        sm[0][w] = mr;
        sm[1][w] = mi;
        sm[2][w] = ma;
        sm[3][w] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < 4; k++) {
            double v = 0.0;
            for (int j = 0; j < WG / 64; j++) v = fmax(v, sm[k][j]);
            part[k * 2048 + blockIdx.x] = v;
        }
        __threadfence();
        last = atomicAdd(ctl + C_T4, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < (int)gridDim.x; i += WG)
        for (int k = 0; k < 4; k++) v[k] = fmax(v[k], part[k * 2048 + i]);
    for (int k = 0; k < 4; k++) v[k] = wave_max(v[k]);
    __syncthreads();
    if (l == 0)
        for (int k = 0; k < 4; k++) sm[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < 4; k++) {
            double t = 0.0;
            for (int j = 0; j < WG / 64; j++) t = fmax(t, sm[k][j]);
            v[k] = t;
        }
        scal[S_RES] = v[0];
        scal[S_ERRN] = v[1];
        scal[S_ERRD] = v[2];
        ctl[C_NAN] = v[3] != 0.0;
        ctl[C_T4] = 0;
    }
}

__global__ void commit_kernel(int64_t n, double acr, const double *__restrict__ dUa, const double *__restrict__ dFin,
                              const double *__restrict__ R, double *__restrict__ U, double *__restrict__ F,
                              double *__restrict__ Rc)
{
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) {
        U[i] += dUa[i];
        F[i] += dFin[i];
        Rc[i] = (1.0 - acr) * Rc[i] + R[i];
    }
}

unsigned vec_grid(const amr_model *m)
{
    int64_t nv = (m->nnodes * 3 + WG - 1) / WG;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(nv, 2048));
}

}  // namespace

int amr_vec_inc_begin(amr_model *m, double dT, double acr, int add)
{
    double **v = m->sv;
    hipLaunchKernelGGL(inc_begin_kernel, dim3(vec_grid(m)), dim3(WG), 0, m->stream, m->nnodes * 3, dT, acr, add,
                       v[SV_UEX], v[SV_FEX], v[SV_RC], v[SV_DFEX], v[SV_R], v[SV_DUA], v[SV_DUI]);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

int amr_vec_fill(amr_model *m, double *x, double val)
{
    hipLaunchKernelGGL(fill_kernel, dim3(vec_grid(m)), dim3(WG), 0, m->stream, m->nnodes * 3, x, val);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

// b = mask .* (R - K (prescribed part of dUi)); u2/ky are scratch
int amr_vec_solve_prep(amr_model *m, const double *dUi, const double *R, double *u2, double *ky, double *b)
{
    const int64_t n = m->nnodes * 3;
    hipLaunchKernelGGL(prescribed_part_kernel, dim3(vec_grid(m)), dim3(WG), 0, m->stream, n, m->mask, dUi, u2);
    AMR_HIP(hipGetLastError());
    int rc = amr_spmv_full(m, u2, ky, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(masked_diff_kernel, dim3(vec_grid(m)), dim3(WG), 0, m->stream, n, m->mask, R, ky, b);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

int amr_vec_combine(amr_model *m, const double *x, double *dUi)
{
    hipLaunchKernelGGL(combine_kernel, dim3(vec_grid(m)), dim3(WG), 0, m->stream, m->nnodes * 3, m->mask, x, dUi);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

int amr_vec_add(amr_model *m, const double *a, const double *b, double *out)
{
    hipLaunchKernelGGL(vadd_kernel, dim3(vec_grid(m)), dim3(WG), 0, m->stream, m->nnodes * 3, a, b, out);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}

int amr_vec_nr_post(amr_model *m, double *res, double *dui_max, double *dua_max, int *nan_flag)
{
    double **v = m->sv;
    int32_t *ctl = (int32_t *)(m->scal + S_NSCAL);
    hipLaunchKernelGGL(nr_post_kernel, dim3(vec_grid(m)), dim3(WG), 0, m->stream, m->nnodes * 3, m->mask,
                       v[SV_DFEX], v[SV_DFIN], v[SV_DUI], v[SV_DUA], v[SV_R], m->part, ctl, m->scal);
    AMR_HIP(hipGetLastError());
    double hs[S_NSCAL];
    int32_t hc[C_NCTL];
    AMR_HIP(hipMemcpyAsync(hs, m->scal, sizeof hs, hipMemcpyDeviceToHost, m->stream));
    AMR_HIP(hipMemcpyAsync(hc, ctl, sizeof hc, hipMemcpyDeviceToHost, m->stream));
    AMR_HIP(hipStreamSynchronize(m->stream));
    *res = hs[S_RES];
    *dui_max = hs[S_ERRN];
    *dua_max = hs[S_ERRD];
    *nan_flag = hc[C_NAN];
    return AMR_OK;
}

int amr_vec_commit(amr_model *m)
{
    double **v = m->sv;
    hipLaunchKernelGGL(commit_kernel, dim3(vec_grid(m)), dim3(WG), 0, m->stream, m->nnodes * 3, m->acr, v[SV_DUA],
                       v[SV_DFIN], v[SV_R], v[SV_U], v[SV_F], v[SV_RC]);
    AMR_HIP(hipGetLastError());
    return AMR_OK;
}
